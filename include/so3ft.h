/*
 * so3ft.h -- C ABI of the MI355X (gfx950) SO(3) Fourier transform library
 * (libso3ft.so, built from pysofft_amd/csrc).
 *
 * This is the drop-in boundary for PySOFFT's f2py module `pysofft._soft.py`
 * (pysofft/py_wrap.f90): each entry point below names the py_wrap procedure it
 * replaces.  Differences of kind, not of meaning:
 *   - every call returns an int status (SO3FT_OK or a negative code) and never
 *     aborts; so3ft_last_error() returns a message for the calling thread;
 *   - arrays are DEVICE pointers (hipMalloc'd or torch CUDA tensors) owned by the
 *     caller; the library owns only its plan data and scratch;
 *   - every transform takes the HIP stream to run on (NULL = default stream) and
 *     returns after enqueueing (stream-ordered, like any HIP call).
 *
 * Layouts are the reference's:
 *   coefficients : complex128 (or complex64 for the *_c64 entry points), length
 *                  NC = (4bw^3-bw)/3, "mnl" order (pysofft/utils.f90:87-201); for a
 *                  handle made with recurrence_type = Risbo and precompute_wigners = 0,
 *                  "lmn" order (utils.f90:228-355), as the reference (soft.py:98-106)
 *   SO(3) function: C-order (gamma, alpha, beta), shape (2bw,2bw,2bw)
 *                  (docs/defs/so3_func.md:5-13), complex128 or float64
 *   *_many       : batch-major, i.e. (n, NC) and (n, 2bw, 2bw, 2bw)
 *   SH coefficients for correlation: complex128, SHTNS LMc order l(l+1)+m, bw^2
 *
 * A handle may be used from one stream at a time (scratch is per handle).
 */
#ifndef SO3FT_H
#define SO3FT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t so3ft_handle;

enum {
    SO3FT_OK = 0,
    SO3FT_EINVAL = -1,    /* bad argument (bandwidth, lmax, size, null pointer) */
    SO3FT_EHANDLE = -2,   /* unknown or destroyed handle                        */
    SO3FT_EHIP = -3,      /* HIP runtime error (message in so3ft_last_error)    */
    SO3FT_ENOMEM = -4,    /* device allocation failed                           */
    SO3FT_EUNSUPPORTED = -5
};

/* recurrence types, pysofft/soft.f08:14-15 (py_wrap.f90:48-63) */
int so3ft_get_kostelec_recurrence_val(void);
int so3ft_get_risbo_recurrence_val(void);

const char *so3ft_last_error(void);

/* ---- lifecycle: py_wrap.f90:79-91 py_init_soft, :172-180 py_destroy,
 *      :152-171 py_reset, :181-188 py_init_fft ---------------------------- */
int so3ft_init(int64_t bw, int64_t lmax, int precompute_wigners, int init_ffts,
               int64_t recurrence_type, int32_t fftw_flags, so3ft_handle *out);
int so3ft_destroy(so3ft_handle h);
int so3ft_reset(so3ft_handle h, int64_t bw, int64_t lmax, int precompute_wigners,
                int init_ffts, int64_t recurrence_type, int32_t fftw_flags);
int so3ft_init_fft(so3ft_handle h, int use_real_fft);

/* ---- getters / setters: py_wrap.f90:100-151 ------------------------------- */
int so3ft_get_bw(so3ft_handle h, int64_t *bw);
int so3ft_get_lmax(so3ft_handle h, int64_t *lmax);
int so3ft_set_lmax(so3ft_handle h, int64_t lmax);
int so3ft_get_recurrence_type(so3ft_handle h, int64_t *rt);
int so3ft_wigners_are_precomputed(so3ft_handle h, int *yes);
int so3ft_get_fftw_flags(so3ft_handle h, int32_t *flags);

/* ---- FP64 DWT engine selection (no counterpart in the reference, which has one
 * CPU GEMV path): SO3FT_DWT_VALU (default) runs the inverse contraction on the FP64
 * vector ALUs (lane = beta) and the forward on v_mfma_f64_16x16x4f64 with lane = beta
 * recurrence chains; SO3FT_DWT_MFMA runs both on v_mfma_f64_16x16x4f64 with 16-degree
 * x 64-beta tiles of d per order pair.  All are FP64 throughout and give the same
 * results to rounding.  Kept across so3ft_reset.  Handles with a d table (made with
 * precompute_wigners, or with recurrence_type = Risbo: the table is filled once at creation
 * by the handle's recurrence, soft.f08:194-219) read d from it on the default engine's
 * kernels whatever the engine; the complex64 entry points compute d on the fly. */
enum { SO3FT_DWT_VALU = 0, SO3FT_DWT_MFMA = 1 };
int so3ft_set_dwt_engine(so3ft_handle h, int engine);
int so3ft_get_dwt_engine(so3ft_handle h, int *engine);

/* ---- complex transforms: py_wrap.f90:236-257, 280-301 --------------------- */
int so3ft_isoft(so3ft_handle h, const void *coeff, void *so3func, void *stream);   /* py_isoft */
int so3ft_soft(so3ft_handle h, const void *so3func, void *coeff, void *stream);    /* py_soft  */
int so3ft_isoft_many(so3ft_handle h, const void *coeffs, void *so3funcs, int64_t n, void *stream);
int so3ft_soft_many(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream);

/* ---- real transforms: py_wrap.f90:258-279, 302-323 ------------------------ */
int so3ft_irsoft(so3ft_handle h, const void *coeff, void *so3func_real, void *stream);  /* py_irsoft */
int so3ft_rsoft(so3ft_handle h, const void *so3func_real, void *coeff, void *stream);   /* py_rsoft  */
int so3ft_irsoft_many(so3ft_handle h, const void *coeffs, void *so3funcs_real, int64_t n, void *stream);
int so3ft_rsoft_many(so3ft_handle h, const void *so3funcs_real, void *coeffs, int64_t n, void *stream);

/* ---- mixed precision (no counterpart in the reference, whose transforms are FP64
 * only): complex64 coefficients (n, NC) and complex64 SO(3) functions (n, 2bw, 2bw,
 * 2bw); the Wigner-d recurrence, quadrature weights and FFT butterflies run in FP64,
 * the DWT contraction on the matrix cores with d split into FP32 hi + lo parts (hi
 * products in FP32, lo products in bfloat16) and FP64 accumulation across 16-degree blocks.  Any bw (FFT
 * lengths 2^a or 3*2^a run the register FFT, others an LDS Stockham pass).  Same semantics as
 * so3ft_isoft_many / so3ft_soft_many. */
int so3ft_isoft_many_c64(so3ft_handle h, const void *coeffs, void *so3funcs, int64_t n, void *stream);
int so3ft_soft_many_c64(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream);

/* ---- zero-allocation forward transform (no counterpart in the reference): for callers
 * that cannot allocate device memory inside the call (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results and semantics
 * as so3ft_soft_many_c64, but its scratch (the fft arrays of up to 256 MiB / (8 (2bw)^3)
 * transforms per pass, plus their coefficient sets on Risbo handles) is a static device
 * array of 256 MiB in the library, one per GPU, held until the process ends; the call
 * allocates nothing.  One transform must fit: bw <= 128, else SO3FT_EINVAL.  Every call on
 * one GPU shares that GPU's array, so calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller. */
int so3ft_soft_many_c64_noalloc(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream);

/* ---- grouped mixed-precision transforms of uneven bandwidths (no counterpart in the
 * reference, whose *_many take one bandwidth per instance).  A group fixes n bandwidths
 * (2 <= bw <= 256); its transforms run as one persistent DWT launch
 * over the ragged work of all bandwidths plus per-bandwidth FFT passes (parallel graph
 * branches), captured into a hipGraph per (input, output) pointer pair on first use and
 * replayed after.  Arrays are PACKED: transform i's complex64 coefficients (NC(bw_i),
 * mnl) start at element coeff_off[i], its complex64 SO(3) function ((2bw_i)^3, C-order
 * (gamma, alpha, beta)) at element grid_off[i] (so3ft_group_layout; host only, no GPU).
 * The first transform call (or so3ft_group_autotune) times the candidate launch shapes on
 * the group's shapes, synchronising the stream; the choice is kept per (device,
 * bandwidth multiset).  Precision and semantics as so3ft_isoft_many_c64 /
 * so3ft_soft_many_c64.  A group may be used from one stream at a time.
 * config (8 int32): inverse TB, inverse grid, forward TB, forward NW (waves per block:
 * 1), forward grid (grid 0 = one block per work item, else a multiple of 8 blocks
 * looping over the work list), tuned flag (get only), inverse schedule, forward schedule;
 * TB = transforms per wave, inverse {1, 2}, forward {1, 2, 4}; schedule 0 = one grouped DWT
 * launch then the bandwidths' FFT passes as parallel graph branches, 1 = one graph branch
 * per bandwidth holding its DWT and FFT passes (its forward DWT with 4 transforms per wave
 * from bw = 64, else 2), 2 = the bandwidths one after another on one stream through the
 * single-bandwidth complex64 kernels (what each bandwidth's *_many_c64 call runs), -1 = all
 * three timed on first use (in a graph), the fastest kept. */
int so3ft_group_layout(const int64_t *bws, int64_t n, int64_t *coeff_off, int64_t *grid_off,
                       int64_t *coeff_total, int64_t *grid_total);
int so3ft_group_create(const int64_t *bws, int64_t n, int64_t *group);
int so3ft_group_destroy(int64_t group);
int so3ft_group_isoft_c64(int64_t group, const void *coeffs, void *so3funcs, void *stream);
int so3ft_group_soft_c64(int64_t group, const void *so3funcs, void *coeffs, void *stream);
int so3ft_group_autotune(int64_t group, void *stream);
int so3ft_group_get_config(int64_t group, int32_t *config);
int so3ft_group_set_config(int64_t group, const int32_t *config);
int so3ft_group_set_graphs(int64_t group, int enable);  /* 0: launch eagerly on the stream */

/* ---- stage-level entry points: py_wrap.f90:190-233 -------------------------
 * fft_array has the SO(3)-function shape; entry (i_n, i_m, beta) holds
 * F(beta, m, n) with i = m>=0 ? m : 2bw+m (pysofft/utils.f90:45-57).
 * The inverse DWT writes every (m,n) column it owns and zeros the others. */
int so3ft_inverse_wigner_trf_cmplx(so3ft_handle h, const void *coeff, void *fft_array, void *stream);
int so3ft_forward_wigner_trf_cmplx(so3ft_handle h, const void *fft_array, void *coeff, void *stream);

/* ---- py_wrap.f90:417-436 py_fft / py_ifft: 2-D DFT over (alpha,gamma) of every
 * beta slice (FFTW_FORWARD / FFTW_BACKWARD sign), scaled by 1/(2bw) ---------- */
int so3ft_fft(so3ft_handle h, const void *f1, void *f2, void *stream);
int so3ft_ifft(so3ft_handle h, const void *f1, void *f2, void *stream);
/* ---- py_wrap.f90:437-456 py_rfft / py_irfft: r2c of a float64 (2bw)^3 array into the
 * half-complex complex128 (2bw, bw+1, 2bw) (C-order (gamma, alpha, beta), alpha halved)
 * and c2r back (FFTW_BACKWARD), both scaled by 1/(2bw) */
int so3ft_rfft(so3ft_handle h, const void *f1, void *f2, void *stream);
int so3ft_irfft(so3ft_handle h, const void *f1, void *f2, void *stream);

/* ---- py_wrap.f90:325-344 integrate_over_so3 (synchronises the stream) ----- */
int so3ft_integrate_over_so3_cmplx(so3ft_handle h, const void *f, double *re, double *im, void *stream);
int so3ft_integrate_over_so3_real(so3ft_handle h, const void *f, double *out, void *stream);

/* ---- rotational cross-correlation: py_wrap.f90:368-378, 379-391 ----------- */
int so3ft_cross_correlation_ylm_cmplx(so3ft_handle h, const void *f_lm, const void *g_lm, void *cc,
                                      void *stream);
/* radial_limits: 0-based inclusive [lo, hi] shell range (py_wrap passes 1-based);
 * radial_sampling_points is a HOST array of length n_radial. */
int so3ft_cross_correlation_ylm_cmplx_3d(so3ft_handle h, const void *f_lms, const void *g_lms, void *cc,
                                         const double *radial_sampling_points, int64_t n_radial,
                                         int64_t r_lo, int64_t r_hi, void *stream);

/* ---- real correlation: py_wrap.f90:392-402, 403-415.  f_ml, g_ml: complex128 MLr
 * coefficients of real functions (m >= 0, f^l_m at m(2bw-1-m)/2 + l; bw(bw+1)/2 entries
 * per set), cc: float64 (2bw, 2bw, 2bw).  3-D: (n_radial, bw(bw+1)/2) sets, limits as for
 * the complex 3-D call. */
int so3ft_cross_correlation_ylm_real(so3ft_handle h, const void *f_ml, const void *g_ml, void *cc, void *stream);
int so3ft_cross_correlation_ylm_real_3d(so3ft_handle h, const void *f_mls, const void *g_mls, void *cc,
                                        const double *radial_sampling_points, int64_t n_radial,
                                        int64_t r_lo, int64_t r_hi, void *stream);

/* ---- stage-level correlation DWTs: py_wrap.f90:346-356 py_inverse_wigner_trf_corr_cmplx
 * (LMc inputs, every valid (m,n) column written, the others zeroed) and :357-367
 * py_inverse_wigner_trf_corr_real (MLr inputs, the half-complex columns m >= 0 and
 * (0, -n) written, the others zeroed); fft_array complex128 (2bw, 2bw, 2bw) as for
 * so3ft_inverse_wigner_trf_cmplx. */
int so3ft_inverse_wigner_trf_corr_cmplx(so3ft_handle h, const void *f_lm, const void *g_lm, void *fft_array,
                                        void *stream);
int so3ft_inverse_wigner_trf_corr_real(so3ft_handle h, const void *f_ml, const void *g_ml, void *fft_array,
                                       void *stream);

/* ---- batched correlation (no counterpart in the reference, whose callers loop over
 * py_cross_correlation_ylm_cmplx): f_lms, g_lms (n, bw^2) LMc; cc (n, 2bw, 2bw, 2bw). */
int so3ft_cross_correlation_ylm_cmplx_many(so3ft_handle h, const void *f_lms, const void *g_lms, void *cc,
                                           int64_t n, void *stream);
/* correlation and peak search without keeping the grids: for each pair, the first
 * flat C-order (gamma, alpha, beta) index of max |C| into idx (DEVICE int64 (n,)) and
 * max |C| into val (DEVICE float64 (n,), may be NULL).  The peak is the relative
 * rotation as in tests/test_fortran_backend.py:695-722 (argmax of |corr|). */
int so3ft_cross_correlation_argmax_many(so3ft_handle h, const void *f_lms, const void *g_lms, int64_t n,
                                        void *idx, void *val, void *stream);

/* ---- Wigner-D rotation of SH coefficient sets (LMc layout, bw^2 each) by ZYZ Euler
 * angles, pysofft/soft.py:713-764 (numpy there; no f2py counterpart):
 *   y'^l_n = e^{-i n gamma} sum_m y^l_m e^{-i m alpha} d^l_{n,m}(beta).
 * euler is a DEVICE (nrot, 3) float64 array.  paired = 0: out (nrot, nset, bw^2), every
 * set by every rotation; paired = 1: nset == nrot, set i by rotation i, out (nrot, bw^2). */
int so3ft_rotate_ylm_cmplx(int64_t bw, const void *ylms, int64_t nset, const void *euler, int64_t nrot,
                           int paired, void *out, void *stream);

/* ---- learned alignment proposal (no counterpart in the reference) ---------------
 * A small MLP (BF16 operands on MFMA, FP32 accumulation, FP32 master weights and Adam)
 * maps pair features to a 6-D rotation representation (Gram-Schmidt to a rotation).
 * dims: nlayers+1 widths, dims[nlayers] == 6; hidden layers use GELU.  Parameters are
 * host float32, layer by layer [W (out x in, row-major), b (out)].  x, R9, T9 are DEVICE
 * float32 arrays (n, dims[0]) / (n, 9) (row-major rotation matrices).  The network
 * must be given parameters (so3ft_mlp_set_params) before its first forward pass.
 * Errors of these calls are reported by so3ft_mlp_last_error(). */
const char *so3ft_mlp_last_error(void);
int so3ft_mlp_create(const int32_t *dims, int nlayers, int64_t *handle);
int so3ft_mlp_destroy(int64_t handle);
int so3ft_mlp_num_params(int64_t handle, int64_t *n);
int so3ft_mlp_set_params(int64_t handle, const float *params);
int so3ft_mlp_get_params(int64_t handle, int which, float *params);  /* 0 weights, 1 gradients */
int so3ft_mlp_forward(int64_t handle, const float *x, int64_t n, float *R9, float *out6, void *stream);
/* forward, loss mean ||R - T||_F^2 (HOST float *loss, may be NULL), backward; one Adam step
 * when lr > 0 (lr == 0 leaves the weights and only computes the gradients) */
int so3ft_mlp_train_step(int64_t handle, const float *x, const float *T9, int64_t n, float lr, float *loss,
                         void *stream);
/* features of n pairs of LMc coefficient sets (complex128, (n, bw^2), 9 <= bw <= 64) into
 * x (DEVICE float32 (n, 452)): the l <= 8 coefficients of f and g (re, im) and the log
 * per-degree powers of f and g for l < 64, scaled by the l <= 8 RMS of f; standardised
 * with mean/std (DEVICE float32 (452)) when given. */
int so3ft_align_features(int64_t bw, const void *f_lms, const void *g_lms, int64_t n, const float *mean,
                         const float *stdv, float *x, void *stream);
/* bf16 GEMM test hook: C (M x N float32) = A (M x K) B (K x N), row-major bf16, DEVICE */
int so3ft_gemm_bf16(const void *A, const void *B, float *C, int M, int N, int K, void *stream);
/* local correlation refinement: for each pair, the correlation |C| (degrees < lmax <= 64) on
 * the 5 x 5 x 5 window of the grid of bandwidth grid_bw around euler_in (DEVICE float64
 * (n, 3)); euler_out (DEVICE (n, 3)) gets the window's argmax, score (DEVICE (n,), may be
 * NULL) its |C|, on the scale of the full correlation grid. */
int so3ft_refine_correlation_peaks(int64_t bw, const void *f_lms, const void *g_lms, int64_t n,
                                   const double *euler_in, double *euler_out, double *score, int64_t grid_bw,
                                   int64_t lmax, void *stream);

/* ---- measurement hook (no counterpart in the reference): launches one kernel of
 * the pipeline `reps` times on `stream` between two HIP events and returns the mean
 * duration of one launch.  which: 0 inverse DWT (a = coeffs, b = T x (2bw)^3 scratch),
 * 1 forward DWT (a = scratch, b = coeffs), 2 FFT pass along alpha (a -> b),
 * 3 FFT pass along gamma (a -> b), 4 / 5 the default engine's FP64 inverse (VALU) /
 * forward (FP64 MFMA, lane = beta), 6 / 7 the mixed-precision (complex64) inverse /
 * forward DWT, 8 / 9 the complex64 FFT passes along alpha / gamma, 12 the FP64 forward
 * on the VALU, 13 a plain streaming copy of T complex64 grids a -> b (the bandwidth
 * reference of the FFT passes).  T transforms per launch. */
int so3ft_bench_kernel(so3ft_handle h, int which, int reps, void *a, void *b, int64_t T, void *stream,
                       float *ms_per_launch);

/* ---- Wigner-d utilities of the reference's make_wigner module (pysofft/__init__.py:12-13
 * exports it as `pysofft.wigner`).  Arrays are DEVICE float64, betas contiguous.
 * so3ft_wigner_mn_recurrence_risbo: wigner.f90:529-585, one step l-1 -> l of Risbo's
 * recurrence for the full small-d matrix at nbeta angles: d_prev (nbeta, (2l-1)^2) ->
 * d_out (nbeta, (2l+1)^2), rows [i][j] row-major per beta; cos_b, sin_b: cos/sin of the
 * half angles (the reference's rotate_ylm passes cos(b/2), sin(-b/2)); sqrts[k] = sqrt(k),
 * nsqrt >= max(3, 2l+1).  d_prev is not read for l < 2. */
int so3ft_wigner_mn_recurrence_risbo(const void *d_prev, int64_t l, const void *cos_b, const void *sin_b,
                                     const void *sqrts, int64_t nsqrt, int64_t nbeta, void *d_out, void *stream);
/* wigner.f90:626-748 wigner_recurrence_risbo_reduced: one step l-1 -> l of the symmetry-reduced
 * recurrence, d^l_{m,n} for 0 <= m <= n <= l at nbeta betas that are closed under
 * beta -> pi - beta in reversed order (the step reads pi - beta as the reversed array):
 * dl_prev (l(l+1)/2, nbeta) -> dl_out ((l+1)(l+2)/2, nbeta), rows tri(m, n) = m(2s-m+1)/2 + n-m
 * on the triangle of side s; normalized: d~ = sqrt((2l+1)/2) d; sqrts only read at l = 1. */
int so3ft_wigner_recurrence_risbo_reduced(const void *dl_prev, int64_t l, const void *cos_b, const void *sin_b,
                                          const void *sqrts, int64_t nsqrt, int64_t nbeta, int normalized,
                                          void *dl_out, void *stream);
/* wigner.f90:784-822 genwig_all_risbo (recurrence_type 1) / :423-465 genwig_all_kostelec (0):
 * every normalised d~^l_{m,n}(beta_k), 0 <= m <= n <= l < bw, at the 2bw beta samples, into out
 * (bw(2 + 3bw + bw^2)/6 rows in mnl order, utils.f90:683-695, x 2bw) */
int so3ft_genwig_all(int64_t bw, int64_t recurrence_type, void *out, void *stream);
/* wigner.f90:844-934 full_big_d_risbo: D^l_{mn}(alpha, beta, gamma) for every (l, m, n) in mnl
 * order into out (complex128, NC(bw)), with that routine's signs; normalized: d~ and / (2 pi).
 * Synchronises the stream. */
int so3ft_full_big_d_risbo(int64_t bw, double alpha, double beta, double gamma, int normalized, void *out,
                           void *stream);

/* ---- utils: utils.f90:471-505 enforce_real_sym_mnl(_many), in place on device */
int so3ft_enforce_real_sym_mnl(int64_t bw, void *coeffs, int64_t n, void *stream);
/* ---- utils.f90:510-560 enforce_real_sym_lmn(_many) on lmn-ordered sets, in place on
 * device (synchronises the stream) */
int so3ft_enforce_real_sym_lmn(int64_t bw, void *coeffs, int64_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SO3FT_H */
