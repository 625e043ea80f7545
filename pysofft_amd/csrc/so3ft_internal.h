// Internal declarations shared by the libso3ft translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace so3 {

typedef double double4_t __attribute__((ext_vector_type(4)));

// Variant table for a pair 0 <= p <= q < bw (pysofft/soft.f08:534-631): the
// (m, n) orders that share d~^l_{p,q}(beta) through the symmetries of
// docs/defs/wigner.md:30-40.  Index v:
//   0 (p,q)   1 (-q,-p)   2 (q,p)  (-1)^(p+q)   3 (-p,-q) (-1)^(p+q)
//   4 (p,-q)  (-1)^(p+l) reversed beta     5 (-p,q)  (-1)^(q+l) reversed
//   6 (q,-p)  (-1)^(p+l) reversed          7 (-q,p)  (-1)^(q+l) reversed
struct Variant {
    int m, n;
    double sign;   // constant sign
    bool lsign;    // extra (-1)^l
    bool rev;      // beta -> pi - beta (index k -> N-1-k)
    bool valid;
};

__host__ __device__ inline Variant variant(int v, int p, int q)
{
    Variant r;
    const double spq = ((p + q) & 1) ? -1.0 : 1.0;
    const double sp = (p & 1) ? -1.0 : 1.0;
    const double sq = (q & 1) ? -1.0 : 1.0;
    switch (v) {
    case 0: r = {p, q, 1.0, false, false, true}; break;
    case 1: r = {-q, -p, 1.0, false, false, !(p == 0 && q == 0)}; break;
    case 2: r = {q, p, spq, false, false, p != q}; break;
    case 3: r = {-p, -q, spq, false, false, p != q}; break;
    case 4: r = {p, -q, sp, true, true, p != 0}; break;
    case 5: r = {-p, q, sq, true, true, p != 0}; break;
    case 6: r = {q, -p, sp, true, true, p != 0 && p != q}; break;
    default: r = {-q, p, sq, true, true, p != 0 && p != q}; break;
    }
    return r;
}

// DWT coefficient/row assembly modes
enum DwtMode : int {
    DWT_COMPLEX = 0,   // isoft/soft (soft.f08:534-631, 937-1023)
    DWT_REAL = 1,      // irsoft/rsoft half-complex semantics (soft.f08:1259-1318, 1620-1707)
    DWT_CORR = 2,      // cross-correlation coefficients from SH pairs (soft.f08:2242-2341)
};

struct DeviceTables {
    double *cosb = nullptr;    // [N]    cos(beta_k)
    double *w = nullptr;       // [N]    Driscoll-Healy weights
    double *seeds = nullptr;   // [ntri][N]  d~^q_{p,q}(beta_k)
    int64_t *slice = nullptr;  // [(2bw-1)^2] mnl start of f^l_{m,n}
    int32_t *pairs = nullptr;  // [npair] p | q << 16, ordered by decreasing bw-q
    double2 *tw = nullptr;     // [N]    exp(-2 pi i k / N)
    double *rc = nullptr;      // [sum over pairs of (bw - q - 1 + REC_PAD)][3] recurrence constants (a, x, b)
    int64_t *rc_off = nullptr; // [ntri]  first entry of pair tri(p, q) in rc
    double *dtab = nullptr;    // [mnl row][N] d~^l_{p,q}(beta_k), 0 <= p <= q <= l < bw (null: on the fly)
};

// first row of pair (p, q) in the d table: rows in mnl order, pair-major then degree
// (pysofft/utils.f90:683-695, zero-based); row of (p, q, l) = dtab_row(p, q) + l - q
__host__ __device__ inline int64_t dtab_row(int64_t p, int64_t q, int64_t bw)
{
    return (p * (3 * bw * bw + p * p + 2 - 3 * bw * (p - 2) - 3 * p)) / 6 + ((q - p) * (2 * bw + 1 - p - q)) / 2;
}
// doubles of a plan's d table, and of the scratch its Risbo fill needs
int64_t dtab_entries(int bw);
int64_t dtab_work_entries(int bw);
// the d table by the reduced Risbo recurrence (wigner.f90:784-822) / by the Kostelec
// recurrence from the plan's seeds and constants (wigner.f90:423-465)
hipError_t fill_dtab_risbo(int bw, double *tab, double *work, hipStream_t s);
hipError_t fill_dtab_kostelec(const DeviceTables &tb, int bw, double *tab, hipStream_t s);
// one step l-1 -> l of the reduced Risbo recurrence at nb betas (closed under beta -> pi -
// beta, reversed): d1 [l(l+1)/2][nb] -> d [(l+1)(l+2)/2][nb], h scratch [l(l+1)/2][nb];
// c, s: half-angle cos / sin; sq2 = sqrt(2) as the caller's table holds it; tab (nb = 2bw):
// also write degree l into a plan table
hipError_t launch_risbo_reduced_step(const double *d1, double *h, double *d, int l, int nb, const double *c,
                                     const double *s, double sq2, int normalized, double *tab, int bw, hipStream_t st);

// recurrence constants of pair (p, q) for the mixed-precision kernels: entry i = (a, x, b)
// of the step from degree q + i to q + i + 1, d^{l+1} = a (cos(beta) - x) d^l - b d^{l-1},
// for 0 <= i < L - 1 (L = bw - q); the next REC_PAD entries are zero, so a chain run past
// the last degree stays at zero with no per-degree test (up to a 32-degree tile past it)
constexpr int REC_PAD = 32;

struct FftDesc {
    int N = 0;
    int nrad = 0;
    int rad[32] = {0};
};

// launches (stream-ordered; return hipError_t of the launch)
hipError_t launch_dwt_inverse(const DeviceTables &t, int bw, int N, int npair, int mode,
                              const double2 *coeff, int64_t c_ts, const double2 *aux,
                              double2 *F, int64_t f_ts, int T, hipStream_t s);
hipError_t launch_dwt_forward(const DeviceTables &t, int bw, int N, int npair, int mode,
                              const double2 *F, int64_t f_ts, double2 *coeff, int64_t c_ts, int T,
                              hipStream_t s);

// FP64 on the vector ALUs (lane = beta for the inverse, lane = degree for the forward)
// (modes DWT_COMPLEX and DWT_REAL; DWT_CORR forms the correlation coefficients of the SH
// pairs coeff (f) and aux (g), LMc sets at stride c_ts, as it loads them, degrees > lcap zero)
hipError_t launch_dwt_inverse_valu(const DeviceTables &t, int bw, int N, int npair, int mode,
                                   const double2 *coeff, int64_t c_ts, const double2 *aux, int lcap,
                                   double2 *F, int64_t f_ts, int T, hipStream_t s);
hipError_t launch_dwt_forward_valu(const DeviceTables &t, int bw, int N, int npair, int mode,
                                   const double2 *F, int64_t f_ts, double2 *coeff, int64_t c_ts, int T,
                                   hipStream_t s);

// mnl correlation coefficients of T pairs of SH sets (transform stride t_ts; layout 0 LMc,
// 1 MLr of real functions), each the w-weighted sum over nr shells at stride r_ts (w ==
// nullptr: nr = 1, weight 1), into out (stride c_ts); degrees l > lcap are set to zero
hipError_t launch_corr_coeffs(const DeviceTables &t, int bw, const double2 *f, const double2 *g, int64_t t_ts,
                              int layout, int nr, int64_t r_ts, const double *w, double2 *out, int64_t c_ts,
                              int lcap, int T, hipStream_t s);

// mixed precision: complex64 coefficients and DWT arrays, FP64 recurrence, FP32 MFMA
// with d split into hi + lo, FP64 accumulation across 16-term blocks
hipError_t launch_dwt_inverse_mixed(const DeviceTables &t, int bw, int N, int npair, int mode,
                                    const float2 *coeff, int64_t c_ts, const float2 *aux, float2 *F, int64_t f_ts,
                                    int T, hipStream_t s);
hipError_t launch_dwt_forward_mixed(const DeviceTables &t, int bw, int N, int npair, int mode, const float2 *F,
                                    int64_t f_ts, float2 *coeff, int64_t c_ts, int T, hipStream_t s);

// one bandwidth of a grouped mixed-precision batch: its plan tables and the element
// offsets of its first transform in the packed coefficient / grid arrays (the segment's
// transforms are contiguous, nc resp. N^3 apart)
struct MixSegment {
    const double *cosb, *w, *seeds, *rc;
    const int64_t *slice, *rc_off;
    int bw, N, T, pad;
    int64_t coff, goff, nc;
};
// work items: x = segment (-1 = padding), y = p | q << 16, z = first transform, w = chunk
// (64 betas for the inverse; 0 for the forward, whose item covers every beta and degree)
hipError_t launch_dwt_inverse_mixed_grouped(int TB, const float2 *coeff, float2 *F, const MixSegment *segs,
                                            const int4 *items, int nitem, int grid, int max_bw, hipStream_t s);
hipError_t launch_dwt_forward_mixed_grouped(int TB, const float2 *F, float2 *coeff, const MixSegment *segs,
                                            const int4 *items, int nitem, int grid, int max_bw, hipStream_t s);
int mixed_grouped_occupancy(int inverse, int TB, int max_bw);
// FP64 forward DWT (modes DWT_COMPLEX and DWT_REAL) on FP64 MFMA with the lane-beta
// structure of the mixed forward (dwt_mixed.hip)
hipError_t launch_dwt_forward_m64(const DeviceTables &tb, int bw, int N, int npair, int mode, const double2 *F,
                                  int64_t f_ts, double2 *coeff, int64_t c_ts, int T, hipStream_t s);

struct FftAxis {
    int64_t in_t, in_o, in_r;      // element strides: transform, outer line, fft row
    int64_t out_t, out_o, out_r;
    int mask_rows;                 // 1: treat rows whose order is Nyquist or |order| > lmax as zero
    int mask_outer;                // 1: whole line zero when the outer order is invalid
    int in_real, out_real;         // input float64 (imag 0) / output real part into float64
    int sign;                      // -1 FFTW_FORWARD, +1 FFTW_BACKWARD
    double scale;
};
hipError_t launch_fft_axis(const FftDesc &fd, const double2 *tw, int bw, int lmax, const void *in,
                           void *out, const FftAxis &ax, int T, hipStream_t s);

bool fft_reg_supported(int N);
// (value, index) of a peak search: the larger value wins, the smaller index breaks ties
__device__ __forceinline__ void amax_better(double &v, int64_t &i, double v2, int64_t i2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}
// block-wide reduction of every thread's (v, i) into thread 0's (blocks of up to 1024
// threads; sv / si hold one entry per wave; every thread of the block calls it)
__device__ __forceinline__ void amax_block_reduce(double &v, int64_t &i, double *sv, int64_t *si)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_down(v, o, 64);
        const int64_t i2 = __shfl_down(i, o, 64);
        amax_better(v, i, v2, i2);
    }
    if (lane == 0) { sv[wid] = v; si[wid] = i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) amax_better(v, i, sv[w], si[w]);
    }
}
// per transform, the best of nblk partial (|z|^2, index) pairs: idx[t], and sqrt of the value
// into val[t] (val may be null)
hipError_t launch_argmax_final(const double *pv, const int64_t *pi, int nblk, int T, int64_t *idx, double *val,
                               hipStream_t s);
// FP64 pass along one axis (complex128 in, complex128 not written) fused with the peak
// search of |out|^2: each block leaves its best (|z|^2, C-order index within the transform)
// in pv / pi [T][*nblk].  hipErrorNotSupported where no fused shape exists (run the pass and
// launch_argmax_abs instead).  Partials per transform: fft_reg_argmax_partials(N).
hipError_t launch_fft_reg_argmax(int N, const double2 *tw, int bw, int lmax, const double2 *in, const FftAxis &ax,
                                 int T, double *pv, int64_t *pi, int *nblk, hipStream_t s);
int fft_reg_argmax_partials(int N);
// the LDS Stockham pass for any N (mixed radices), complex64 or complex128 storage
hipError_t launch_fft_stockham(int N, const double2 *tw, int bw, int lmax, const void *in, void *out,
                               const FftAxis &ax, int T, int in_c64, int out_c64, int col0, int ncols, hipStream_t s);
// plain device copy of nvec 16-byte vectors (16-byte aligned pointers): bandwidth reference
hipError_t launch_copy16(const void *a, void *b, int64_t nvec, hipStream_t s);
// register two-step pass for N = 2^a or 3 * 2^a (other N: the Stockham pass); in_c64 /
// out_c64 select complex64 storage; only beta columns [col0, col0+ncols) are processed
hipError_t launch_fft_reg(int N, const double2 *tw, int bw, int lmax, const void *in, void *out,
                          const FftAxis &ax, int T, int in_c64, int out_c64, int col0, int ncols, hipStream_t s);

// zero the fft-array columns whose order is Nyquist or beyond lmax; half: also every
// column with m < 0 (the half-complex array of the real transforms)
hipError_t launch_zero_invalid_columns(double2 *F, int64_t f_ts, int bw, int N, int lmax, int T, int half,
                                       hipStream_t s);
hipError_t launch_scale(double2 *X, int64_t n, double scale, hipStream_t s);
hipError_t launch_integrate(const void *f, int is_real, const double *w, int N, double *partials,
                            int nblocks, double *result2, hipStream_t s);
// f^l_{-m,-n} = (-1)^{m+n} conj f^l_{m,n} on mnl arrays; m0_neg_src selects the m = 0
// source side (1: n < 0 as utils.f90:471-488, 0: n > 0 as the lmn routine utils.f90:510-550)
hipError_t launch_enforce_real_sym(double2 *c, int64_t c_ts, const int64_t *slice, int bw, int T, int m0_neg_src,
                                   hipStream_t s);
// half-complex (N, N/2+1, N) <-> full (N, N, N) complex arrays over the alpha axis
hipError_t launch_halfcomplex(const double2 *in, double2 *out, int N, int to_full, hipStream_t s);
// lmn <-> mnl reordering of T coefficient sets (esz 16: complex128, 8: complex64); only
// the first nkeep lmn positions (the degrees l < L for nkeep = NC(L)) are carried: to mnl
// the others become zero, from mnl they are not written
hipError_t launch_permute_lmn(const int32_t *perm, int64_t nc, int64_t nkeep, const void *in, int64_t in_ts,
                              void *out, int64_t out_ts, int esz, int to_mnl, int T, hipStream_t s);

// per transform: first flat index (and value) of max |x| over n_per complex values;
// pv / pi hold argmax_partials(n_per) * T partials
int argmax_partials(int64_t n_per);
hipError_t launch_argmax_abs(const double2 *x, int64_t n_per, int T, double *pv, int64_t *pi, int64_t *idx,
                             double *val, hipStream_t s);

// shared with the other translation units of the C ABI: set the calling thread's
// so3ft_last_error() message (returns code); device tables of a plan handle
int set_error(int code, const std::string &msg);
int plan_device_tables(int64_t h, DeviceTables *tb, int64_t *bw, int64_t *lmax);

int dwt_lds_bytes_inverse(int bw, int N);
int dwt_lds_bytes_forward(int bw, int N);
hipError_t dwt_set_attributes(int bw, int N);

}  // namespace so3
