/*
 * so3_oracle.c -- CPU restatement of PySOFFT's SO(3) Fourier transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it; the product library (pysofft_amd) never links, loads or calls it.
 *
 * Everything is FP64, written from the reference's algorithm description; each
 * function cites the reference file:line it follows (paths relative to the
 * reference checkout of TimBerberich/pysofft v0.9.4).  Parity status of each
 * function (what pins it) is recorded in tests/test_oracle_*.py.
 *
 * Layouts (identical to the reference):
 *   coefficients : "mnl" order, NC(bw) = (4bw^3-bw)/3 complex128
 *                  (utils.f90:87-116 slice starts, utils.f90:140-201 loop order)
 *   so3 function : C-order (gamma, alpha, beta), i.e. index (jg*N + ja)*N + kb,
 *                  N = 2*bw  (docs/defs/so3_func.md:5-13)
 *   fft array    : same layout, entry (i_n, i_m, beta) with i = m>=0 ? m : N+m
 *                  (utils.f90:45-57 order_to_ids)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;
typedef int64_t i64;

static const double PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------------- */
/* index helpers                                                              */
/* ------------------------------------------------------------------------- */

/* utils.f90:32-36 */
i64 or_total_num_coeffs(i64 bw) { return (4 * bw * bw * bw - bw) / 3; }

/* utils.f90:652-656 (0-based) */
static inline i64 tri(i64 p, i64 q, i64 bw) { return (p * (2 * bw - p + 1)) / 2 + q - p; }

/* utils.f90:45-57 (0-based) */
static inline i64 wrap(i64 m, i64 N) { return m >= 0 ? m : N + m; }

static inline i64 mnidx(i64 m, i64 n, i64 bw) { return (m + bw - 1) * (2 * bw - 1) + (n + bw - 1); }

/*
 * Slice-start table for the mnl layout.  Rather than restating the closed form of
 * coeff_slice_mnl (utils.f90:87-116), it walks the loop order that layout is
 * defined by (utils.f90:140-201, docs/defs/coeff.md "mnl" example): for each
 * 0<=m1<=m2<bw the variants (m1,m2),(-m2,-m1),(m2,m1),(-m1,-m2),(m1,-m2),(-m1,m2),
 * (m2,-m1),(-m2,m1) (duplicates skipped) each own bw-m2 consecutive entries.
 * table[mnidx(m,n)] = 0-based start of f^l_{m,n}, l = max(|m|,|n|) .. bw-1.
 */
void or_slice_table(i64 bw, i64 *table)
{
    i64 pos = 0;
    for (i64 m1 = 0; m1 < bw; ++m1) {
        for (i64 m2 = m1; m2 < bw; ++m2) {
            i64 L = bw - m2;
            table[mnidx(m1, m2, bw)] = pos; pos += L;
            if (m1 == 0 && m2 == 0) continue;
            table[mnidx(-m2, -m1, bw)] = pos; pos += L;
            if (m1 != m2) {
                table[mnidx(m2, m1, bw)] = pos; pos += L;
                table[mnidx(-m1, -m2, bw)] = pos; pos += L;
            }
            if (m1 == 0 || m2 == 0) continue;
            table[mnidx(m1, -m2, bw)] = pos; pos += L;
            table[mnidx(-m1, m2, bw)] = pos; pos += L;
            if (m1 == m2) continue;
            table[mnidx(m2, -m1, bw)] = pos; pos += L;
            table[mnidx(-m2, m1, bw)] = pos; pos += L;
        }
    }
}

/* utils.f90:140-201: (l,m,n) of every entry in storage order, out is NC x 3 */
void or_coeff_degrees(i64 bw, i64 *lmn)
{
    i64 *tab = malloc(sizeof(i64) * (2 * bw - 1) * (2 * bw - 1));
    or_slice_table(bw, tab);
    for (i64 m = -(bw - 1); m < bw; ++m)
        for (i64 n = -(bw - 1); n < bw; ++n) {
            i64 a = labs(m) > labs(n) ? labs(m) : labs(n);
            i64 s = tab[mnidx(m, n, bw)];
            for (i64 l = a; l < bw; ++l) {
                lmn[3 * (s + l - a) + 0] = l;
                lmn[3 * (s + l - a) + 1] = m;
                lmn[3 * (s + l - a) + 2] = n;
            }
        }
    free(tab);
}

/* ------------------------------------------------------------------------- */
/* sampling and quadrature                                                    */
/* ------------------------------------------------------------------------- */

/* wigner.f90:34-39: (2i+1)pi/(2n) */
void or_beta_samples(i64 n, double *b)
{
    double f = PI / (double)(2 * n);
    for (i64 i = 0; i < n; ++i) b[i] = (double)(2 * i + 1) * f;
}

/* utils.f90:389-404: Driscoll-Healy weights for the 2bw Chebyshev nodes */
void or_legendre_weights(i64 bw, double *w)
{
    double xi = PI / (double)(4 * bw);
    for (i64 k = 0; k < 2 * bw; ++k) {
        double ko = (double)(2 * k + 1), s = 0.0;
        for (i64 j = 0; j < bw; ++j) {
            double jo = (double)(2 * j + 1);
            s += (1.0 / jo) * sin(ko * jo * xi);
        }
        w[k] = s * ((2.0 / (double)bw) * sin(ko * xi));
    }
}

/*
 * Seeds d~^{q}_{p,q}(beta_k) for 0<=p<=q<bw by the stair recursion:
 * wigner.f90:130-157 (dlml_recursion_l_contiguous, normalized: nc1=2,nc2=3) driven
 * by wigner.f90:164-187 (compute_all_dlml_l_contiguous).  out is [tri(p,q)][k].
 */
void or_seeds(i64 bw, i64 N, const double *sincos, const double *cos2, double *out)
{
    double *t = malloc(sizeof(double) * bw * N);
    for (i64 k = 0; k < N; ++k) t[k] = sqrt(0.5);
    for (i64 m = 0; m < bw; ++m) {
        i64 sw = m + 1, sw2 = (m == 0) ? m : sw;
        i64 lim = (bw - m) < sw2 ? (bw - m) : sw2;
        for (i64 i = 1; i <= lim; ++i) {              /* (l,m-1) -> (l+1,m) */
            i64 l = m + (i - 1) - 1;
            double f = sqrt((double)((2 * l + 2) * (2 * l + 3)) / (double)((l + m) * (l + m + 1)));
            for (i64 k = 0; k < N; ++k) t[(i - 1) * N + k] = t[(i - 1) * N + k] * f * cos2[k];
        }
        for (i64 i = sw; i <= bw - m - 1; ++i) {      /* (l,m) -> (l+1,m)   */
            i64 l = m + (i - 1);
            double f = sqrt((double)((2 * l + 2) * (2 * l + 3)) / (double)((l + m + 1) * (l - m + 1)));
            for (i64 k = 0; k < N; ++k) t[i * N + k] = t[(i - 1) * N + k] * f * sincos[k];
        }
        memcpy(out + tri(m, m, bw) * N, t, sizeof(double) * (bw - m) * N);
    }
    free(t);
}

/*
 * d~^l_{p,q}(beta_k) for l = q .. bw-1 into wig[(l-q)*N + k] (l-major, like
 * genwig_l2_trsp wigner.f90:317-345), three-term recurrence wigner.f90:227-260.
 */
void or_genwig(i64 p, i64 q, i64 bw, i64 N, const double *cosb, const double *seed, double *wig)
{
    i64 L = bw - q;
    memcpy(wig, seed, sizeof(double) * N);
    for (i64 i = 1; i < L; ++i) {
        i64 l = q + i - 1;
        const double *dl = wig + (i - 1) * N;
        const double *dm = (i >= 2) ? wig + (i - 2) * N : NULL;
        double *dn = wig + i * N;
        double cc = (double)(l + 1) *
                    sqrt(1.0 / (double)(((l + 1) * (l + 1) - p * p) * ((l + 1) * (l + 1) - q * q)));
        cc = cc * sqrt((double)(2 * l + 3));
        double cl = sqrt((double)(2 * l + 1));
        if (l == 0) {
            for (i64 k = 0; k < N; ++k) dn[k] = cc * cl * cosb[k] * dl[k];
        } else {
            double x = (double)(p * q) / (double)(l * (l + 1));
            for (i64 k = 0; k < N; ++k) dn[k] = cc * cl * (cosb[k] - x) * dl[k];
            if (l != q) {
                double c1 = sqrt((double)((l * l - p * p) * (l * l - q * q))) / (double)l;
                c1 = c1 / sqrt((double)(2 * l - 1));
                for (i64 k = 0; k < N; ++k) dn[k] = dn[k] - cc * c1 * dm[k];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* mixed-radix Stockham FFT along one axis of a row-batched array             */
/* ------------------------------------------------------------------------- */

typedef struct {
    i64 N, nrad, rad[64];
    zc *wf, *wb; /* exp(-2 pi i k/N), exp(+2 pi i k/N), k<N */
} offt;

static void offt_init(offt *f, i64 N)
{
    f->N = N; f->nrad = 0;
    i64 n = N;
    while (n % 4 == 0) { f->rad[f->nrad++] = 4; n /= 4; }
    while (n % 2 == 0) { f->rad[f->nrad++] = 2; n /= 2; }
    for (i64 p = 3; n > 1; p += 2)
        while (n % p == 0) { f->rad[f->nrad++] = p; n /= p; }
    f->wf = malloc(sizeof(zc) * N);
    f->wb = malloc(sizeof(zc) * N);
    for (i64 k = 0; k < N; ++k) {
        long double th = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
        double c = (double)cosl(th), s = (double)sinl(th);
        f->wf[k] = c - I * s;
        f->wb[k] = c + I * s;
    }
}

static void offt_free(offt *f) { free(f->wf); free(f->wb); }

/*
 * In-place DFT of length N along rows: X[r*rs + b], r<N, b<nb (b contiguous).
 * X'[k] = sum_r X[r] exp(sign*2 pi i r k/N).  tmp holds 2*N*nb complex.
 */
static void offt_rows(const offt *f, zc *X, i64 rs, i64 nb, int sign, zc *tmp)
{
    i64 N = f->N;
    if (N == 1) return;
    const zc *W = sign < 0 ? f->wf : f->wb;
    zc *A = tmp, *Bf = tmp + N * nb;
    for (i64 r = 0; r < N; ++r) memcpy(A + r * nb, X + r * rs, sizeof(zc) * nb);
    i64 Ns = 1;
    zc v[64], u[64];
    for (i64 s = 0; s < f->nrad; ++s) {
        i64 R = f->rad[s], NR = N / R;
        for (i64 j = 0; j < NR; ++j) {
            i64 k = j % Ns, dst = (j / Ns) * Ns * R + k;
            zc tw[64];
            for (i64 r = 0; r < R; ++r) tw[r] = W[(r * k * (N / (Ns * R))) % N];
            for (i64 b = 0; b < nb; ++b) {
                for (i64 r = 0; r < R; ++r) v[r] = A[(j + r * NR) * nb + b] * tw[r];
                if (R == 2) {
                    u[0] = v[0] + v[1]; u[1] = v[0] - v[1];
                } else if (R == 4) {
                    zc a0 = v[0] + v[2], a1 = v[0] - v[2], a2 = v[1] + v[3], a3 = v[1] - v[3];
                    zc ja3 = (sign < 0) ? (-I * a3) : (I * a3);
                    u[0] = a0 + a2; u[2] = a0 - a2; u[1] = a1 + ja3; u[3] = a1 - ja3;
                } else {
                    for (i64 qq = 0; qq < R; ++qq) {
                        zc acc = 0;
                        for (i64 r = 0; r < R; ++r) acc += v[r] * W[((r * qq) % R) * NR];
                        u[qq] = acc;
                    }
                }
                for (i64 r = 0; r < R; ++r) Bf[(dst + r * Ns) * nb + b] = u[r];
            }
        }
        zc *t = A; A = Bf; Bf = t;
        Ns *= R;
    }
    for (i64 r = 0; r < N; ++r) memcpy(X + r * rs, A + r * nb, sizeof(zc) * nb);
}

/* 2-D DFT over (m, n) of an N^3 array [i_n][i_m][beta] for every beta (the
 * FFTW plan of soft.f08:242-291: rank 2, n=[2bw,2bw], howmany=2bw, stride 2bw). */
static void offt2(const offt *f, zc *X, int sign, int use_mp)
{
    i64 N = f->N;
#pragma omp parallel if (use_mp)
    {
        zc *tmp = malloc(sizeof(zc) * 2 * N * N);
#pragma omp for schedule(static)
        for (i64 in = 0; in < N; ++in) offt_rows(f, X + in * N * N, N, N, sign, tmp); /* along m */
#pragma omp for schedule(static)
        for (i64 im = 0; im < N; ++im) offt_rows(f, X + im * N, N * N, N, sign, tmp); /* along n */
        free(tmp);
    }
}

/* ------------------------------------------------------------------------- */
/* plan                                                                        */
/* ------------------------------------------------------------------------- */

typedef struct {
    i64 bw, lmax, N, nc;
    double *beta, *cosb, *sincos, *cos2, *w, *seeds;
    i64 *slice;
    offt fft;
} oplan;

/* soft.f08:344-384 (init with kostelec on-the-fly recurrence) */
oplan *or_plan_create(i64 bw, i64 lmax)
{
    oplan *P = calloc(1, sizeof(oplan));
    i64 N = 2 * bw;
    P->bw = bw; P->lmax = lmax; P->N = N; P->nc = or_total_num_coeffs(bw);
    P->beta = malloc(sizeof(double) * N);
    P->cosb = malloc(sizeof(double) * N);
    P->sincos = malloc(sizeof(double) * N);
    P->cos2 = malloc(sizeof(double) * N);
    P->w = malloc(sizeof(double) * N);
    or_beta_samples(N, P->beta);
    for (i64 k = 0; k < N; ++k) {           /* wigner.f90:51-58 */
        P->cosb[k] = cos(P->beta[k]);
        P->sincos[k] = cos(P->beta[k] / 2) * sin(P->beta[k] / 2);
        P->cos2[k] = pow(cos(P->beta[k] / 2), 2);
    }
    or_legendre_weights(bw, P->w);
    P->seeds = malloc(sizeof(double) * (bw * (bw + 1) / 2) * N);
    or_seeds(bw, N, P->sincos, P->cos2, P->seeds);
    P->slice = malloc(sizeof(i64) * (2 * bw - 1) * (2 * bw - 1));
    or_slice_table(bw, P->slice);
    offt_init(&P->fft, N);
    return P;
}

void or_plan_destroy(oplan *P)
{
    if (!P) return;
    free(P->beta); free(P->cosb); free(P->sincos); free(P->cos2); free(P->w);
    free(P->seeds); free(P->slice); offt_free(&P->fft); free(P);
}

i64 or_plan_bw(const oplan *P) { return P->bw; }
void or_plan_weights(const oplan *P, double *w) { memcpy(w, P->w, sizeof(double) * P->N); }
void or_plan_seeds(const oplan *P, double *s) { memcpy(s, P->seeds, sizeof(double) * (P->bw * (P->bw + 1) / 2) * P->N); }

#define FIDX(in_, im_, k_) ((((in_) * N) + (im_)) * N + (k_))

/* ------------------------------------------------------------------------- */
/* discrete Wigner transforms, complex                                        */
/* ------------------------------------------------------------------------- */

/* one variant (m,n) of the inverse DWT: F(beta',m,n) = s0 * sum_l c_l [(-1)^l] d~^l_{pq}(beta)
 * (l outer, beta inner: the row-wise form of the reference's matmul(coeff_part, wig_mat)) */
static void inv_variant(const oplan *P, const zc *c, zc *F, i64 m, i64 n, const double *wig, i64 L,
                        double s0, int lsign, int rev)
{
    i64 N = P->N, bw = P->bw;
    const zc *cs = c + P->slice[mnidx(m, n, bw)];
    i64 im = wrap(m, N), in = wrap(n, N);
    double *ar = calloc(2 * N, sizeof(double)), *ai = ar + N;
    for (i64 i = 0; i < L; ++i) {
        i64 l = bw - L + i;
        double sl = (lsign && (l & 1)) ? -1.0 : 1.0;
        double cr = creal(cs[i]) * (sl * s0), ci = cimag(cs[i]) * (sl * s0);
        const double *wr = wig + i * N;
        for (i64 k = 0; k < N; ++k) { ar[k] += cr * wr[k]; ai[k] += ci * wr[k]; }
    }
    zc *Fc = F + FIDX(in, im, 0);
    for (i64 k = 0; k < N; ++k) Fc[rev ? N - 1 - k : k] = ar[k] + I * ai[k];
    free(ar);
}

/* soft.f08:534-631 inverse_wigner_loop_body_cmplx_kostelec */
static void inv_pair(const oplan *P, const zc *c, zc *F, i64 m1, i64 m2, double *wig)
{
    i64 bw = P->bw, N = P->N, L = bw - m2;
    or_genwig(m1, m2, bw, N, P->cosb, P->seeds + tri(m1, m2, bw) * N, wig);
    double s1 = (m1 & 1) ? -1.0 : 1.0, s2 = (m2 & 1) ? -1.0 : 1.0;
    inv_variant(P, c, F, m1, m2, wig, L, 1.0, 0, 0);
    if (m1 == 0 && m2 == 0) return;
    inv_variant(P, c, F, -m2, -m1, wig, L, 1.0, 0, 0);
    if (m1 != m2) {
        inv_variant(P, c, F, m2, m1, wig, L, s1 * s2, 0, 0);
        inv_variant(P, c, F, -m1, -m2, wig, L, s1 * s2, 0, 0);
    }
    if (m1 == 0 || m2 == 0) return;
    inv_variant(P, c, F, m1, -m2, wig, L, s1, 1, 1);
    inv_variant(P, c, F, -m1, m2, wig, L, s2, 1, 1);
    if (m1 == m2) return;
    inv_variant(P, c, F, m2, -m1, wig, L, s1, 1, 1);
    inv_variant(P, c, F, -m2, m1, wig, L, s2, 1, 1);
}

/* soft.f08:632-675 inverse_wigner_trf_cmplx_kostelec (F must be zeroed by caller) */
void or_inverse_dwt(const oplan *P, const zc *coeff, zc *F, int use_mp)
{
    i64 bw = P->bw, Lm = P->lmax, N = P->N;
    i64 npair = (Lm + 1) * (Lm + 2) / 2;
#pragma omp parallel if (use_mp)
    {
        double *wig = malloc(sizeof(double) * bw * N);
#pragma omp for schedule(dynamic)
        for (i64 t = 0; t < npair; ++t) {
            i64 m1 = 0, r = t;
            while (r >= Lm + 1 - m1) { r -= Lm + 1 - m1; ++m1; }
            inv_pair(P, coeff, F, m1, m1 + r, wig);
        }
        free(wig);
    }
}

/* one variant of the forward DWT: c_l = s0 [(-1)^l] sum_k w_k F(beta_k',m,n) d~^l(beta_k) */
static void fwd_variant(const oplan *P, const zc *F, zc *c, i64 m, i64 n, const double *wig, i64 L,
                        double s0, int lsign, int rev)
{
    i64 N = P->N, bw = P->bw;
    zc *cs = c + P->slice[mnidx(m, n, bw)];
    i64 im = wrap(m, N), in = wrap(n, N);
    double *wr = malloc(sizeof(double) * 2 * N), *wi = wr + N;
    const zc *Fc = F + FIDX(in, im, 0);
    for (i64 k = 0; k < N; ++k) {
        zc v = Fc[rev ? N - 1 - k : k];
        wr[k] = P->w[k] * creal(v); wi[k] = P->w[k] * cimag(v);
    }
    for (i64 i = 0; i < L; ++i) {
        i64 l = bw - L + i;
        const double *d = wig + i * N;
        double ar = 0, ai = 0;
        for (i64 k = 0; k < N; ++k) { ar += wr[k] * d[k]; ai += wi[k] * d[k]; }
        double sl = (lsign && (l & 1)) ? -1.0 : 1.0;
        cs[i] = (ar + I * ai) * (s0 * sl);
    }
    free(wr);
}

/* soft.f08:937-1023 forward_wigner_loop_body_cmplx_kostelec */
static void fwd_pair(const oplan *P, const zc *F, zc *c, i64 m1, i64 m2, double *wig)
{
    i64 bw = P->bw, N = P->N, L = bw - m2;
    or_genwig(m1, m2, bw, N, P->cosb, P->seeds + tri(m1, m2, bw) * N, wig);
    double s1 = (m1 & 1) ? -1.0 : 1.0, s2 = (m2 & 1) ? -1.0 : 1.0;
    fwd_variant(P, F, c, m1, m2, wig, L, 1.0, 0, 0);
    if (m1 == 0 && m2 == 0) return;
    fwd_variant(P, F, c, -m2, -m1, wig, L, 1.0, 0, 0);
    if (m1 != m2) {
        fwd_variant(P, F, c, m2, m1, wig, L, s1 * s2, 0, 0);
        fwd_variant(P, F, c, -m1, -m2, wig, L, s1 * s2, 0, 0);
    }
    if (m1 == 0 || m2 == 0) return;
    fwd_variant(P, F, c, m1, -m2, wig, L, s1, 1, 1);
    fwd_variant(P, F, c, -m1, m2, wig, L, s2, 1, 1);
    if (m1 == m2) return;
    fwd_variant(P, F, c, m2, -m1, wig, L, s1, 1, 1);
    fwd_variant(P, F, c, -m2, m1, wig, L, s2, 1, 1);
}

/* soft.f08:1024-1065 forward_wigner_trf_cmplx_kostelec */
void or_forward_dwt(const oplan *P, const zc *F, zc *coeff, int use_mp)
{
    i64 bw = P->bw, Lm = P->lmax, N = P->N;
    i64 npair = (Lm + 1) * (Lm + 2) / 2;
#pragma omp parallel if (use_mp)
    {
        double *wig = malloc(sizeof(double) * bw * N);
#pragma omp for schedule(dynamic)
        for (i64 t = 0; t < npair; ++t) {
            i64 m1 = 0, r = t;
            while (r >= Lm + 1 - m1) { r -= Lm + 1 - m1; ++m1; }
            fwd_pair(P, F, coeff, m1, m1 + r, wig);
        }
        free(wig);
    }
}

/* ------------------------------------------------------------------------- */
/* discrete Wigner transforms, real (half-complex)                            */
/* ------------------------------------------------------------------------- */

/* soft.f08:1259-1318 inverse_wigner_loop_body_real_kostelec: only m>=0 rows */
static void inv_pair_real(const oplan *P, const zc *c, zc *F, i64 m1, i64 m2, double *wig)
{
    i64 bw = P->bw, N = P->N, L = bw - m2;
    or_genwig(m1, m2, bw, N, P->cosb, P->seeds + tri(m1, m2, bw) * N, wig);
    double s1 = (m1 & 1) ? -1.0 : 1.0;
    inv_variant(P, c, F, m1, m2, wig, L, 1.0, 0, 0);
    if (m1 == 0 && m2 == 0) return;
    if (m1 != m2) {
        /* F(m2,m1) = conj( sum_l f^l_{-m2,-m1} d~^l_{m1,m2} ) */
        const zc *cs = c + P->slice[mnidx(-m2, -m1, bw)];
        i64 im = wrap(m2, N), in = wrap(m1, N);
        for (i64 k = 0; k < N; ++k) {
            zc acc = 0;
            for (i64 i = 0; i < L; ++i) acc += cs[i] * wig[i * N + k];
            F[FIDX(in, im, k)] = conj(acc);
        }
    }
    if (m1 == 0 || m2 == 0) return;
    inv_variant(P, c, F, m1, -m2, wig, L, s1, 1, 1);
    if (m1 == m2) return;
    inv_variant(P, c, F, m2, -m1, wig, L, s1, 1, 1);
}

/*
 * soft.f08:1319-1379 inverse_wigner_trf_real_kostelec, written into a full N^3
 * complex array: rows m>=0 as the reference, the (0,-n) fill of soft.f08:1372-1377,
 * and rows m<0 as the Hermitian completion the c2r plan implies.
 */
void or_inverse_dwt_real(const oplan *P, const zc *coeff, zc *F, int use_mp)
{
    i64 bw = P->bw, Lm = P->lmax, N = P->N;
    i64 npair = (Lm + 1) * (Lm + 2) / 2;
#pragma omp parallel if (use_mp)
    {
        double *wig = malloc(sizeof(double) * bw * N);
#pragma omp for schedule(dynamic)
        for (i64 t = 0; t < npair; ++t) {
            i64 m1 = 0, r = t;
            while (r >= Lm + 1 - m1) { r -= Lm + 1 - m1; ++m1; }
            inv_pair_real(P, coeff, F, m1, m1 + r, wig);
        }
        free(wig);
    }
    for (i64 n = 1; n <= Lm; ++n)
        for (i64 k = 0; k < N; ++k) F[FIDX(wrap(-n, N), 0, k)] = conj(F[FIDX(n, 0, k)]);
    for (i64 m = 1; m < bw; ++m)
        for (i64 in = 0; in < N; ++in)
            for (i64 k = 0; k < N; ++k)
                F[FIDX((N - in) % N, N - m, k)] = conj(F[FIDX(in, m, k)]);
}

/* soft.f08:1620-1707 forward_wigner_loop_body_real_kostelec (reads rows m>=0 only) */
static void fwd_pair_real(const oplan *P, const zc *F, zc *c, i64 m1, i64 m2, double *wig)
{
    i64 bw = P->bw, N = P->N, L = bw - m2;
    or_genwig(m1, m2, bw, N, P->cosb, P->seeds + tri(m1, m2, bw) * N, wig);
    double s1 = (m1 & 1) ? -1.0 : 1.0, s2 = (m2 & 1) ? -1.0 : 1.0;
    zc *cpp = c + P->slice[mnidx(m1, m2, bw)];
    fwd_variant(P, F, c, m1, m2, wig, L, 1.0, 0, 0);
    if (m2 == 0) return;
    /* (-m2,-m1) from the half-complex row m2 (soft.f08:3286-3300) */
    zc *cnn = c + P->slice[mnidx(-m2, -m1, bw)];
    {
        i64 im = wrap(m2, N), in = wrap(m1, N);
        for (i64 i = 0; i < L; ++i) {
            zc acc = 0;
            for (i64 k = 0; k < N; ++k) acc += (P->w[k] * conj(F[FIDX(in, im, k)])) * wig[i * N + k];
            cnn[i] = acc;
        }
    }
    if (m1 != m2) {
        zc *a = c + P->slice[mnidx(m2, m1, bw)], *b = c + P->slice[mnidx(-m1, -m2, bw)];
        for (i64 i = 0; i < L; ++i) { a[i] = (s1 * s2) * conj(cnn[i]); b[i] = (s1 * s2) * conj(cpp[i]); }
    }
    if (m1 == 0 || m2 == 0) return;
    fwd_variant(P, F, c, m1, -m2, wig, L, s1, 1, 1);
    {
        zc *a = c + P->slice[mnidx(-m1, m2, bw)], *b = c + P->slice[mnidx(m1, -m2, bw)];
        for (i64 i = 0; i < L; ++i) a[i] = (s1 * s2) * conj(b[i]);
    }
    if (m1 == m2) return;
    fwd_variant(P, F, c, m2, -m1, wig, L, s1, 1, 1);
    {
        zc *a = c + P->slice[mnidx(-m2, m1, bw)], *b = c + P->slice[mnidx(m2, -m1, bw)];
        for (i64 i = 0; i < L; ++i) a[i] = (s1 * s2) * conj(b[i]);
    }
}

/* soft.f08:1708-1749 forward_wigner_trf_real_kostelec */
void or_forward_dwt_real(const oplan *P, const zc *F, zc *coeff, int use_mp)
{
    i64 bw = P->bw, Lm = P->lmax, N = P->N;
    i64 npair = (Lm + 1) * (Lm + 2) / 2;
#pragma omp parallel if (use_mp)
    {
        double *wig = malloc(sizeof(double) * bw * N);
#pragma omp for schedule(dynamic)
        for (i64 t = 0; t < npair; ++t) {
            i64 m1 = 0, r = t;
            while (r >= Lm + 1 - m1) { r -= Lm + 1 - m1; ++m1; }
            fwd_pair_real(P, F, coeff, m1, m1 + r, wig);
        }
        free(wig);
    }
}

/* ------------------------------------------------------------------------- */
/* full transforms                                                             */
/* ------------------------------------------------------------------------- */

/* py_wrap.f90:417-426 / soft.f08:3256-3262 use a scaled variant; this is the raw
 * 2-D DFT used by the transforms: sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD. */
void or_fft2(const oplan *P, zc *X, int sign, int use_mp) { offt2(&P->fft, X, sign, use_mp); }

/* soft.f08:1887-1898 isoft_ */
void or_isoft(const oplan *P, const zc *coeff, zc *out, int use_mp)
{
    i64 N = P->N, G = N * N * N;
    memset(out, 0, sizeof(zc) * G);
    or_inverse_dwt(P, coeff, out, use_mp);
    offt2(&P->fft, out, -1, use_mp);
    double s = 1.0 / (2.0 * PI);
    for (i64 i = 0; i < G; ++i) out[i] *= s;
}

/* soft.f08:1910-1924 soft_ (input is not modified) */
void or_soft(const oplan *P, const zc *so3func, zc *coeff, int use_mp)
{
    i64 N = P->N, G = N * N * N;
    zc *F = malloc(sizeof(zc) * G);
    memcpy(F, so3func, sizeof(zc) * G);
    offt2(&P->fft, F, +1, use_mp);
    double s = 2.0 * PI / ((double)N * (double)N);
    for (i64 i = 0; i < G; ++i) F[i] *= s;
    or_forward_dwt(P, F, coeff, use_mp);
    free(F);
}

/* soft.f08:1937-1949 irsoft_: c2r of the half-complex DWT output */
void or_irsoft(const oplan *P, const zc *coeff, double *out, int use_mp)
{
    i64 N = P->N, G = N * N * N;
    zc *F = calloc(G, sizeof(zc));
    or_inverse_dwt_real(P, coeff, F, use_mp);
    offt2(&P->fft, F, -1, use_mp);
    double s = 1.0 / (2.0 * PI);
    for (i64 i = 0; i < G; ++i) out[i] = creal(F[i]) * s;
    free(F);
}

/* soft.f08:1961-1973 rsoft_: r2c (+conj) then the half-complex forward DWT */
void or_rsoft(const oplan *P, const double *so3func, zc *coeff, int use_mp)
{
    i64 N = P->N, G = N * N * N;
    zc *F = malloc(sizeof(zc) * G);
    for (i64 i = 0; i < G; ++i) F[i] = so3func[i];
    offt2(&P->fft, F, +1, use_mp);
    double s = 2.0 * PI / ((double)N * (double)N);
    for (i64 i = 0; i < G; ++i) F[i] *= s;
    or_forward_dwt_real(P, F, coeff, use_mp);
    free(F);
}

/* soft.f08:1985-2119 *_many: OpenMP over transforms, each transform serial */
void or_isoft_many(const oplan *P, const zc *c, zc *f, i64 n, int use_mp)
{
    i64 N = P->N, G = N * N * N;
#pragma omp parallel for schedule(dynamic) if (use_mp)
    for (i64 i = 0; i < n; ++i) or_isoft(P, c + i * P->nc, f + i * G, 0);
}
void or_soft_many(const oplan *P, const zc *f, zc *c, i64 n, int use_mp)
{
    i64 N = P->N, G = N * N * N;
#pragma omp parallel for schedule(dynamic) if (use_mp)
    for (i64 i = 0; i < n; ++i) or_soft(P, f + i * G, c + i * P->nc, 0);
}
void or_irsoft_many(const oplan *P, const zc *c, double *f, i64 n, int use_mp)
{
    i64 N = P->N, G = N * N * N;
#pragma omp parallel for schedule(dynamic) if (use_mp)
    for (i64 i = 0; i < n; ++i) or_irsoft(P, c + i * P->nc, f + i * G, 0);
}
void or_rsoft_many(const oplan *P, const double *f, zc *c, i64 n, int use_mp)
{
    i64 N = P->N, G = N * N * N;
#pragma omp parallel for schedule(dynamic) if (use_mp)
    for (i64 i = 0; i < n; ++i) or_rsoft(P, f + i * G, c + i * P->nc, 0);
}

/* soft.f08:2122-2145 integrate_over_so3_{cmplx,real} (real part only for complex) */
void or_integrate(const oplan *P, const zc *f, double *re, double *im)
{
    i64 N = P->N;
    zc tot = 0;
    for (i64 k = 0; k < N; ++k) {
        zc s = 0;
        for (i64 j = 0; j < N * N; ++j) s += f[j * N + k];
        tot += s * P->w[k];
    }
    tot *= pow(PI / (double)P->bw, 2);
    *re = creal(tot); *im = cimag(tot);
}

/* utils.f90:471-488 enforce_real_sym_mnl (same loop order, so the same half wins) */
void or_enforce_real_sym(i64 bw, zc *c)
{
    i64 *tab = malloc(sizeof(i64) * (2 * bw - 1) * (2 * bw - 1));
    or_slice_table(bw, tab);
    for (i64 i = 0; i < bw; ++i) c[tab[mnidx(0, 0, bw)] + i] = creal(c[tab[mnidx(0, 0, bw)] + i]);
    for (i64 m1 = 0; m1 < bw; ++m1)
        for (i64 m2 = -bw + 1; m2 < bw; ++m2) {
            double s = (((m1 + m2) & 1) ? -1.0 : 1.0);
            i64 a = labs(m1) > labs(m2) ? labs(m1) : labs(m2), L = bw - a;
            zc *src = c + tab[mnidx(m1, m2, bw)], *dst = c + tab[mnidx(-m1, -m2, bw)];
            for (i64 i = 0; i < L; ++i) dst[i] = s * conj(src[i]);
        }
    free(tab);
}

/* ------------------------------------------------------------------------- */
/* rotational cross-correlation                                               */
/* ------------------------------------------------------------------------- */

/*
 * Wigner coefficients of C(R) = int f(x) conj(g(Rx)) dx from SH coefficients in
 * the SHTNS LMc layout l(l+1)+m (utils.f90:735-741):
 * cc^l_{mn} = 2 pi sqrt(2/(2l+1)) (-1)^{m+n} f^l_m conj(g^l_n)
 * (soft.f08:2242-2341 loop body, wig_norm soft.f08:2367, transpose :2398-2403).
 */
void or_corr_coeffs(i64 bw, const zc *flm, const zc *glm, zc *cc)
{
    i64 *tab = malloc(sizeof(i64) * (2 * bw - 1) * (2 * bw - 1));
    or_slice_table(bw, tab);
    for (i64 m = -(bw - 1); m < bw; ++m)
        for (i64 n = -(bw - 1); n < bw; ++n) {
            i64 a = labs(m) > labs(n) ? labs(m) : labs(n);
            double s = (((m + n) & 1) ? -1.0 : 1.0);
            for (i64 l = a; l < bw; ++l) {
                double wn = 2.0 * PI * sqrt(2.0 / (double)(2 * l + 1));
                cc[tab[mnidx(m, n, bw)] + l - a] = wn * flm[l * (l + 1) + m] * conj(glm[l * (l + 1) + n]) * s;
            }
        }
    free(tab);
}

/* soft.f08:2613-2656 cross_correlation_ylm_cmplx: inverse DWT of cc, FFTW forward, /2pi */
void or_cross_correlation_cmplx(const oplan *P, const zc *flm, const zc *glm, zc *cc, int use_mp)
{
    zc *c = malloc(sizeof(zc) * P->nc);
    or_corr_coeffs(P->bw, flm, glm, c);
    or_isoft(P, c, cc, use_mp);
    free(c);
}

/* soft.f08:2657-2738 cross_correlation_ylm_cmplx_3d:
 * sum over shells r in [r_lo, r_hi] (0-based, inclusive) of r^2 C_r, / (r_hi-r_lo+dr) */
void or_cross_correlation_cmplx_3d(const oplan *P, const zc *flms, const zc *glms, i64 nr,
                                   const double *rad, i64 r_lo, i64 r_hi, zc *cc, int use_mp)
{
    i64 N = P->N, G = N * N * N, nb = P->bw * P->bw;
    zc *tmp = malloc(sizeof(zc) * G);
    memset(cc, 0, sizeof(zc) * G);
    double step = rad[1] - rad[0];
    double inv = 1.0 / (rad[r_hi] - rad[r_lo] + step);
    for (i64 r = r_lo; r <= r_hi; ++r) {
        or_cross_correlation_cmplx(P, flms + r * nb, glms + r * nb, tmp, use_mp);
        double r2 = rad[r] * rad[r];
        for (i64 i = 0; i < G; ++i) cc[i] += tmp[i] * r2;
    }
    for (i64 i = 0; i < G; ++i) cc[i] *= inv;
    free(tmp);
    (void)nr;
}

/* ------------------------------------------------------------------------- */
/* Wigner-D rotation of spherical-harmonic coefficients                      */
/* ------------------------------------------------------------------------- */

/*
 * Full small-d matrix recurrence of Risbo, degree l-1 -> l, wigner.f90:529-585.
 * dprev is (2l-1)^2, dout is (2l+1)^2, both row-major [i][j] (= Fortran index
 * i*(2l+1)+j+1).  c = cos(beta/2), s = sin(beta/2); sq[k] = sqrt(k).
 */
void or_risbo_step(const double *dprev, i64 l, double c, double s, const double *sq, double *dout)
{
    if (l == 0) { dout[0] = 1.0; return; }
    if (l == 1) {
        dout[0] = c * c; dout[1] = sq[2] * c * s; dout[2] = s * s;
        dout[3] = -dout[1]; dout[4] = dout[0] - dout[2]; dout[5] = dout[1];
        dout[6] = dout[2]; dout[7] = -dout[1]; dout[8] = dout[0];
        return;
    }
    i64 deg = 2 * l - 1, pd = deg + 1;
    double *t = calloc((size_t)pd * pd, sizeof(double));
    double inv = 1.0 / (double)deg;
    for (i64 i = 0; i < deg; ++i)
        for (i64 j = 0; j < deg; ++j) {
            double x = dprev[i * deg + j];
            t[i * pd + j]           += inv * sq[deg - i] * sq[deg - j] * x * c;
            t[(i + 1) * pd + j]     -= inv * sq[i + 1] * sq[deg - j] * x * s;
            t[i * pd + j + 1]       += inv * sq[deg - i] * sq[j + 1] * x * s;
            t[(i + 1) * pd + j + 1] += inv * sq[i + 1] * sq[j + 1] * x * c;
        }
    deg = 2 * l; pd = deg + 1; inv = 1.0 / (double)deg;
    memset(dout, 0, sizeof(double) * pd * pd);
    for (i64 i = 0; i < deg; ++i)
        for (i64 j = 0; j < deg; ++j) {
            double x = t[i * deg + j];
            dout[i * pd + j]           += inv * sq[deg - i] * sq[deg - j] * x * c;
            dout[(i + 1) * pd + j]     -= inv * sq[i + 1] * sq[deg - j] * x * s;
            dout[i * pd + j + 1]       += inv * sq[deg - i] * sq[j + 1] * x * s;
            dout[(i + 1) * pd + j + 1] += inv * sq[i + 1] * sq[j + 1] * x * c;
        }
    free(t);
}

/*
 * Symmetry-reduced Risbo recurrence, wigner.f90:626-748: d^{l-1}_{m,n} (0 <= m <= n <= l-1)
 * -> d^l_{m,n} (0 <= m <= n <= l) at nb betas, through the half degree l - 1/2.  Arrays
 * are [tri][nb] with tri(m, n) over the triangle of side l (input) / l + 1 (output) and
 * the betas contiguous; the set of betas must be closed under beta -> pi - beta in
 * reversed order (the step reads the half-degree values at pi - beta as the reversed
 * array).  c[k] = cos(beta_k / 2), s[k] = sin(beta_k / 2); sq[k] = sqrt(k) (only sq[2] is
 * read).  normalized: d~ = sqrt((2l+1)/2) d.
 */
void or_risbo_reduced_step(const double *d1, i64 l, i64 nb, const double *c, const double *s, const double *sq,
                           int normalized, double *d)
{
#define T1(m, n) ((((m) * (2 * l - (m) + 1)) / 2 + (n) - (m)) * nb)            /* side l     */
#define T2(m, n) ((((m) * (2 * (l + 1) - (m) + 1)) / 2 + (n) - (m)) * nb)      /* side l + 1 */
    if (l == 0) {
        for (i64 k = 0; k < nb; ++k) d[k] = normalized ? sqrt(0.5) : 1.0;
        return;
    }
    if (l == 1) {
        const double nc = sqrt(1.5);
        for (i64 k = 0; k < nb; ++k) {
            const double a = c[k] * c[k] - s[k] * s[k], b = sq[2] * c[k] * s[k], e = c[k] * c[k];
            d[k] = normalized ? a * nc : a;
            d[nb + k] = normalized ? b * nc : b;
            d[2 * nb + k] = normalized ? e * nc : e;
        }
        return;
    }
    const i64 nh = (l * (l + 1)) / 2;
    double *h = calloc((size_t)(nh * nb), sizeof(double));
    /* l -> l + 1/2 (the half-degree values live on the side-l triangle) */
    {
        const double inv = 1.0 / (double)(2 * l - 1);
        const i64 ll = l - 1;
        const double nc = normalized ? inv * sqrt((double)(2 * l) / (double)(2 * l - 1)) : inv;
        for (i64 m = 0; m <= l - 2; ++m) {
            for (i64 k = 0; k < nb; ++k)
                h[T1(m, m) + k] = (d1[T1(m, m) + k] * (double)(l + m) * c[k]
                                   - d1[T1(m, m + 1) + k] * sqrt((double)((l + m) * (ll - m))) * s[k]
                                   - d1[T1(m, m + 1) + k] * sqrt((double)((ll - m) * (l + m))) * s[k]
                                   + d1[T1(m + 1, m + 1) + k] * (double)(ll - m) * c[k]) * nc;
            for (i64 n = m + 1; n <= l - 2; ++n)
                for (i64 k = 0; k < nb; ++k)
                    h[T1(m, n) + k] = (d1[T1(m, n) + k] * sqrt((double)((l + m) * (l + n))) * c[k]
                                       - d1[T1(m, n + 1) + k] * sqrt((double)((l + m) * (ll - n))) * s[k]
                                       + d1[T1(m + 1, n) + k] * sqrt((double)((ll - m) * (l + n))) * s[k]
                                       + d1[T1(m + 1, n + 1) + k] * sqrt((double)((ll - m) * (ll - n))) * c[k]) * nc;
            for (i64 k = 0; k < nb; ++k)
                h[T1(m, l - 1) + k] = (d1[T1(m, l - 1) + k] * sqrt((double)((l + m) * (2 * l - 1))) * c[k]
                                       + d1[T1(m + 1, l - 1) + k] * sqrt((double)((ll - m) * (2 * l - 1))) * s[k]) * nc;
        }
        for (i64 k = 0; k < nb; ++k)
            h[T1(l - 1, l - 1) + k] = d1[T1(l - 1, l - 1) + k] * (double)(2 * l - 1) * c[k] * nc;
    }
    /* l + 1/2 -> l + 1; hr(idx, k) = the half-degree value at pi - beta_k */
#define HR(idx, k) h[(idx) + (nb - 1 - (k))]
    {
        const double inv = 1.0 / (double)(2 * l);
        const double nc = normalized ? inv * sqrt((double)(2 * l + 1) / (double)(2 * l)) : inv;
        const double nc2 = normalized ? 0.5 * sqrt((double)(2 * l + 1) / (double)(2 * l)) : 0.5;
        const double lsym = (l & 1) ? -1.0 : 1.0, lpsym = -lsym;
        for (i64 k = 0; k < nb; ++k)
            d[T2(0, 0) + k] = (2.0 * h[T1(0, 0) + k] * c[k] - lpsym * HR(T1(0, 0), k) * s[k]
                               + lsym * HR(T1(0, 0), k) * s[k]) * nc2;
        for (i64 n = 1; n <= l - 1; ++n) {
            const double nsym = (n & 1) ? -1.0 : 1.0;
            for (i64 k = 0; k < nb; ++k)
                d[T2(0, n) + k] = (lsym * nsym * HR(T1(0, n - 1), k) * sqrt((double)(l * (l + n))) * c[k]
                                   - lpsym * nsym * HR(T1(0, n), k) * sqrt((double)(l * (l - n))) * s[k]
                                   + h[T1(0, n - 1) + k] * sqrt((double)(l * (l + n))) * s[k]
                                   + h[T1(0, n) + k] * sqrt((double)(l * (l - n))) * c[k]) * nc;
        }
        for (i64 k = 0; k < nb; ++k)
            d[T2(0, l) + k] = (HR(T1(0, l - 1), k) * c[k] + h[T1(0, l - 1) + k] * s[k]) * nc2 * sqrt(2.0);
        for (i64 m = 1; m <= l - 1; ++m) {
            for (i64 k = 0; k < nb; ++k)
                d[T2(m, m) + k] = (h[T1(m - 1, m - 1) + k] * (double)(l + m) * c[k]
                                   - h[T1(m - 1, m) + k] * sqrt((double)((l + m) * (l - m))) * s[k]
                                   - h[T1(m - 1, m) + k] * sqrt((double)((l - m) * (l + m))) * s[k]
                                   + h[T1(m, m) + k] * (double)(l - m) * c[k]) * nc;
            for (i64 n = m + 1; n <= l - 1; ++n)
                for (i64 k = 0; k < nb; ++k)
                    d[T2(m, n) + k] = (h[T1(m - 1, n - 1) + k] * sqrt((double)((l + m) * (l + n))) * c[k]
                                       - h[T1(m - 1, n) + k] * sqrt((double)((l + m) * (l - n))) * s[k]
                                       + h[T1(m, n - 1) + k] * sqrt((double)((l - m) * (l + n))) * s[k]
                                       + h[T1(m, n) + k] * sqrt((double)((l - m) * (l - n))) * c[k]) * nc;
            for (i64 k = 0; k < nb; ++k)
                d[T2(m, l) + k] = (h[T1(m - 1, l - 1) + k] * sqrt((double)((l + m) * (2 * l))) * c[k]
                                   + h[T1(m, l - 1) + k] * sqrt((double)((l - m) * (2 * l))) * s[k]) * nc;
        }
        for (i64 k = 0; k < nb; ++k)
            d[T2(l, l) + k] = h[T1(l - 1, l - 1) + k] * (double)(2 * l) * c[k] * nc;
    }
#undef HR
#undef T1
#undef T2
    free(h);
}

/* number of (p, q, l) entries 0 <= p <= q <= l < bw: the d table's rows */
i64 or_dtab_rows(i64 bw) { return (bw * (2 + 3 * bw + bw * bw)) / 6; }

/* first row of pair (p, q) in the mnl d table (utils.f90:683-695 minus one) */
static i64 dtab_row(i64 p, i64 q, i64 bw)
{
    return (p * (3 * bw * bw + p * p + 2 - 3 * bw * (p - 2) - 3 * p)) / 6 + ((q - p) * (2 * bw + 1 - p - q)) / 2;
}

/*
 * genwig_all_risbo, wigner.f90:784-822: d~^l_{p,q}(beta_k) for 0 <= p <= q <= l < bw at the
 * 2bw beta samples, by the reduced Risbo recurrence, stored [mnl row][2bw] (row of (p,q,l)
 * = dtab_row(p, q) + l - q).
 */
void or_genwig_all_risbo(i64 bw, int normalized, double *out)
{
    const i64 N = 2 * bw;
    double *c = malloc(sizeof(double) * N), *s = malloc(sizeof(double) * N), *sq = malloc(sizeof(double) * (2 * N + 1));
    for (i64 k = 0; k < N; ++k) {
        const double b = (double)(2 * k + 1) * (PI / (double)(2 * N));
        c[k] = cos(b / 2);
        s[k] = sin(b / 2);
    }
    for (i64 k = 0; k <= 2 * N; ++k) sq[k] = sqrt((double)k);
    const i64 nt = (bw * (bw + 1)) / 2;
    double *a = calloc((size_t)(nt * N), sizeof(double)), *b2 = calloc((size_t)(nt * N), sizeof(double));
    for (i64 l = 0; l < bw; ++l) {
        or_risbo_reduced_step(a, l, N, c, s, sq, normalized, b2);
        for (i64 m = 0; m <= l; ++m)
            for (i64 n = m; n <= l; ++n) {
                const i64 src = ((m * (2 * (l + 1) - m + 1)) / 2 + n - m) * N;
                memcpy(out + (dtab_row(m, n, bw) + l - n) * N, b2 + src, sizeof(double) * N);
            }
        double *t = a; a = b2; b2 = t;
    }
    free(a); free(b2); free(c); free(s); free(sq);
}

/*
 * full_big_d_risbo, wigner.f90:844-934: D^l_{m,n}(alpha, beta, gamma) =
 * e^{-i m alpha} d^l_{m,n}(beta) e^{-i n gamma} for every (l, m, n) in mnl order, the reduced
 * Risbo values at beta and pi - beta spread over the eight (m, n) variants with the signs
 * of that routine; normalized: d~ and an overall 1 / (2 pi).
 */
void or_full_big_d_risbo(i64 bw, double alpha, double beta, double gamma, int normalized, zc *out)
{
    double c[2], s[2];
    const double bb[2] = {beta, PI - beta};
    for (int k = 0; k < 2; ++k) {
        c[k] = cos(bb[k] / 2);
        s[k] = sin(bb[k] / 2);
    }
    double *sq = malloc(sizeof(double) * (2 * bw + 1));
    for (i64 k = 0; k <= 2 * bw; ++k) sq[k] = sqrt((double)k);
    i64 *sl = malloc(sizeof(i64) * (2 * bw - 1) * (2 * bw - 1));
    or_slice_table(bw, sl);
    const i64 nt = (bw * (bw + 1)) / 2;
    double *a = calloc((size_t)(2 * nt), sizeof(double)), *b2 = calloc((size_t)(2 * nt), sizeof(double));
#define LOC(m, n, l) (sl[mnidx((m), (n), bw)] + (l) - ((llabs(m) > llabs(n)) ? llabs(m) : llabs(n)))
#define EX(ma, ng) (cexp(I * ((double)(ma) * alpha)) * cexp(I * ((double)(ng) * gamma)))
    for (i64 l = 0; l < bw; ++l) {
        or_risbo_reduced_step(a, l, 2, c, s, sq, normalized, b2);
        const double sl_ = (l & 1) ? -1.0 : 1.0;
        for (i64 m = 0; m <= l; ++m) {
            const double sm = (m & 1) ? -1.0 : 1.0;
            for (i64 n = m; n <= l; ++n) {
                const double sn = (n & 1) ? -1.0 : 1.0;
                const i64 t = ((m * (2 * (l + 1) - m + 1)) / 2 + n - m) * 2;
                const double d1 = b2[t], d2 = b2[t + 1];
                out[LOC(m, n, l)] = EX(-m, -n) * d1;
                if (m == 0 && n == 0) continue;
                out[LOC(-n, -m, l)] = EX(n, m) * d1;
                if (m != n) {
                    out[LOC(n, m, l)] = (sm * sn) * EX(-n, -m) * d1;
                    out[LOC(-m, -n, l)] = (sm * sn) * EX(m, n) * d1;
                }
                if (m == 0 || n == 0) continue;
                out[LOC(-m, n, l)] = (sm * sl_) * EX(m, -n) * d2;
                out[LOC(m, -n, l)] = (sn * sl_) * EX(-m, n) * d2;
                if (m == n) continue;
                out[LOC(-n, m, l)] = (sm * sl_) * EX(n, -m) * d2;
                out[LOC(n, -m, l)] = (sn * sl_) * EX(-n, m) * d2;
            }
        }
        double *tmp = a; a = b2; b2 = tmp;
    }
#undef LOC
#undef EX
    if (normalized) {
        const i64 nc = or_total_num_coeffs(bw);
        for (i64 i = 0; i < nc; ++i) out[i] *= 1.0 / (2.0 * PI);
    }
    free(a); free(b2); free(sq); free(sl);
}

/*
 * genwig_all_kostelec, wigner.f90:423-465: the same table by the Kostelec recurrence
 * (stair seeds, then or_genwig per pair), [mnl row][2bw].
 */
void or_genwig_all_kostelec(i64 bw, double *out)
{
    const i64 N = 2 * bw;
    double *cosb = malloc(sizeof(double) * N), *sc = malloc(sizeof(double) * N), *c2 = malloc(sizeof(double) * N);
    for (i64 k = 0; k < N; ++k) {
        const double b = (double)(2 * k + 1) * (PI / (double)(2 * N));
        cosb[k] = cos(b);
        sc[k] = cos(b / 2) * sin(b / 2);
        c2[k] = pow(cos(b / 2), 2);
    }
    double *seeds = malloc(sizeof(double) * ((bw * (bw + 1)) / 2) * N);
    or_seeds(bw, N, sc, c2, seeds);
    for (i64 p = 0; p < bw; ++p)
        for (i64 q = p; q < bw; ++q)
            or_genwig(p, q, bw, N, cosb, seeds + tri(p, q, bw) * N, out + dtab_row(p, q, bw) * N);
    free(cosb); free(sc); free(c2); free(seeds);
}

/*
 * soft.py:713-764 rotate_ylm_cmplx: for every rotation (alpha,beta,gamma) and every
 * coefficient set y (LMc layout, bw^2 entries):
 *   y'^l_n = sum_m y^l_m d^l(-beta)[m][n] e^{-i m alpha} e^{-i n gamma}.
 * out is [rot][set][bw^2].
 */
void or_rotate_ylm_cmplx(i64 bw, const zc *y, i64 nset, const double *euler, i64 nrot, zc *out, int use_mp)
{
    i64 nb = bw * bw;
#pragma omp parallel for schedule(dynamic) if (use_mp)
    for (i64 r = 0; r < nrot; ++r) {
        double al = euler[3 * r], be = euler[3 * r + 1], ga = euler[3 * r + 2];
        double c = cos(be / 2), s = sin(-be / 2);
        double *sq = malloc(sizeof(double) * (2 * bw + 1));
        for (i64 k = 0; k <= 2 * bw; ++k) sq[k] = sqrt((double)k);
        double *d0 = calloc((size_t)(2 * bw + 1) * (2 * bw + 1), sizeof(double));
        double *d1 = calloc((size_t)(2 * bw + 1) * (2 * bw + 1), sizeof(double));
        zc *ea = malloc(sizeof(zc) * (2 * bw - 1)), *eg = malloc(sizeof(zc) * (2 * bw - 1));
        for (i64 m = -(bw - 1); m < bw; ++m) {
            ea[m + bw - 1] = cexp(-I * al * (double)m);
            eg[m + bw - 1] = cexp(-I * ga * (double)m);
        }
        for (i64 l = 0; l < bw; ++l) {
            or_risbo_step(d0, l, c, s, sq, d1);
            double *t = d0; d0 = d1; d1 = t;
            i64 w = 2 * l + 1;
            for (i64 st = 0; st < nset; ++st) {
                const zc *ys = y + st * nb + l * l;
                zc *os = out + (r * nset + st) * nb + l * l;
                for (i64 j = 0; j < w; ++j) {
                    zc acc = 0;
                    for (i64 i = 0; i < w; ++i)
                        acc += ys[i] * (d0[i * w + j] * eg[j - l + bw - 1] * ea[i - l + bw - 1]);
                    os[j] = acc;
                }
            }
        }
        free(sq); free(d0); free(d1); free(ea); free(eg);
    }
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
