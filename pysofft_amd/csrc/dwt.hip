// Discrete Wigner transforms (DWT) over the 2bw beta samples, FP64, on MFMA.
//
// One workgroup owns one order pair 0 <= p <= q <= lmax of one transform and the
// (up to) eight (m,n) variants that share d~^l_{p,q}(beta) (pysofft/soft.f08:534-631
// inverse, :937-1023 forward).  d~^l_{p,q}(beta_k) is generated on the fly by the
// FP64 three-term recurrence in l (pysofft/wigner.f90:227-260), one beta per lane,
// seeded with d~^q_{p,q}(beta_k) (wigner.f90:130-187, stored per plan), and staged
// through LDS in blocks of l.  The contraction over l (inverse) or beta (forward)
// runs on v_mfma_f64_16x16x4_f64 with the 8 variants x (re, im) = 16 rows as the
// MFMA M dimension, so every pair with 0 < p < q fills whole 16x16 tiles.
#include "so3ft_internal.h"

namespace so3 {

static constexpr int DWT_THREADS = 256;

__device__ __forceinline__ double4_t mfma64(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int wrap_order(int m, int N) { return m >= 0 ? m : N + m; }

// v_mfma_f64_16x16x4f64 returns, in item r of lane l, row (l >> 4) + 4 r of the 16x16
// tile (column l & 15).  Variant v's real part goes to row rowRe(v) and its imaginary
// part to rowRe(v) + 4, so lane group g = l >> 4 holds (re, im) of variant 2g in items
// 0, 1 and of variant 2g + 1 in items 2, 3.
__device__ __forceinline__ int rowRe(int v) { return (v >> 1) + 8 * (v & 1); }

// Constants of the step l -> l+1 of the normalised Kostelec recurrence
// (pysofft/wigner.f90:243-257):  d^{l+1} = a (cos b - x) d^l - b d^{l-1}.
__device__ inline void rec_consts(int l, int p, int q, double &a, double &x, double &b)
{
    const int64_t lp = l + 1;
    const int64_t den = (lp * lp - (int64_t)p * p) * (lp * lp - (int64_t)q * q);
    double cc = (double)lp * sqrt(1.0 / (double)den);
    cc = cc * sqrt((double)(2 * l + 3));
    const double cl = sqrt((double)(2 * l + 1));
    a = cc * cl;
    x = (l == 0) ? 0.0 : (double)((int64_t)p * q) / (double)((int64_t)l * (l + 1));
    if (l == q) {
        b = 0.0;
    } else {
        double c1 = sqrt((double)(((int64_t)l * l - (int64_t)p * p) * ((int64_t)l * l - (int64_t)q * q))) /
                    (double)l;
        c1 = c1 / sqrt((double)(2 * l - 1));
        b = cc * c1;
    }
}

__device__ __forceinline__ int64_t slice_of(const int64_t *slice, int m, int n, int bw)
{
    return slice[(int64_t)(m + bw - 1) * (2 * bw - 1) + (n + bw - 1)];
}

__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// Coefficient of row-variant v at degree l for the inverse DWT, sign included.
__device__ inline double2 inverse_coeff(int mode, int v, int p, int q, int l, int j, int bw,
                                        const double2 *__restrict__ c, const double2 *__restrict__ aux,
                                        const int64_t *__restrict__ slice)
{
    const Variant vr = variant(v, p, q);
    if (mode == DWT_COMPLEX) {
        double2 x = c[slice_of(slice, vr.m, vr.n, bw) + j];
        const double s = vr.sign * ((vr.lsign && (l & 1)) ? -1.0 : 1.0);
        return make_double2(s * x.x, s * x.y);
    }
    if (mode == DWT_CORR) {
        // cc^l_{mn} = 2 pi sqrt(2/(2l+1)) (-1)^{m+n} f^l_m conj(g^l_n)   (soft.f08:2277, :2367)
        const double2 f = c[l * (l + 1) + vr.m];
        const double2 g = aux[l * (l + 1) + vr.n];
        const double wn = 2.0 * 3.14159265358979323846 * sqrt(2.0 / (double)(2 * l + 1));
        const double s = wn * ((((vr.m + vr.n) & 1) != 0) ? -1.0 : 1.0) * vr.sign *
                         ((vr.lsign && (l & 1)) ? -1.0 : 1.0);
        // f * conj(g)
        return make_double2(s * (f.x * g.x + f.y * g.y), s * (f.y * g.x - f.x * g.y));
    }
    // DWT_REAL: rows m >= 0 from the coefficients as soft.f08:1259-1318 reads them,
    // rows m < 0 as the Hermitian completion F(-m,-n) = conj F(m,n) of the c2r plan.
    const double spl = (((p + l) & 1) != 0) ? -1.0 : 1.0;
    switch (v) {
    case 0: return c[slice_of(slice, p, q, bw) + j];
    case 1: {
        if (p == q) return conj2(c[slice_of(slice, p, q, bw) + j]);
        return c[slice_of(slice, -q, -p, bw) + j];
    }
    case 2: return conj2(c[slice_of(slice, -q, -p, bw) + j]);
    case 3: return conj2(c[slice_of(slice, p, q, bw) + j]);
    case 4: { double2 x = c[slice_of(slice, p, -q, bw) + j]; return make_double2(spl * x.x, spl * x.y); }
    case 5: { double2 x = c[slice_of(slice, p, -q, bw) + j]; return make_double2(spl * x.x, -spl * x.y); }
    case 6: { double2 x = c[slice_of(slice, q, -p, bw) + j]; return make_double2(spl * x.x, spl * x.y); }
    default: { double2 x = c[slice_of(slice, q, -p, bw) + j]; return make_double2(spl * x.x, -spl * x.y); }
    }
}

// LDS (doubles): cst[3*LpM] | A[16*LpM] | D[2][KL][N16]
template <int TPW, int KL>
__global__ __launch_bounds__(DWT_THREADS) void dwt_inverse_kernel(
    const double2 *__restrict__ coeff, int64_t c_ts, const double2 *__restrict__ aux,
    double2 *__restrict__ F, int64_t f_ts, const double *__restrict__ cosb,
    const double *__restrict__ seeds, const int64_t *__restrict__ slice,
    const int32_t *__restrict__ pairs, int bw, int N, int mode)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pq = pairs[blockIdx.x];
    const int p = pq & 0xffff, q = pq >> 16;
    const int64_t t = blockIdx.y;
    const int L = bw - q;
    const int Lp = (L + KL - 1) / KL * KL;
    const int LpM = (bw + KL - 1) / KL * KL;
    const int N16 = (N + 15) & ~15;
    const int ntile = N16 >> 4;
    double *cst = smem;
    double *A = cst + 3 * LpM;
    double *Dt = A + 16 * LpM;

    const double2 *ct = coeff + t * c_ts;
    const double2 *at = aux ? aux + t * c_ts : nullptr;

    for (int i = tid; i < L - 1; i += DWT_THREADS) {
        double a, x, b;
        rec_consts(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }
    for (int idx = tid; idx < 8 * Lp; idx += DWT_THREADS) {
        const int v = idx / Lp, j = idx - v * Lp;
        double2 c = make_double2(0.0, 0.0);
        if (j < L && variant(v, p, q).valid) c = inverse_coeff(mode, v, p, q, q + j, j, bw, ct, at, slice);
        A[rowRe(v) * LpM + j] = c.x;
        A[(rowRe(v) + 4) * LpM + j] = c.y;
    }

    // recurrence state: betas tid, tid+256, ... (N16 <= 4*256)
    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    double dcur[4], dprev[4], cb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int beta = tid + r * DWT_THREADS;
        dcur[r] = (beta < N) ? seeds[tri * N + beta] : 0.0;
        cb[r] = (beta < N) ? cosb[beta] : 0.0;
        dprev[r] = 0.0;
    }
    double4_t acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = double4_t{0.0, 0.0, 0.0, 0.0};
    __syncthreads();

    const int nblk = Lp / KL;
    for (int blk = 0; blk < nblk; ++blk) {
        double *D = Dt + (blk & 1) * KL * N16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int beta = tid + r * DWT_THREADS;
            if (beta < N16) {
#pragma unroll
                for (int j = 0; j < KL; ++j) {
                    const int li = blk * KL + j;
                    D[j * N16 + beta] = (li < L && beta < N) ? dcur[r] : 0.0;
                    if (li < L - 1) {
                        const double a = cst[3 * li], x = cst[3 * li + 1], b = cst[3 * li + 2];
                        const double dn = (a * (cb[r] - x)) * dcur[r] - b * dprev[r];
                        dprev[r] = dcur[r];
                        dcur[r] = dn;
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < KL / 4; ++s) {
            const int kr = 4 * s + (lane >> 4);
            const double a = A[(lane & 15) * LpM + blk * KL + kr];
#pragma unroll
            for (int i = 0; i < TPW; ++i) {
                const int tile = wave + 4 * i;
                if (tile < ntile) {
                    const double b = D[kr * N16 + tile * 16 + (lane & 15)];
                    acc[i] = mfma64(a, b, acc[i]);
                }
            }
        }
    }

    // epilogue: lane group g holds (re, im) of variants 2g and 2g+1 for column lane&15
    const int g = lane >> 4;
    double2 *Ft = F + t * f_ts;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int tile = wave + 4 * i;
        if (tile >= ntile) continue;
        const int beta = tile * 16 + (lane & 15);
        if (beta >= N) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const Variant vr = variant(2 * g + h, p, q);
            if (!vr.valid) continue;
            const double re = h ? acc[i][2] : acc[i][0];
            const double im = h ? acc[i][3] : acc[i][1];
            const int bo = vr.rev ? N - 1 - beta : beta;
            Ft[((int64_t)wrap_order(vr.n, N) * N + wrap_order(vr.m, N)) * N + bo] = make_double2(re, im);
        }
    }
}

// LDS (doubles): cst[3*LpM] | Aw[16][N16] | D[16][N16] | red[4][256]
__global__ __launch_bounds__(DWT_THREADS) void dwt_forward_kernel(
    const double2 *__restrict__ F, int64_t f_ts, double2 *__restrict__ coeff, int64_t c_ts,
    const double *__restrict__ w, const double *__restrict__ cosb, const double *__restrict__ seeds,
    const int64_t *__restrict__ slice, const int32_t *__restrict__ pairs, int bw, int N, int mode)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pq = pairs[blockIdx.x];
    const int p = pq & 0xffff, q = pq >> 16;
    const int64_t t = blockIdx.y;
    const int L = bw - q;
    const int LpM = (bw + 15) / 16 * 16;
    const int N16 = (N + 15) & ~15;
    double *cst = smem;
    double *Aw = cst + 3 * LpM;
    double *D = Aw + 16 * N16;
    double *red = D + 16 * N16;

    const double2 *Ft = F + t * f_ts;
    double2 *ct = coeff + t * c_ts;

    for (int i = tid; i < L - 1; i += DWT_THREADS) {
        double a, x, b;
        rec_consts(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }
    // weighted F rows (soft.f08:962): row 2v = Re, 2v+1 = Im; reversed variants read pi - beta
    for (int idx = tid; idx < 8 * N16; idx += DWT_THREADS) {
        const int v = idx / N16, beta = idx - v * N16;
        double2 x = make_double2(0.0, 0.0);
        const Variant vr = variant(v, p, q);
        if (vr.valid && beta < N) {
            const int bi = vr.rev ? N - 1 - beta : beta;
            int m = vr.m, n = vr.n;
            bool cj = false;
            if (mode == DWT_REAL && (v & 1)) {
                // rows the half-complex forward takes as conj of an m >= 0 row
                // (soft.f08:1650, 3286-3300 and the symmetry fills of :1662-1706)
                const int sv = (v == 1) ? ((p == q) ? 0 : 2) : (v == 3 ? 0 : v - 1);
                const Variant s2 = variant(sv, p, q);
                m = s2.m; n = s2.n; cj = true;
            }
            x = Ft[((int64_t)wrap_order(n, N) * N + wrap_order(m, N)) * N + bi];
            if (cj) x.y = -x.y;
            const double wb = w[beta];
            x.x *= wb; x.y *= wb;
        }
        Aw[rowRe(v) * N16 + beta] = x.x;
        Aw[(rowRe(v) + 4) * N16 + beta] = x.y;
    }

    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    double dcur[4], dprev[4], cb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int beta = tid + r * DWT_THREADS;
        dcur[r] = (beta < N) ? seeds[tri * N + beta] : 0.0;
        cb[r] = (beta < N) ? cosb[beta] : 0.0;
        dprev[r] = 0.0;
    }
    __syncthreads();

    const int nks = N16 >> 2;
    for (int lt = 0; lt * 16 < L; ++lt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int beta = tid + r * DWT_THREADS;
            if (beta < N16) {
                for (int j = 0; j < 16; ++j) {
                    const int li = lt * 16 + j;
                    D[j * N16 + beta] = (li < L && beta < N) ? dcur[r] : 0.0;
                    if (li < L - 1) {
                        const double a = cst[3 * li], x = cst[3 * li + 1], b = cst[3 * li + 2];
                        const double dn = (a * (cb[r] - x)) * dcur[r] - b * dprev[r];
                        dprev[r] = dcur[r];
                        dcur[r] = dn;
                    }
                }
            }
        }
        __syncthreads();
        double4_t acc = double4_t{0.0, 0.0, 0.0, 0.0};
        for (int s = wave; s < nks; s += 4) {
            const int kb = 4 * s + (lane >> 4);
            const double a = Aw[(lane & 15) * N16 + kb];
            const double b = D[(lane & 15) * N16 + kb];
            acc = mfma64(a, b, acc);
        }
        red[wave * 256 + lane * 4 + 0] = acc[0];
        red[wave * 256 + lane * 4 + 1] = acc[1];
        red[wave * 256 + lane * 4 + 2] = acc[2];
        red[wave * 256 + lane * 4 + 3] = acc[3];
        __syncthreads();
        {
            const double sum = red[tid] + red[256 + tid] + red[512 + tid] + red[768 + tid];
            const int ln = tid >> 2, rr = tid & 3;  // item rr of lane ln: row (ln >> 4) + 4 rr
            const int col = ln & 15;
            const int v = 2 * (ln >> 4) + (rr >> 1);
            const int l = q + lt * 16 + col;
            const Variant vr = variant(v, p, q);
            if (vr.valid && l < bw) {
                const double s = vr.sign * ((vr.lsign && (l & 1)) ? -1.0 : 1.0);
                double *dst = reinterpret_cast<double *>(ct + slice_of(slice, vr.m, vr.n, bw) + (l - q));
                dst[rr & 1] = s * sum;
            }
        }
        __syncthreads();
    }
}

int dwt_lds_bytes_inverse(int bw, int N)
{
    const int KL = (N > 512) ? 4 : 8;
    const int LpM = (bw + KL - 1) / KL * KL;
    const int N16 = (N + 15) & ~15;
    return (int)sizeof(double) * (3 * LpM + 16 * LpM + 2 * KL * N16);
}

int dwt_lds_bytes_forward(int bw, int N)
{
    const int LpM = (bw + 15) / 16 * 16;
    const int N16 = (N + 15) & ~15;
    return (int)sizeof(double) * (3 * LpM + 32 * N16 + 1024);
}

template <int TPW, int KL>
static hipError_t set_inv_attr(int bytes)
{
    return hipFuncSetAttribute((const void *)dwt_inverse_kernel<TPW, KL>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// The dynamic-LDS limit is a per-kernel attribute shared by every plan, so it is set
// once to the largest size any supported bandwidth (bw <= 256) needs.
hipError_t dwt_set_attributes(int, int)
{
    static bool done = false;
    if (done) return hipSuccess;
    const int bi = dwt_lds_bytes_inverse(256, 512), bf = dwt_lds_bytes_forward(256, 512);
    hipError_t e;
    if ((e = set_inv_attr<1, 8>(bi)) != hipSuccess) return e;
    if ((e = set_inv_attr<2, 8>(bi)) != hipSuccess) return e;
    if ((e = set_inv_attr<4, 8>(bi)) != hipSuccess) return e;
    if ((e = set_inv_attr<8, 8>(bi)) != hipSuccess) return e;
    if ((e = hipFuncSetAttribute((const void *)dwt_forward_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bf)) !=
        hipSuccess)
        return e;
    done = true;
    return hipSuccess;
}

hipError_t launch_dwt_inverse(const DeviceTables &tb, int bw, int N, int npair, int mode,
                              const double2 *coeff, int64_t c_ts, const double2 *aux, double2 *F,
                              int64_t f_ts, int T, hipStream_t s)
{
    if (npair <= 0 || T <= 0) return hipSuccess;
    const int bytes = dwt_lds_bytes_inverse(bw, N);
    const int ntile = ((N + 15) & ~15) >> 4;
    dim3 grid(npair, T), block(DWT_THREADS);
#define SO3_INV(TPW, KL)                                                                            \
    hipLaunchKernelGGL((dwt_inverse_kernel<TPW, KL>), grid, block, bytes, s, coeff, c_ts, aux, F, f_ts, \
                       tb.cosb, tb.seeds, tb.slice, tb.pairs, bw, N, mode)
    if (ntile <= 4) SO3_INV(1, 8);
    else if (ntile <= 8) SO3_INV(2, 8);
    else if (ntile <= 16) SO3_INV(4, 8);
    else SO3_INV(8, 8);
#undef SO3_INV
    return hipGetLastError();
}

hipError_t launch_dwt_forward(const DeviceTables &tb, int bw, int N, int npair, int mode,
                              const double2 *F, int64_t f_ts, double2 *coeff, int64_t c_ts, int T,
                              hipStream_t s)
{
    if (npair <= 0 || T <= 0) return hipSuccess;
    const int bytes = dwt_lds_bytes_forward(bw, N);
    dim3 grid(npair, T), block(DWT_THREADS);
    hipLaunchKernelGGL(dwt_forward_kernel, grid, block, bytes, s, F, f_ts, coeff, c_ts, tb.w, tb.cosb,
                       tb.seeds, tb.slice, tb.pairs, bw, N, mode);
    return hipGetLastError();
}

}  // namespace so3
