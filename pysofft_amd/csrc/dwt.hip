// Discrete Wigner transforms (DWT) over the 2bw beta samples on MFMA.
//
// Work unit: one order pair 0 <= p <= q <= lmax and the (up to) eight (m,n)
// variants that share d~^l_{p,q}(beta) (pysofft/soft.f08:534-631 inverse,
// :937-1023 forward), for a batch of TB transforms.  d~^l_{p,q}(beta_k) is generated
// on the fly by the FP64 three-term recurrence in l (pysofft/wigner.f90:227-260),
// one beta per lane, from the seeds d~^q_{p,q}(beta_k) (wigner.f90:130-187) kept in
// the plan.  A 16(l) x 64(beta) tile of d in LDS feeds the MFMAs of all TB
// transforms, so the recurrence is paid once per batch.  The 8 variants x (re, im)
// are the 16 MFMA rows.
//
// Inverse: one 64-lane wave per (pair, 64-beta range, transform batch); the wave
// contracts over l with the coefficients as the A operand.
// Forward: the waves of a block cover all betas of a pair; A = w_k F(beta_k') stays
// in registers for the whole pair, each wave contracts its own 64 betas and the
// block sums the wave partials per 16-l tile.
//
// MF = 0: v_mfma_f64_16x16x4f64 (FP64 operands, FP64 accumulate: the plain method).
// MF = 1: v_mfma_f32_16x16x4f32 with d split into FP32 hi + lo parts (two MFMAs),
//         FP32 accumulation over 16 terms flushed into FP64 accumulators.
#include "so3ft_internal.h"

namespace so3 {

typedef float float4_t __attribute__((ext_vector_type(4)));

template <int MF> struct Mma;
template <> struct Mma<0> {
    typedef double T;
    typedef double4_t Acc;
    __device__ __forceinline__ static Acc mma(double a, double b, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // v_mfma_f64_16x16x4f64 returns row (l>>4) + 4r of the tile in item r of lane l,
    // so A row i carries variant 2*(i&3) + ((i>>2)>>1), part (i>>2)&1
    __device__ __forceinline__ static void row_of(int i, int &v, int &reim)
    {
        const int g = i & 3, r = i >> 2;
        v = 2 * g + (r >> 1); reim = r & 1;
    }
};
template <> struct Mma<1> {
    typedef float T;
    typedef float4_t Acc;
    __device__ __forceinline__ static Acc mma(float a, float b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // v_mfma_f32_16x16x4f32 returns row 4(l>>4) + r in item r
    __device__ __forceinline__ static void row_of(int i, int &v, int &reim)
    {
        const int g = i >> 2, r = i & 3;
        v = 2 * g + (r >> 1); reim = r & 1;
    }
};
// Either way, item r of lane group g = l>>4 holds variant 2g + (r>>1), part r&1.

__device__ __forceinline__ int wrap_order(int m, int N) { return m >= 0 ? m : N + m; }

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

template <class T> struct Cplx;
template <> struct Cplx<double2> { typedef double S; };
template <> struct Cplx<float2> { typedef float S; };

// Where one lane's A-row coefficient comes from, for the inverse DWT.
struct CoefSrc {
    int64_t base;   // mnl offset of the variant's slice (COMPLEX / REAL)
    int m, n;       // orders (CORR)
    double sign;
    bool lsign, conj, valid;
};

__device__ inline CoefSrc coef_src(int mode, int v, int p, int q, int bw, const int64_t *slice)
{
    const Variant vr = variant(v, p, q);
    CoefSrc s;
    s.m = vr.m; s.n = vr.n; s.valid = vr.valid; s.conj = false;
    s.sign = vr.sign; s.lsign = vr.lsign; s.base = 0;
    if (!vr.valid) return s;
    if (mode == DWT_COMPLEX) {
        s.base = slice_of(slice, vr.m, vr.n, bw);
    } else if (mode == DWT_REAL) {
        // rows m >= 0 read the coefficients as soft.f08:1259-1318 does; rows m < 0 are
        // the Hermitian completion F(-m,-n) = conj F(m,n) of the c2r plan
        const double sp = (p & 1) ? -1.0 : 1.0;
        switch (v) {
        case 0: s.base = slice_of(slice, p, q, bw); s.sign = 1.0; s.lsign = false; break;
        case 1:
            if (p == q) { s.base = slice_of(slice, p, q, bw); s.conj = true; }
            else s.base = slice_of(slice, -q, -p, bw);
            s.sign = 1.0; s.lsign = false; break;
        case 2: s.base = slice_of(slice, -q, -p, bw); s.conj = true; s.sign = 1.0; s.lsign = false; break;
        case 3: s.base = slice_of(slice, p, q, bw); s.conj = true; s.sign = 1.0; s.lsign = false; break;
        case 4: s.base = slice_of(slice, p, -q, bw); s.sign = sp; s.lsign = true; break;
        case 5: s.base = slice_of(slice, p, -q, bw); s.conj = true; s.sign = sp; s.lsign = true; break;
        case 6: s.base = slice_of(slice, q, -p, bw); s.sign = sp; s.lsign = true; break;
        default: s.base = slice_of(slice, q, -p, bw); s.conj = true; s.sign = sp; s.lsign = true; break;
        }
    }
    return s;
}

// one real component of the A operand (coefficient row) at degree l
template <class TC>
__device__ __forceinline__ double coef_value(int mode, const CoefSrc &s, int reim, int l, int q, const TC *ct,
                                             const TC *at)
{
    typedef typename Cplx<TC>::S S;
    if (!s.valid) return 0.0;
    double sg = s.sign * ((s.lsign && (l & 1)) ? -1.0 : 1.0);
    if (mode == DWT_CORR) {
        // cc^l_{mn} = 2 pi sqrt(2/(2l+1)) (-1)^{m+n} f^l_m conj(g^l_n)   (soft.f08:2277, :2367)
        const TC f = ct[l * (l + 1) + s.m];
        const TC g = at[l * (l + 1) + s.n];
        const double wn = 2.0 * 3.14159265358979323846 * sqrt(2.0 / (double)(2 * l + 1));
        sg *= wn * ((((s.m + s.n) & 1) != 0) ? -1.0 : 1.0);
        const double fx = f.x, fy = f.y, gx = g.x, gy = g.y;
        return reim ? sg * (fy * gx - fx * gy) : sg * (fx * gx + fy * gy);
    }
    const S *ptr = reinterpret_cast<const S *>(ct + s.base + (l - q));
    const double x = (double)ptr[reim];
    return (reim && s.conj) ? -sg * x : sg * x;
}

template <class TF> __device__ __forceinline__ void store_c(TF *p, double re, double im);
template <> __device__ __forceinline__ void store_c<double2>(double2 *p, double re, double im) { *p = make_double2(re, im); }
template <> __device__ __forceinline__ void store_c<float2>(float2 *p, double re, double im)
{
    *p = make_float2((float)re, (float)im);
}

// ---------------------------------------------------------------------------
// inverse
// ---------------------------------------------------------------------------
template <int MF, int TB, class TC, class TF>
__global__ __launch_bounds__(64) void dwt_inverse_kernel(const TC *__restrict__ coeff, int64_t c_ts,
                                                         const TC *__restrict__ aux, TF *__restrict__ F,
                                                         int64_t f_ts, int T, const double *__restrict__ cosb,
                                                         const double *__restrict__ seeds,
                                                         const int64_t *__restrict__ slice,
                                                         const int32_t *__restrict__ pairs, int bw, int N, int mode)
{
    typedef Mma<MF> M;
    typedef typename M::T E;
    constexpr int NS = (MF == 1) ? 2 : 1;  // d as hi (+ lo) parts
    __shared__ double cst[3 * 256];
    __shared__ E Dw[NS][16][64];
    const int lane = threadIdx.x;
    const int pq = pairs[blockIdx.x];
    const int p = pq & 0xffff, q = pq >> 16;
    const int L = bw - q;
    const int b0 = blockIdx.y * 64;
    const int t0 = blockIdx.z * TB;
    const int nt = min(TB, T - t0);
    const int N16 = (N + 15) & ~15;
    const int ntile = min(4, (N16 - b0) >> 4);

    for (int i = lane; i < L - 1; i += 64) {
        double a, x, b;
        rec_consts(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }
    int v, reim;
    M::row_of(lane & 15, v, reim);
    const CoefSrc src = coef_src(mode, v, p, q, bw, slice);

    const int beta = b0 + lane;
    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    double dcur = (beta < N) ? seeds[tri * N + beta] : 0.0;
    const double cb = (beta < N) ? cosb[beta] : 0.0;
    double dprev = 0.0;

    typename M::Acc acc[TB][4];
    double4_t accd[TB][4];
#pragma unroll
    for (int t = 0; t < TB; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[t][j] = typename M::Acc{0, 0, 0, 0};
            accd[t][j] = double4_t{0.0, 0.0, 0.0, 0.0};
        }

    // A operands of the 4 k-steps of an l block for every transform; the next block's
    // operands are fetched while the current block runs (register double buffer)
    E a[TB][4], an[TB][4];
    auto fetch = [&](E (&dst)[TB][4], int l0) {
#pragma unroll
        for (int t = 0; t < TB; ++t)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int li = l0 + 4 * s + (lane >> 4);
                double x = 0.0;
                if (t < nt && li < L)
                    x = coef_value(mode, src, reim, q + li, q, coeff + (int64_t)(t0 + t) * c_ts,
                                   aux ? aux + (int64_t)(t0 + t) * c_ts : nullptr);
                dst[t][s] = (E)x;
            }
    };
    fetch(a, 0);
    for (int l0 = 0; l0 < L; l0 += 16) {
        if (l0 + 16 < L) fetch(an, l0 + 16);
        __syncthreads();  // previous block's MFMAs have read Dw
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int li = l0 + j;
            const double d = (li < L && beta < N) ? dcur : 0.0;
            if (MF == 1) {
                const float hi = (float)d;
                Dw[0][j][lane] = (E)hi;
                Dw[NS - 1][j][lane] = (E)(float)(d - (double)hi);
            } else {
                Dw[0][j][lane] = (E)d;
            }
            if (li < L - 1) {
                const double ca = cst[3 * li], cx = cst[3 * li + 1], cbb = cst[3 * li + 2];
                const double dn = (ca * (cb - cx)) * dcur - cbb * dprev;
                dprev = dcur;
                dcur = dn;
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kr = 4 * s + (lane >> 4);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j < ntile) {
                    const E bh = Dw[0][kr][16 * j + (lane & 15)];
#pragma unroll
                    for (int t = 0; t < TB; ++t)
                        if (t < nt) acc[t][j] = M::mma(a[t][s], bh, acc[t][j]);
                    if (MF == 1) {
                        const E bl = Dw[NS - 1][kr][16 * j + (lane & 15)];
#pragma unroll
                        for (int t = 0; t < TB; ++t)
                            if (t < nt) acc[t][j] = M::mma(a[t][s], bl, acc[t][j]);
                    }
                }
            }
        }
        if (MF == 1) {  // flush the 16-term FP32 partial sums into FP64
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) accd[t][j][r] += (double)acc[t][j][r];
                    acc[t][j] = typename M::Acc{0, 0, 0, 0};
                }
        }
#pragma unroll
        for (int t = 0; t < TB; ++t)
#pragma unroll
            for (int s = 0; s < 4; ++s) a[t][s] = an[t][s];
    }

    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < TB; ++t) {
        if (t >= nt) continue;
        TF *Ft = F + (int64_t)(t0 + t) * f_ts;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= ntile) continue;
            const int bt = b0 + 16 * j + (lane & 15);
            if (bt >= N) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const Variant vr = variant(2 * g + h, p, q);
                if (!vr.valid) continue;
                double re, im;
                if (MF == 1) { re = accd[t][j][2 * h]; im = accd[t][j][2 * h + 1]; }
                else { re = (double)acc[t][j][2 * h]; im = (double)acc[t][j][2 * h + 1]; }
                const int bo = vr.rev ? N - 1 - bt : bt;
                store_c<TF>(Ft + ((int64_t)wrap_order(vr.n, N) * N + wrap_order(vr.m, N)) * N + bo, re, im);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <class TF> __device__ __forceinline__ double2 load_c(const TF *p);
template <> __device__ __forceinline__ double2 load_c<double2>(const double2 *p) { return *p; }
template <> __device__ __forceinline__ double2 load_c<float2>(const float2 *p)
{
    const float2 v = *p;
    return make_double2((double)v.x, (double)v.y);
}

template <class TC> __device__ __forceinline__ void store_part(TC *p, int reim, double x);
template <> __device__ __forceinline__ void store_part<double2>(double2 *p, int reim, double x)
{
    reinterpret_cast<double *>(p)[reim] = x;
}
template <> __device__ __forceinline__ void store_part<float2>(float2 *p, int reim, double x)
{
    reinterpret_cast<float *>(p)[reim] = (float)x;
}

// dynamic LDS: cst[3*256] doubles | Dw[nw][NS][16][64] E | red[nw][TB][256] doubles | obase[256] int64
template <int MF, int TB, class TF, class TC>
__global__ __launch_bounds__(512) void dwt_forward_kernel(const TF *__restrict__ F, int64_t f_ts,
                                                          TC *__restrict__ coeff, int64_t c_ts, int T,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ cosb,
                                                          const double *__restrict__ seeds,
                                                          const int64_t *__restrict__ slice,
                                                          const int32_t *__restrict__ pairs, int bw, int N, int mode)
{
    typedef Mma<MF> M;
    typedef typename M::T E;
    constexpr int NS = (MF == 1) ? 2 : 1;
    extern __shared__ double smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    double *cst = smem;
    E *Dall = reinterpret_cast<E *>(cst + 3 * 256);
    E *Dw = Dall + (int64_t)wave * NS * 16 * 64;
    double *red = reinterpret_cast<double *>(Dall + (int64_t)nw * NS * 16 * 64);

    const int pq = pairs[blockIdx.x];
    const int p = pq & 0xffff, q = pq >> 16;
    const int L = bw - q;
    const int t0 = blockIdx.y * TB;
    const int nt = min(TB, T - t0);
    const int b0 = 64 * wave;

    for (int i = tid; i < L - 1; i += blockDim.x) {
        double a, x, b;
        rec_consts(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }

    // A operands: row lane&15 -> (variant, part); k-step s covers betas b0 + 4s + (lane>>4)
    int v, reim;
    M::row_of(lane & 15, v, reim);
    const Variant vr = variant(v, p, q);
    int mm = vr.m, nn = vr.n;
    bool cj = false;
    if (mode == DWT_REAL && (v & 1)) {
        // rows the half-complex forward takes as conj of an m >= 0 row
        // (soft.f08:1650, 3286-3300 and the symmetry fills of :1662-1706)
        const int sv = (v == 1) ? ((p == q) ? 0 : 2) : (v == 3 ? 0 : v - 1);
        const Variant s2 = variant(sv, p, q);
        mm = s2.m; nn = s2.n; cj = true;
    }
    const int64_t col = ((int64_t)wrap_order(nn, N) * N + wrap_order(mm, N)) * N;
    E a[TB][16];
#pragma unroll
    for (int t = 0; t < TB; ++t)
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int bt = b0 + 4 * s + (lane >> 4);
            double x = 0.0;
            if (t < nt && vr.valid && bt < N) {
                const double2 f = load_c<TF>(F + (int64_t)(t0 + t) * f_ts + col + (vr.rev ? N - 1 - bt : bt));
                x = reim ? (cj ? -f.y : f.y) : f.x;
                x *= w[bt];
            }
            a[t][s] = (E)x;
        }

    const int beta = b0 + lane;
    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    double dcur = (beta < N) ? seeds[tri * N + beta] : 0.0;
    const double cb = (beta < N) ? cosb[beta] : 0.0;
    double dprev = 0.0;
    // output element e of every tile is item e&3 of lane e>>2: variant 2((e>>2)>>4) + ((e&3)>>1),
    // part e&1, l offset (e>>2)&15; its coefficient slice offset is looked up once per pair
    int64_t *obase = reinterpret_cast<int64_t *>(red + (int64_t)nw * TB * 256);
    for (int e = tid; e < 256; e += blockDim.x) {
        const Variant ovr = variant(2 * ((e >> 2) >> 4) + ((e & 3) >> 1), p, q);
        obase[e] = ovr.valid ? slice_of(slice, ovr.m, ovr.n, bw) : -1;
    }
    __syncthreads();

    for (int lt = 0; lt * 16 < L; ++lt) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int li = lt * 16 + j;
            const double d = (li < L && beta < N) ? dcur : 0.0;
            if (MF == 1) {
                const float hi = (float)d;
                Dw[(0 * 16 + j) * 64 + lane] = (E)hi;
                Dw[((NS - 1) * 16 + j) * 64 + lane] = (E)(float)(d - (double)hi);
            } else {
                Dw[j * 64 + lane] = (E)d;
            }
            if (li < L - 1) {
                const double ca = cst[3 * li], cx = cst[3 * li + 1], cbb = cst[3 * li + 2];
                const double dn = (ca * (cb - cx)) * dcur - cbb * dprev;
                dprev = dcur;
                dcur = dn;
            }
        }
        __syncthreads();
        double4_t part[TB];
#pragma unroll
        for (int t = 0; t < TB; ++t) part[t] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s0 = 0; s0 < 16; s0 += 4) {
            typename M::Acc acc[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) acc[t] = typename M::Acc{0, 0, 0, 0};
#pragma unroll
            for (int s = s0; s < s0 + 4; ++s) {
                // B[k][n] = d(beta = b0 + 4s + k, l = lt*16 + n)
                const E bh = Dw[(0 * 16 + (lane & 15)) * 64 + 4 * s + (lane >> 4)];
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    if (t < nt) acc[t] = M::mma(a[t][s], bh, acc[t]);
                if (MF == 1) {
                    const E bl = Dw[((NS - 1) * 16 + (lane & 15)) * 64 + 4 * s + (lane >> 4)];
#pragma unroll
                    for (int t = 0; t < TB; ++t)
                        if (t < nt) acc[t] = M::mma(a[t][s], bl, acc[t]);
                }
            }
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) part[t][r] += (double)acc[t][r];
        }
#pragma unroll
        for (int t = 0; t < TB; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[((int64_t)wave * TB + t) * 256 + lane * 4 + r] = part[t][r];
        __syncthreads();
        for (int idx = tid; idx < nt * 256; idx += blockDim.x) {
            const int t = idx >> 8, e = idx & 255;
            const int l = q + lt * 16 + ((e >> 2) & 15);
            const int64_t ob = obase[e];
            if (ob >= 0 && l < bw) {
                double sum = 0.0;
                for (int ww = 0; ww < nw; ++ww) sum += red[((int64_t)ww * TB + t) * 256 + e];
                const Variant ovr = variant(2 * ((e >> 2) >> 4) + ((e & 3) >> 1), p, q);
                const double sg = ovr.sign * ((ovr.lsign && (l & 1)) ? -1.0 : 1.0);
                store_part<TC>(coeff + (int64_t)(t0 + t) * c_ts + ob + (l - q), e & 1, sg * sum);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static int fwd_waves(int N) { return (((N + 15) & ~15) + 63) / 64; }

template <int MF, int TB>
static int fwd_lds_bytes(int N)
{
    const int nw = fwd_waves(N);
    const int E = (MF == 1) ? 4 : 8, NS = (MF == 1) ? 2 : 1;
    return 3 * 256 * 8 + nw * NS * 16 * 64 * E + nw * TB * 256 * 8 + 256 * 8;
}

int dwt_lds_bytes_inverse(int, int) { return 0; }
int dwt_lds_bytes_forward(int, int N) { return fwd_lds_bytes<0, 2>(N); }

// transforms per block: enough to amortise the recurrence, few enough to stay spill-free
template <int MF> constexpr int inv_tb() { return MF == 1 ? 2 : 4; }

template <int MF, int TB, class TF, class TC>
static hipError_t set_fwd_attr()
{
    return hipFuncSetAttribute((const void *)dwt_forward_kernel<MF, TB, TF, TC>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

hipError_t dwt_set_attributes(int, int)
{
    static bool done = false;
    if (done) return hipSuccess;
    hipError_t e;
    if ((e = set_fwd_attr<0, 2, double2, double2>()) != hipSuccess) return e;
    if ((e = set_fwd_attr<0, 1, double2, double2>()) != hipSuccess) return e;
    done = true;
    return hipSuccess;
}

template <int MF, class TC, class TF>
static hipError_t inv_typed(const DeviceTables &tb, int bw, int N, int npair, int mode, const TC *coeff,
                            int64_t c_ts, const TC *aux, TF *F, int64_t f_ts, int T, hipStream_t s)
{
    constexpr int TB = inv_tb<MF>();
    const int N16 = (N + 15) & ~15;
    dim3 grid(npair, (N16 + 63) / 64, (T + TB - 1) / TB), block(64);
    hipLaunchKernelGGL((dwt_inverse_kernel<MF, TB, TC, TF>), grid, block, 0, s, coeff, c_ts, aux, F, f_ts, T,
                       tb.cosb, tb.seeds, tb.slice, tb.pairs, bw, N, mode);
    return hipGetLastError();
}

template <int MF, int TB, class TF, class TC>
static hipError_t fwd_launch(const DeviceTables &tb, int bw, int N, int npair, int mode, const TF *F, int64_t f_ts,
                             TC *coeff, int64_t c_ts, int T, hipStream_t s)
{
    const int nw = fwd_waves(N);
    const int bytes = fwd_lds_bytes<MF, TB>(N);
    dim3 grid(npair, (T + TB - 1) / TB), block(64 * nw);
    hipLaunchKernelGGL((dwt_forward_kernel<MF, TB, TF, TC>), grid, block, bytes, s, F, f_ts, coeff, c_ts, T, tb.w,
                       tb.cosb, tb.seeds, tb.slice, tb.pairs, bw, N, mode);
    return hipGetLastError();
}

template <int MF, class TF, class TC>
static hipError_t fwd_typed(const DeviceTables &tb, int bw, int N, int npair, int mode, const TF *F, int64_t f_ts,
                            TC *coeff, int64_t c_ts, int T, hipStream_t s)
{
    if constexpr (MF == 0) {
        if (N <= 256) return fwd_launch<MF, 2>(tb, bw, N, npair, mode, F, f_ts, coeff, c_ts, T, s);
        return fwd_launch<MF, 1>(tb, bw, N, npair, mode, F, f_ts, coeff, c_ts, T, s);
    } else {
        if (N <= 256) return fwd_launch<MF, 4>(tb, bw, N, npair, mode, F, f_ts, coeff, c_ts, T, s);
        return fwd_launch<MF, 2>(tb, bw, N, npair, mode, F, f_ts, coeff, c_ts, T, s);
    }
}

hipError_t launch_dwt_inverse(const DeviceTables &tb, int bw, int N, int npair, int mode, const double2 *coeff,
                              int64_t c_ts, const double2 *aux, double2 *F, int64_t f_ts, int T, hipStream_t s)
{
    if (bw > 256) return hipErrorInvalidValue;   // the recurrence constants' LDS holds bw <= 256
    if (npair <= 0 || T <= 0) return hipSuccess;
    return inv_typed<0, double2, double2>(tb, bw, N, npair, mode, coeff, c_ts, aux, F, f_ts, T, s);
}

hipError_t launch_dwt_forward(const DeviceTables &tb, int bw, int N, int npair, int mode, const double2 *F,
                              int64_t f_ts, double2 *coeff, int64_t c_ts, int T, hipStream_t s)
{
    if (bw > 256) return hipErrorInvalidValue;
    if (npair <= 0 || T <= 0) return hipSuccess;
    return fwd_typed<0, double2, double2>(tb, bw, N, npair, mode, F, f_ts, coeff, c_ts, T, s);
}

}  // namespace so3
