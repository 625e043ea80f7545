// Mixed-precision discrete Wigner transforms: complex64 coefficients and fft arrays,
// FP64 Kostelec recurrence (pysofft/wigner.f90:227-260), contraction on
// v_mfma_f32_16x16x4f32 with d split into FP32 hi + lo parts (two MFMAs per product,
// products of an exact FP32 operand with hi and lo), FP32 accumulation over 16 terms
// and FP64 accumulation across those blocks.  Work unit: an order pair 0 <= p <= q with
// its up to 8 symmetric variants (pysofft/soft.f08:534-631, 937-1023); the MFMA N
// dimension is the 16 columns (8 variants x re/im) of one transform.
//
// MFMA operand layouts (v_mfma_f32_16x16x4f32, probed on gfx950): lane i + 16k holds
// A[i][k] and B[k][i]; item r of the 4-float result of lane l is row 4(l>>4) + r, column
// l & 15.
//
// inverse: one wave per (pair, 64-beta chunk, TB transforms).  Lane = beta runs the
//   recurrence; every 4 degrees the wave writes d (hi, lo) for its 64 betas to a small
//   LDS tile and reads it back as the A operands of four 16-beta tiles (M = beta,
//   K = degree); B = the coefficients (K = degree, N = variant x re/im).
// forward: one wave per (pair, TB transforms), lane = beta within 64-beta chunks; for
//   each 16-degree tile and chunk the lanes advance their recurrence and store
//   w(beta) d (hi, lo) in LDS as [degree][beta]; A = that tile (M = degree, K = beta),
//   B = the fft-array samples (K = beta, N = variant x re/im).
#include "so3ft_internal.h"

namespace so3 {

typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrap_m(int m, int N) { return m >= 0 ? m : N + m; }

__device__ inline void rec_consts_m(int l, int p, int q, double &a, double &x, double &b)
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

__device__ __forceinline__ int64_t slice_m(const int64_t *slice, int m, int n, int bw)
{
    return slice[(int64_t)(m + bw - 1) * (2 * bw - 1) + (n + bw - 1)];
}

__device__ __forceinline__ f4_t mma32(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split(double d, float &hi, float &lo)
{
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

// ---------------------------------------------------------------------------
// inverse
// ---------------------------------------------------------------------------
constexpr int INV_DS = 80;  // padded row of the inverse d tile (float2 entries)

// one inverse work item: order pair (p, q), transforms t0 .. t0+TB-1 (< T), betas
// 64 chunk .. 64 chunk + 63; cst / Dl are the block's LDS
template <int TB>
__device__ __forceinline__ void inv_mix_item(const float2 *__restrict__ coeff, int64_t c_ts,
                                             float2 *__restrict__ F, int64_t f_ts, int T,
                                             const double *__restrict__ cosb, const double *__restrict__ seeds,
                                             const int64_t *__restrict__ slice, int bw, int N, int p, int q, int t0,
                                             int chunk, double *cst, float2 (*Dl)[4][INV_DS])
{
    const int lane = threadIdx.x;
    const int L = bw - q;
    const int b0 = chunk * 64;
    const int beta = b0 + lane;
    const int nt = min(TB, T - t0);

    // B operand of this lane: column j = lane & 15 (variant j >> 1, part j & 1), degree
    // offset k = lane >> 4 within each 4-degree step
    const int j = lane & 15, kk = lane >> 4;
    const int vj = j >> 1, partj = j & 1;
    const Variant vr = variant(vj, p, q);
    const int64_t offj = vr.valid ? slice_m(slice, vr.m, vr.n, bw) : slice_m(slice, p, q, bw);
    const float sgn = vr.valid ? (float)vr.sign : 0.0f;
    const bool lsg = vr.lsign;
    const float *src[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t)
        src[t] = reinterpret_cast<const float *>(coeff + (int64_t)(t0 + min(t, nt - 1)) * c_ts + offj) + partj;

    for (int i = lane; i < L - 1; i += 64) {
        double a, x, b;
        rec_consts_m(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }
    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    double dcur = (beta < N) ? seeds[tri * N + beta] : 0.0;
    const double cb = (beta < N) ? cosb[beta] : 0.0;
    double dprev = 0.0;

    f4_t acc[TB][4];
    double accd[TB][4][4];
#pragma unroll
    for (int t = 0; t < TB; ++t)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            acc[t][g] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) accd[t][g][r] = 0.0;
        }
    __syncthreads();

    // B operand of degree step l0: the signed coefficient of column j at degree l0 + kk;
    // kept in a ring three steps deep so the loads are in flight while earlier steps run
    auto loadb = [&](float (&dst)[TB], int l0) {
        const int lk = l0 + kk;
        const float s = (lsg && ((q + lk) & 1)) ? -sgn : sgn;
        const int lc = min(lk, L - 1);   // clamped: past-the-end degrees enter with a zero sign
        const float sz = (lk < L) ? s : 0.0f;
#pragma unroll
        for (int t = 0; t < TB; ++t) dst[t] = sz * src[t][2 * (int64_t)lc];
    };
    float bq0[TB], bq1[TB], bq2[TB];
    loadb(bq0, 0);
    loadb(bq1, 4);
    loadb(bq2, 8);

    // d for degrees l0..l0+3 of this lane's beta -> tile buf (then advanced past them)
    auto dtile = [&](int buf, int l0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int li = l0 + k;
            float hi, lo;
            split(li < L ? dcur : 0.0, hi, lo);
            Dl[buf][k][lane] = make_float2(hi, lo);
            if (li < L - 1) {
                const double ca = cst[3 * li], cx = cst[3 * li + 1], cbb = cst[3 * li + 2];
                const double dn = (ca * (cb - cx)) * dcur - cbb * dprev;
                dprev = dcur;
                dcur = dn;
            }
        }
    };
    dtile(0, 0);
    for (int l0 = 0, buf = 0; l0 < L; l0 += 4, buf ^= 1) {
        // LDS operations of one wave execute in order; the empty asm keeps the compiler
        // from moving tile reads above the writes that produced them
        asm volatile("" ::: "memory");
        float2 a[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) a[g] = Dl[buf][kk][16 * g + j];
        float bv[TB];
#pragma unroll
        for (int t = 0; t < TB; ++t) { bv[t] = bq0[t]; bq0[t] = bq1[t]; bq1[t] = bq2[t]; }
        loadb(bq2, l0 + 12);
        // all hi products, then all lo products: no back-to-back dependent MFMAs
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < TB; ++t) acc[t][g] = mma32(a[g].x, bv[t], acc[t][g]);
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < TB; ++t) acc[t][g] = mma32(a[g].y, bv[t], acc[t][g]);
        // the next step's recurrence runs on the VALU while the MFMAs execute
        asm volatile("" ::: "memory");
        if (l0 + 4 < L) dtile(buf ^ 1, l0 + 4);
        if (((l0 >> 2) & 3) == 3 || l0 + 4 >= L) {  // flush every 16 degrees
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) accd[t][g][r] += (double)acc[t][g][r];
                    acc[t][g] = f4_t{0.f, 0.f, 0.f, 0.f};
                }
        }
    }
    // item r of lane l: beta row 4(l>>4) + r of tile g, column j = l & 15
    if (!vr.valid) return;
    const int64_t colbase = ((int64_t)wrap_m(vr.n, N) * N + wrap_m(vr.m, N)) * N;
#pragma unroll
    for (int t = 0; t < TB; ++t) {
        if (t >= nt) continue;
        float *Ft = reinterpret_cast<float *>(F + (int64_t)(t0 + t) * f_ts + colbase) + partj;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int bt = b0 + 16 * g + 4 * kk + r;
                if (bt < N) Ft[2 * (int64_t)(vr.rev ? N - 1 - bt : bt)] = (float)accd[t][g][r];
            }
    }
}

template <int TB>
__global__ __launch_bounds__(64) void dwt_inverse_mix(const float2 *__restrict__ coeff, int64_t c_ts,
                                                      float2 *__restrict__ F, int64_t f_ts, int T,
                                                      const double *__restrict__ cosb, const double *__restrict__ seeds,
                                                      const int64_t *__restrict__ slice,
                                                      const int32_t *__restrict__ pairs, int bw, int N, int npair,
                                                      int nch, int nunit)
{
    __shared__ double cst[3 * 258];
    __shared__ float2 Dl[2][4][INV_DS];    // [step parity][degree in step][beta] = (hi, lo)
    const int id = blockIdx.x;
    const int chunk = (id >> 3) % nch;
    const int unit = ((id >> 3) / nch) * 8 + (id & 7);  // chunks of a unit share an XCD
    if (unit >= nunit) return;
    const int pq = pairs[unit % npair];
    inv_mix_item<TB>(coeff, c_ts, F, f_ts, T, cosb, seeds, slice, bw, N, pq & 0xffff, pq >> 16,
                     (unit / npair) * TB, chunk, cst, Dl);
}

// grouped launch over transforms of several bandwidths: a persistent grid walks the
// work list (item = segment, p | q << 16, first transform, chunk; segment -1 = padding)
template <int TB>
__global__ __launch_bounds__(64) void dwt_inverse_mix_grouped(const float2 *__restrict__ coeff,
                                                              float2 *__restrict__ F,
                                                              const MixSegment *__restrict__ segs,
                                                              const int4 *__restrict__ items, int nitem)
{
    __shared__ double cst[3 * 258];
    __shared__ float2 Dl[2][4][INV_DS];
    for (int it = blockIdx.x; it < nitem; it += gridDim.x) {
        const int4 w = items[it];
        if (w.x < 0) continue;
        const MixSegment sg = segs[w.x];
        __syncthreads();  // the previous item's LDS reads are done
        inv_mix_item<TB>(coeff + sg.coff, sg.nc, F + sg.goff, (int64_t)sg.N * sg.N * sg.N, sg.T, sg.cosb, sg.seeds,
                         sg.slice, sg.bw, sg.N, w.y & 0xffff, w.y >> 16, w.z, w.w, cst, Dl);
    }
}

// ---------------------------------------------------------------------------
// forward, one wave per (pair, TB transforms): lane = beta within a 64-beta chunk; the
// lane carries one recurrence chain per chunk (beta = 64 c + lane), its state kept in
// LDS between tiles.  Degrees go in 16-degree tiles; for each (tile, chunk) the lanes
// advance their chain by 16 degrees, writing w(beta) d (hi, lo) to an LDS tile
// [degree][beta], and the wave runs 16 K-steps of v_mfma_f32_16x16x4f32 with A = that
// tile (M = degree, K = beta) and B = the fft-array samples (K = beta, N = variant x
// re/im), in four independent accumulators per transform.  FP32 accumulation over 16
// betas, FP64 across them and across chunks: a tile's coefficients are complete after
// its last chunk, so there is no cross-wave reduction and no degree is recomputed.
// ---------------------------------------------------------------------------
constexpr int FT_DS = 66;   // padded row of the forward tile (float2 entries)

template <int TB, int MT>
__device__ __forceinline__ void fwd_mix2_item(const float2 *__restrict__ F, int64_t f_ts,
                                              float2 *__restrict__ coeff, int64_t c_ts, int T,
                                              const double *__restrict__ w, const double *__restrict__ cosb,
                                              const double *__restrict__ seeds, const int64_t *__restrict__ slice,
                                              int bw, int N, int p, int q, int t0, double *cst,
                                              float2 (*Dt)[FT_DS], double2 (*rs)[64])
{
    constexpr int TD = 16 * MT;   // degrees per tile: MT row tiles of 16 share each B operand
    constexpr int KS = MT == 1 ? 4 : 2;   // independent accumulators per (row tile, transform)
    const int lane = threadIdx.x & 63;
    const int L = bw - q;
    const int nc = (N + 63) >> 6;
    const int nt = min(TB, T - t0);
    const int j = lane & 15, kk = lane >> 4;
    const int vj = j >> 1, partj = j & 1;
    const Variant vr = variant(vj, p, q);
    const int64_t colj = ((int64_t)wrap_m(vr.n, N) * N + wrap_m(vr.m, N)) * N;
    const float *src[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t)
        src[t] = reinterpret_cast<const float *>(F + (int64_t)max(0, min(t0 + t, T - 1)) * f_ts + colj) + partj;

    for (int i = lane; i < L - 1; i += 64) {
        double a, x, b;
        rec_consts_m(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }
    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    // recurrence state (d^l, d^{l-1}) of this lane's chain in each chunk
    for (int c = 0; c < nc; ++c) {
        const int beta = 64 * c + lane;
        rs[c][lane] = make_double2(beta < N ? seeds[tri * N + beta] : 0.0, 0.0);
    }
    __syncthreads();   // cst, rs (one wave: orders the LDS writes before the reads)

#pragma unroll 1
    for (int l0 = 0; l0 < L; l0 += TD) {
        double accd[MT][TB][4];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) accd[m][t][r] = 0.0;
#pragma unroll 1
        for (int c = 0; c < nc; ++c) {
            // B operands of this chunk: the sample of column j at beta 64 c + 4 s + kk
            float bv[16][TB];
#pragma unroll
            for (int s4 = 0; s4 < 16; ++s4) {
                const int bt = min(64 * c + 4 * s4 + kk, N - 1);   // rows past N meet zero A columns
                const int64_t o = 2 * (int64_t)(vr.rev ? N - 1 - bt : bt);
#pragma unroll
                for (int t = 0; t < TB; ++t) bv[s4][t] = src[t][o];
            }
            // w d (hi, lo) for degrees l0 .. l0 + TD - 1 at this lane's beta of chunk c
            const int beta = 64 * c + lane;
            const bool ok = beta < N;
            const double cb = ok ? cosb[beta] : 0.0, wt = ok ? w[beta] : 0.0;
            asm volatile("" ::: "memory");
            double2 st = rs[c][lane];
            double dcur = st.x, dprev = st.y;
#pragma unroll
            for (int k = 0; k < TD; ++k) {
                const int li = l0 + k;
                float hi, lo;
                split(li < L ? wt * dcur : 0.0, hi, lo);
                Dt[k][lane] = make_float2(hi, lo);
                if (li < L - 1) {
                    const double ca = cst[3 * li], cx = cst[3 * li + 1], cbb = cst[3 * li + 2];
                    const double dn = (ca * (cb - cx)) * dcur - cbb * dprev;
                    dprev = dcur;
                    dcur = dn;
                }
            }
            rs[c][lane] = make_double2(dcur, dprev);
            asm volatile("" ::: "memory");
            // independent accumulators per (row tile, transform): K-steps split KS ways,
            // 16 betas each at most: no back-to-back dependent MFMAs
            f4_t acc[MT][TB][KS];
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int t = 0; t < TB; ++t)
#pragma unroll
                    for (int g = 0; g < KS; ++g) acc[m][t][g] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s4 = 0; s4 < 16; ++s4) {
                float2 a[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) a[m] = Dt[16 * m + j][4 * s4 + kk];
                const int g = s4 % KS;
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int t = 0; t < TB; ++t) acc[m][t][g] = mma32(a[m].x, bv[s4][t], acc[m][t][g]);
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int t = 0; t < TB; ++t) acc[m][t][g] = mma32(a[m].y, bv[s4][t], acc[m][t][g]);
                if (KS == 2 && (s4 & 7) == 7) {   // KS = 2: flush every 8 K-steps (16 betas each)
#pragma unroll
                    for (int m = 0; m < MT; ++m)
#pragma unroll
                        for (int t = 0; t < TB; ++t)
#pragma unroll
                            for (int gg = 0; gg < KS; ++gg) {
#pragma unroll
                                for (int r = 0; r < 4; ++r) accd[m][t][r] += (double)acc[m][t][gg][r];
                                acc[m][t][gg] = f4_t{0.f, 0.f, 0.f, 0.f};
                            }
                }
            }
            if (KS == 4) {
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int t = 0; t < TB; ++t)
#pragma unroll
                        for (int gg = 0; gg < KS; ++gg)
#pragma unroll
                            for (int r = 0; r < 4; ++r) accd[m][t][r] += (double)acc[m][t][gg][r];
            }
        }
        // item r of lane l: degree row 4(l>>4) + r of row tile m, column j = l & 15
        if (vr.valid) {
            const int64_t o = slice_m(slice, vr.m, vr.n, bw);
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                if (t >= nt) continue;
                float *ct = reinterpret_cast<float *>(coeff + (int64_t)(t0 + t) * c_ts + o) + partj;
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int li = l0 + 16 * m + 4 * kk + r;
                        if (li < L) {
                            const int l = q + li;
                            const double sg = vr.sign * ((vr.lsign && (l & 1)) ? -1.0 : 1.0);
                            ct[2 * (int64_t)li] = (float)(sg * accd[m][t][r]);
                        }
                    }
            }
        }
    }
}

template <int TB, int MT>
__global__ __launch_bounds__(64) void dwt_forward_mix2(const float2 *__restrict__ F, int64_t f_ts,
                                                       float2 *__restrict__ coeff, int64_t c_ts, int T,
                                                       const double *__restrict__ w, const double *__restrict__ cosb,
                                                       const double *__restrict__ seeds,
                                                       const int64_t *__restrict__ slice,
                                                       const int32_t *__restrict__ pairs, int bw, int N, int npair,
                                                       int nunit)
{
    // dynamic LDS: Dt [16 MT][FT_DS] float2 | rs [nc][64] double2 | cst [3 (bw - 1)] double
    extern __shared__ double smx[];
    float2 (*Dt)[FT_DS] = reinterpret_cast<float2 (*)[FT_DS]>(smx);
    double2 (*rs)[64] = reinterpret_cast<double2 (*)[64]>(smx + 16 * MT * FT_DS);
    double *cst = smx + 16 * MT * FT_DS + 2 * 64 * ((N + 63) >> 6);
    const int unit = blockIdx.x;
    if (unit >= nunit) return;
    const int pq = pairs[unit % npair];
    fwd_mix2_item<TB, MT>(F, f_ts, coeff, c_ts, T, w, cosb, seeds, slice, bw, N, pq & 0xffff, pq >> 16,
                          (unit / npair) * TB, cst, Dt, rs);
}

// grouped forward over transforms of several bandwidths (see dwt_inverse_mix_grouped);
// dynamic LDS sized for the largest bandwidth of the group
template <int TB, int MT>
__global__ __launch_bounds__(64) void dwt_forward_mix2_grouped(const float2 *__restrict__ F,
                                                               float2 *__restrict__ coeff,
                                                               const MixSegment *__restrict__ segs,
                                                               const int4 *__restrict__ items, int nitem)
{
    extern __shared__ double smx[];
    for (int it = blockIdx.x; it < nitem; it += gridDim.x) {
        const int4 wk = items[it];
        if (wk.x < 0) continue;
        const MixSegment sg = segs[wk.x];
        __syncthreads();  // the previous item's LDS reads are done
        float2 (*Dt)[FT_DS] = reinterpret_cast<float2 (*)[FT_DS]>(smx);
        double2 (*rs)[64] = reinterpret_cast<double2 (*)[64]>(smx + 16 * MT * FT_DS);
        double *cst = smx + 16 * MT * FT_DS + 2 * 64 * ((sg.N + 63) >> 6);
        fwd_mix2_item<TB, MT>(F + sg.goff, (int64_t)sg.N * sg.N * sg.N, coeff + sg.coff, sg.nc, sg.T, sg.w, sg.cosb,
                              sg.seeds, sg.slice, sg.bw, sg.N, wk.y & 0xffff, wk.y >> 16, wk.z, cst, Dt, rs);
    }
}

static int grid_m(int nunit, int nch) { return (nunit + 7) / 8 * 8 * nch; }

// dynamic LDS bytes of dwt_forward_mix2 (A tile of 16 MT degrees, chunk states,
// recurrence constants)
static size_t fwd2_lds(int bw, int N, int MT)
{
    return sizeof(double) * (16 * MT * FT_DS + 2 * 64 * ((N + 63) / 64) + 3 * (bw > 1 ? bw - 1 : 1));
}

hipError_t launch_dwt_inverse_mixed(const DeviceTables &tb, int bw, int N, int npair, int mode, const float2 *coeff,
                                    int64_t c_ts, const float2 *aux, float2 *F, int64_t f_ts, int T, hipStream_t s)
{
    if (mode != DWT_COMPLEX || aux) return hipErrorInvalidValue;
    if (npair <= 0 || T <= 0) return hipSuccess;
    const int nch = (N + 63) / 64;
#define SO3_MINV(TBV)                                                                                          \
    do {                                                                                                       \
        const int nunit = npair * ((T + TBV - 1) / TBV);                                                       \
        hipLaunchKernelGGL((dwt_inverse_mix<TBV>), dim3(grid_m(nunit, nch)), dim3(64), 0, s, coeff, c_ts, F, f_ts, \
                           T, tb.cosb, tb.seeds, tb.slice, tb.pairs, bw, N, npair, nch, nunit);                  \
    } while (0)
    if (T >= 2) SO3_MINV(2);  // TB = 4 needs 288 registers (one wave per SIMD)
    else SO3_MINV(1);
#undef SO3_MINV
    return hipGetLastError();
}

hipError_t launch_dwt_forward_mixed(const DeviceTables &tb, int bw, int N, int npair, int mode, const float2 *F,
                                    int64_t f_ts, float2 *coeff, int64_t c_ts, int T, hipStream_t s)
{
    if (mode != DWT_COMPLEX) return hipErrorInvalidValue;
    if (npair <= 0 || T <= 0) return hipSuccess;
#define SO3_MFWD2(TBV, MTV)                                                                                    \
    do {                                                                                                       \
        const int nunit = npair * ((T + TBV - 1) / TBV);                                                       \
        hipLaunchKernelGGL((dwt_forward_mix2<TBV, MTV>), dim3(nunit), dim3(64), fwd2_lds(bw, N, MTV), s, F, f_ts, \
                           coeff, c_ts, T, tb.w, tb.cosb, tb.seeds, tb.slice, tb.pairs, bw, N, npair, nunit);   \
    } while (0)
    // four transforms per wave from bw = 96 (more MFMA work per recurrence step), else two;
    // one 16-degree row tile per B operand load (measured faster than two)
    if (bw >= 96 && T >= 4) SO3_MFWD2(4, 1);
    else if (T >= 2) SO3_MFWD2(2, 1);
    else SO3_MFWD2(1, 1);
#undef SO3_MFWD2
    return hipGetLastError();
}

// ---- grouped launches: TB (transforms per wave) and grid from the tuned configuration
hipError_t launch_dwt_inverse_mixed_grouped(int TB, const float2 *coeff, float2 *F, const MixSegment *segs,
                                            const int4 *items, int nitem, int grid, hipStream_t s)
{
    if (nitem <= 0) return hipSuccess;
    if (grid <= 0) return hipErrorInvalidValue;
    if (TB == 2)
        hipLaunchKernelGGL((dwt_inverse_mix_grouped<2>), dim3(grid), dim3(64), 0, s, coeff, F, segs, items, nitem);
    else if (TB == 1)
        hipLaunchKernelGGL((dwt_inverse_mix_grouped<1>), dim3(grid), dim3(64), 0, s, coeff, F, segs, items, nitem);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_dwt_forward_mixed_grouped(int TB, const float2 *F, float2 *coeff, const MixSegment *segs,
                                            const int4 *items, int nitem, int grid, int max_bw, hipStream_t s)
{
    if (nitem <= 0) return hipSuccess;
    if (grid <= 0) return hipErrorInvalidValue;
    const size_t lds = fwd2_lds(max_bw, 2 * max_bw, 1);
    if (TB == 4)
        hipLaunchKernelGGL((dwt_forward_mix2_grouped<4, 1>), dim3(grid), dim3(64), lds, s, F, coeff, segs, items, nitem);
    else if (TB == 2)
        hipLaunchKernelGGL((dwt_forward_mix2_grouped<2, 1>), dim3(grid), dim3(64), lds, s, F, coeff, segs, items, nitem);
    else if (TB == 1)
        hipLaunchKernelGGL((dwt_forward_mix2_grouped<1, 1>), dim3(grid), dim3(64), lds, s, F, coeff, segs, items, nitem);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// resident blocks per CU of a grouped kernel (forward: LDS for bandwidth max_bw)
int mixed_grouped_occupancy(int inverse, int TB, int max_bw)
{
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (inverse) {
        if (TB == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_inverse_mix_grouped<2>, 64, 0);
        else if (TB == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_inverse_mix_grouped<1>, 64, 0);
    } else {
        const size_t lds = fwd2_lds(max_bw, 2 * max_bw, 1);
        if (TB == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_forward_mix2_grouped<4, 1>, 64, lds);
        else if (TB == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_forward_mix2_grouped<2, 1>, 64, lds);
        else if (TB == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_forward_mix2_grouped<1, 1>, 64, lds);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

}  // namespace so3
