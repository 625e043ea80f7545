// FP64 discrete Wigner transforms on the vector ALUs (the plain method).
//
// On gfx950 a wave issuing v_fma_f64 sustains more FP64 throughput than
// v_mfma_f64_16x16x4f64, so the FP64 DWTs run on the VALU.  One work unit is an
// order pair 0 <= p <= q with its up to 8 symmetric (m, n) variants
// (pysofft/soft.f08:534-631, 937-1023), which share d~^l_{p,q}(beta) up to the
// signs, (-1)^l factors and beta reversals of the variant table.
//
// Operand delivery: one contraction operand is per lane (d~ for the lane's beta or
// degree); the other (a coefficient, or a sample of F) is the same for all lanes.
// Those shared operands are loaded with ordinary vector loads, 16 per 16-lane row
// (lane k of a row holds variant k & 7 at step offset k >> 3, as a complex), and
// reach every lane of the row through the 64-bit DPP row_newbcast:k source modifier
// of v_fmac_f64_dpp.  One row load thus feeds 32 FMAs per lane with no LDS traffic,
// and vector loads can be issued two steps ahead (they complete in order).
//  - inverse: lane = beta.  Each lane runs the FP64 Kostelec recurrence in l for its
//    beta (pysofft/wigner.f90:227-260) and accumulates F(beta) = sum_l c_l d~^l(beta)
//    for 8 variants x TB transforms in registers.
//  - forward: lane = degree.  For each 64-beta tile the waves of a block (as betas)
//    run the recurrence and store w(beta) d~^l(beta) transposed in LDS; each wave (as
//    64 degrees) then accumulates f^l = sum_beta F(beta) w d~^l(beta) for its TB
//    transforms.
// Constant signs, (-1)^l factors and conjugations are applied to d or to the sums.
#include "so3ft_internal.h"

namespace so3 {

__device__ __forceinline__ int wrap_o(int m, int N) { return m >= 0 ? m : N + m; }

// d~^{l+1} = a (cos(beta) - x) d~^l - b d~^{l-1} (pysofft/wigner.f90:227-260)
__device__ inline void rec_consts_v(int l, int p, int q, double &a, double &x, double &b)
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

// the plan's recurrence-constant table, read-only while a kernel runs: in the constant
// address space a load whose index is the same in every lane is a scalar load
typedef const __attribute__((address_space(4))) double *cdptr;
typedef const __attribute__((address_space(4))) int64_t *ci64ptr;

__device__ __forceinline__ int64_t slice_v(const int64_t *slice, int m, int n, int bw)
{
    return slice[(int64_t)(m + bw - 1) * (2 * bw - 1) + (n + bw - 1)];
}

// Where the inverse DWT of variant v reads its coefficients: the mnl slice, the
// constant sign and whether the coefficient enters conjugated; the (-1)^l factor
// applies to v >= 4 in both modes.
//   complex (soft.f08:534-631): slice (m_v, n_v), sign of the variant table.
//   real    (soft.f08:1259-1318): rows m >= 0 read directly, rows m < 0 as the
//           Hermitian completion F(-m,-n) = conj F(m,n) of the c2r plan.
struct InvSrc {
    int64_t off;
    double sign;
    bool conj;
};

__device__ inline InvSrc inv_src(int mode, int v, int p, int q, int bw, const int64_t *slice)
{
    const Variant vr = variant(v, p, q);
    if (mode != DWT_REAL) return {slice_v(slice, vr.m, vr.n, bw), vr.sign, false};
    const double sp = (p & 1) ? -1.0 : 1.0;
    switch (v) {
    case 0: return {slice_v(slice, p, q, bw), 1.0, false};
    case 1:
        if (p == q) return {slice_v(slice, p, q, bw), 1.0, true};
        return {slice_v(slice, -q, -p, bw), 1.0, false};
    case 2: return {slice_v(slice, -q, -p, bw), 1.0, true};
    case 3: return {slice_v(slice, p, q, bw), 1.0, true};
    case 4: return {slice_v(slice, p, -q, bw), sp, false};
    case 5: return {slice_v(slice, p, -q, bw), sp, true};
    case 6: return {slice_v(slice, q, -p, bw), sp, false};
    default: return {slice_v(slice, q, -p, bw), sp, true};
    }
}

// v_fmac_f64_dpp with lane K of each 16-lane row as src0 (64-bit DPP on gfx950 allows
// only row_newbcast).  s_nop 1 gives the two wait states a DPP read needs after a VALU
// write of its source.
#define SO3_FM(A, C, D, K) "v_fmac_f64_dpp %" #A ", %" #C ", %" #D " row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n"
#define SO3_FV(RE, IM, D, K) SO3_FM(RE, 16, D, K) SO3_FM(IM, 17, D, K)

// acc[2v] += Re(c from lane v + 8h) * d_h, acc[2v+1] += Im(...) * d_h for v < 8,
// h < 2; variants 4..7 use dv (the (-1)^l-signed d) when SPLIT is set.
template <bool SPLIT>
__device__ __forceinline__ void fma_block(double (&a)[16], double2 c, double d0, double d1, double dv0, double dv1)
{
    if (SPLIT) {
        asm volatile("s_nop 1\n"
                     SO3_FV(0, 1, 18, 0) SO3_FV(2, 3, 18, 1) SO3_FV(4, 5, 18, 2) SO3_FV(6, 7, 18, 3)
                     SO3_FV(8, 9, 20, 4) SO3_FV(10, 11, 20, 5) SO3_FV(12, 13, 20, 6) SO3_FV(14, 15, 20, 7)
                     SO3_FV(0, 1, 19, 8) SO3_FV(2, 3, 19, 9) SO3_FV(4, 5, 19, 10) SO3_FV(6, 7, 19, 11)
                     SO3_FV(8, 9, 21, 12) SO3_FV(10, 11, 21, 13) SO3_FV(12, 13, 21, 14) SO3_FV(14, 15, 21, 15)
                     : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]),
                       "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]),
                       "+v"(a[14]), "+v"(a[15])
                     : "v"(c.x), "v"(c.y), "v"(d0), "v"(d1), "v"(dv0), "v"(dv1));
    } else {
        asm volatile("s_nop 1\n"
                     SO3_FV(0, 1, 18, 0) SO3_FV(2, 3, 18, 1) SO3_FV(4, 5, 18, 2) SO3_FV(6, 7, 18, 3)
                     SO3_FV(8, 9, 18, 4) SO3_FV(10, 11, 18, 5) SO3_FV(12, 13, 18, 6) SO3_FV(14, 15, 18, 7)
                     SO3_FV(0, 1, 19, 8) SO3_FV(2, 3, 19, 9) SO3_FV(4, 5, 19, 10) SO3_FV(6, 7, 19, 11)
                     SO3_FV(8, 9, 19, 12) SO3_FV(10, 11, 19, 13) SO3_FV(12, 13, 19, 14) SO3_FV(14, 15, 19, 15)
                     : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]),
                       "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]),
                       "+v"(a[14]), "+v"(a[15])
                     : "v"(c.x), "v"(c.y), "v"(d0), "v"(d1), "v"(dv0), "v"(dv1));
    }
}
#undef SO3_FV
#undef SO3_FM

// Block index -> (work unit, beta chunk) so that the nch chunks of one unit run on
// the same XCD (block ids that agree mod 8 share an XCD and its L2) and read the same
// coefficients from there.
__device__ __forceinline__ bool decode_unit(int nch, int nunit, int &unit, int &chunk)
{
    const int id = blockIdx.x;
    const int k = id >> 3;
    chunk = k % nch;
    unit = (k / nch) * 8 + (id & 7);
    return unit < nunit;
}

// Inverse.  Block = one wave, lanes = 64 betas of one chunk; TB transforms per wave.
// Work unit = (pair, transform group), pairs fastest.  TAB: d~ is read from the plan's
// d table (rows of the pair's degrees, the lane's beta) instead of the recurrence.
// CORR: the coefficients are the correlation products cc^l_{mn} of SH pairs (coeff = f,
// aux = g, LMc sets at stride c_ts; soft.f08:2242-2341), formed as each is loaded instead
// of being staged in mnl order first (degrees l > lcap are zero).
template <int TB, bool TAB, bool CORR>
__global__ __launch_bounds__(64) void dwt_inverse_valu(const double2 *__restrict__ coeff, int64_t c_ts,
                                                       double2 *__restrict__ F, int64_t f_ts, int T,
                                                       const double *__restrict__ cosb,
                                                       const double *__restrict__ seeds,
                                                       const int64_t *__restrict__ slice,
                                                       const int32_t *__restrict__ pairs, int bw, int N, int mode,
                                                       int npair, int nch, int nunit,
                                                       const double *__restrict__ dtab,
                                                       const double2 *__restrict__ aux, int lcap,
                                                       const double *__restrict__ rc,
                                                       const int64_t *__restrict__ rc_off)
{
    // CORR: the degree weights (dynamic LDS, bw doubles)
    extern __shared__ double cstw[];
    int unit, chunk;
    if (!decode_unit(nch, nunit, unit, chunk)) return;
    const int lane = threadIdx.x;
    // on the fly: pairs fastest.  TAB: the transform groups of one pair are consecutive
    // units of one XCD (unit = 8 slot + xcd), so the pair's table rows are read from HBM
    // once and then from that XCD's L2.
    int pidx, grp;
    if (TAB) {
        const int ng = (T + TB - 1) / TB, slot = unit >> 3;
        pidx = (slot / ng) * 8 + (unit & 7);
        grp = slot % ng;
        if (pidx >= npair) return;
    } else {
        pidx = unit % npair;
        grp = unit / npair;
    }
    const int pq = pairs[pidx];
    const int p = pq & 0xffff, q = pq >> 16;
    const int L = bw - q;
    const int beta = chunk * 64 + lane;
    const int t0 = grp * TB;
    const int nt = min(TB, T - t0);

    // the coefficient this lane holds for the row broadcasts: variant k & 7, degree
    // offset k >> 3 (k = lane & 15); an invalid variant reads variant 0's slice and an
    // out-of-range transform the last valid one (neither is stored)
    const int kv = lane & 7, kh = (lane >> 3) & 1;
    const int64_t off0 = inv_src(mode, 0, p, q, bw, slice).off;
    const int64_t offk = variant(kv, p, q).valid ? inv_src(mode, kv, p, q, bw, slice).off : off0;
    const double2 *src[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t)
        src[t] = coeff + (int64_t)(t0 + min(t, nt - 1)) * c_ts + (CORR ? 0 : offk + kh);
    // CORR: the lane's variant (variant 0 for an invalid one) and the g sets
    const Variant vk = variant(variant(kv, p, q).valid ? kv : 0, p, q);
    const double2 *gsr[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) gsr[t] = CORR ? aux + (int64_t)(t0 + min(t, nt - 1)) * c_ts : nullptr;
    // 2 pi sqrt(2/(2l+1)) of the pair's degrees in cstw
    // degrees past the slice are read clamped to its last entry and enter with d = 0
    auto load = [&](double2 (&dst)[TB], int j0) {
        if constexpr (CORR) {
            const int lr = min(j0 + kh, L - 1), l = q + lr, ix = l * (l + 1);
            const double wl = l <= lcap ? cstw[lr] : 0.0;
            const double wn = ((vk.m + vk.n) & 1) ? -wl : wl;
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                const double2 a = src[t][ix + vk.m], b = gsr[t][ix + vk.n];
                const double re = a.x * b.x + a.y * b.y, im = a.y * b.x - a.x * b.y;
                dst[t] = make_double2(wn * re, wn * im);
            }
        } else {
            const int j = min(j0 + kh, L - 1) - kh;
#pragma unroll
            for (int t = 0; t < TB; ++t) dst[t] = src[t][j];
        }
    };

    if (CORR) {
        for (int i = lane; i < L; i += 64)
            cstw[i] = 2.0 * 3.14159265358979323846 * sqrt(2.0 / (double)(2 * (q + i) + 1));
        __syncthreads();   // the first loads read them
    }
    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    // the pair's recurrence constants (a, x, b) from the plan table, zero past the last
    // degree: the same entry in every lane, so scalar loads
    const cdptr rcp = TAB ? nullptr : (cdptr)rc + 3 * ((ci64ptr)rc_off)[tri];
    double dcur = (!TAB && beta < N) ? seeds[tri * N + beta] : 0.0;
    const double cb = (!TAB && beta < N) ? cosb[beta] : 0.0;
    double dprev = 0.0;
    const double sq = (q & 1) ? -1.0 : 1.0;  // (-1)^l for l = q + j, j even
    // table rows of this pair at the lane's beta (lanes past N read row entry 0 and use 0)
    const double *trow = TAB ? dtab + dtab_row(p, q, bw) * N + (beta < N ? beta : 0) : nullptr;
    const double tsc = beta < N ? 1.0 : 0.0;
    double dring[3][2];   // table values of the three buffered steps
    auto load_d = [&](double (&dd)[2], int j0) {
        if (!TAB) return;
        dd[0] = j0 < L ? tsc * trow[(int64_t)j0 * N] : 0.0;
        dd[1] = j0 + 1 < L ? tsc * trow[(int64_t)(j0 + 1) * N] : 0.0;
    };

    double acc[TB][16];
#pragma unroll
    for (int t = 0; t < TB; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0;
    // three-step ring of row loads: degrees j0, j0+1 sit in buf[(j0 / 2) % 3]
    double2 buf[3][TB];
    load(buf[0], 0);
    load(buf[1], 2);
    load(buf[2], 4);
    load_d(dring[0], 0);
    load_d(dring[1], 2);
    load_d(dring[2], 4);
    __syncthreads();

    // one step = degrees j0, j0+1 from buf[u]; the recurrence always advances twice
    // (past the last degree it reads the table's zero padding and the result is never
    // used)
    auto step = [&](int u, int j0, bool tail) {
        if (TAB) {
            const double d0 = dring[u][0], d1 = dring[u][1];
            const double da0 = sq * d0, da1 = -sq * d1;
#pragma unroll
            for (int t = 0; t < TB; ++t) fma_block<true>(acc[t], buf[u][t], d0, d1, da0, da1);
            return;
        }
        const double d0 = dcur;
        {
            const double ca = rcp[3 * j0], cx = rcp[3 * j0 + 1], cbb = rcp[3 * j0 + 2];
            const double dn = (ca * (cb - cx)) * dcur - cbb * dprev;
            dprev = dcur;
            dcur = dn;
        }
        const double d1 = (!tail || j0 + 1 < L) ? dcur : 0.0;
        {
            const double ca = rcp[3 * j0 + 3], cx = rcp[3 * j0 + 4], cbb = rcp[3 * j0 + 5];
            const double dn = (ca * (cb - cx)) * dcur - cbb * dprev;
            dprev = dcur;
            dcur = dn;
        }
        const double da0 = sq * d0, da1 = -sq * d1;
#pragma unroll
        for (int t = 0; t < TB; ++t) fma_block<true>(acc[t], buf[u][t], d0, d1, da0, da1);
    };
    int j6 = 0;
    for (; j6 + 6 <= L; j6 += 6) {
        step(0, j6, false);
        load(buf[0], j6 + 6);
        load_d(dring[0], j6 + 6);
        step(1, j6 + 2, false);
        load(buf[1], j6 + 8);
        load_d(dring[1], j6 + 8);
        step(2, j6 + 4, false);
        load(buf[2], j6 + 10);
        load_d(dring[2], j6 + 10);
    }
    if (j6 < L) step(0, j6, true);
    if (j6 + 2 < L) step(1, j6 + 2, true);
    if (j6 + 4 < L) step(2, j6 + 4, true);
    if (beta >= N) return;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const Variant vr = variant(v, p, q);
        if (!vr.valid) continue;
        const InvSrc sr = inv_src(mode, v, p, q, bw, slice);
        const double sy = sr.conj ? -sr.sign : sr.sign;
        const int bo = vr.rev ? N - 1 - beta : beta;
        const int64_t col = ((int64_t)wrap_o(vr.n, N) * N + wrap_o(vr.m, N)) * N + bo;
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            if (t >= nt) continue;
            // written once, read by the next pass from HBM: a non-temporal store
            typedef double d2v __attribute__((ext_vector_type(2)));
            const d2v x = {sr.sign * acc[t][2 * v], sy * acc[t][2 * v + 1]};
            __builtin_nontemporal_store(x, reinterpret_cast<d2v *>(F + (int64_t)(t0 + t) * f_ts + col));
        }
    }
}

// Forward.  Block = NW waves sharing one 64-beta x 64-degree tile of w d~ in LDS;
// wave w owns transforms [t0 + w TB, t0 + (w+1) TB) and lanes = degrees
// l = q + lc0 + lane.  Each wave runs the recurrence for 64/NW of a tile's betas.
// Work unit = (pair, group of NW*TB transforms); chunk = 64-degree range.
template <int TB, int NW>
__global__ __launch_bounds__(64 * NW) void dwt_forward_valu(const double2 *__restrict__ F, int64_t f_ts,
                                                            double2 *__restrict__ coeff, int64_t c_ts, int T,
                                                            const double *__restrict__ w,
                                                            const double *__restrict__ cosb,
                                                            const double *__restrict__ seeds,
                                                            const int64_t *__restrict__ slice,
                                                            const int32_t *__restrict__ pairs, int bw, int N,
                                                            int mode, int npair, int nch, int nunit)
{
    constexpr int BPW = 64 / NW;    // betas of a tile whose recurrence one wave runs
    __shared__ double cst[3 * 256];
    __shared__ double Dt[64][65];   // [beta in tile][degree in chunk] = w d~
    __shared__ double sd[3][512];   // seed d~^q_{p,q}, cos(beta), w(beta) of every beta
    int unit, chunk;
    if (!decode_unit(nch, nunit, unit, chunk)) return;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index, uniform
    const int pq = pairs[unit % npair];
    const int p = pq & 0xffff, q = pq >> 16;
    const int L = bw - q;
    const int lc0 = chunk * 64;
    if (lc0 >= L) return;
    const int lend = min(L, lc0 + 64);  // recurrence runs l = q .. q+lend-1
    const int t0 = ((unit / npair) * NW + wv) * TB;
    const int nt = min(TB, T - t0);     // <= 0 for spare waves of the last group

    // the F sample this lane holds for the row broadcasts: variant k & 7 at beta
    // offset k >> 3; the half-complex forward takes the odd variants' rows as
    // conjugates of m >= 0 rows (soft.f08:1650, 3286-3300, 1662-1706)
    const int kv = lane & 7, kh = (lane >> 3) & 1;
    int64_t colk;
    {
        const Variant vr = variant(kv, p, q);
        int mm = vr.m, nn = vr.n;
        if (mode == DWT_REAL && (kv & 1)) {
            const int sv = (kv == 1) ? ((p == q) ? 0 : 2) : (kv == 3 ? 0 : kv - 1);
            const Variant s2 = variant(sv, p, q);
            mm = s2.m; nn = s2.n;
        }
        colk = ((int64_t)wrap_o(nn, N) * N + wrap_o(mm, N)) * N;
    }
    const bool rev = kv >= 4;
    const double2 *src[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) src[t] = F + (int64_t)max(0, min(t0 + t, T - 1)) * f_ts + colk;
    const int NP = N >> 1;  // beta pairs
    auto load = [&](double2 (&dst)[TB], int bp) {
        const int b = 2 * min(bp, NP - 1) + kh;  // prefetches past the end re-read the last pair
#pragma unroll
        for (int t = 0; t < TB; ++t) dst[t] = src[t][rev ? N - 1 - b : b];
    };

    for (int i = threadIdx.x; i < lend - 1; i += 64 * NW) {
        double a, x, b;
        rec_consts_v(q + i, p, q, a, x, b);
        cst[3 * i] = a; cst[3 * i + 1] = x; cst[3 * i + 2] = b;
    }
    {
        const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
        for (int i = threadIdx.x; i < N; i += 64 * NW) {
            sd[0][i] = seeds[tri * N + i];
            sd[1][i] = cosb[i];
            sd[2][i] = w[i];
        }
    }

    double acc[TB][16];
#pragma unroll
    for (int t = 0; t < TB; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0;
    double2 buf[4][TB];  // ring of row loads: pair bp in buf[bp & 3], refilled 4 pairs ahead
    if (nt > 0) {
        load(buf[0], 0);
        load(buf[1], 1);
        load(buf[2], 2);
        load(buf[3], 3);
    }

    auto step = [&](int u, int bl) {
        const double d0 = Dt[bl][lane], d1 = Dt[bl + 1][lane];
#pragma unroll
        for (int t = 0; t < TB; ++t) fma_block<false>(acc[t], buf[u][t], d0, d1, d0, d1);
    };
    for (int b0 = 0; b0 < N; b0 += 64) {
        __syncthreads();  // cst ready / previous tile consumed
        if (lane < BPW) {
            const int bl = wv * BPW + lane;
            const int beta = b0 + bl;
            const bool ok = beta < N;
            double dc = ok ? sd[0][beta] : 0.0, dp = 0.0;
            const double cb = ok ? sd[1][beta] : 0.0;
            const double wb = ok ? sd[2][beta] : 0.0;
            // the constants do not depend on d, so unrolling lets their LDS reads run
            // ahead of the two-op dependency chain through d
            auto adv = [&](int li) {
                const double ca = cst[3 * li], cx = cst[3 * li + 1], cbb = cst[3 * li + 2];
                const double dn = (ca * (cb - cx)) * dc - cbb * dp;
                dp = dc;
                dc = dn;
            };
            int li = 0;
#pragma unroll 8
            for (; li < lc0; ++li) adv(li);
#pragma unroll 8
            for (; li < lend - 1; ++li) {
                Dt[bl][li - lc0] = wb * dc;
                adv(li);
            }
            Dt[bl][li - lc0] = wb * dc;
        }
        __syncthreads();
        if (nt <= 0) continue;
        // beta pairs of this tile; every tile but the last has 32, a multiple of 4, so
        // the ring stays aligned (pair bp always sits in buf[bp & 3])
        const int npt = min(32, (N - b0) >> 1), gb = b0 >> 1;
        int i = 0;
        for (; i + 4 <= npt; i += 4) {
            step(0, 2 * i);
            load(buf[0], gb + i + 4);
            step(1, 2 * i + 2);
            load(buf[1], gb + i + 5);
            step(2, 2 * i + 4);
            load(buf[2], gb + i + 6);
            step(3, 2 * i + 6);
            load(buf[3], gb + i + 7);
        }
        if (i < npt) step(0, 2 * i);
        if (i + 1 < npt) step(1, 2 * i + 2);
        if (i + 2 < npt) step(2, 2 * i + 4);
    }
    const int li = lc0 + lane;
    if (li >= L || nt <= 0) return;
    const int l = q + li;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const Variant vr = variant(v, p, q);
        if (!vr.valid) continue;
        const double s = vr.sign * ((vr.lsign && (l & 1)) ? -1.0 : 1.0);
        const double sy = (mode == DWT_REAL && (v & 1)) ? -s : s;
        const int64_t o = slice_v(slice, vr.m, vr.n, bw) + li;
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            if (t >= nt) continue;
            coeff[(int64_t)(t0 + t) * c_ts + o] = make_double2(s * acc[t][2 * v], sy * acc[t][2 * v + 1]);
        }
    }
}

// f^l_m of one SH coefficient set: layout 0 = SHTNS LMc (index l(l+1)+m); layout 1 =
// MLr of a real function (pysofft/utils.f90:758-772, m >= 0 at m(2bw-1-m)/2 + l), the
// negative orders following from f^l_{-m} = (-1)^m conj(f^l_m)
__device__ __forceinline__ double2 sh_coeff(const double2 *__restrict__ x, int layout, int l, int m, int bw)
{
    if (layout == 0) return x[l * (l + 1) + m];
    const int am = m < 0 ? -m : m;
    double2 v = x[(int64_t)am * (2 * bw - 1 - am) / 2 + l];
    if (m < 0) {
        v.y = -v.y;
        if (am & 1) { v.x = -v.x; v.y = -v.y; }
    }
    return v;
}

// cc^l_{mn} = 2 pi sqrt(2/(2l+1)) (-1)^{m+n} sum_r w_r f^l_{m,r} conj(g^l_{n,r}) in mnl order
// (pysofft/soft.f08:2242-2341; real inputs soft.f08:2806-2872, whose half-complex
// branches are this formula with the negative orders completed).  Transform t = blockIdx.y
// reads its sets at t * t_ts; nr shells r at stride r_ts are summed with weights w (w ==
// nullptr: one shell of weight 1).  Degrees l > lcap are zero (the Risbo loops of
// soft.f08:2564-2580, 3046-3074 stop at l = lmax).  One wave per order pair and transform.
__global__ __launch_bounds__(256) void corr_coeffs_kernel(const double2 *__restrict__ f, const double2 *__restrict__ g,
                                                          int64_t t_ts, int layout, int nr, int64_t r_ts,
                                                          const double *__restrict__ w, double2 *__restrict__ out,
                                                          int64_t c_ts, const int64_t *__restrict__ slice, int bw,
                                                          int lcap)
{
    // wave = order pair 0 <= m1 <= m2 < bw (triangular index, m1 outer; four per block) of
    // transform blockIdx.y: its variant slices are contiguous in the mnl layout, so the
    // wave's lanes write consecutive entries
    const int P = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (P >= bw * (bw + 1) / 2) return;
    const int64_t t = blockIdx.y;
    const double B2 = 2.0 * bw + 1.0;
    int m1 = (int)((B2 - sqrt(B2 * B2 - 8.0 * P)) * 0.5);
    m1 = max(0, min(m1, bw - 1));
    while (m1 > 0 && m1 * (2 * bw - m1 + 1) / 2 > P) --m1;
    while (m1 + 1 < bw && (m1 + 1) * (2 * bw - m1) / 2 <= P) ++m1;
    const int m2 = m1 + (P - m1 * (2 * bw - m1 + 1) / 2);
    const int L = bw - m2;
    // the pair's variants in the order of the mnl layout (pysofft/utils.f90:87-201):
    // k = 0..7: (m1,m2), (-m2,-m1), (m2,m1), (-m1,-m2), (m1,-m2), (-m1,m2), (m2,-m1),
    // (-m2,m1); m1 = m2 keeps k = 0, 1, 4, 5, m1 = 0 keeps k = 0..3, (0, 0) only k = 0
    const int nv = (m1 == 0 && m2 == 0) ? 1 : ((m1 == 0 || m1 == m2) ? 4 : 8);
    const int64_t base = slice[(int64_t)(m1 + bw - 1) * (2 * bw - 1) + (m2 + bw - 1)];
    const float rL = 1.0f / (float)L;
    for (int o = lane; o < nv * L; o += 64) {
        int v = (int)(((float)o + 0.5f) * rL), li = o - v * L;   // o = v L + li, 0 <= li < L
        if (li < 0) { --v; li += L; } else if (li >= L) { ++v; li -= L; }
        const int l = m2 + li;
        const int k = (m1 == m2 && v >= 2) ? v + 2 : v;
        int m, n;
        switch (k) {
        case 0: m = m1; n = m2; break;
        case 1: m = -m2; n = -m1; break;
        case 2: m = m2; n = m1; break;
        case 3: m = -m1; n = -m2; break;
        case 4: m = m1; n = -m2; break;
        case 5: m = -m1; n = m2; break;
        case 6: m = m2; n = -m1; break;
        default: m = -m2; n = m1; break;
        }
        double re = 0.0, im = 0.0;
        const int nrl = l <= lcap ? nr : 0;
        for (int r = 0; r < nrl; ++r) {
            const double2 a = sh_coeff(f + t * t_ts + r * r_ts, layout, l, m, bw);
            const double2 b = sh_coeff(g + t * t_ts + r * r_ts, layout, l, n, bw);
            const double wr = w ? w[r] : 1.0;
            re += wr * (a.x * b.x + a.y * b.y);
            im += wr * (a.y * b.x - a.x * b.y);
        }
        const double wl = 2.0 * 3.14159265358979323846 * sqrt(2.0 / (double)(2 * l + 1));   // 2 pi sqrt(2/(2l+1))
        const double wn = ((m + n) & 1) ? -wl : wl;
        out[t * c_ts + base + o] = make_double2(wn * re, wn * im);
    }
}

hipError_t launch_corr_coeffs(const DeviceTables &tb, int bw, const double2 *f, const double2 *g, int64_t t_ts,
                              int layout, int nr, int64_t r_ts, const double *w, double2 *out, int64_t c_ts,
                              int lcap, int T, hipStream_t s)
{
    const int npq = bw * (bw + 1) / 2;
    hipLaunchKernelGGL(corr_coeffs_kernel, dim3((npq + 3) / 4, T), dim3(256), 0, s, f, g, t_ts, layout, nr, r_ts, w,
                       out, c_ts, tb.slice, bw, lcap);
    return hipGetLastError();
}

static int valu_tb_env()
{
    static int tb = [] {
        const char *e = getenv("SO3FT_VALU_TB");
        return e ? atoi(e) : 0;
    }();
    return tb;
}

static int grid_of(int nunit, int nch) { return (nunit + 7) / 8 * 8 * nch; }

hipError_t launch_dwt_inverse_valu(const DeviceTables &tb, int bw, int N, int npair, int mode, const double2 *coeff,
                                   int64_t c_ts, const double2 *aux, int lcap, double2 *F, int64_t f_ts, int T,
                                   hipStream_t s)
{
    int TBv = T >= 4 ? 4 : (T >= 2 ? 2 : 1);
    if (valu_tb_env()) TBv = valu_tb_env();
    const int nch = (N + 63) / 64;
    // dynamic LDS: the degree weights of a correlation
    const bool corr = mode == DWT_CORR;
    const size_t cst_bytes = corr ? sizeof(double) * (size_t)bw : 0;
    const size_t tab_bytes = cst_bytes;
    const int kmode = corr ? (int)DWT_COMPLEX : mode;   // the products land in the complex slices
#define SO3_INV(TBV, CR)                                                                                          \
    do {                                                                                                          \
        const int ngr = (T + TBV - 1) / TBV;                                                                      \
        const int nunit = npair * ngr;                                                                            \
        const int nunit_t = (npair + 7) / 8 * 8 * ngr;   /* table order: 8 pairs x groups per slot run */       \
        if (tb.dtab)                                                                                              \
            hipLaunchKernelGGL((dwt_inverse_valu<TBV, true, CR>), dim3(grid_of(nunit_t, nch)), dim3(64),          \
                               tab_bytes, s, coeff, c_ts, F, f_ts, T, tb.cosb, tb.seeds, tb.slice, tb.pairs, bw,  \
                               N, kmode, npair, nch, nunit_t, tb.dtab, aux, lcap, tb.rc, tb.rc_off);              \
        else                                                                                                      \
            hipLaunchKernelGGL((dwt_inverse_valu<TBV, false, CR>), dim3(grid_of(nunit, nch)), dim3(64),           \
                               cst_bytes, s, coeff, c_ts, F, f_ts, T, tb.cosb, tb.seeds, tb.slice, tb.pairs, bw,  \
                               N, kmode, npair, nch, nunit, nullptr, aux, lcap, tb.rc, tb.rc_off);                \
    } while (0)
    switch (TBv * 2 + corr) {
    case 2: SO3_INV(1, false); break;
    case 3: SO3_INV(1, true); break;
    case 4: SO3_INV(2, false); break;
    case 5: SO3_INV(2, true); break;
    case 9: SO3_INV(4, true); break;
    default: SO3_INV(4, false); break;
    }
#undef SO3_INV
    return hipGetLastError();
}

hipError_t launch_dwt_forward_valu(const DeviceTables &tb, int bw, int N, int npair, int mode, const double2 *F,
                                   int64_t f_ts, double2 *coeff, int64_t c_ts, int T, hipStream_t s)
{
    if (mode == DWT_CORR || bw > 256) return hipErrorInvalidValue;   // LDS sized for bw <= 256
    // TB = 4 needs 238 VGPRs (2 waves per SIMD) and measured slower than TB = 2
    int TBv = T >= 8 ? 2 : 1;
    if (valu_tb_env()) TBv = valu_tb_env();
    const int NWv = (TBv > 1 || T >= 3) ? 4 : T;
    const int nch = (bw + 63) / 64;
#define SO3_FWD(TBV, NWV)                                                                                         \
    do {                                                                                                          \
        const int nunit = npair * ((T + TBV * NWV - 1) / (TBV * NWV));                                            \
        hipLaunchKernelGGL((dwt_forward_valu<TBV, NWV>), dim3(grid_of(nunit, nch)), dim3(64 * NWV), 0, s, F, f_ts, \
                           coeff, c_ts, T, tb.w, tb.cosb, tb.seeds, tb.slice, tb.pairs, bw, N, mode, npair, nch,  \
                           nunit);                                                                                \
    } while (0)
    if (TBv == 4) SO3_FWD(4, 4);
    else if (TBv == 2) SO3_FWD(2, 4);
    else if (NWv == 4) SO3_FWD(1, 4);
    else if (NWv == 2) SO3_FWD(1, 2);
    else SO3_FWD(1, 1);
#undef SO3_FWD
    return hipGetLastError();
}

}  // namespace so3
