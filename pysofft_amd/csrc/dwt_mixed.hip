// Mixed-precision discrete Wigner transforms: complex64 coefficients and fft arrays,
// FP64 Kostelec recurrence (pysofft/wigner.f90:227-260), contraction on the matrix cores.
// The inverse splits d into FP32 hi + lo parts: the hi products on v_mfma_f32_16x16x4f32
// (exact FP32 operands), the lo products, four K-steps at a time, on
// v_mfma_f32_16x16x16_bf16 (they are below 2^-24 of the result, so bfloat16 operands
// suffice); the forward uses the FP32 value of w d.  FP32 accumulation over at most 16
// terms and FP64 accumulation across those blocks.  Work unit: an order pair
// 0 <= p <= q with its up to 8 symmetric variants (pysofft/soft.f08:534-631, 937-1023);
// the MFMA N dimension is 16 columns = 8 variants x 2 transforms, real and imaginary
// parts in separate MFMAs, so a lane always loads or stores whole complex values.
//
// MFMA operand layouts (v_mfma_f32_16x16x4f32, probed on gfx950): lane i + 16k holds
// A[i][k] and B[k][i]; item r of the 4-float result of lane l is row 4(l>>4) + r, column
// l & 15.
//
// inverse: one wave per (pair, 64-beta chunk, 2 transforms).  Lane = beta runs the
//   recurrence; every 8 degrees the wave writes d for its 64 betas to LDS (hi parts as
//   an FP32 tile, lo parts as bfloat16 in the slot order of the bfloat16 MFMA) and reads
//   them back as the A operands of four 16-beta tiles (M = beta, K = degree); B = the
//   coefficients (K = degree, N = columns).
// forward: one wave per (pair, TB transforms), lane = beta within 64-beta chunks; for
//   each 32-degree tile and chunk the lanes advance their recurrence and store
//   w(beta) d in LDS as [degree][beta]; A = that tile's two 16-degree halves (M = degree,
//   K = beta), B = the fft-array samples (K = beta, N = columns).
#include "so3ft_internal.h"

#include <type_traits>

namespace so3 {

typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrap_m(int m, int N) { return m >= 0 ? m : N + m; }

// the plan's recurrence-constant table, read-only for a kernel's lifetime: addressed as
// constant memory, a load whose index is the same in every lane becomes a scalar load
// (also when the table pointer itself was read from memory, as in the grouped kernels)
typedef const __attribute__((address_space(4))) double *const_dptr;
typedef const __attribute__((address_space(4))) int64_t *const_i64ptr;

// one step of the chain: (dcur, dprev) <- (d^{l+1}, d^l) with entry i of the pair's
// recurrence constants (zero past the last degree), from the constant table or from LDS
template <class P>
__device__ __forceinline__ void rec_step(P rcp, int i, double cb, double &dcur, double &dprev)
{
    const double t = rcp[3 * i] * (cb - rcp[3 * i + 1]);   // one scalar operand per instruction
    const double dn = fma(t, dcur, -rcp[3 * i + 2] * dprev);
    dprev = dcur;
    dcur = dn;
}

// plan tables live in global memory; the grouped kernels read their pointers from the
// segment table, so the address space is restated for the compiler (no flat accesses)
template <class T> using gptr = const __attribute__((address_space(1))) T *;

__device__ __forceinline__ int64_t slice_m(const int64_t *slice, int m, int n, int bw)
{
    return ((gptr<int64_t>)slice)[(int64_t)(m + bw - 1) * (2 * bw - 1) + (n + bw - 1)];
}

__device__ __forceinline__ f4_t mma32(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// the lo products on the bfloat16 matrix path: v_mfma_f32_16x16x16_bf16, lane i + 16k
// holding A[i][4k + r] and B[4k + r][i] in slot r (probed on gfx950), result layout as
// mma32.  A lo part is below 2^-24 of its d, so rounding it and the other operand to
// bfloat16 (8-bit significands) changes a product by under 2^-32 of d times that operand:
// four K-steps of lo products cost one instruction instead of four FP32 ones.
typedef __bf16 bf4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
typedef short s4_t __attribute__((ext_vector_type(4)));
// two floats as bfloat16 (rounded toward zero) in one word, first in the low half.  Bit
// operations on purpose: a conversion of a float that was itself rounded from a double is
// folded by the compiler into a correctly rounded double -> bfloat16 conversion, about ten
// instructions per value; the lo parts need no such rounding.
__device__ __forceinline__ uint32_t bf16x2_trunc(float a, float b)
{
    // v_perm_b32: result byte i = byte sel_i of (b : a) (selectors 0-3 pick from a);
    // = (a >> 16) | (b & 0xffff0000)
    return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, b), __builtin_bit_cast(uint32_t, a), 0x07060302u);
}
__device__ __forceinline__ bf4_t bf4_of(uint32_t lo, uint32_t hi)
{
    return __builtin_bit_cast(bf4_t, make_uint2(lo, hi));
}
__device__ __forceinline__ f4_t mma_bf16(bf4_t a, bf4_t b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4_t, a), __builtin_bit_cast(s4_t, b), c, 0,
                                                     0, 0);
}

__device__ __forceinline__ void split(double d, float &hi, float &lo)
{
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

// ---------------------------------------------------------------------------
// inverse
// ---------------------------------------------------------------------------
constexpr int INV_DS = 72;  // padded row of the inverse hi tile (floats)
constexpr int INV_LS = 80;  // padded row of the inverse lo table (4 bfloat16 per beta)
constexpr int INV_TD = 8;   // degrees per tile: two K-steps of 4

// dynamic LDS of the inverse: Dh [2][INV_TD][INV_DS] float | Dlo [4][INV_LS][2] uint32
constexpr int INV_LDS_D = (2 * INV_TD * INV_DS + 4 * INV_LS * 2) / 2;   // doubles
static size_t inv_lds(int)
{
    return sizeof(double) * INV_LDS_D;
}
typedef float (*InvHi)[INV_TD][INV_DS];
typedef uint32_t (*InvLo)[INV_LS][2];

// one inverse work item: order pair (p, q), transforms t0 .. t0+TB-1 (< T), betas
// 64 chunk .. 64 chunk + 63.  MFMA columns j = variant j & 7 of transform t0 + (j >> 3),
// real and imaginary parts in separate MFMAs; degrees go in tiles of 8, K-step s of a
// tile taking degree l0 + 2 k + s in lane row k, so that a lane's coefficients of both
// K-steps are adjacent in memory.  rc / rc_off: recurrence constants; Dh / Dlo: the
// block's LDS.
template <int TB>
__device__ __forceinline__ void inv_mix_item(const float2 *__restrict__ coeff, int64_t c_ts,
                                             float2 *__restrict__ F, int64_t f_ts, int T,
                                             const double *__restrict__ cosb, const double *__restrict__ seeds,
                                             const double *__restrict__ rc, const int64_t *__restrict__ rc_off,
                                             const int64_t *__restrict__ slice, int bw, int N, int p, int q, int t0,
                                             int chunk, InvHi Dh, InvLo Dlo)
{
    static_assert(TB == 1 || TB == 2, "one transform pair per wave");
    const int lane = threadIdx.x;
    const int L = bw - q;
    const int b0 = chunk * 64;
    const int beta = b0 + lane;
    const int nt = min(TB, T - t0);

    // B operands of this lane: column j = lane & 15, degree row k = lane >> 4
    const int j = lane & 15, kk = lane >> 4;
    const int vj = j & 7, tt = j >> 3;
    const Variant vr = variant(vj, p, q);
    const int64_t offj = vr.valid ? slice_m(slice, vr.m, vr.n, bw) : slice_m(slice, p, q, bw);
    // signs of K-steps 0 and 1: degree l = q + l0 + 2 k + s with l0 a multiple of 8, so
    // the variant's (-1)^l factor depends on s only
    const float sg0 = vr.valid ? (float)vr.sign * ((vr.lsign && (q & 1)) ? -1.0f : 1.0f) : 0.0f;
    const float sg1 = vr.lsign ? -sg0 : sg0;
    const float2 *src = coeff + (int64_t)(t0 + min(tt, nt - 1)) * c_ts + offj;

    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    // the pair's recurrence constants: scalar loads (the same in every lane), each one
    // operand of a step's instructions
    const const_dptr rcp = (const_dptr)rc + 3 * ((const_i64ptr)rc_off)[tri];
    double dcur = (beta < N) ? ((gptr<double>)seeds)[tri * N + beta] : 0.0;
    const double cb = (beta < N) ? ((gptr<double>)cosb)[beta] : 0.0;
    double dprev = 0.0;

    f4_t ar[4], ai[4];
    double dr[4][4], di[4][4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        ar[g] = ai[g] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 4; ++r) dr[g][r] = di[g][r] = 0.0;
    }
    // coefficients of degrees l0 + 2k and l0 + 2k + 1 (re, im, re, im), each clamped into
    // the slice (past-the-end degrees enter with a zero sign); no branch, so that the
    // ring's loads stay in flight across tiles (a wait after a branch merge is for all)
    auto loadb = [&](int l0, float2 &x0, float2 &x1) {
        const int lk = l0 + 2 * kk;
        x0 = src[min(lk, L - 1)];
        x1 = src[min(lk + 1, L - 1)];
    };
    // d for degrees l0 .. l0 + 7 of this lane's beta (then advanced past them; zero past
    // the last degree): hi parts -> tile buf, lo parts as bfloat16 -> the lo table, slots
    // 2 h, 2 h + 1 of row k for degrees l0 + 2 k, l0 + 2 k + 1, h = the tile's place in
    // its flush block
    auto dtile = [&](auto hc, int buf, int l0) {
        constexpr int h = decltype(hc)::value;
#pragma unroll
        for (int k = 0; k < INV_TD / 2; ++k) {
            float hi0, lo0, hi1, lo1;
            split(dcur, hi0, lo0);
            Dh[buf][2 * k][lane] = hi0;
            rec_step(rcp, l0 + 2 * k, cb, dcur, dprev);
            split(dcur, hi1, lo1);
            Dh[buf][2 * k + 1][lane] = hi1;
            rec_step(rcp, l0 + 2 * k + 1, cb, dcur, dprev);
            Dlo[k][lane][h] = bf16x2_trunc(lo0, lo1);
        }
    };
    // lo products of a 16-degree flush block: bfloat16 slot 2 h + s holds K-step s of
    // the block's tile h (degree l0 + 8 h + 2 k + s in lane row k), for A and B alike
    uint32_t blr[2], bli[2];
    // one 8-degree tile (h = 0 / 1: first / second of its flush block): 2 K-steps x 4
    // beta tiles x (re, im) FP32 MFMAs of the hi parts with the ring slot (ra, rb)
    // (loaded two tiles earlier, reloaded here for tile l0 + 16); the second tile adds
    // the block's lo products, 8 bfloat16 MFMAs; interleaved with the recurrence that
    // fills the other LDS tile for the next one
    auto tile = [&](auto hc, int l0, float2 &ra, float2 &rb) {
        constexpr int h = decltype(hc)::value;
        const int buf = (l0 >> 3) & 1;
        float a[2][4];
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
            for (int g = 0; g < 4; ++g) a[s2][g] = Dh[buf][2 * kk + s2][16 * g + j];
        // the block's lo parts, read before this tile's dtile rewrites slots 0, 1 (the
        // LDS operations of a wave execute in order)
        bf4_t alo[4];
        if (h == 1) {
#pragma unroll
            for (int g = 0; g < 4; ++g)
                alo[g] = __builtin_bit_cast(bf4_t, *reinterpret_cast<const uint2 *>(&Dlo[kk][16 * g + j][0]));
        }
        const int lk = l0 + 2 * kk;
        const float z0 = lk < L ? sg0 : 0.0f, z1 = lk + 1 < L ? sg1 : 0.0f;
        const float br0 = z0 * ra.x, bi0 = z0 * ra.y, br1 = z1 * rb.x, bi1 = z1 * rb.y;
        loadb(l0 + 16, ra, rb);
        __builtin_amdgcn_sched_barrier(0);   // the reload is issued here, two tiles ahead
        blr[h] = bf16x2_trunc(br0, br1);
        bli[h] = bf16x2_trunc(bi0, bi1);
        if (h == 1) {
            const bf4_t lr = bf4_of(blr[0], blr[1]), li = bf4_of(bli[0], bli[1]);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                ar[g] = mma_bf16(alo[g], lr, ar[g]);
                ai[g] = mma_bf16(alo[g], li, ai[g]);
            }
        }
        // per K-step the hi products: eight independent accumulators between dependent
        // MFMAs; the first K-step of a flush block starts them from zero (an inline
        // constant operand, no register clearing)
        constexpr f4_t zero4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            ar[g] = mma32(a[0][g], br0, h == 0 ? zero4 : ar[g]);
            ai[g] = mma32(a[0][g], bi0, h == 0 ? zero4 : ai[g]);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            ar[g] = mma32(a[1][g], br1, ar[g]);
            ai[g] = mma32(a[1][g], bi1, ai[g]);
        }
        dtile(std::integral_constant<int, 1 - h>{}, buf ^ 1, l0 + INV_TD);
        // issue order: each MFMA followed by a few VALU instructions of the recurrence,
        // which run while the matrix core works
#pragma unroll
        for (int i = 0; i < (h ? 24 : 16); ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, h ? 3 : 4, 0);
        }
    };
    // the lo products of a flush block without a second tile (its B slots 2, 3 zero)
    auto lo_only = [&]() {
        const bf4_t lr = bf4_of(blr[0], 0u), li = bf4_of(bli[0], 0u);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const bf4_t alo = __builtin_bit_cast(bf4_t, *reinterpret_cast<const uint2 *>(&Dlo[kk][16 * g + j][0]));
            ar[g] = mma_bf16(alo, lr, ar[g]);
            ai[g] = mma_bf16(alo, li, ai[g]);
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    // FP32 sums of (at most) 16 degrees into the FP64 totals (the next block's first
    // MFMAs restart the FP32 sums)
    auto flush = [&]() {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                dr[g][r] += (double)ar[g][r];
                di[g][r] += (double)ai[g][r];
            }
        }
    };

    float2 r0a, r0b, r1a, r1b;
    loadb(0, r0a, r0b);
    loadb(8, r1a, r1b);
    dtile(std::integral_constant<int, 0>{}, 0, 0);
    // LDS operations of one wave execute in order, so a tile written during one step is
    // read by the next with no barrier; two ring slots, one per tile of a 16-degree flush.
    // No exit between a reload and the tiles after it: a load whose value one exit path
    // does not need would be moved past the exit test, next to its use.
    int l0 = 0;
    for (; l0 + 16 < L; l0 += 16) {
        tile(I0{}, l0, r0a, r0b);
        tile(I1{}, l0 + 8, r1a, r1b);
        flush();
    }
    tile(I0{}, l0, r0a, r0b);
    if (l0 + 8 < L) tile(I1{}, l0 + 8, r1a, r1b);
    else lo_only();
    flush();

    // item r of lane l: beta row 4(l>>4) + r of tile g, column j = l & 15: betas
    // b0 + 16 g + 4 k .. + 3 of one column, stored as two 16-byte writes (reversed in
    // memory for the variants at pi - beta)
    if (!vr.valid || tt >= nt) return;
    const int64_t colbase = ((int64_t)wrap_m(vr.n, N) * N + wrap_m(vr.m, N)) * N;
    float2 *Ft = F + (int64_t)(t0 + tt) * f_ts + colbase;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int bt = b0 + 16 * g + 4 * kk;
        float2 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = make_float2((float)dr[g][r], (float)di[g][r]);
        if (bt + 3 < N) {
            const int m0 = vr.rev ? N - 4 - bt : bt;
            const float4 w0 = vr.rev ? make_float4(v[3].x, v[3].y, v[2].x, v[2].y)
                                     : make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
            const float4 w1 = vr.rev ? make_float4(v[1].x, v[1].y, v[0].x, v[0].y)
                                     : make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
            // written once, read by the next pass from HBM: non-temporal stores
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v x0 = {w0.x, w0.y, w0.z, w0.w}, x1 = {w1.x, w1.y, w1.z, w1.w};
            __builtin_nontemporal_store(x0, reinterpret_cast<f4v *>(Ft + m0));
            __builtin_nontemporal_store(x1, reinterpret_cast<f4v *>(Ft + m0 + 2));
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (bt + r < N) Ft[vr.rev ? N - 1 - bt - r : bt + r] = v[r];
        }
    }
}

template <int TB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void dwt_inverse_mix(
    const float2 *__restrict__ coeff, int64_t c_ts, float2 *__restrict__ F, int64_t f_ts, int T,
    const double *__restrict__ cosb, const double *__restrict__ seeds, const double *__restrict__ rc,
    const int64_t *__restrict__ rc_off, const int64_t *__restrict__ slice, const int32_t *__restrict__ pairs, int bw,
    int N, int npair, int nch, int nunit)
{
    extern __shared__ double smx[];   // hi tiles | lo table
    const InvHi Dh = reinterpret_cast<InvHi>(smx);
    const InvLo Dlo = reinterpret_cast<InvLo>(smx + INV_TD * INV_DS);
    const int id = blockIdx.x;
    const int chunk = (id >> 3) % nch;
    const int unit = ((id >> 3) / nch) * 8 + (id & 7);  // chunks of a unit share an XCD
    if (unit >= nunit) return;
    const int pq = pairs[unit % npair];
    inv_mix_item<TB>(coeff, c_ts, F, f_ts, T, cosb, seeds, rc, rc_off, slice, bw, N, pq & 0xffff, pq >> 16,
                     (unit / npair) * TB, chunk, Dh, Dlo);
}

// grouped launch over transforms of several bandwidths: a persistent grid walks the
// work list (item = segment, p | q << 16, first transform, chunk; segment -1 = padding)
template <int TB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void dwt_inverse_mix_grouped(
    const float2 *__restrict__ coeff, float2 *__restrict__ F, const MixSegment *__restrict__ segs,
    const int4 *__restrict__ items, int nitem)
{
    extern __shared__ double smx[];
    const InvHi Dh = reinterpret_cast<InvHi>(smx);
    const InvLo Dlo = reinterpret_cast<InvLo>(smx + INV_TD * INV_DS);
    for (int it = blockIdx.x; it < nitem; it += gridDim.x) {
        const int4 w = items[it];
        if (w.x < 0) continue;
        const MixSegment sg = segs[w.x];
        __syncthreads();  // the previous item's LDS reads are done
        inv_mix_item<TB>(coeff + sg.coff, sg.nc, F + sg.goff, (int64_t)sg.N * sg.N * sg.N, sg.T, sg.cosb, sg.seeds,
                         sg.rc, sg.rc_off, sg.slice, sg.bw, sg.N, w.y & 0xffff, w.y >> 16, w.z, w.w, Dh, Dlo);
    }
}

// ---------------------------------------------------------------------------
// forward, one wave per (pair, TB transforms): v_mfma_f32_16x16x4f32 with A = w(beta) d
// tiles [degree][beta] in LDS (M = degree, K = beta) and B = the fft-array samples (K =
// beta, N = 8 variants x 2 transforms), one MFMA for the real and one for the imaginary
// parts.  FP32 accumulation over a few K-steps, FP64 across them and across chunks: a
// tile's coefficients are complete after its last chunk, so there is no cross-wave
// reduction and no degree is recomputed.
//
// B loads: K-step s of lane (column j, k) takes beta 8 (s >> 1) + 2 k + (s & 1) of the
// chunk, so one 16-byte load gives a lane both samples (re, im, re, im) of two K-steps
// and the four k lanes of a column read 64 contiguous bytes: 8 loads per chunk and
// transform pair.  The columns are sample-contiguous in beta (reversed for the variants
// at pi - beta), so the fft array must be 16-byte aligned with an even transform stride.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fwd_stage_sampling(double2 (*rs)[64], const double *cosb, const double *w, int N);
constexpr int FT32_DS = 65;   // padded row of the 32-degree hi tile (floats)
constexpr int FT32_TD = 32;   // degrees per tile

template <int TB>
__device__ __forceinline__ void fwd_mix_item32(const float2 *__restrict__ F, int64_t f_ts,
                                               float2 *__restrict__ coeff, int64_t c_ts, int T,
                                               const double *__restrict__ seeds, const double *__restrict__ rc,
                                               const int64_t *__restrict__ rc_off,
                                               const int64_t *__restrict__ slice, int bw, int N, int p, int q,
                                               int t0, float (*Dt)[FT32_DS], double2 (*rs)[64])
{
    constexpr int TG = (TB + 1) / 2;   // transform pairs: column j = variant j & 7 of transform 2 g + (j >> 3)
    constexpr int KS = 2;              // accumulators per (pair, part, M-tile): K-step s uses s % KS
    const int lane = threadIdx.x & 63;
    const int L = bw - q;
    const int nc = (N + 63) >> 6;
    const int nt = min(TB, T - t0);
    const int j = lane & 15, kk = lane >> 4;
    const int vj = j & 7, tt = j >> 3;
    const Variant vr = variant(vj, p, q);
    const int64_t colj = ((int64_t)wrap_m(vr.n, N) * N + wrap_m(vr.m, N)) * N;
    const float2 *src[TG];
#pragma unroll
    for (int g = 0; g < TG; ++g)
        src[g] = F + (int64_t)max(0, min(t0 + 2 * g + tt, T - 1)) * f_ts + colj;

    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    const const_dptr rcp = (const_dptr)rc + 3 * ((const_i64ptr)rc_off)[tri];
    for (int c = 0; c < nc; ++c) {
        const int beta = 64 * c + lane;
        rs[c][lane] = make_double2(beta < N ? ((gptr<double>)seeds)[tri * N + beta] : 0.0, 0.0);
    }
    __syncthreads();   // rs (one wave: orders the LDS writes before the reads)

    // B operands of K-steps 2h, 2h+1 of chunk c for transform pair g, as in the inverse
    float4 raw[8][TG];
    auto load_b = [&](int c, int h) {
        const int b0 = min(64 * c + 8 * h + 2 * kk, N - 2);
        const int m0 = vr.rev ? N - 2 - b0 : b0;
#pragma unroll
        for (int g = 0; g < TG; ++g) raw[h][g] = *reinterpret_cast<const float4 *>(src[g] + m0);
    };
#pragma unroll
    for (int h = 0; h < 8; ++h) load_b(0, h);

    // one degree tile of MT 16-degree M-tiles starting at degree l0 (chain states in rs)
    auto tile = [&](int l0, auto mtc) {
        constexpr int MT = decltype(mtc)::value;
        double accr[TG][2][4], acci[TG][2][4];
#pragma unroll
        for (int g = 0; g < TG; ++g)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) accr[g][mt][r] = acci[g][mt][r] = 0.0;
#pragma unroll 1
        for (int c = 0; c < nc; ++c) {
            const int cn = c + 1 == nc ? 0 : c + 1;   // chunk whose operands the K-steps reload
            const double2 cw = rs[nc + c][lane];
            const double cb = cw.x, wt = cw.y;
            asm volatile("" ::: "memory");
            double2 st = rs[c][lane];
            double dcur = st.x, dprev = st.y;
#pragma unroll
            for (int k = 0; k < 16 * MT; ++k) {   // the chain is zero past the last degree
                Dt[k][lane] = (float)(wt * dcur);
                rec_step(rcp, l0 + k, cb, dcur, dprev);
            }
            rs[c][lane] = make_double2(dcur, dprev);
            asm volatile("" ::: "memory");
            f4_t ar[TG][2][KS], ai[TG][2][KS];
            // FP32 sums of at most 4 K-steps (16 products) per accumulator, then FP64: the
            // accumulators are flushed after each half of the chunk's K-steps
#pragma unroll
            for (int s4 = 0; s4 < 16; ++s4) {
                // the first K-step of each accumulator in a half starts it from zero (an
                // inline constant operand, no register clearing)
                const bool first = (s4 & 7) < KS;
                constexpr f4_t zero4 = {0.f, 0.f, 0.f, 0.f};
                const int col = 8 * (s4 >> 1) + 2 * kk + (s4 & 1);   // beta of this K-step in the chunk
                const float a0 = Dt[j][col], a1 = MT > 1 ? Dt[16 + j][col] : 0.0f;
                const int u = s4 % KS;
                const bool second = (s4 & 1) != vr.rev;   // this K-step's sample is the load's second
#pragma unroll
                for (int g = 0; g < TG; ++g) {
                    const float4 v = raw[s4 >> 1][g];
                    const float bre = second ? v.z : v.x, bim = second ? v.w : v.y;
                    ar[g][0][u] = mma32(a0, bre, first ? zero4 : ar[g][0][u]);
                    ai[g][0][u] = mma32(a0, bim, first ? zero4 : ai[g][0][u]);
                    if (MT > 1) {
                        ar[g][1][u] = mma32(a1, bre, first ? zero4 : ar[g][1][u]);
                        ai[g][1][u] = mma32(a1, bim, first ? zero4 : ai[g][1][u]);
                    }
                }
                if (s4 & 1) {
                    load_b(cn, s4 >> 1);
                    __builtin_amdgcn_sched_barrier(0);   // issued here, not after the last K-step
                }
                if ((s4 & 7) == 7) {
#pragma unroll
                    for (int g = 0; g < TG; ++g)
#pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                            for (int u = 0; u < KS; ++u)
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    accr[g][mt][r] += (double)ar[g][mt][u][r];
                                    acci[g][mt][r] += (double)ai[g][mt][u][r];
                                }
                }
            }
        }
        // item r of lane l in M-tile mt: degree row 16 mt + 4(l>>4) + r, column j = l & 15
        if (vr.valid) {
            const int64_t o = slice_m(slice, vr.m, vr.n, bw);
#pragma unroll
            for (int g = 0; g < TG; ++g) {
                const int t = 2 * g + tt;
                if (t >= nt) continue;
                float2 *ct = coeff + (int64_t)(t0 + t) * c_ts + o;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int li = l0 + 16 * mt + 4 * kk + r;
                        if (li < L) {
                            const double sg = vr.sign * ((vr.lsign && ((q + li) & 1)) ? -1.0 : 1.0);
                            ct[li] = make_float2((float)(sg * accr[g][mt][r]), (float)(sg * acci[g][mt][r]));
                        }
                    }
            }
        }
    };
    // 32-degree tiles; a pair whose degree count is 1 to 16 past a multiple of 32 takes a
    // 16-degree tile first, so that no tile computes more than 15 degrees past the last (the
    // padding of whole tiles cost 1.39x the useful degrees at bw = 128, this 1.18x).  The
    // M-tiles of a degree give the same sums in either grouping.
    int l0 = 0;
    if ((L & 31) != 0 && (L & 31) <= 16) {
        tile(0, std::integral_constant<int, 1>{});
        l0 = 16;
    }
#pragma unroll 1
    for (; l0 < L; l0 += FT32_TD) tile(l0, std::integral_constant<int, 2>{});
}

template <int TB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void dwt_forward_mix32(
    const float2 *__restrict__ F, int64_t f_ts, float2 *__restrict__ coeff, int64_t c_ts, int T,
    const double *__restrict__ w, const double *__restrict__ cosb, const double *__restrict__ seeds,
    const double *__restrict__ rc, const int64_t *__restrict__ rc_off, const int64_t *__restrict__ slice,
    const int32_t *__restrict__ pairs, int bw, int N, int npair, int nunit)
{
    // dynamic LDS: Dt [32][FT32_DS] float | rs [2 nc][64] double2
    extern __shared__ double smx[];
    float (*Dt)[FT32_DS] = reinterpret_cast<float (*)[FT32_DS]>(smx);
    double2 (*rs)[64] = reinterpret_cast<double2 (*)[64]>(smx + (FT32_TD * FT32_DS + 1) / 2);
    const int unit = blockIdx.x;
    if (unit >= nunit) return;
    const int pq = pairs[unit % npair];
    fwd_stage_sampling(rs, cosb, w, N);
    fwd_mix_item32<TB>(F, f_ts, coeff, c_ts, T, seeds, rc, rc_off, slice, bw, N, pq & 0xffff, pq >> 16,
                       (unit / npair) * TB, Dt, rs);
}

// grouped forward over transforms of several bandwidths (see dwt_inverse_mix_grouped);
// dynamic LDS sized for the largest bandwidth of the group
template <int TB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void dwt_forward_mix32_grouped(
    const float2 *__restrict__ F, float2 *__restrict__ coeff, const MixSegment *__restrict__ segs,
    const int4 *__restrict__ items, int nitem)
{
    extern __shared__ double smx[];
    float (*Dt)[FT32_DS] = reinterpret_cast<float (*)[FT32_DS]>(smx);
    double2 (*rs)[64] = reinterpret_cast<double2 (*)[64]>(smx + (FT32_TD * FT32_DS + 1) / 2);
    for (int it = blockIdx.x; it < nitem; it += gridDim.x) {
        const int4 wk = items[it];
        if (wk.x < 0) continue;
        const MixSegment sg = segs[wk.x];
        __syncthreads();  // the previous item's LDS reads are done
        fwd_stage_sampling(rs, sg.cosb, sg.w, sg.N);
        fwd_mix_item32<TB>(F + sg.goff, (int64_t)sg.N * sg.N * sg.N, coeff + sg.coff, sg.nc, sg.T, sg.seeds,
                           sg.rc, sg.rc_off, sg.slice, sg.bw, sg.N, wk.y & 0xffff, wk.y >> 16, wk.z, Dt, rs);
    }
}

// (cos beta, w(beta)) rows of the chain state table
__device__ __forceinline__ void fwd_stage_sampling(double2 (*rs)[64], const double *cosb, const double *w, int N)
{
    const int lane = threadIdx.x & 63, nc = (N + 63) >> 6;
    for (int c = 0; c < nc; ++c) {
        const int beta = 64 * c + lane;
        rs[nc + c][lane] = beta < N ? make_double2(((gptr<double>)cosb)[beta], ((gptr<double>)w)[beta])
                                    : make_double2(0.0, 0.0);
    }
}

// ---------------------------------------------------------------------------
// FP64 forward DWT on v_mfma_f64_16x16x4f64 with the structure of the mixed forward:
// one wave per (pair, 2 transforms), lane = beta within 64-beta chunks, chain state in
// LDS between 16-degree tiles, A = w d (FP64) tile [degree][beta] in LDS, B = the
// complex128 samples (8 variants x 2 transforms; real and imaginary parts in separate
// MFMAs), FP64 accumulation in the MFMA accumulators across every chunk.  K-step s of
// lane row k takes beta 4 s + k of the chunk (the four rows of a column read 64
// contiguous bytes).  f64 16x16x4 result item r of lane l: row (l >> 4) + 4 r, column
// l & 15 (probed on gfx950).
// ---------------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t mma64(double a, double b, d4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int F64_DS = 66;   // padded row of the FP64 A tile (doubles)

template <int TB, bool TAB>
__device__ __forceinline__ void fwd_m64_item(const double2 *__restrict__ F, int64_t f_ts,
                                             double2 *__restrict__ coeff, int64_t c_ts, int T,
                                             const double *__restrict__ seeds, const double *__restrict__ rc,
                                             const int64_t *__restrict__ rc_off, const int64_t *__restrict__ slice,
                                             int bw, int N, int p, int q, int t0, int real, double (*Dt)[F64_DS],
                                             double2 (*rs)[64], const double *__restrict__ dtab)
{
    static_assert(TB == 1 || TB == 2, "one transform pair per wave");
    const int lane = threadIdx.x & 63;
    const int L = bw - q;
    const int nc = (N + 63) >> 6;
    const int nt = min(TB, T - t0);
    const int j = lane & 15, kk = lane >> 4;
    const int vj = j & 7, tt = j >> 3;
    const Variant vr = variant(vj, p, q);
    // real mode (rsoft, half-complex fft array): the odd variants read the conjugate of an
    // m >= 0 column (soft.f08:1650, 1662-1706)
    int mm = vr.m, nn = vr.n;
    if (real && (vj & 1)) {
        const Variant s2 = variant(vj == 1 ? (p == q ? 0 : 2) : (vj == 3 ? 0 : vj - 1), p, q);
        mm = s2.m;
        nn = s2.n;
    }
    const int64_t colj = ((int64_t)wrap_m(nn, N) * N + wrap_m(mm, N)) * N;
    const double2 *src = F + (int64_t)max(0, min(t0 + tt, T - 1)) * f_ts + colj;

    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    const const_dptr rcp = (const_dptr)rc + 3 * ((const_i64ptr)rc_off)[tri];
    for (int c = 0; c < nc; ++c) {
        const int beta = 64 * c + lane;
        rs[c][lane] = make_double2(beta < N ? ((gptr<double>)seeds)[tri * N + beta] : 0.0, 0.0);
    }
    __syncthreads();   // rs (one wave: orders the LDS writes before the reads)

    // B operand of K-step s of chunk c: the sample at beta 64 c + 4 s + k (clamped into
    // the column; rows past N meet zero A columns)
    double2 bv[16];
    auto load_b = [&](int c, int s4) {
        const int bt = min(64 * c + 4 * s4 + kk, N - 1);
        bv[s4] = src[vr.rev ? N - 1 - bt : bt];
    };
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) load_b(0, s4);

#pragma unroll 1
    for (int l0 = 0; l0 < L; l0 += 16) {
        d4_t are[2], aim[2];   // two independent chains per part
#pragma unroll
        for (int u = 0; u < 2; ++u) are[u] = aim[u] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int c = 0; c < nc; ++c) {
            const int cn = c + 1 == nc ? 0 : c + 1;   // chunk whose operands the K-steps reload
            const double2 cw = rs[nc + c][lane];
            const double cb = cw.x, wt = cw.y;
            asm volatile("" ::: "memory");
            if (TAB) {   // d~ from the table rows of this pair (zero past the last degree / beta)
                const int beta = 64 * c + lane;
                const double *tr = dtab + dtab_row(p, q, bw) * N + (beta < N ? beta : 0);
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    Dt[k][lane] = (l0 + k < L && beta < N) ? wt * tr[(int64_t)(l0 + k) * N] : 0.0;
            } else {
                double2 st = rs[c][lane];
                double dcur = st.x, dprev = st.y;
#pragma unroll
                for (int k = 0; k < 16; ++k) {   // the chain is zero past the last degree
                    Dt[k][lane] = wt * dcur;
                    rec_step(rcp, l0 + k, cb, dcur, dprev);
                }
                rs[c][lane] = make_double2(dcur, dprev);
            }
            asm volatile("" ::: "memory");
#pragma unroll
            for (int s4 = 0; s4 < 16; ++s4) {
                const double a = Dt[j][4 * s4 + kk];
                const int u = s4 & 1;
                are[u] = mma64(a, bv[s4].x, are[u]);
                aim[u] = mma64(a, bv[s4].y, aim[u]);
                load_b(cn, s4);   // consumed: reload for the next chunk (once more, unused, at the end)
                if ((s4 & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        // item r of lane l: degree row (l >> 4) + 4 r, column j = l & 15
        if (vr.valid && tt < nt) {
            const int64_t o = slice_m(slice, vr.m, vr.n, bw);
            double2 *ct = coeff + (int64_t)(t0 + tt) * c_ts + o;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int li = l0 + kk + 4 * r;
                if (li < L) {
                    const double sg = vr.sign * ((vr.lsign && ((q + li) & 1)) ? -1.0 : 1.0);
                    const double sy = (real && (vj & 1)) ? -sg : sg;
                    ct[li] = make_double2(sg * (are[0][r] + are[1][r]), sy * (aim[0][r] + aim[1][r]));
                }
            }
        }
    }
}

template <int TB, bool TAB>
__global__ __launch_bounds__(64) void dwt_forward_m64(const double2 *__restrict__ F, int64_t f_ts,
                                                      double2 *__restrict__ coeff, int64_t c_ts, int T,
                                                      const double *__restrict__ w, const double *__restrict__ cosb,
                                                      const double *__restrict__ seeds,
                                                      const double *__restrict__ rc,
                                                      const int64_t *__restrict__ rc_off,
                                                      const int64_t *__restrict__ slice,
                                                      const int32_t *__restrict__ pairs, int bw, int N, int npair,
                                                      int nunit, int real, const double *__restrict__ dtab)
{
    // dynamic LDS: Dt [16][F64_DS] double | rs [2 nc][64] double2 (chain states, then (cos beta, w))
    extern __shared__ double smx[];
    double (*Dt)[F64_DS] = reinterpret_cast<double (*)[F64_DS]>(smx);
    double2 (*rs)[64] = reinterpret_cast<double2 (*)[64]>(smx + 16 * F64_DS);
    const int unit = blockIdx.x;
    if (unit >= nunit) return;
    const int pq = pairs[unit % npair];
    fwd_stage_sampling(rs, cosb, w, N);
    fwd_m64_item<TB, TAB>(F, f_ts, coeff, c_ts, T, seeds, rc, rc_off, slice, bw, N, pq & 0xffff, pq >> 16,
                          (unit / npair) * TB, real, Dt, rs, dtab);
}

// FP64 forward, one block of nc = N / 64 waves per (pair, 2 transforms): wave c owns beta
// chunk c, its recurrence chains and its B operands (the complex128 samples of its chunk,
// 16 registers' worth of K-steps) stay in registers across all degree tiles, so every
// fft-array byte is read once (the single-wave kernel re-reads each column once per
// 16-degree tile and is bound by those bytes).  Per tile each wave writes its w d tile
// to its own LDS tile, runs its 16 K-steps, and leaves its chunk's partial sums in that
// tile; after a barrier the block adds the partials in chunk order and writes the
// tile's coefficients.
template <int TB, bool TAB>
__global__ __launch_bounds__(512) void dwt_forward_m64w(
    const double2 *__restrict__ F, int64_t f_ts, double2 *__restrict__ coeff, int64_t c_ts, int T,
    const double *__restrict__ w, const double *__restrict__ cosb, const double *__restrict__ seeds,
    const double *__restrict__ rc, const int64_t *__restrict__ rc_off, const int64_t *__restrict__ slice,
    const int32_t *__restrict__ pairs, int bw, int N, int npair, int nunit, int real, const double *__restrict__ dtab)
{
    static_assert(TB == 1 || TB == 2, "one transform pair per wave");
    constexpr int NV = 2 * 4;   // partials per lane: (re, im) x result item
    static_assert(NV * 64 <= 16 * F64_DS, "partials fit the wave's tile");
    // dynamic LDS: Dt [nc][16][F64_DS] double; after a tile's K-steps each wave's tile
    // holds its partials [NV][64] double
    extern __shared__ double smx[];
    const int unit = blockIdx.x;
    if (unit >= nunit) return;
    const int nc = blockDim.x >> 6;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double (*Dt)[F64_DS] = reinterpret_cast<double (*)[F64_DS]>(smx) + 16 * wv;
    auto red = [&](int c, int v) -> double * { return smx + (int64_t)c * 16 * F64_DS + v * 64; };
    // on the fly: pairs fastest.  TAB: the transform groups of one pair are consecutive
    // blocks of one XCD (block = 8 slot + xcd), so its table rows come from L2 after the
    // first read
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
    const int pq = pairs[pidx], p = pq & 0xffff, q = pq >> 16;
    const int t0 = grp * TB;
    const int L = bw - q;
    const int nt = min(TB, T - t0);
    const int j = lane & 15, kk = lane >> 4;
    const int vj = j & 7, tt = j >> 3;
    const Variant vr = variant(vj, p, q);
    int mm = vr.m, nn = vr.n;   // real mode: odd variants read the conjugate of an m >= 0 column
    if (real && (vj & 1)) {
        const Variant s2 = variant(vj == 1 ? (p == q ? 0 : 2) : (vj == 3 ? 0 : vj - 1), p, q);
        mm = s2.m;
        nn = s2.n;
    }
    const int64_t colj = ((int64_t)wrap_m(nn, N) * N + wrap_m(mm, N)) * N;
    const double2 *src = F + (int64_t)max(0, min(t0 + tt, T - 1)) * f_ts + colj;

    const int64_t tri = (int64_t)p * (2 * bw - p + 1) / 2 + (q - p);
    const const_dptr rcp = (const_dptr)rc + 3 * ((const_i64ptr)rc_off)[tri];
    const int beta = 64 * wv + lane;
    const bool ok = beta < N;
    const double cb = ok ? ((gptr<double>)cosb)[beta] : 0.0, wt = ok ? ((gptr<double>)w)[beta] : 0.0;
    double dcur = (!TAB && ok) ? ((gptr<double>)seeds)[tri * N + beta] : 0.0, dprev = 0.0;
    const double *trow = TAB ? dtab + dtab_row(p, q, bw) * N + (ok ? beta : 0) : nullptr;
    double2 bv[16];   // K-step s: the sample at beta 64 c + 4 s + k (clamped; zero A rows past N)
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) {
        const int bt = min(64 * wv + 4 * s4 + kk, N - 1);
        bv[s4] = src[vr.rev ? N - 1 - bt : bt];
    }
    // this thread's output column (lane l = threadIdx.x & 63 of the result layout)
    const int out_l = threadIdx.x & 63, out_j = out_l & 15, out_tt = out_j >> 3;
    const Variant vo = variant(out_j & 7, p, q);
    const int64_t out_off = vo.valid ? slice_m(slice, vo.m, vo.n, bw) : 0;
    const bool out_ok = vo.valid && out_tt < nt;

#pragma unroll 1
    for (int l0 = 0; l0 < L; l0 += 16) {
        if (TAB) {
#pragma unroll
            for (int k = 0; k < 16; ++k)
                Dt[k][lane] = (l0 + k < L && ok) ? wt * trow[(int64_t)(l0 + k) * N] : 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) {   // the chain is zero past the last degree
                Dt[k][lane] = wt * dcur;
                rec_step(rcp, l0 + k, cb, dcur, dprev);
            }
        }
        d4_t are[2], aim[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) are[u] = aim[u] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < 16; ++s4) {
            const double a = Dt[j][4 * s4 + kk];
            const int u = s4 & 1;
            are[u] = mma64(a, bv[s4].x, are[u]);
            aim[u] = mma64(a, bv[s4].y, aim[u]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            red(wv, r)[lane] = are[0][r] + are[1][r];
            red(wv, 4 + r)[lane] = aim[0][r] + aim[1][r];
        }
        __syncthreads();
        // item r of lane l: degree row (l >> 4) + 4 r, column l & 15; thread t of the block
        // takes lane t & 63 and rows r = t >> 6, t >> 6 + nc, ...
        if (out_ok) {
            for (int r = threadIdx.x >> 6; r < 4; r += nc) {
                const int li = l0 + (out_l >> 4) + 4 * r;
                if (li >= L) continue;
                double sr = 0.0, si = 0.0;
                for (int c = 0; c < nc; ++c) {
                    sr += red(c, r)[out_l];
                    si += red(c, 4 + r)[out_l];
                }
                const double sg = vo.sign * ((vo.lsign && ((q + li) & 1)) ? -1.0 : 1.0);
                const double sy = (real && (out_j & 1)) ? -sg : sg;
                coeff[(int64_t)(t0 + out_tt) * c_ts + out_off + li] = make_double2(sg * sr, sy * si);
            }
        }
        __syncthreads();   // the tiles (holding the partials) are rewritten by the next tile
    }
}

static int grid_m(int nunit, int nch) { return (nunit + 7) / 8 * 8 * nch; }

// dynamic LDS bytes of dwt_forward_mix32 (32-degree tile and chunk states)
static size_t fwd32_lds(int N)
{
    return sizeof(double) * ((FT32_TD * FT32_DS + 1) / 2 + 4 * 64 * ((N + 63) / 64));
}

// the 16-byte fft-array accesses of both directions
static bool fwd_aligned(const float2 *F, int64_t f_ts)
{
    return (reinterpret_cast<uintptr_t>(F) & 15) == 0 && (f_ts & 1) == 0;
}

hipError_t launch_dwt_inverse_mixed(const DeviceTables &tb, int bw, int N, int npair, int mode, const float2 *coeff,
                                    int64_t c_ts, const float2 *aux, float2 *F, int64_t f_ts, int T, hipStream_t s)
{
    if (mode != DWT_COMPLEX || aux || !fwd_aligned(F, f_ts)) return hipErrorInvalidValue;
    if (npair <= 0 || T <= 0) return hipSuccess;
    const int nch = (N + 63) / 64;
#define SO3_MINV(TBV)                                                                                          \
    do {                                                                                                       \
        const int nunit = npair * ((T + TBV - 1) / TBV);                                                       \
        hipLaunchKernelGGL((dwt_inverse_mix<TBV>), dim3(grid_m(nunit, nch)), dim3(64), inv_lds(bw), s, coeff,    \
                           c_ts, F, f_ts, T, tb.cosb, tb.seeds, tb.rc, tb.rc_off, tb.slice, tb.pairs, bw, N,     \
                           npair, nch, nunit);                                                                  \
    } while (0)
    if (T >= 2) SO3_MINV(2);  // one transform pair per wave
    else SO3_MINV(1);
#undef SO3_MINV
    return hipGetLastError();
}

hipError_t launch_dwt_forward_mixed(const DeviceTables &tb, int bw, int N, int npair, int mode, const float2 *F,
                                    int64_t f_ts, float2 *coeff, int64_t c_ts, int T, hipStream_t s)
{
    if (mode != DWT_COMPLEX || !fwd_aligned(F, f_ts)) return hipErrorInvalidValue;
    if (npair <= 0 || T <= 0) return hipSuccess;
#define SO3_MFWD(TBV)                                                                                          \
    do {                                                                                                       \
        const int nunit = npair * ((T + TBV - 1) / TBV);                                                       \
        hipLaunchKernelGGL((dwt_forward_mix32<TBV>), dim3(nunit), dim3(64), fwd32_lds(N), s, F, f_ts, coeff, c_ts, \
                           T, tb.w, tb.cosb, tb.seeds, tb.rc, tb.rc_off, tb.slice, tb.pairs, bw, N, npair, nunit);  \
    } while (0)
    // four transforms per wave from bw = 64 (more MFMA work per recurrence step; measured
    // 8% faster than two at bw = 64, batch 16), else two
    if (bw >= 64 && T >= 4) SO3_MFWD(4);
    else if (T >= 2) SO3_MFWD(2);
    else SO3_MFWD(1);
#undef SO3_MFWD
    return hipGetLastError();
}

// FP64 forward on FP64 MFMA (modes DWT_COMPLEX and DWT_REAL); with the plan's d table
// (tb.dtab) the A tiles are read from it instead of the recurrence
template <int TB, bool TAB>
static void launch_m64(const DeviceTables &tb, int bw, int N, int npair, const double2 *F, int64_t f_ts,
                       double2 *coeff, int64_t c_ts, int T, int nunit, int real, hipStream_t s)
{
    const int nc = (N + 63) / 64;
    if (nc >= 2 && nc <= 8) {   // one wave per 64-beta chunk: fft-array samples read once
        const size_t ldsw = sizeof(double) * (size_t)nc * 16 * F64_DS;
        // table order: 8 pairs x transform groups per run of slots (see the kernel)
        const int nb = TAB ? (npair + 7) / 8 * 8 * ((T + TB - 1) / TB) : nunit;
        hipLaunchKernelGGL((dwt_forward_m64w<TB, TAB>), dim3(nb), dim3(64 * nc), ldsw, s, F, f_ts, coeff, c_ts, T,
                           tb.w, tb.cosb, tb.seeds, tb.rc, tb.rc_off, tb.slice, tb.pairs, bw, N, npair, nb, real,
                           tb.dtab);
        return;
    }
    const size_t lds = sizeof(double) * (16 * F64_DS + 4 * 64 * ((N + 63) / 64));
    hipLaunchKernelGGL((dwt_forward_m64<TB, TAB>), dim3(nunit), dim3(64), lds, s, F, f_ts, coeff, c_ts, T, tb.w,
                       tb.cosb, tb.seeds, tb.rc, tb.rc_off, tb.slice, tb.pairs, bw, N, npair, nunit, real, tb.dtab);
}

hipError_t launch_dwt_forward_m64(const DeviceTables &tb, int bw, int N, int npair, int mode, const double2 *F,
                                  int64_t f_ts, double2 *coeff, int64_t c_ts, int T, hipStream_t s)
{
    if (mode != DWT_COMPLEX && mode != DWT_REAL) return hipErrorInvalidValue;
    const int real = mode == DWT_REAL;
    if (npair <= 0 || T <= 0) return hipSuccess;
    const int TB = T >= 2 ? 2 : 1;
    const int nunit = npair * ((T + TB - 1) / TB);
    if (tb.dtab) {
        if (TB == 2) launch_m64<2, true>(tb, bw, N, npair, F, f_ts, coeff, c_ts, T, nunit, real, s);
        else launch_m64<1, true>(tb, bw, N, npair, F, f_ts, coeff, c_ts, T, nunit, real, s);
    } else {
        if (TB == 2) launch_m64<2, false>(tb, bw, N, npair, F, f_ts, coeff, c_ts, T, nunit, real, s);
        else launch_m64<1, false>(tb, bw, N, npair, F, f_ts, coeff, c_ts, T, nunit, real, s);
    }
    return hipGetLastError();
}

// ---- grouped launches: TB (transforms per wave) and grid from the tuned configuration
hipError_t launch_dwt_inverse_mixed_grouped(int TB, const float2 *coeff, float2 *F, const MixSegment *segs,
                                            const int4 *items, int nitem, int grid, int max_bw, hipStream_t s)
{
    if (nitem <= 0) return hipSuccess;
    if (grid <= 0 || !fwd_aligned(F, 0)) return hipErrorInvalidValue;
    const size_t lds = inv_lds(max_bw);
    if (TB == 2)
        hipLaunchKernelGGL((dwt_inverse_mix_grouped<2>), dim3(grid), dim3(64), lds, s, coeff, F, segs, items, nitem);
    else if (TB == 1)
        hipLaunchKernelGGL((dwt_inverse_mix_grouped<1>), dim3(grid), dim3(64), lds, s, coeff, F, segs, items, nitem);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_dwt_forward_mixed_grouped(int TB, const float2 *F, float2 *coeff, const MixSegment *segs,
                                            const int4 *items, int nitem, int grid, int max_bw, hipStream_t s)
{
    if (nitem <= 0) return hipSuccess;
    if (grid <= 0 || !fwd_aligned(F, 0)) return hipErrorInvalidValue;
    const size_t lds = fwd32_lds(2 * max_bw);
    if (TB == 4)
        hipLaunchKernelGGL((dwt_forward_mix32_grouped<4>), dim3(grid), dim3(64), lds, s, F, coeff, segs, items, nitem);
    else if (TB == 2)
        hipLaunchKernelGGL((dwt_forward_mix32_grouped<2>), dim3(grid), dim3(64), lds, s, F, coeff, segs, items, nitem);
    else if (TB == 1)
        hipLaunchKernelGGL((dwt_forward_mix32_grouped<1>), dim3(grid), dim3(64), lds, s, F, coeff, segs, items, nitem);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// resident blocks per CU of a grouped kernel (LDS for bandwidth max_bw)
int mixed_grouped_occupancy(int inverse, int TB, int max_bw)
{
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (inverse) {
        const size_t lds = inv_lds(max_bw);
        if (TB == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_inverse_mix_grouped<2>, 64, lds);
        else if (TB == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_inverse_mix_grouped<1>, 64, lds);
    } else {
        const size_t lds = fwd32_lds(2 * max_bw);
        if (TB == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_forward_mix32_grouped<4>, 64, lds);
        else if (TB == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_forward_mix32_grouped<2>, 64, lds);
        else if (TB == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dwt_forward_mix32_grouped<1>, 64, lds);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

}  // namespace so3
