// Register-resident two-step FFT pass for line lengths N = R1 * R2, each factor 2^a or
// 3 * 2^a (every N = 2 bw with bw = 2^a or 3 * 2^a up to 256).
//
// For one line family (rows at element stride in_r / out_r, beta columns contiguous)
// a workgroup owns C beta columns.  Step 1: thread (n2, c) loads the R1 elements
// x[R2*n1 + n2], n1 < R1, straight into registers, runs a length-R1 DFT, multiplies
// by W_N^{n2 k1} and writes y[k1][n2] to LDS.  Step 2 (one barrier later): thread
// (k1, c) reads y[k1][n2], n2 < R2, runs a length-R2 DFT and stores
// X[k1 + R1*k2] directly from registers.  All arithmetic is FP64; the storage type
// T is double2 (complex128) or float2 (complex64) so the same kernel serves the FP64
// path and the mixed-precision path (FP64 butterflies on complex64 data).
#include "so3ft_internal.h"

#include <algorithm>

namespace so3 {

struct FftRegK {
    int N, C, bw, lmax;
    int64_t in_t, in_o, in_r, out_t, out_o, out_r;
    int mask_rows, mask_outer, in_real, out_real, sign;
    int col0, ncols;   // beta column window [col0, col0+ncols) (slab processing)
    double scale;
    double *pv;        // peak-search partials of the fused pass (fft_reg_lean<..., true>)
    int64_t *pi;
};

__device__ __forceinline__ bool order_ok2(int i, int N, int bw, int lmax)
{
    if (i == bw) return false;
    const int m = (i < bw) ? i : i - N;
    return (m < 0 ? -m : m) <= lmax;
}

__device__ __forceinline__ double2 cmul2(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// W_R^k = exp(sign 2 pi i k / R): compile-time constants when R divides 16 (the DFTs'
// inner twiddles; fewer registers than table loads), else the sign-adjusted table
__device__ __forceinline__ double c16(int j)   // cos(2 pi j / 16), 0 <= j <= 4
{
    return j == 0 ? 1.0 : j == 1 ? 0.9238795325112867 : j == 2 ? 0.7071067811865476 : j == 3 ? 0.3826834323650898 : 0.0;
}
template <int R>
__device__ __forceinline__ double2 twc(int k, const double2 *tw, int NR, int sign)
{
    if constexpr (16 % R == 0) {
        const int j = k * (16 / R);   // W_16^j, 0 < j < 8
        const double c = j <= 4 ? c16(j) : -c16(8 - j);
        const double sn = j <= 4 ? c16(4 - j) : c16(j - 4);
        return make_double2(c, sign < 0 ? -sn : sn);
    } else {
        return tw[k * NR];
    }
}

// In-register DFT of length R = 2^a or 3 * 2^a: recursive radix-2 decimation in time
// down to a radix-3 (or trivial) base; W_R^j = tw[j * (N/R)] with tw the
// sign-adjusted length-N table in LDS.
template <int R>
struct Dft {
    __device__ __forceinline__ static void run(double2 *v, const double2 *tw, int NR, int sign)
    {
        if constexpr (R == 3) {
            // X_k = sum_j x_j W_3^{jk};  W_3 = tw[NR], W_3^2 = tw[2 NR]
            const double2 w1 = tw[NR], w2 = tw[2 * NR];
            const double2 a = v[0], b = v[1], c = v[2];
            const double2 b1 = cmul2(b, w1), c1 = cmul2(c, w2), b2 = cmul2(b, w2), c2 = cmul2(c, w1);
            v[0] = make_double2(a.x + b.x + c.x, a.y + b.y + c.y);
            v[1] = make_double2(a.x + b1.x + c1.x, a.y + b1.y + c1.y);
            v[2] = make_double2(a.x + b2.x + c2.x, a.y + b2.y + c2.y);
        } else {
            static_assert(R % 2 == 0, "length must be 2^a or 3*2^a");
            double2 e[R / 2], o[R / 2];
#pragma unroll
            for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
            Dft<R / 2>::run(e, tw, NR * 2, sign);
            Dft<R / 2>::run(o, tw, NR * 2, sign);
#pragma unroll
            for (int k = 0; k < R / 2; ++k) {
                double2 t;
                if (k == 0) t = o[0];
                else if (4 * k == R) t = (sign < 0) ? make_double2(o[k].y, -o[k].x) : make_double2(-o[k].y, o[k].x);
                else t = cmul2(o[k], twc<R>(k, tw, NR, sign));
                v[k] = make_double2(e[k].x + t.x, e[k].y + t.y);
                v[k + R / 2] = make_double2(e[k].x - t.x, e[k].y - t.y);
            }
        }
    }
};
template <>
struct Dft<1> {
    __device__ __forceinline__ static void run(double2 *, const double2 *, int, int) {}
};

template <class T> struct Scalar;
template <> struct Scalar<double2> { typedef double type; };
template <> struct Scalar<float2> { typedef float type; };

template <class T> __device__ __forceinline__ double2 ld(const T *p, int64_t off);
template <> __device__ __forceinline__ double2 ld<double2>(const double2 *p, int64_t off) { return p[off]; }
template <> __device__ __forceinline__ double2 ld<float2>(const float2 *p, int64_t off)
{
    const float2 v = p[off];
    return make_double2((double)v.x, (double)v.y);
}
template <class T> __device__ __forceinline__ void st(T *p, int64_t off, double2 v);
// the passes read and write every element once: non-temporal accesses (streamed past the
// caches; measured 5% faster on complex64 passes)
template <> __device__ __forceinline__ void st<double2>(double2 *p, int64_t off, double2 v)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<d2v *>(p + off));
}
template <> __device__ __forceinline__ void st<float2>(float2 *p, int64_t off, double2 v)
{
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v t = {(float)v.x, (float)v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<f2v *>(p + off));
}
__device__ __forceinline__ float2 ld_nt(const float2 *p)
{
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v t = __builtin_nontemporal_load(reinterpret_cast<const f2v *>(p));
    return make_float2(t.x, t.y);
}
__device__ __forceinline__ double2 ld_nt(const double2 *p)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v t = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
    return make_double2(t.x, t.y);
}

template <class T> __device__ __forceinline__ T zero_of();
template <> __device__ __forceinline__ double2 zero_of<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zero_of<float2>() { return make_float2(0.f, 0.f); }
__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return make_double2((double)v.x, (double)v.y); }

// TI: input element type (double2/float2; in_real reads the .x-sized scalar instead)
// TO: output element type
//
// Complex64 input (TI = float2): a block walks the column tiles
// cb = col0 + (blockIdx.x + j * gridDim.x) * C and loads the rows of tile j+1, in their
// 8-byte storage type, before transforming tile j.  The pass is latency-bound: the LDS
// exchange caps a CU at two blocks, so without the prefetch a complex64 pass has half the
// bytes in flight of a complex128 one.  Complex128 input keeps one tile per block (a
// second FP64 row set would not fit the register budget).
template <int R1, int R2, class TI, class TO>
__global__ __launch_bounds__(256) void fft_reg_kernel(const void *__restrict__ in, void *__restrict__ out,
                                                      const double2 *__restrict__ twg, FftRegK k)
{
    constexpr int N = R1 * R2;
    constexpr int RM = R1 > R2 ? R1 : R2;
    constexpr bool PF = sizeof(TI) == 8;
    extern __shared__ double2 sm[];
    double2 *tw = sm;           // N
    double2 *Y = sm + N;        // [k1][n2][c]  N * C
    const int C = k.C, tid = threadIdx.x;
    const int c = tid % C, grp = tid / C;
    const int64_t o = blockIdx.y, t = blockIdx.z;
    if (k.mask_outer && !order_ok2((int)o, N, k.bw, k.lmax)) return;
    const int ntile = (k.ncols + C - 1) / C;

    for (int i = tid; i < N; i += blockDim.x) {
        double2 w = twg[i];
        if (k.sign > 0) w.y = -w.y;
        tw[i] = w;
    }
    // step-1 rows of tile jt for thread (n2 = grp, c): rows R2*n1 + n2, masked rows zero.
    // Row n1 sits at element n1 * (R2 in_r) past the thread's first row: 32-bit offsets
    // from one per-tile pointer (in_r N < 2^31 for N <= 512), the row mask computed once.
    uint32_t rowmask = 0;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
        const int r = R2 * n1 + grp;
        if (grp < R2 && (!k.mask_rows || order_ok2(r, N, k.bw, k.lmax))) rowmask |= 1u << n1;
    }
    const uint32_t rstep = (uint32_t)(R2 * k.in_r);
    auto load = [&](TI (&dst)[R1], int jt) {
        const int cb = k.col0 + jt * C;
        const bool colok = jt < ntile && (cb + c) < (k.col0 + k.ncols);
        const int64_t ib = t * k.in_t + o * k.in_o + cb + c + (int64_t)grp * k.in_r;
        const TI *pb = reinterpret_cast<const TI *>(in) + ib;
        const typename Scalar<TI>::type *ps = reinterpret_cast<const typename Scalar<TI>::type *>(in) + ib;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
            TI x = zero_of<TI>();
            if (colok && ((rowmask >> n1) & 1u)) {
                const uint32_t off = (uint32_t)n1 * rstep;
                if (k.in_real) x.x = ps[off];
                else x = ld_nt(pb + off);
            }
            dst[n1] = x;
        }
    };
    const int G = PF ? gridDim.x : ntile;
    int jt = blockIdx.x;
    TI raw[R1];
    load(raw, jt);
    for (; jt < ntile; jt += G) {
        double2 v[RM];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = widen(raw[n1]);
        if (PF) load(raw, jt + G);
        const int cb = k.col0 + jt * C;
        const bool colok = (cb + c) < (k.col0 + k.ncols);
        __syncthreads();  // twiddles visible / previous tile's exchange consumed
        if (grp < R2) {
            const int n2 = grp;
            Dft<R1>::run(v, tw, N / R1, k.sign);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) {
                const double2 y = (k1 == 0 || n2 == 0) ? v[k1] : cmul2(v[k1], tw[n2 * k1]);
                Y[(k1 * R2 + n2) * C + c] = y;
            }
        }
        __syncthreads();
        if (grp < R1) {
            const int k1 = grp;
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) v[n2] = Y[(k1 * R2 + n2) * C + c];
            Dft<R2>::run(v, tw, N / R2, k.sign);
            if (colok) {
                // output row k1 + R1 k2: 32-bit offsets from the thread's first row
                const int64_t ob = t * k.out_t + o * k.out_o + cb + c + (int64_t)k1 * k.out_r;
                const uint32_t ostep = (uint32_t)(R1 * k.out_r);
                TO *po = reinterpret_cast<TO *>(out) + ob;
                typename Scalar<TO>::type *pr = reinterpret_cast<typename Scalar<TO>::type *>(out) + ob;
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {
                    const uint32_t off = (uint32_t)k2 * ostep;
                    const double2 x = make_double2(v[k2].x * k.scale, v[k2].y * k.scale);
                    if (k.out_real) pr[off] = (typename Scalar<TO>::type)x.x;
                    else st<TO>(po, off, x);
                }
            }
        }
        if (!PF) break;
    }
}

// streaming copy, one 16-byte vector per thread with non-temporal accesses: the bandwidth
// reference for the FFT passes (the fastest of the copy shapes measured on MI355X; more
// vectors per thread ran slower)
__global__ __launch_bounds__(256) void copy16_kernel(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n)
{
    typedef float f4v __attribute__((ext_vector_type(4)));
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        __builtin_nontemporal_store(__builtin_nontemporal_load(reinterpret_cast<const f4v *>(a) + i),
                                    reinterpret_cast<f4v *>(b) + i);
}

hipError_t launch_copy16(const void *a, void *b, int64_t nvec, hipStream_t s)
{
    if (nvec <= 0) return hipSuccess;
    if ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) return hipErrorInvalidValue;
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)((nvec + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(a), reinterpret_cast<float4 *>(b), nvec);
    return hipGetLastError();
}

// One column tile per block, no prefetch, the step exchange through LDS in two FP64 halves
// (real, then imaginary parts): 36 KB of LDS and at most 168 registers, so three blocks
// share a CU (the kernel above holds two).  Measured faster for complex128 passes (11% at
// N = 256, 31% at N = 128) and complex64 passes up to N = 192 (10% at N = 128, 2% at
// N = 192); the prefetching kernel stays ahead for complex64 at N = 256.
//
// AMAX: the pass's output is not stored; each block leaves the largest |out|^2 of its tile
// and that element's C-order index within the transform (the smallest on ties) in
// k.pv / k.pi [transform][block], for a peak search that never writes the grid.
//
// Up to N = 192 the whole complex step fits (52 KB with 16 columns, still three blocks per
// CU): one exchange barrier instead of three.
__host__ __device__ constexpr bool lean_fullx(int N) { return N <= 192; }
__host__ __device__ constexpr int lean_lds_bytes(int N, int C)
{
    return (int)(sizeof(double2) * N + (lean_fullx(N) ? sizeof(double2) : sizeof(double)) * N * C);
}

template <int R1, int R2, class TI, class TO, bool AMAX = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void fft_reg_lean(
    const void *__restrict__ in, void *__restrict__ out, const double2 *__restrict__ twg, FftRegK k)
{
    constexpr int N = R1 * R2;
    constexpr int RM = R1 > R2 ? R1 : R2;
    constexpr bool FX = lean_fullx(N);
    extern __shared__ double2 sm[];
    double2 *tw = sm;
    double *Yd = reinterpret_cast<double *>(sm + N);   // [k1][n2][c]
    double2 *Y = sm + N;                               // FX: [k1][n2][c] complex
    const int C = k.C, tid = threadIdx.x;
    const int c = tid % C, grp = tid / C;
    const int64_t o = blockIdx.y, t = blockIdx.z;
    if (k.mask_outer && !order_ok2((int)o, N, k.bw, k.lmax)) return;
    const int jt = blockIdx.x;
    const int cb = k.col0 + jt * C;
    const bool colok = (cb + c) < (k.col0 + k.ncols);
    for (int i = tid; i < N; i += blockDim.x) {
        double2 w = twg[i];
        if (k.sign > 0) w.y = -w.y;
        tw[i] = w;
    }
    double2 v[RM];
    {
        const int64_t ib = t * k.in_t + o * k.in_o + cb + c + (int64_t)grp * k.in_r;
        const TI *pb = reinterpret_cast<const TI *>(in) + ib;
        const uint32_t rstep = (uint32_t)(R2 * k.in_r);
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
            const int r = R2 * n1 + grp;
            TI x = zero_of<TI>();
            if (colok && grp < R2 && (!k.mask_rows || order_ok2(r, N, k.bw, k.lmax))) x = ld_nt(pb + (uint32_t)n1 * rstep);
            v[n1] = widen(x);
        }
    }
    __syncthreads();   // twiddles
    if (grp < R2) {
        const int n2 = grp;
        Dft<R1>::run(v, tw, N / R1, k.sign);
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1)
            if (n2 != 0) v[k1] = cmul2(v[k1], tw[n2 * k1]);
        if constexpr (FX) {
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) Y[(k1 * R2 + n2) * C + c] = v[k1];
        } else {
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) Yd[(k1 * R2 + n2) * C + c] = v[k1].x;
        }
    }
    __syncthreads();
    double re[R2];
    if constexpr (!FX) {
        if (grp < R1) {
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) re[n2] = Yd[(grp * R2 + n2) * C + c];
        }
        __syncthreads();
        if (grp < R2) {
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) Yd[(k1 * R2 + grp) * C + c] = v[k1].y;
        }
        __syncthreads();
    }
    double bv = -1.0;
    int64_t bi = INT64_MAX;
    if (grp < R1) {
        const int k1 = grp;
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2)
            v[n2] = FX ? Y[(k1 * R2 + n2) * C + c] : make_double2(re[n2], Yd[(k1 * R2 + n2) * C + c]);
        Dft<R2>::run(v, tw, N / R2, k.sign);
        if (AMAX && colok) {
            const int64_t ib = o * k.out_o + cb + c + (int64_t)k1 * k.out_r;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const double2 x = make_double2(v[k2].x * k.scale, v[k2].y * k.scale);
                amax_better(bv, bi, x.x * x.x + x.y * x.y, ib + (int64_t)(R1 * k2) * k.out_r);
            }
        } else if (colok) {
            const int64_t ob = t * k.out_t + o * k.out_o + cb + c + (int64_t)k1 * k.out_r;
            const uint32_t ostep = (uint32_t)(R1 * k.out_r);
            TO *po = reinterpret_cast<TO *>(out) + ob;
            typename Scalar<TO>::type *pr = reinterpret_cast<typename Scalar<TO>::type *>(out) + ob;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const uint32_t off = (uint32_t)k2 * ostep;
                const double2 x = make_double2(v[k2].x * k.scale, v[k2].y * k.scale);
                if (k.out_real) pr[off] = (typename Scalar<TO>::type)x.x;
                else st<TO>(po, off, x);
            }
        }
    }
    if constexpr (AMAX) {
        __shared__ double sv[16];
        __shared__ int64_t si[16];
        amax_block_reduce(bv, bi, sv, si);
        if (threadIdx.x == 0) {
            const int64_t nb = (int64_t)gridDim.x * gridDim.y;
            k.pv[t * nb + (int64_t)blockIdx.y * gridDim.x + blockIdx.x] = bv;
            k.pi[t * nb + (int64_t)blockIdx.y * gridDim.x + blockIdx.x] = bi;
        }
    }
}

// The pass fused with the peak search, complex128 in, nothing stored: a block walks AW
// consecutive outer indices of one column tile, loading the rows of the next while
// transforming the current (a read-only pass has no stores to overlap its loads with), and
// reduces its best (|out|^2, C-order index; the smallest index on ties) once, into
// k.pv / k.pi [transform][blockIdx.y * gridDim.x + blockIdx.x].
constexpr int AW = 8;          // outer indices per block
constexpr int AW_MAXN = 128;   // longer passes keep one index per block (fft_reg_lean): the walk's
                               // prefetched rows would spill

template <int R1, int R2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void fft_amax_walk(const double2 *__restrict__ in, const double2 *__restrict__ twg,
                                                     FftRegK k)
{
    constexpr int N = R1 * R2;
    constexpr int RM = R1 > R2 ? R1 : R2;
    extern __shared__ double2 sm[];
    double2 *tw = sm;
    double2 *Y = sm + N;   // [k1][n2][c]: the whole complex exchange (N C 16 bytes, 32 KB at N = 128)
    const int C = k.C, tid = threadIdx.x;
    const int c = tid % C, grp = tid / C;
    const int64_t t = blockIdx.z;
    const int cb = k.col0 + (int)blockIdx.x * C;
    const bool colok = (cb + c) < (k.col0 + k.ncols);
    const int ob = (int)blockIdx.y * AW;
    for (int i = tid; i < N; i += blockDim.x) {
        double2 w = twg[i];
        if (k.sign > 0) w.y = -w.y;
        tw[i] = w;
    }
    // step-1 rows of outer index o for thread (n2 = grp, c): rows R2 n1 + n2, masked rows zero
    uint32_t rowmask = 0;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
        const int r = R2 * n1 + grp;
        if (colok && grp < R2 && (!k.mask_rows || order_ok2(r, N, k.bw, k.lmax))) rowmask |= 1u << n1;
    }
    const uint32_t rstep = (uint32_t)(R2 * k.in_r);
    auto load = [&](double2 (&dst)[R1], int o) {
        const bool live = o < N;
        const double2 *pb = in + (t * k.in_t + (int64_t)min(o, N - 1) * k.in_o + cb + c + (int64_t)grp * k.in_r);
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
            double2 x = make_double2(0.0, 0.0);
            if (live && ((rowmask >> n1) & 1u)) x = ld_nt(pb + (uint32_t)n1 * rstep);
            dst[n1] = x;
        }
    };
    double2 raw[R1];
    load(raw, ob);
    double bv = -1.0;
    int64_t bi = INT64_MAX;
#pragma unroll 1
    for (int w = 0; w < AW; ++w) {
        const int o = ob + w;
        double2 v[RM];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = raw[n1];
        if (w + 1 < AW) load(raw, o + 1);
        __syncthreads();   // twiddles visible / the previous index's exchange consumed
        if (grp < R2) {
            const int n2 = grp;
            Dft<R1>::run(v, tw, N / R1, k.sign);
            // the twiddles are read from LDS here, on every index (held in registers across
            // the walk they would cost 4 registers each)
            asm volatile("" ::: "memory");
#pragma unroll
            for (int k1 = 1; k1 < R1; ++k1)
                if (n2 != 0) v[k1] = cmul2(v[k1], tw[n2 * k1]);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) Y[(k1 * R2 + n2) * C + c] = v[k1];
        }
        __syncthreads();
        if (grp < R1 && o < N) {
            const int k1 = grp;
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) v[n2] = Y[(k1 * R2 + n2) * C + c];
            Dft<R2>::run(v, tw, N / R2, k.sign);
            if (colok) {
                const int64_t ib = (int64_t)o * k.out_o + cb + c + (int64_t)k1 * k.out_r;
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {
                    const double2 x = make_double2(v[k2].x * k.scale, v[k2].y * k.scale);
                    amax_better(bv, bi, x.x * x.x + x.y * x.y, ib + (int64_t)(R1 * k2) * k.out_r);
                }
            }
        }
    }
    __shared__ double sv[16];
    __shared__ int64_t si[16];
    amax_block_reduce(bv, bi, sv, si);
    if (threadIdx.x == 0) {
        const int64_t nb = (int64_t)gridDim.x * gridDim.y;
        k.pv[t * nb + (int64_t)blockIdx.y * gridDim.x + blockIdx.x] = bv;
        k.pi[t * nb + (int64_t)blockIdx.y * gridDim.x + blockIdx.x] = bi;
    }
}

static bool split_pow2(int N, int &R1, int &R2)
{
    switch (N) {
    case 4: R1 = 2; R2 = 2; return true;
    case 6: R1 = 3; R2 = 2; return true;
    case 8: R1 = 4; R2 = 2; return true;
    case 12: R1 = 4; R2 = 3; return true;
    case 16: R1 = 4; R2 = 4; return true;
    case 24: R1 = 6; R2 = 4; return true;
    case 32: R1 = 8; R2 = 4; return true;
    case 48: R1 = 8; R2 = 6; return true;
    case 64: R1 = 8; R2 = 8; return true;
    case 96: R1 = 12; R2 = 8; return true;
    case 128: R1 = 16; R2 = 8; return true;
    case 192: R1 = 16; R2 = 12; return true;
    case 256: R1 = 16; R2 = 16; return true;
    case 384: R1 = 24; R2 = 16; return true;
    case 512: R1 = 32; R2 = 16; return true;
    default: return false;
    }
}

bool fft_reg_supported(int N)
{
    int a, b;
    return split_pow2(N, a, b);
}

static FftRegK make_k(int N, int R1, int bw, int lmax, const FftAxis &ax, int col0, int ncols)
{
    FftRegK k{};
    k.N = N;
    k.C = (R1 >= 24) ? 8 : 16;
    if (k.C > N) k.C = N;
    k.bw = bw; k.lmax = lmax;
    k.in_t = ax.in_t; k.in_o = ax.in_o; k.in_r = ax.in_r;
    k.out_t = ax.out_t; k.out_o = ax.out_o; k.out_r = ax.out_r;
    k.mask_rows = ax.mask_rows; k.mask_outer = ax.mask_outer;
    k.in_real = ax.in_real; k.out_real = ax.out_real; k.sign = ax.sign; k.scale = ax.scale;
    k.col0 = col0; k.ncols = ncols;
    k.pv = nullptr; k.pi = nullptr;
    return k;
}

// the one-tile kernel exists for N <= 256 only (larger lengths would spill)
template <int R1, int R2, class TI, class TO>
static void launch_lean(dim3 grid, dim3 block, int bytes, hipStream_t s, const void *in, void *out,
                        const double2 *tw, const FftRegK &k)
{
    if constexpr (R1 * R2 <= 256) hipLaunchKernelGGL((fft_reg_lean<R1, R2, TI, TO>), grid, block, bytes, s, in, out, tw, k);
}

template <class TI, class TO>
static hipError_t launch_reg_typed(int N, const double2 *tw, int bw, int lmax, const void *in, void *out,
                                   const FftAxis &ax, int T, int col0, int ncols, hipStream_t s)
{
    int R1 = 0, R2 = 0;
    if (!split_pow2(N, R1, R2)) return hipErrorInvalidValue;
    const FftRegK k = make_k(N, R1, bw, lmax, ax, col0, ncols);
    const int threads = (R1 > R2 ? R1 : R2) * k.C;
    const int bytes = (int)sizeof(double2) * (N + N * k.C);
    // complex64 input: each block walks up to 8 column tiles (prefetching the next; a
    // block's first tile is the only one whose load latency is not hidden), as long as
    // the grid keeps about 4096 blocks (8 tiles per block measured 2-4% faster than 16
    // at N = 256); complex128 input: one tile per block
    const int ntile = (ncols + k.C - 1) / k.C;
    const int64_t lines = (int64_t)N * T;
    const int gfill = (int)std::min<int64_t>(ntile, (4096 + lines - 1) / lines);
    const int gx = (sizeof(TI) == 8) ? std::max((ntile + 7) / 8, gfill) : ntile;
    dim3 grid(gx, N, T), block(threads);
    // one tile per block with the exchange in halves: complex128 input up to N = 256,
    // complex64 input up to N = 192 (not for real input, read as scalars by the kernel above)
    const bool lean = !ax.in_real && (sizeof(TI) == 16 ? N <= 256 : N <= 192);
    const int lean_bytes = lean_lds_bytes(N, k.C);
#define SO3_REG(A, B)                                                                                   \
    do {                                                                                                \
        static bool attr = false;                                                                       \
        if (!attr) {                                                                                    \
            hipError_t e = hipFuncSetAttribute((const void *)fft_reg_kernel<A, B, TI, TO>,              \
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);  \
            if (e != hipSuccess) return e;                                                              \
            attr = true;                                                                                \
        }                                                                                               \
        if (lean)                                                                                       \
            launch_lean<A, B, TI, TO>(dim3(ntile, N, T), block, lean_bytes, s, in, out, tw, k);         \
        else                                                                                            \
            hipLaunchKernelGGL((fft_reg_kernel<A, B, TI, TO>), grid, block, bytes, s, in, out, tw, k); \
    } while (0)
    switch (N) {
    case 4: SO3_REG(2, 2); break;
    case 6: SO3_REG(3, 2); break;
    case 8: SO3_REG(4, 2); break;
    case 12: SO3_REG(4, 3); break;
    case 16: SO3_REG(4, 4); break;
    case 24: SO3_REG(6, 4); break;
    case 32: SO3_REG(8, 4); break;
    case 48: SO3_REG(8, 6); break;
    case 64: SO3_REG(8, 8); break;
    case 96: SO3_REG(12, 8); break;
    case 128: SO3_REG(16, 8); break;
    case 192: SO3_REG(16, 12); break;
    case 256: SO3_REG(16, 16); break;
    case 384: SO3_REG(24, 16); break;
    default: SO3_REG(32, 16); break;
    }
#undef SO3_REG
    return hipGetLastError();
}

hipError_t launch_fft_reg(int N, const double2 *tw, int bw, int lmax, const void *in, void *out,
                          const FftAxis &ax, int T, int in_c64, int out_c64, int col0, int ncols, hipStream_t s)
{
    // lengths without a register-kernel factorisation: the LDS Stockham kernel
    if (!fft_reg_supported(N)) return launch_fft_stockham(N, tw, bw, lmax, in, out, ax, T, in_c64, out_c64, col0, ncols, s);
    if (!in_c64 && !out_c64) return launch_reg_typed<double2, double2>(N, tw, bw, lmax, in, out, ax, T, col0, ncols, s);
    if (in_c64 && out_c64) return launch_reg_typed<float2, float2>(N, tw, bw, lmax, in, out, ax, T, col0, ncols, s);
    if (in_c64) return launch_reg_typed<float2, double2>(N, tw, bw, lmax, in, out, ax, T, col0, ncols, s);
    return launch_reg_typed<double2, float2>(N, tw, bw, lmax, in, out, ax, T, col0, ncols, s);
}

// the fused pass exists where the one-tile complex128 kernel does (N = 2^a or 3 * 2^a up to
// 256), for whole lines (no outer mask) into a complex128 result
int fft_reg_argmax_partials(int N)
{
    int R1 = 0, R2 = 0;
    if (!split_pow2(N, R1, R2) || N > 256) return 0;
    const int C = std::min((R1 >= 24) ? 8 : 16, N);
    return (N + C - 1) / C * (N <= AW_MAXN ? (N + AW - 1) / AW : N);
}

hipError_t launch_fft_reg_argmax(int N, const double2 *tw, int bw, int lmax, const double2 *in, const FftAxis &ax,
                                 int T, double *pv, int64_t *pi, int *nblk, hipStream_t s)
{
    int R1 = 0, R2 = 0;
    if (!split_pow2(N, R1, R2) || N > 256 || ax.in_real || ax.out_real || ax.mask_outer || !pv || !pi)
        return hipErrorNotSupported;
    FftRegK k = make_k(N, R1, bw, lmax, ax, 0, N);
    k.pv = pv;
    k.pi = pi;
    const int ntile = (N + k.C - 1) / k.C;
    const int bytes = lean_lds_bytes(N, k.C);
    const dim3 block((R1 > R2 ? R1 : R2) * k.C);
    if (N > AW_MAXN) {
        *nblk = ntile * N;
        const dim3 grid(ntile, N, T);
        if (N == 192)
            hipLaunchKernelGGL((fft_reg_lean<16, 12, double2, double2, true>), grid, block, bytes, s, in, nullptr, tw, k);
        else
            hipLaunchKernelGGL((fft_reg_lean<16, 16, double2, double2, true>), grid, block, bytes, s, in, nullptr, tw, k);
        return hipGetLastError();
    }
    const int nouter = (N + AW - 1) / AW;
    *nblk = ntile * nouter;
    const dim3 grid(ntile, nouter, T);
    const int wbytes = (int)(sizeof(double2) * N * (1 + k.C));   // twiddles + complex exchange
#define SO3_AMAX(A, B)                                                                                       \
    do {                                                                                                    \
        if constexpr (A * B <= AW_MAXN)                                                                     \
            hipLaunchKernelGGL((fft_amax_walk<A, B>), grid, block, wbytes, s, in, tw, k);                    \
    } while (0)
    // the smaller radix first: every thread of the block loads rows (R2 >= R1 groups), and
    // the prefetched rows take fewer registers
    switch (N) {
    case 4: SO3_AMAX(2, 2); break;
    case 6: SO3_AMAX(2, 3); break;
    case 8: SO3_AMAX(2, 4); break;
    case 12: SO3_AMAX(3, 4); break;
    case 16: SO3_AMAX(4, 4); break;
    case 24: SO3_AMAX(4, 6); break;
    case 32: SO3_AMAX(4, 8); break;
    case 48: SO3_AMAX(6, 8); break;
    case 64: SO3_AMAX(8, 8); break;
    case 96: SO3_AMAX(8, 12); break;
    default: SO3_AMAX(8, 16); break;
    }
#undef SO3_AMAX
    return hipGetLastError();
}

}  // namespace so3
