// alpha-gamma 2-D DFT of every beta slice, done as two LDS-resident passes of
// batched 1-D mixed-radix Stockham FFTs (replaces the FFTW plan of
// pysofft/soft.f08:220-292: rank 2, n=[2bw,2bw], howmany=2bw, stride 2bw).
//
// Both the SO(3)-function array (gamma, alpha, beta) and the DWT array
// (i_n, i_m, beta) keep beta contiguous, so one pass loads an (N rows x C beta
// columns) tile of one line family into LDS with contiguous C-element row
// segments, runs the FFT down the rows for all C columns, and stores the tile.
// The pass along m (stride N) and the pass along n (stride N^2) are the same
// kernel with different strides.  Orders that carry no data (the Nyquist order
// bw and |order| > lmax) are skipped on load.
#include <cstdlib>

#include "so3ft_internal.h"

namespace so3 {

static constexpr int FFT_THREADS = 256;

struct FftK {
    int N, C, nrad, bw, lmax;
    int rad[16];
    int64_t in_t, in_o, in_r, out_t, out_o, out_r;
    int mask_rows, mask_outer, in_real, out_real, sign;
    int in_c64, out_c64;   // complex64 storage of the input / output (FP64 arithmetic)
    int col0, ncols;       // beta column window [col0, col0 + ncols)
    double scale;
};

__device__ __forceinline__ bool order_ok(int i, int N, int bw, int lmax)
{
    if (i == bw) return false;  // Nyquist order, never written (soft.f08:1894)
    const int m = (i < bw) ? i : i - N;
    return (m < 0 ? -m : m) <= lmax;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(FFT_THREADS) void fft_axis_kernel(const void *__restrict__ in,
                                                               void *__restrict__ out,
                                                               const double2 *__restrict__ twg, FftK k)
{
    extern __shared__ double2 sm[];
    const int N = k.N, C = k.C, tid = threadIdx.x;
    double2 *tw = sm;
    double2 *X0 = tw + N;
    double2 *X1 = X0 + N * C;
    const int c0 = k.col0 + blockIdx.x * C;
    const int cend = k.col0 + k.ncols;
    const int64_t o = blockIdx.y, t = blockIdx.z;

    if (k.mask_outer && !order_ok((int)o, N, k.bw, k.lmax)) return;  // consumer masks this line

    for (int i = tid; i < N; i += FFT_THREADS) {
        double2 w = twg[i];
        if (k.sign > 0) w.y = -w.y;
        tw[i] = w;
    }
    const int64_t ib = t * k.in_t + o * k.in_o + c0;
    for (int idx = tid; idx < N * C; idx += FFT_THREADS) {
        const int r = idx / C, c = idx - r * C;
        double2 v = make_double2(0.0, 0.0);
        if (c0 + c < cend && (!k.mask_rows || order_ok(r, N, k.bw, k.lmax))) {
            const int64_t off = ib + r * k.in_r + c;
            if (k.in_real) {
                v.x = reinterpret_cast<const double *>(in)[off];
            } else if (k.in_c64) {
                const float2 f = reinterpret_cast<const float2 *>(in)[off];
                v = make_double2((double)f.x, (double)f.y);
            } else {
                v = reinterpret_cast<const double2 *>(in)[off];
            }
        }
        X0[idx] = v;
    }
    __syncthreads();

    double2 *A = X0, *B = X1;
    int Ns = 1;
    for (int s = 0; s < k.nrad; ++s) {
        const int R = k.rad[s], nb = N / R, st = N / (Ns * R);
        if (R == 4) {
            for (int idx = tid; idx < nb * C; idx += FFT_THREADS) {
                const int j = idx / C, c = idx - j * C, kk = j % Ns;
                const double2 v0 = A[j * C + c];
                const double2 v1 = cmul(A[(j + nb) * C + c], tw[kk * st]);
                const double2 v2 = cmul(A[(j + 2 * nb) * C + c], tw[2 * kk * st]);
                const double2 v3 = cmul(A[(j + 3 * nb) * C + c], tw[3 * kk * st]);
                const double2 a0 = make_double2(v0.x + v2.x, v0.y + v2.y);
                const double2 a1 = make_double2(v0.x - v2.x, v0.y - v2.y);
                const double2 a2 = make_double2(v1.x + v3.x, v1.y + v3.y);
                const double2 a3 = make_double2(v1.x - v3.x, v1.y - v3.y);
                // j3 = sign * i * a3
                const double2 j3 = (k.sign < 0) ? make_double2(a3.y, -a3.x) : make_double2(-a3.y, a3.x);
                const int d = (j / Ns) * Ns * 4 + kk;
                B[d * C + c] = make_double2(a0.x + a2.x, a0.y + a2.y);
                B[(d + Ns) * C + c] = make_double2(a1.x + j3.x, a1.y + j3.y);
                B[(d + 2 * Ns) * C + c] = make_double2(a0.x - a2.x, a0.y - a2.y);
                B[(d + 3 * Ns) * C + c] = make_double2(a1.x - j3.x, a1.y - j3.y);
            }
        } else if (R == 2) {
            for (int idx = tid; idx < nb * C; idx += FFT_THREADS) {
                const int j = idx / C, c = idx - j * C, kk = j % Ns;
                const double2 v0 = A[j * C + c];
                const double2 v1 = cmul(A[(j + nb) * C + c], tw[kk * st]);
                const int d = (j / Ns) * Ns * 2 + kk;
                B[d * C + c] = make_double2(v0.x + v1.x, v0.y + v1.y);
                B[(d + Ns) * C + c] = make_double2(v0.x - v1.x, v0.y - v1.y);
            }
        } else {
            // generic radix: one output of one butterfly per work item, O(R) each
            for (int idx = tid; idx < N * C; idx += FFT_THREADS) {
                const int c = idx % C, rest = idx / C;
                const int j = rest % nb, qq = rest / nb, kk = j % Ns;
                double2 acc = make_double2(0.0, 0.0);
                for (int r = 0; r < R; ++r) {
                    const double2 v = cmul(A[(j + r * nb) * C + c], tw[r * kk * st]);
                    const double2 wr = tw[((r * qq) % R) * nb];
                    acc.x += v.x * wr.x - v.y * wr.y;
                    acc.y += v.x * wr.y + v.y * wr.x;
                }
                const int d = (j / Ns) * Ns * R + kk;
                B[(d + qq * Ns) * C + c] = acc;
            }
        }
        __syncthreads();
        double2 *tmp = A; A = B; B = tmp;
        Ns *= R;
    }

    const int64_t ob = t * k.out_t + o * k.out_o + c0;
    for (int idx = tid; idx < N * C; idx += FFT_THREADS) {
        const int r = idx / C, c = idx - r * C;
        if (c0 + c >= cend) continue;
        const double2 v = A[idx];
        const int64_t off = ob + r * k.out_r + c;
        if (k.out_real) reinterpret_cast<double *>(out)[off] = v.x * k.scale;
        else if (k.out_c64) reinterpret_cast<float2 *>(out)[off] = make_float2((float)(v.x * k.scale), (float)(v.y * k.scale));
        else reinterpret_cast<double2 *>(out)[off] = make_double2(v.x * k.scale, v.y * k.scale);
    }
}

static int fft_cols(int N)
{
    int C = 2048 / N;
    if (C < 1) C = 1;
    if (C > N) C = N;
    return C;
}

// radices of N: 4s, then 2s, then odd factors
static int radices(int N, int *rad)
{
    int nr = 0, n = N;
    while (n % 4 == 0) { rad[nr++] = 4; n /= 4; }
    while (n % 2 == 0) { rad[nr++] = 2; n /= 2; }
    for (int p = 3; n > 1; p += 2)
        while (n % p == 0) { rad[nr++] = p; n /= p; }
    return nr;
}

hipError_t launch_fft_stockham(int N, const double2 *tw, int bw, int lmax, const void *in, void *out,
                               const FftAxis &ax, int T, int in_c64, int out_c64, int col0, int ncols, hipStream_t s)
{
    FftK k;
    k.N = N;
    k.C = fft_cols(N);
    int rad[32];
    k.nrad = radices(N, rad);
    if (k.nrad > 16) return hipErrorInvalidValue;
    for (int i = 0; i < 16; ++i) k.rad[i] = (i < k.nrad) ? rad[i] : 1;
    k.bw = bw; k.lmax = lmax;
    k.in_t = ax.in_t; k.in_o = ax.in_o; k.in_r = ax.in_r;
    k.out_t = ax.out_t; k.out_o = ax.out_o; k.out_r = ax.out_r;
    k.mask_rows = ax.mask_rows; k.mask_outer = ax.mask_outer;
    k.in_real = ax.in_real; k.out_real = ax.out_real;
    k.in_c64 = in_c64; k.out_c64 = out_c64;
    k.col0 = col0; k.ncols = ncols;
    k.sign = ax.sign; k.scale = ax.scale;
    const int bytes = (int)sizeof(double2) * (N + 2 * N * k.C);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)fft_axis_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           96 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    dim3 grid((ncols + k.C - 1) / k.C, N, T), block(FFT_THREADS);
    hipLaunchKernelGGL(fft_axis_kernel, grid, block, bytes, s, in, out, tw, k);
    return hipGetLastError();
}

hipError_t launch_fft_axis(const FftDesc &fd, const double2 *tw, int bw, int lmax, const void *in, void *out,
                           const FftAxis &ax, int T, hipStream_t s)
{
    if (fft_reg_supported(fd.N) && !getenv("SO3FT_FFT_STOCKHAM"))
        return launch_fft_reg(fd.N, tw, bw, lmax, in, out, ax, T, 0, 0, 0, fd.N, s);
    return launch_fft_stockham(fd.N, tw, bw, lmax, in, out, ax, T, 0, 0, 0, fd.N, s);
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

__global__ void zero_invalid_kernel(double2 *F, int64_t f_ts, int bw, int N, int lmax, int half)
{
    const int64_t t = blockIdx.y;
    const int64_t n3 = (int64_t)N * N * N;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const int im = (int)((i / N) % N), in = (int)(i / ((int64_t)N * N));
        if (!order_ok(im, N, bw, lmax) || !order_ok(in, N, bw, lmax) || (half && im > bw))
            F[t * f_ts + i] = make_double2(0.0, 0.0);
    }
}

hipError_t launch_zero_invalid_columns(double2 *F, int64_t f_ts, int bw, int N, int lmax, int T, int half,
                                       hipStream_t s)
{
    const int64_t n3 = (int64_t)N * N * N;
    int nb = (int)((n3 + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(zero_invalid_kernel, dim3(nb, T), dim3(256), 0, s, F, f_ts, bw, N, lmax, half);
    return hipGetLastError();
}

__global__ void scale_kernel(double2 *X, int64_t n, double sc)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = X[i];
        X[i] = make_double2(v.x * sc, v.y * sc);
    }
}

hipError_t launch_scale(double2 *X, int64_t n, double sc, hipStream_t s)
{
    int nb = (int)((n + 255) / 256);
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(256), 0, s, X, n, sc);
    return hipGetLastError();
}

// sum_k w_k sum_{alpha,gamma} f(gamma,alpha,beta_k): partial sums per block, then one block
__global__ void integrate_partial_kernel(const void *f, int is_real, const double *w, int N, double *partials)
{
    __shared__ double sre[256], sim[256];
    const int64_t n3 = (int64_t)N * N * N;
    double re = 0.0, im = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const double wk = w[i % N];
        if (is_real) {
            re += wk * reinterpret_cast<const double *>(f)[i];
        } else {
            const double2 v = reinterpret_cast<const double2 *>(f)[i];
            re += wk * v.x; im += wk * v.y;
        }
    }
    sre[threadIdx.x] = re; sim[threadIdx.x] = im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) { sre[threadIdx.x] += sre[threadIdx.x + st]; sim[threadIdx.x] += sim[threadIdx.x + st]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = sre[0]; partials[2 * blockIdx.x + 1] = sim[0]; }
}

__global__ void integrate_final_kernel(const double *partials, int nb, double *res)
{
    __shared__ double sre[256], sim[256];
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) { re += partials[2 * i]; im += partials[2 * i + 1]; }
    sre[threadIdx.x] = re; sim[threadIdx.x] = im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) { sre[threadIdx.x] += sre[threadIdx.x + st]; sim[threadIdx.x] += sim[threadIdx.x + st]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { res[0] = sre[0]; res[1] = sim[0]; }
}

hipError_t launch_integrate(const void *f, int is_real, const double *w, int N, double *partials, int nblocks,
                            double *result2, hipStream_t s)
{
    hipLaunchKernelGGL(integrate_partial_kernel, dim3(nblocks), dim3(256), 0, s, f, is_real, w, N, partials);
    hipLaunchKernelGGL(integrate_final_kernel, dim3(1), dim3(256), 0, s, partials, nblocks, result2);
    return hipGetLastError();
}

// utils.f90:471-488: f^l_{-m,-n} = (-1)^{m+n} conj f^l_{m,n}; for m = 0 the n < 0 side is
// the source (the reference's loop order lets it win), f^l_{0,0} is made real.
// m0_neg_src: for m = 0 the n < 0 entries are the sources (utils.f90:471-488, mnl);
// else the n > 0 entries (utils.f90:510-550, lmn)
__global__ void enforce_real_sym_kernel(double2 *c, int64_t c_ts, const int64_t *slice, int bw, int m0_neg_src)
{
    const int64_t t = blockIdx.y;
    const int W = 2 * bw - 1;
    double2 *ct = c + t * c_ts;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < W * W; idx += gridDim.x * blockDim.x) {
        const int m = idx / W - (bw - 1), n = idx % W - (bw - 1);
        const int a = (m < 0 ? -m : m) > (n < 0 ? -n : n) ? (m < 0 ? -m : m) : (n < 0 ? -n : n);
        const int L = bw - a;
        double2 *dst = ct + slice[idx];
        if (m == 0 && n == 0) {
            for (int i = 0; i < L; ++i) dst[i].y = 0.0;
            continue;
        }
        const bool is_dst = (m < 0) || (m == 0 && (m0_neg_src ? n > 0 : n < 0));
        if (!is_dst) continue;
        const double2 *src = ct + slice[(int64_t)(-m + bw - 1) * W + (-n + bw - 1)];
        const double sg = (((m + n) & 1) != 0) ? -1.0 : 1.0;
        for (int i = 0; i < L; ++i) { const double2 v = src[i]; dst[i] = make_double2(sg * v.x, -sg * v.y); }
    }
}

hipError_t launch_enforce_real_sym(double2 *c, int64_t c_ts, const int64_t *slice, int bw, int T, int m0_neg_src,
                                   hipStream_t s)
{
    const int W = 2 * bw - 1;
    int nb = (W * W + 255) / 256;
    for (int t0 = 0; t0 < T; t0 += 32768) {  // grid y extent < 65536
        const int nt = T - t0 < 32768 ? T - t0 : 32768;
        hipLaunchKernelGGL(enforce_real_sym_kernel, dim3(nb, nt), dim3(256), 0, s, c + (int64_t)t0 * c_ts, c_ts, slice,
                           bw, m0_neg_src);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// lmn <-> mnl coefficient order: perm[i] = mnl position of lmn position i.  to_mnl:
// out[perm[i]] = in[i] for i < nkeep and 0 beyond; else out[i] = in[perm[i]] for i < nkeep,
// entries beyond left as they are.  The lmn order is degree-major, so nkeep = NC(L) keeps
// exactly the degrees l < L.  T sets at strides in_ts / out_ts.
template <class E>
__global__ void permute_kernel(const E *__restrict__ in, E *__restrict__ out, const int32_t *__restrict__ perm,
                               int64_t nc, int64_t nkeep, int64_t in_ts, int64_t out_ts, int to_mnl)
{
    const int64_t t = blockIdx.y;
    const int64_t lim = to_mnl ? nc : nkeep;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < lim; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[i];
        if (to_mnl) {
            E v = {};
            if (i < nkeep) v = in[t * in_ts + i];
            out[t * out_ts + j] = v;
        } else {
            out[t * out_ts + i] = in[t * in_ts + j];
        }
    }
}

hipError_t launch_permute_lmn(const int32_t *perm, int64_t nc, int64_t nkeep, const void *in, int64_t in_ts,
                              void *out, int64_t out_ts, int esz, int to_mnl, int T, hipStream_t s)
{
    if (nkeep < 0 || nkeep > nc) return hipErrorInvalidValue;
    int nb = (int)((nc + 255) / 256);
    if (nb > 2048) nb = 2048;
    for (int t0 = 0; t0 < T; t0 += 32768) {
        const int nt = T - t0 < 32768 ? T - t0 : 32768;
        if (esz == 16)
            hipLaunchKernelGGL(permute_kernel<double2>, dim3(nb, nt), dim3(256), 0, s,
                               reinterpret_cast<const double2 *>(in) + (int64_t)t0 * in_ts,
                               reinterpret_cast<double2 *>(out) + (int64_t)t0 * out_ts, perm, nc, nkeep, in_ts,
                               out_ts, to_mnl);
        else if (esz == 8)
            hipLaunchKernelGGL(permute_kernel<float2>, dim3(nb, nt), dim3(256), 0, s,
                               reinterpret_cast<const float2 *>(in) + (int64_t)t0 * in_ts,
                               reinterpret_cast<float2 *>(out) + (int64_t)t0 * out_ts, perm, nc, nkeep, in_ts,
                               out_ts, to_mnl);
        else
            return hipErrorInvalidValue;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// half-complex <-> full complex over the alpha (m) axis, C-order arrays:
// half (N, B+1, N) = (i_n, i_m <= B, beta), full (N, N, N) = (i_n, i_m, beta).
// to_full: entries i_m > B take conj(half(-n, -m)) (the c2r input's Hermitian completion)
__global__ void halfcomplex_kernel(const double2 *__restrict__ in, double2 *__restrict__ out, int N, int to_full)
{
    const int B = N / 2;
    const int64_t n3 = (int64_t)N * N * N;
    const int64_t nh = (int64_t)N * (B + 1) * N;
    const int64_t total = to_full ? n3 : nh;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int beta = (int)(i % N);
        if (to_full) {
            const int im = (int)((i / N) % N), in_ = (int)(i / ((int64_t)N * N));
            if (im <= B) {
                out[i] = in[((int64_t)in_ * (B + 1) + im) * N + beta];
            } else {
                const double2 v = in[((int64_t)((N - in_) % N) * (B + 1) + (N - im)) * N + beta];
                out[i] = make_double2(v.x, -v.y);
            }
        } else {
            const int im = (int)((i / N) % (B + 1)), in_ = (int)(i / ((int64_t)N * (B + 1)));
            out[i] = in[((int64_t)in_ * N + im) * N + beta];
        }
    }
}

hipError_t launch_halfcomplex(const double2 *in, double2 *out, int N, int to_full, hipStream_t s)
{
    const int64_t total = to_full ? (int64_t)N * N * N : (int64_t)N * (N / 2 + 1) * N;
    int nb = (int)((total + 255) / 256);
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(halfcomplex_kernel, dim3(nb), dim3(256), 0, s, in, out, N, to_full);
    return hipGetLastError();
}

}  // namespace so3
