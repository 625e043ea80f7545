// Alignment-proposal MLP: features -> 3 x (Linear + GELU) -> Linear -> 6-D rotation
// representation (two 3-vectors, Gram-Schmidt to a rotation matrix).  No counterpart in
// the reference; the network, its training kernels and its inference kernels are this
// library's own.  BF16 weights and activations feed v_mfma_f32_16x16x16bf16_1k with FP32
// accumulation; the optimiser keeps FP32 master weights (Adam).
//
// MFMA operand layout (v_mfma_f32_16x16x16bf16_1k): lane l holds A[l & 15][4 (l >> 4) + i]
// and B[4 (l >> 4) + i][l & 15], i < 4; item r of the 4-float result is row 4 (l >> 4) + r,
// column l & 15.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace so3 {

typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f4m_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x)
{
    // round to nearest even (NaN stays NaN)
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// C[m][n] = sum_k A(m,k) B(k,n) with A(m,k) = A[m sam + k sak], B(k,n) = B[k sbk + n sbn]
// (bf16), FP32 accumulate.  Epilogue: z = C + bias[n] (if bias) stored to Cf (FP32, if
// given) and act(z) stored to Cb (bf16, if given); act = GELU when gelu, else identity.
// Block tile 64 x 64, K step 16; 4 waves, each a 32 x 32 sub-tile of 2 x 2 MFMA tiles.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const uint16_t *__restrict__ A, int64_t sam, int64_t sak,
                                                        const uint16_t *__restrict__ B, int64_t sbk, int64_t sbn,
                                                        int M, int N, int K, const float *__restrict__ bias,
                                                        float *__restrict__ Cf, uint16_t *__restrict__ Cb, int gelu,
                                                        int accumulate)
{
    __shared__ uint16_t As[64][20];  // [m][k], padded
    __shared__ uint16_t Bs[64][20];  // [n][k], padded
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    f4m_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = f4m_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + 256 * e;        // 1024 entries of each tile
            const int r = idx >> 4, kk = idx & 15;
            const int m = m0 + r, n = n0 + r, k = k0 + kk;
            As[r][kk] = (m < M && k < K) ? A[(int64_t)m * sam + (int64_t)k * sak] : (uint16_t)0;
            Bs[r][kk] = (n < N && k < K) ? B[(int64_t)k * sbk + (int64_t)n * sbn] : (uint16_t)0;
        }
        __syncthreads();
        const int row = lane & 15, kg = 4 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            s4_t a;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = (short)As[wm + 16 * i + row][kg + q];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                s4_t b;
#pragma unroll
                for (int q = 0; q < 4; ++q) b[q] = (short)Bs[wn + 16 * jj + row][kg + q];
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][jj], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * jj + (lane & 15);
                if (m >= M || n >= N) continue;
                float z = acc[i][jj][r] + (bias ? bias[n] : 0.f);
                const int64_t o = (int64_t)m * N + n;
                if (Cf) Cf[o] = accumulate ? Cf[o] + z : z;
                if (Cb) {
                    const float y = gelu ? 0.5f * z * (1.f + erff(z * 0.70710678118654752f)) : z;
                    Cb[o] = f2bf(y);
                }
            }
}

// dZ = dA * GELU'(Z) (bf16 out); GELU(z) = z Phi(z), GELU'(z) = Phi(z) + z phi(z)
__global__ void gelu_bwd_kernel(const float *__restrict__ dA, const float *__restrict__ Z, uint16_t *__restrict__ dZb,
                                float *__restrict__ dZf, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i];
    const float Phi = 0.5f * (1.f + erff(z * 0.70710678118654752f));
    const float phi = 0.3989422804014327f * expf(-0.5f * z * z);
    const float g = dA[i] * (Phi + z * phi);
    dZb[i] = f2bf(g);
    if (dZf) dZf[i] = g;
}

// bias gradient: db[n] = sum_m dZ[m][n].  A 1024-thread block per 64 columns: wave w
// sums rows w, w + 16, ... of its 64 columns (256-byte rows), then the 16 partial sums
// are added in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void colsum_kernel(const float *__restrict__ dZ, int M, int N,
                                                       float *__restrict__ db)
{
    __shared__ float part[16][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    float s = 0.f;
    if (n < N) {
#pragma unroll 4
        for (int m = g; m < M; m += 16) s += dZ[(int64_t)m * N + n];
    }
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) {
        float t = 0.f;
        for (int k = 0; k < 16; ++k) t += part[k][c];
        db[n] = t;
    }
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// 6-D representation -> rotation R (columns x, y, z):
//   x = a/|a|, b' = b - (x.b) x, y = b'/|b'|, z = x cross y
__device__ inline void gs6(const float *o, float *R, float *x, float *y, float *bp, float &na, float &nb)
{
    float a[3] = {o[0], o[1], o[2]}, b[3] = {o[3], o[4], o[5]};
    na = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) + 1e-12f;
    for (int i = 0; i < 3; ++i) x[i] = a[i] / na;
    const float xb = x[0] * b[0] + x[1] * b[1] + x[2] * b[2];
    for (int i = 0; i < 3; ++i) bp[i] = b[i] - xb * x[i];
    nb = sqrtf(bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]) + 1e-12f;
    for (int i = 0; i < 3; ++i) y[i] = bp[i] / nb;
    const float z[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
    for (int i = 0; i < 3; ++i) { R[3 * i] = x[i]; R[3 * i + 1] = y[i]; R[3 * i + 2] = z[i]; }
}

// rotation matrices (row-major 3x3) from the outputs
__global__ void rot6d_kernel(const float *__restrict__ out6, int n, float *__restrict__ R9)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float R[9], x[3], y[3], bp[3], na, nb;
    gs6(out6 + 6 * (int64_t)s, R, x, y, bp, na, nb);
    for (int i = 0; i < 9; ++i) R9[9 * (int64_t)s + i] = R[i];
}

// loss = mean_s ||R_s - T_s||_F^2 and its gradient with respect to the 6 outputs
__global__ void rot6d_loss_kernel(const float *__restrict__ out6, const float *__restrict__ T9, int n,
                                  float *__restrict__ g6, float *__restrict__ loss_part)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (s < n) {
        float R[9], x[3], y[3], bp[3], na, nb;
        const float *o = out6 + 6 * (int64_t)s;
        gs6(o, R, x, y, bp, na, nb);
        float gR[9];
        for (int i = 0; i < 9; ++i) {
            const float d = R[i] - T9[9 * (int64_t)s + i];
            l += d * d;
            gR[i] = 2.f * d / (float)n;
        }
        float gx[3], gy[3], gz[3];
        for (int i = 0; i < 3; ++i) { gx[i] = gR[3 * i]; gy[i] = gR[3 * i + 1]; gz[i] = gR[3 * i + 2]; }
        // z = x cross y: gx += y cross gz, gy += gz cross x
        gx[0] += y[1] * gz[2] - y[2] * gz[1]; gx[1] += y[2] * gz[0] - y[0] * gz[2]; gx[2] += y[0] * gz[1] - y[1] * gz[0];
        gy[0] += gz[1] * x[2] - gz[2] * x[1]; gy[1] += gz[2] * x[0] - gz[0] * x[2]; gy[2] += gz[0] * x[1] - gz[1] * x[0];
        // y = b'/|b'|
        const float gyy = gy[0] * y[0] + gy[1] * y[1] + gy[2] * y[2];
        float gbp[3];
        for (int i = 0; i < 3; ++i) gbp[i] = (gy[i] - gyy * y[i]) / nb;
        // b' = b - (x.b) x
        const float b[3] = {o[3], o[4], o[5]};
        const float xb = x[0] * b[0] + x[1] * b[1] + x[2] * b[2];
        const float gbx = gbp[0] * x[0] + gbp[1] * x[1] + gbp[2] * x[2];
        float gb[3];
        for (int i = 0; i < 3; ++i) {
            gb[i] = gbp[i] - gbx * x[i];
            gx[i] += -gbx * b[i] - xb * gbp[i];
        }
        // x = a/|a|
        const float gxx = gx[0] * x[0] + gx[1] * x[1] + gx[2] * x[2];
        float* go = g6 + 6 * (int64_t)s;
        for (int i = 0; i < 3; ++i) {
            go[i] = (gx[i] - gxx * x[i]) / na;
            go[3 + i] = gb[i];
        }
    }
    // per-block partial sums of the loss
    __shared__ float red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Adam on FP32 master weights; writes the bf16 copy used by the GEMMs
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, uint16_t *__restrict__ wb, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float wn = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[i] = wn;
    if (wb) wb[i] = f2bf(wn);
}

// Features of a pair: f, g LMc (stride nb = bw^2 complex128), bw >= 9.
//   [Re f_lm, Im f_lm, Re g_lm, Im g_lm for l <= 8 (4 x 81), log P_f(l), log P_g(l) for
//    l < 64 (2 x 64)], coefficients divided by s = sqrt(sum_{l<=8} |f_lm|^2 / 81), powers by
//   s^2, then standardised with (mean, std) when given.  P(l) = sum_m |y_lm|^2; degrees
//   >= bw enter as log of a tiny floor.
__global__ __launch_bounds__(256) void align_features_kernel(const double2 *__restrict__ f,
                                                             const double2 *__restrict__ g, int64_t nb, int bw, int n,
                                                             const float *__restrict__ mean,
                                                             const float *__restrict__ stdv, float *__restrict__ x)
{
    const int s = blockIdx.x;
    if (s >= n) return;
    const double2 *fs = f + (int64_t)s * nb, *gs = g + (int64_t)s * nb;
    __shared__ double pw[2][64];
    const int t = threadIdx.x, lane = t & 63, nw = blockDim.x >> 6;
    // per-degree powers: wave w takes degrees w, w + nw, ...; its lanes read the degree's
    // 2l+1 consecutive coefficients and a shuffle reduction sums them
    for (int l = t >> 6; l < 64; l += nw) {
        double pf = 0.0, pg = 0.0;
        if (l < bw) {
            for (int m = lane - l; m <= l; m += 64) {
                const double2 a = fs[l * (l + 1) + m], b = gs[l * (l + 1) + m];
                pf += a.x * a.x + a.y * a.y;
                pg += b.x * b.x + b.y * b.y;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            pf += __shfl_down(pf, o, 64);
            pg += __shfl_down(pg, o, 64);
        }
        if (lane == 0) {
            pw[0][l] = pf;
            pw[1][l] = pg;
        }
    }
    __syncthreads();
    double e = 0.0;   // sum_{l <= 8} |f_lm|^2
#pragma unroll
    for (int l = 0; l <= 8; ++l) e += pw[0][l];
    const double sc = sqrt(e / 81.0) + 1e-300;
    float *xs = x + (int64_t)s * 452;
    auto put = [&](int k, double v) {
        float y = (float)v;
        if (mean) y = (y - mean[k]) / stdv[k];
        xs[k] = y;
    };
    for (int i = t; i < 81; i += blockDim.x) {
        put(i, fs[i].x / sc);
        put(81 + i, fs[i].y / sc);
        put(162 + i, gs[i].x / sc);
        put(243 + i, gs[i].y / sc);
    }
    for (int l = t; l < 64; l += blockDim.x) {
        put(324 + l, log(pw[0][l] / (sc * sc) + 1e-30));
        put(388 + l, log(pw[1][l] / (sc * sc) + 1e-30));
    }
}

// ---------------------------------------------------------------------------
// the network object
// ---------------------------------------------------------------------------
struct Mlp {
    std::vector<int> dims;               // layer widths, dims[0] = inputs, back = 6
    std::vector<float *> W, b, Wm, Wv, bm, bv, gW, gb;  // FP32 master, Adam moments, grads
    std::vector<uint16_t *> Wb;          // bf16 copies
    int64_t cap = 0;                     // batch capacity of the activation buffers
    std::vector<float *> Z;              // pre-activations (FP32) per layer
    std::vector<uint16_t *> Ab;          // activations (bf16) per layer, Ab[0] = input
    float *dA = nullptr, *dZf = nullptr, *out6 = nullptr, *g6 = nullptr, *lossp = nullptr;
    uint16_t *dZb = nullptr, *dAb = nullptr;
    int64_t step = 0;
    int device = 0;
};

std::mutex g_mlp_mu;
std::map<int64_t, Mlp *> g_mlps;
int64_t g_mlp_next = 1;
thread_local std::string g_mlp_err;

static int mfail(int code, const std::string &m)
{
    g_mlp_err = m;
    return code;
}
#define MHIP(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return mfail(SO3FT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static Mlp *mget(int64_t h)
{
    std::lock_guard<std::mutex> lk(g_mlp_mu);
    auto it = g_mlps.find(h);
    return it == g_mlps.end() ? nullptr : it->second;
}

static int nparams(const Mlp *P)
{
    int n = 0;
    for (size_t l = 0; l + 1 < P->dims.size(); ++l) n += P->dims[l] * P->dims[l + 1] + P->dims[l + 1];
    return n;
}

static void free_acts(Mlp *P)
{
    for (auto p : P->Z) hipFree(p);
    for (auto p : P->Ab) hipFree(p);
    P->Z.clear();
    P->Ab.clear();
    hipFree(P->dA); hipFree(P->dZf); hipFree(P->out6); hipFree(P->g6); hipFree(P->lossp);
    hipFree(P->dZb); hipFree(P->dAb);
    P->dA = P->dZf = P->out6 = P->g6 = P->lossp = nullptr;
    P->dZb = P->dAb = nullptr;
    P->cap = 0;
}

static int ensure_acts(Mlp *P, int64_t n)
{
    if (P->cap >= n) return SO3FT_OK;
    free_acts(P);
    const int L = (int)P->dims.size() - 1;
    int wmax = 0;
    for (int d : P->dims) wmax = std::max(wmax, d);
    P->Z.assign(L, nullptr);
    P->Ab.assign(L + 1, nullptr);
    for (int l = 0; l < L; ++l) MHIP(hipMalloc((void **)&P->Z[l], sizeof(float) * n * P->dims[l + 1]));
    for (int l = 0; l <= L; ++l) MHIP(hipMalloc((void **)&P->Ab[l], sizeof(uint16_t) * n * P->dims[l]));
    MHIP(hipMalloc((void **)&P->dA, sizeof(float) * n * wmax));
    MHIP(hipMalloc((void **)&P->dZf, sizeof(float) * n * wmax));
    MHIP(hipMalloc((void **)&P->dZb, sizeof(uint16_t) * n * wmax));
    MHIP(hipMalloc((void **)&P->dAb, sizeof(uint16_t) * n * wmax));
    MHIP(hipMalloc((void **)&P->out6, sizeof(float) * n * 6));
    MHIP(hipMalloc((void **)&P->g6, sizeof(float) * n * 6));
    MHIP(hipMalloc((void **)&P->lossp, sizeof(float) * ((n + 255) / 256)));
    P->cap = n;
    return SO3FT_OK;
}

static hipError_t gemm(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn, int M,
                       int N, int K, const float *bias, float *Cf, uint16_t *Cb, int gelu, int accumulate, hipStream_t s)
{
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, M, N, K, bias, Cf, Cb, gelu,
                       accumulate);
    return hipGetLastError();
}

// forward of n samples whose bf16 inputs are in P->Ab[0]; outputs in P->out6 (FP32)
static int forward(Mlp *P, int64_t n, hipStream_t s)
{
    const int L = (int)P->dims.size() - 1;
    for (int l = 0; l < L; ++l) {
        const int K = P->dims[l], N = P->dims[l + 1];
        const bool last = l == L - 1;
        // Z = A W^T + b: A(m,k) = Ab[m K + k], B(k,n) = W[n K + k]
        MHIP(gemm(P->Ab[l], K, 1, P->Wb[l], 1, K, (int)n, N, K, P->b[l], last ? P->out6 : P->Z[l],
                  last ? nullptr : P->Ab[l + 1], last ? 0 : 1, 0, s));
    }
    return SO3FT_OK;
}

}  // namespace so3

using namespace so3;

extern "C" {

const char *so3ft_mlp_last_error(void) { return g_mlp_err.c_str(); }

int so3ft_mlp_create(const int32_t *dims, int nlayers, int64_t *out)
{
    if (!dims || nlayers < 1 || nlayers > 16 || !out) return mfail(SO3FT_EINVAL, "bad layer description");
    if (dims[nlayers] != 6) return mfail(SO3FT_EINVAL, "the last layer must have 6 outputs");
    Mlp *P = new Mlp;
    P->dims.assign(dims, dims + nlayers + 1);
    hipGetDevice(&P->device);
    auto alloc = [&](float **p, int64_t n) -> bool {
        if (hipMalloc((void **)p, sizeof(float) * n) != hipSuccess) return false;
        return hipMemset(*p, 0, sizeof(float) * n) == hipSuccess;
    };
    for (int l = 0; l < nlayers; ++l) {
        const int64_t nw = (int64_t)dims[l] * dims[l + 1], nbias = dims[l + 1];
        if (dims[l] < 1 || dims[l + 1] < 1) { delete P; return mfail(SO3FT_EINVAL, "bad layer width"); }
        float *w, *bb, *wm, *wv, *bm, *bv2, *gw, *gbb;
        uint16_t *wbf;
        if (!alloc(&w, nw) || !alloc(&bb, nbias) || !alloc(&wm, nw) || !alloc(&wv, nw) || !alloc(&bm, nbias) ||
            !alloc(&bv2, nbias) || !alloc(&gw, nw) || !alloc(&gbb, nbias) ||
            hipMalloc((void **)&wbf, sizeof(uint16_t) * nw) != hipSuccess) {
            hipGetLastError();
            delete P;
            return mfail(SO3FT_ENOMEM, "weight allocation failed");
        }
        P->W.push_back(w); P->b.push_back(bb); P->Wm.push_back(wm); P->Wv.push_back(wv);
        P->bm.push_back(bm); P->bv.push_back(bv2); P->gW.push_back(gw); P->gb.push_back(gbb); P->Wb.push_back(wbf);
    }
    std::lock_guard<std::mutex> lk(g_mlp_mu);
    *out = g_mlp_next++;
    g_mlps[*out] = P;
    return SO3FT_OK;
}

int so3ft_mlp_destroy(int64_t h)
{
    Mlp *P;
    {
        std::lock_guard<std::mutex> lk(g_mlp_mu);
        auto it = g_mlps.find(h);
        if (it == g_mlps.end()) return mfail(SO3FT_EHANDLE, "unknown network handle");
        P = it->second;
        g_mlps.erase(it);
    }
    hipDeviceSynchronize();
    free_acts(P);
    for (auto v : {&P->W, &P->b, &P->Wm, &P->Wv, &P->bm, &P->bv, &P->gW, &P->gb})
        for (auto p : *v) hipFree(p);
    for (auto p : P->Wb) hipFree(p);
    delete P;
    return SO3FT_OK;
}

int so3ft_mlp_num_params(int64_t h, int64_t *n)
{
    Mlp *P = mget(h);
    if (!P || !n) return mfail(SO3FT_EHANDLE, "unknown network handle");
    *n = nparams(P);
    return SO3FT_OK;
}

// params: HOST float32, layer by layer [W (out x in, row-major), b (out)]
int so3ft_mlp_set_params(int64_t h, const float *params)
{
    Mlp *P = mget(h);
    if (!P || !params) return mfail(SO3FT_EHANDLE, "unknown network handle");
    int64_t o = 0;
    for (size_t l = 0; l + 1 < P->dims.size(); ++l) {
        const int64_t nw = (int64_t)P->dims[l] * P->dims[l + 1], nbias = P->dims[l + 1];
        MHIP(hipMemcpy(P->W[l], params + o, sizeof(float) * nw, hipMemcpyHostToDevice));
        o += nw;
        MHIP(hipMemcpy(P->b[l], params + o, sizeof(float) * nbias, hipMemcpyHostToDevice));
        o += nbias;
        hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((nw + 255) / 256), dim3(256), 0, 0, P->W[l], P->Wb[l], nw);
        MHIP(hipGetLastError());
        MHIP(hipMemset(P->Wm[l], 0, sizeof(float) * nw));
        MHIP(hipMemset(P->Wv[l], 0, sizeof(float) * nw));
        MHIP(hipMemset(P->bm[l], 0, sizeof(float) * nbias));
        MHIP(hipMemset(P->bv[l], 0, sizeof(float) * nbias));
    }
    P->step = 0;
    MHIP(hipDeviceSynchronize());
    return SO3FT_OK;
}

// which = 0: FP32 master parameters, 1: gradients of the last backward pass
int so3ft_mlp_get_params(int64_t h, int which, float *params)
{
    Mlp *P = mget(h);
    if (!P || !params) return mfail(SO3FT_EHANDLE, "unknown network handle");
    MHIP(hipDeviceSynchronize());
    int64_t o = 0;
    for (size_t l = 0; l + 1 < P->dims.size(); ++l) {
        const int64_t nw = (int64_t)P->dims[l] * P->dims[l + 1], nbias = P->dims[l + 1];
        MHIP(hipMemcpy(params + o, which ? P->gW[l] : P->W[l], sizeof(float) * nw, hipMemcpyDeviceToHost));
        o += nw;
        MHIP(hipMemcpy(params + o, which ? P->gb[l] : P->b[l], sizeof(float) * nbias, hipMemcpyDeviceToHost));
        o += nbias;
    }
    return SO3FT_OK;
}

// x: DEVICE float32 (n, dims[0]); out: DEVICE float32 (n, 9) rotation matrices (row-major),
// out6 (optional, DEVICE float32 (n, 6)) the raw outputs
int so3ft_mlp_forward(int64_t h, const float *x, int64_t n, float *R9, float *out6, void *stream)
{
    Mlp *P = mget(h);
    if (!P) return mfail(SO3FT_EHANDLE, "unknown network handle");
    if (!x || n < 1) return mfail(SO3FT_EINVAL, "bad input");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (int rc = ensure_acts(P, n)) return rc;
    const int64_t nin = n * P->dims[0];
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((nin + 255) / 256), dim3(256), 0, s, x, P->Ab[0], nin);
    MHIP(hipGetLastError());
    if (int rc = forward(P, n, s)) return rc;
    if (R9) {
        hipLaunchKernelGGL(rot6d_kernel, dim3((n + 255) / 256), dim3(256), 0, s, P->out6, (int)n, R9);
        MHIP(hipGetLastError());
    }
    if (out6) MHIP(hipMemcpyAsync(out6, P->out6, sizeof(float) * n * 6, hipMemcpyDeviceToDevice, s));
    return SO3FT_OK;
}

// forward + loss + backward (gradients kept on device); if lr > 0 also one Adam step.
// x: DEVICE (n, dims[0]) float32, T9: DEVICE (n, 9) target rotations; loss: HOST float.
int so3ft_mlp_train_step(int64_t h, const float *x, const float *T9, int64_t n, float lr, float *loss, void *stream)
{
    Mlp *P = mget(h);
    if (!P) return mfail(SO3FT_EHANDLE, "unknown network handle");
    if (!x || !T9 || n < 1) return mfail(SO3FT_EINVAL, "bad input");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (int rc = ensure_acts(P, n)) return rc;
    const int L = (int)P->dims.size() - 1;
    const int64_t nin = n * P->dims[0];
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((nin + 255) / 256), dim3(256), 0, s, x, P->Ab[0], nin);
    MHIP(hipGetLastError());
    if (int rc = forward(P, n, s)) return rc;
    const int nbl = (int)((n + 255) / 256);
    hipLaunchKernelGGL(rot6d_loss_kernel, dim3(nbl), dim3(256), 0, s, P->out6, T9, (int)n, P->g6, P->lossp);
    MHIP(hipGetLastError());
    // backward: the last layer's dZ is g6 (no activation)
    const float *dZf = P->g6;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((n * 6 + 255) / 256), dim3(256), 0, s, P->g6, P->dZb, n * 6);
    MHIP(hipGetLastError());
    for (int l = L - 1; l >= 0; --l) {
        const int K = P->dims[l], N = P->dims[l + 1];
        // dW (N x K) = dZ^T A: A(m = out, k = sample) = dZ[k N + m], B(k = sample, n = in) = Ab[k K + n]
        MHIP(gemm(P->dZb, 1, N, P->Ab[l], K, 1, N, K, (int)n, nullptr, P->gW[l], nullptr, 0, 0, s));
        hipLaunchKernelGGL(colsum_kernel, dim3((N + 63) / 64), dim3(1024), 0, s, dZf, (int)n, N, P->gb[l]);
        MHIP(hipGetLastError());
        if (l == 0) break;
        // dA (n x K) = dZ W: A(m, k) = dZ[m N + k], B(k = out, n = in) = Wb[k K + n]
        MHIP(gemm(P->dZb, N, 1, P->Wb[l], K, 1, (int)n, K, N, nullptr, P->dA, nullptr, 0, 0, s));
        const int64_t cnt = n * K;
        hipLaunchKernelGGL(gelu_bwd_kernel, dim3((cnt + 255) / 256), dim3(256), 0, s, P->dA, P->Z[l - 1], P->dZb,
                           P->dZf, cnt);
        MHIP(hipGetLastError());
        dZf = P->dZf;
    }
    if (lr > 0.f) {
        P->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)P->step), c2 = 1.f - powf(b2, (float)P->step);
        for (int l = 0; l < L; ++l) {
            const int64_t nw = (int64_t)P->dims[l] * P->dims[l + 1], nbias = P->dims[l + 1];
            hipLaunchKernelGGL(adam_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, P->W[l], P->gW[l], P->Wm[l],
                               P->Wv[l], P->Wb[l], nw, lr, b1, b2, 1e-8f, c1, c2);
            hipLaunchKernelGGL(adam_kernel, dim3((nbias + 255) / 256), dim3(256), 0, s, P->b[l], P->gb[l], P->bm[l],
                               P->bv[l], (uint16_t *)nullptr, nbias, lr, b1, b2, 1e-8f, c1, c2);
            MHIP(hipGetLastError());
        }
    }
    if (loss) {
        std::vector<float> part(nbl);
        MHIP(hipMemcpyAsync(part.data(), P->lossp, sizeof(float) * nbl, hipMemcpyDeviceToHost, s));
        MHIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (float v : part) sum += v;
        *loss = (float)(sum / (double)n);
    }
    return SO3FT_OK;
}

// features of n pairs (see align_features_kernel); mean/stdv DEVICE float32 (452) or NULL
int so3ft_align_features(int64_t bw, const void *f_lms, const void *g_lms, int64_t n, const float *mean,
                         const float *stdv, float *x, void *stream)
{
    if (bw < 9 || bw > 64 || !f_lms || !g_lms || !x || n < 0) return mfail(SO3FT_EINVAL, "bad argument (9 <= bw <= 64)");
    if ((mean == nullptr) != (stdv == nullptr)) return mfail(SO3FT_EINVAL, "mean and std go together");
    if (n == 0) return SO3FT_OK;
    hipLaunchKernelGGL(align_features_kernel, dim3(n), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2 *>(f_lms), reinterpret_cast<const double2 *>(g_lms), bw * bw,
                       (int)bw, (int)n, mean, stdv, x);
    MHIP(hipGetLastError());
    return SO3FT_OK;
}

// generic bf16 GEMM (test hook): C (M x N, FP32) = A (M x K) B (K x N), row-major bf16
int so3ft_gemm_bf16(const void *A, const void *B, float *C, int M, int N, int K, void *stream)
{
    if (!A || !B || !C || M < 1 || N < 1 || K < 1) return mfail(SO3FT_EINVAL, "bad argument");
    MHIP(gemm(reinterpret_cast<const uint16_t *>(A), K, 1, reinterpret_cast<const uint16_t *>(B), N, 1, M, N, K,
              nullptr, C, nullptr, 0, 0, reinterpret_cast<hipStream_t>(stream)));
    return SO3FT_OK;
}

}  // extern "C"
