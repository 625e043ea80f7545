// Local correlation refinement: the rotational cross-correlation of pairs (f, g) of SH
// coefficient sets (LMc order l(l+1)+m) evaluated on a 5 x 5 x 5 window of a level grid
// (alpha, gamma steps pi/Bg, beta samples (2k+1) pi/(4Bg)) around a current estimate,
// moving the estimate to the window's max |C|.  C is the same function the full
// correlation samples on its grid (pysofft/soft.f08:2242-2420):
//   C(a, b, g) = (1/2pi) sum_{l,m,n} cc^l_mn d~^l_mn(b) e^{-i m a} e^{-i n g},
//   cc^l_mn = 2 pi sqrt(2/(2l+1)) (-1)^(m+n) f^l_m conj(g^l_n),
// evaluated directly in FP64 for degrees l < Lc.  d~ at the window's betas comes from the
// Kostelec recurrence (pysofft/wigner.f90:227-260) started from the closed form
// d~^q_{p,q}(b) = sqrt((2q+1)/2) sqrt((2q)!/((q+p)!(q-p)!)) cos(b/2)^(q+p) sin(b/2)^(q-p)
// (the seeds of wigner.f90:130-187 for 0 < b < pi), and the 8-variant symmetry table.
//
// One block (768 threads) per pair.  Work items (p, q, variant pair) run the recurrence for the 5 betas
// (or pi - beta for the reversed variants), accumulate F(beta, m, n) over l in registers
// and fold e^{-i n gamma_c} into LDS sums G[beta][m][c]; the alpha sums and the argmax
// follow.  No counterpart in the reference (its callers evaluate the whole grid).
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace so3 {

constexpr int RL_MAX = 64;   // largest degree limit Lc

// recurrence constants of the step l -> l+1 for (p, q), l = q .. RL_MAX-2, at
// off[q(q+1)/2 + p] + (l - q)
__device__ __host__ inline int64_t rl_pair(int p, int q) { return (int64_t)q * (q + 1) / 2 + p; }

__device__ __forceinline__ double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// dynamic LDS: G [W beta][NM m][W gamma] | eg [W gamma][NM n] | ea [W alpha][NM m] (double2)
__global__ __launch_bounds__(768) void refine_kernel(const double2 *__restrict__ f, const double2 *__restrict__ g,
                                                     int64_t nb, int n, const double *__restrict__ ein,
                                                     double *__restrict__ eout, double *__restrict__ score, int Bg,
                                                     int Lc, const double *__restrict__ rc,
                                                     const int64_t *__restrict__ rcoff)
{
    const int s = blockIdx.x;
    if (s >= n) return;
    constexpr int W = 5;
    const int NM = 2 * Lc - 1;
    extern __shared__ double sm[];
    double2 *G = reinterpret_cast<double2 *>(sm);    // [W beta][NM m][W gamma]
    double2 *eg = G + W * NM * W;                     // e^{-i n gamma_c}, [c][n + Lc - 1]
    double2 *ea = eg + W * NM;                        // e^{-i m alpha_a}, [a][m + Lc - 1]
    __shared__ double Cabs[W * W * W];
    __shared__ double wl[RL_MAX];                     // 2 pi sqrt(2/(2l+1))
    __shared__ double lfac[2 * RL_MAX + 1];           // log k!
    __shared__ double lq[RL_MAX];                     // log sqrt((2q+1)/2)
    __shared__ double lcb[2 * W], lsb[2 * W], cbv[2 * W];  // log cos(b/2), log sin(b/2), cos b; b, pi - b
    __shared__ double bet[W];
    __shared__ int win[3][W];                         // grid indices of the window
    const int tid = threadIdx.x;
    const double pi = 3.14159265358979323846;
    const int N2 = 2 * Bg;
    const double h = pi / Bg;
    if (tid == 0) {
        // nearest level-grid sample to the estimate; beta samples are (2k+1) pi/(4Bg)
        const double a0 = ein[3 * (int64_t)s], b0 = ein[3 * (int64_t)s + 1], g0 = ein[3 * (int64_t)s + 2];
        const int ja = (int)llround(a0 / h), jg = (int)llround(g0 / h);
        int kb = (int)llround(b0 / (pi / N2) - 0.5);
        kb = min(max(kb, 2), N2 - 3);   // keep the window inside (0, pi)
        for (int i = 0; i < W; ++i) {
            win[0][i] = ((ja + i - 2) % N2 + N2) % N2;
            win[1][i] = kb + i - 2;
            win[2][i] = ((jg + i - 2) % N2 + N2) % N2;
            bet[i] = (2.0 * win[1][i] + 1.0) * pi / (2.0 * N2);
        }
    }
    __syncthreads();
    for (int i = tid; i < W * NM * W; i += blockDim.x) G[i] = make_double2(0.0, 0.0);
    for (int l = tid; l < Lc; l += blockDim.x) {
        wl[l] = 2.0 * pi * sqrt(2.0 / (2.0 * l + 1.0));
        lq[l] = 0.5 * log((2.0 * l + 1.0) * 0.5);
    }
    for (int k = tid; k <= 2 * Lc; k += blockDim.x) lfac[k] = lgamma((double)k + 1.0);
    if (tid < 2 * W) {
        const double be = tid < W ? bet[tid] : pi - bet[tid - W];
        lcb[tid] = log(cos(0.5 * be));
        lsb[tid] = log(sin(0.5 * be));
        cbv[tid] = cos(be);
    }
    for (int i = tid; i < W * NM; i += blockDim.x) {
        const int c = i / NM, m = i % NM - (Lc - 1);
        const double tg = (double)m * win[2][c] * h, ta = (double)m * win[0][c] * h;
        eg[i] = make_double2(cos(tg), -sin(tg));
        ea[i] = make_double2(cos(ta), -sin(ta));
    }
    __syncthreads();

    const double2 *fs = f + (int64_t)s * nb, *gs = g + (int64_t)s * nb;
    // items: (p <= q < Lc) x 4 variant pairs {0,1} {2,3} {4,5} {6,7}
    const int npq = Lc * (Lc + 1) / 2;
    for (int it = tid; it < 4 * npq; it += blockDim.x) {
        const int vg = it & 3, P = it >> 2;
        int q = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
        while (q * (q + 1) / 2 > P) --q;
        while ((q + 1) * (q + 2) / 2 <= P) ++q;
        const int p = P - q * (q + 1) / 2;
        const Variant v0 = variant(2 * vg, p, q), v1 = variant(2 * vg + 1, p, q);
        if (!v0.valid && !v1.valid) continue;
        const int ro = vg >= 2 ? W : 0;   // reversed variants use pi - beta
        // seeds d~^q_{p,q} at the 5 betas (closed form, log domain)
        const double lc = 0.5 * (lfac[2 * q] - lfac[q + p] - lfac[q - p]) + lq[q];
        double d[W], dp[W];
#pragma unroll
        for (int b = 0; b < W; ++b) {
            d[b] = exp(lc + (double)(q + p) * lcb[ro + b] + (double)(q - p) * lsb[ro + b]);
            dp[b] = 0.0;
        }
        double2 F0[W], F1[W];
#pragma unroll
        for (int b = 0; b < W; ++b) { F0[b] = make_double2(0.0, 0.0); F1[b] = make_double2(0.0, 0.0); }
        const double *cst = rc + rcoff[rl_pair(p, q)];
        const double s0 = (((v0.m + v0.n) & 1) ? -1.0 : 1.0) * v0.sign, s1 = (((v1.m + v1.n) & 1) ? -1.0 : 1.0) * v1.sign;
        // operands of degree l (coefficients of both variants, recurrence constants of the
        // step l -> l+1), loaded one degree ahead of their use
        const int m0 = v0.valid ? v0.m : 0, n0 = v0.valid ? v0.n : 0;
        const int m1 = v1.valid ? v1.m : 0, n1 = v1.valid ? v1.n : 0;
        struct Ops {
            double2 f0, g0, f1, g1;
            double ca, cx, cbb;
        };
        auto load = [&](Ops &o, int l) {
            const int r = l * (l + 1);
            o.f0 = fs[r + m0]; o.g0 = gs[r + n0];
            o.f1 = fs[r + m1]; o.g1 = gs[r + n1];
            const int k = max(0, min(l, Lc - 2) - q);   // the last degree has no step: reread entry 0
            o.ca = cst[3 * k]; o.cx = cst[3 * k + 1]; o.cbb = cst[3 * k + 2];
        };
        Ops nx;
        load(nx, q);
        for (int l = q; l < Lc; ++l) {
            const Ops cu = nx;
            if (l + 1 < Lc) load(nx, l + 1);
            const double wn = wl[l];
            const double sl0 = v0.valid ? wn * s0 * ((v0.lsign && (l & 1)) ? -1.0 : 1.0) : 0.0;
            const double sl1 = v1.valid ? wn * s1 * ((v1.lsign && (l & 1)) ? -1.0 : 1.0) : 0.0;
            const double2 z0 = cmulc(cu.f0, cu.g0), z1 = cmulc(cu.f1, cu.g1);
            const double2 c0 = make_double2(sl0 * z0.x, sl0 * z0.y), c1 = make_double2(sl1 * z1.x, sl1 * z1.y);
#pragma unroll
            for (int b = 0; b < W; ++b) {
                F0[b].x = fma(c0.x, d[b], F0[b].x); F0[b].y = fma(c0.y, d[b], F0[b].y);
                F1[b].x = fma(c1.x, d[b], F1[b].x); F1[b].y = fma(c1.y, d[b], F1[b].y);
            }
            if (l < Lc - 1) {
#pragma unroll
                for (int b = 0; b < W; ++b) {
                    const double dn = (cu.ca * (cbv[ro + b] - cu.cx)) * d[b] - cu.cbb * dp[b];
                    dp[b] = d[b];
                    d[b] = dn;
                }
            }
        }
        // fold e^{-i n gamma_c} into G[b][m][c]
        auto fold = [&](const Variant &vv, const double2 (&Fk)[W]) {
#pragma unroll 1
            for (int c = 0; c < W; ++c) {
                const double2 e = eg[c * NM + vv.n + Lc - 1];
#pragma unroll
                for (int b = 0; b < W; ++b) {
                    const double2 z = Fk[b];
                    double2 *dst = &G[(b * NM + (vv.m + Lc - 1)) * W + c];
                    atomicAdd(&dst->x, z.x * e.x - z.y * e.y);
                    atomicAdd(&dst->y, z.x * e.y + z.y * e.x);
                }
            }
        };
        if (v0.valid) fold(v0, F0);
        if (v1.valid) fold(v1, F1);
    }
    __syncthreads();
    // alpha sums: C(a, b, c) = (1/2pi) sum_m G[b][m][c] e^{-i m alpha_a}
    for (int o = tid; o < W * W * W; o += blockDim.x) {
        const int a = o / (W * W), b = (o / W) % W, c = o % W;
        double re = 0.0, im = 0.0;
        for (int mi = 0; mi < NM; ++mi) {
            const double2 z = G[(b * NM + mi) * W + c], e = ea[a * NM + mi];
            re += z.x * e.x - z.y * e.y;
            im += z.x * e.y + z.y * e.x;
        }
        Cabs[o] = sqrt(re * re + im * im) / (2.0 * pi);
    }
    __syncthreads();
    if (tid == 0) {
        int best = 62;  // window centre
        for (int o = 0; o < W * W * W; ++o)
            if (Cabs[o] > Cabs[best]) best = o;
        const int a = best / (W * W), b = (best / W) % W, c = best % W;
        eout[3 * (int64_t)s] = win[0][a] * h;
        eout[3 * (int64_t)s + 1] = bet[b];
        eout[3 * (int64_t)s + 2] = win[2][c] * h;
        if (score) score[s] = Cabs[best];
    }
}

__global__ void rl_consts_kernel(double *rc, const int64_t *off, int Lmax)
{
    const int P = blockIdx.x;  // pair index q(q+1)/2 + p
    int q = 0;
    while ((q + 1) * (q + 2) / 2 <= P) ++q;
    const int p = P - q * (q + 1) / 2;
    for (int l = q + threadIdx.x; l < Lmax - 1; l += blockDim.x) {
        const int64_t lp = l + 1;
        const int64_t den = (lp * lp - (int64_t)p * p) * (lp * lp - (int64_t)q * q);
        double cc = (double)lp * sqrt(1.0 / (double)den);
        cc = cc * sqrt((double)(2 * l + 3));
        const double a = cc * sqrt((double)(2 * l + 1));
        const double x = (l == 0) ? 0.0 : (double)((int64_t)p * q) / (double)((int64_t)l * (l + 1));
        double b = 0.0;
        if (l != q) {
            double c1 = sqrt((double)(((int64_t)l * l - (int64_t)p * p) * ((int64_t)l * l - (int64_t)q * q))) /
                        (double)l;
            b = cc * c1 / sqrt((double)(2 * l - 1));
        }
        double *e = rc + off[P] + 3 * (l - q);
        e[0] = a; e[1] = x; e[2] = b;
    }
}

struct RlTables {
    double *rc = nullptr;
    int64_t *off = nullptr;
};
std::mutex g_rl_mu;
std::map<int, RlTables> g_rl;   // per device

static int rl_tables(RlTables &t)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SO3FT_EHIP;
    std::lock_guard<std::mutex> lk(g_rl_mu);
    auto it = g_rl.find(dev);
    if (it != g_rl.end()) { t = it->second; return SO3FT_OK; }
    const int npair = RL_MAX * (RL_MAX + 1) / 2;
    std::vector<int64_t> off(npair);
    int64_t tot = 0;
    for (int q = 0; q < RL_MAX; ++q)
        for (int p = 0; p <= q; ++p) {
            off[rl_pair(p, q)] = tot;
            tot += 3 * (RL_MAX - q);
        }
    RlTables nt;
    if (hipMalloc((void **)&nt.rc, sizeof(double) * tot) != hipSuccess ||
        hipMalloc((void **)&nt.off, sizeof(int64_t) * npair) != hipSuccess) {
        hipGetLastError();
        return SO3FT_ENOMEM;
    }
    if (hipMemcpy(nt.off, off.data(), sizeof(int64_t) * npair, hipMemcpyHostToDevice) != hipSuccess) return SO3FT_EHIP;
    hipLaunchKernelGGL(rl_consts_kernel, dim3(npair), dim3(64), 0, 0, nt.rc, nt.off, RL_MAX);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return SO3FT_EHIP;
    g_rl[dev] = nt;
    t = nt;
    return SO3FT_OK;
}

}  // namespace so3

using namespace so3;

extern "C" {

int so3ft_refine_correlation_peaks(int64_t bw, const void *f_lms, const void *g_lms, int64_t n, const double *euler_in,
                                   double *euler_out, double *score, int64_t grid_bw, int64_t lmax, void *stream)
{
    if (bw < 2 || !f_lms || !g_lms || !euler_in || !euler_out || n < 0) return SO3FT_EINVAL;
    if (grid_bw < 4 || grid_bw > 4096 || lmax < 1 || lmax > RL_MAX || lmax > bw) return SO3FT_EINVAL;
    if (n == 0) return SO3FT_OK;
    RlTables t;
    if (int rc = rl_tables(t)) return rc;
    const int NM = 2 * (int)lmax - 1;
    const size_t bytes = sizeof(double2) * 35 * NM;   // G (25 NM) + eg, ea (5 NM each)
    static bool attr = false;
    if (!attr) {
        if (hipFuncSetAttribute((const void *)refine_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024) !=
            hipSuccess)
            return SO3FT_EHIP;
        attr = true;
    }
    hipLaunchKernelGGL(refine_kernel, dim3(n), dim3(768), bytes, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2 *>(f_lms), reinterpret_cast<const double2 *>(g_lms), bw * bw,
                       (int)n, euler_in, euler_out, score, (int)grid_bw, (int)lmax, t.rc, t.off);
    return hipGetLastError() == hipSuccess ? SO3FT_OK : SO3FT_EHIP;
}

}  // extern "C"
