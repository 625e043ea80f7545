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
// Four blocks per pair (quarters of the symmetric variants; see refine_kernel), then a
// finalize kernel for |C| and the argmax.  No counterpart in the reference (its callers
// evaluate the whole grid).
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace so3 {

constexpr int RL_MAX = 64;   // largest degree limit Lc

__device__ __forceinline__ double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

constexpr int RW = 5;        // window points per axis

// Step constants of the Kostelec recurrence (wigner.f90:243-256) for 0 <= p <= q <= l < RL_MAX - 1,
// d^{l+1} = (a cos(b) - a x) d^l - b d^{l-1} with a = A_l / sqrt(P_{l+1}),
// b = B_l sqrt(P_l) / sqrt(P_{l+1}), x = p q / (l (l+1)), P_l = (l^2 - p^2)(l^2 - q^2),
// A_l = (l+1) sqrt((2l+1)(2l+3)), B_l = (l+1) sqrt(2l+3) / (l sqrt(2l-1)); stored as
// (a, -a x, -b) in rows [q][l - q][p] so that the lanes of a wave (consecutive p, one q)
// read consecutive entries.  Degrees l in [q, RL_MAX - 2] per q.
__host__ __device__ inline int64_t rc_base(int q)
{
    // sum over q' < q of (q' + 1)(RL_MAX - 1 - q') = RL_MAX q(q+1)/2 - q(q+1)(2q+1)/6
    const int64_t t = (int64_t)q * (q + 1);
    return RL_MAX * (t / 2) - t * (2 * q + 1) / 6;
}

// window of pair s at grid bandwidth Bg: alpha / gamma grid indices and the beta samples
// (2k+1) pi/(4Bg), centred on the nearest grid sample to the estimate (beta kept inside
// (0, pi))
__device__ inline void refine_window(const double *ein, int64_t s, int Bg, int (&win)[3][RW], double (&bet)[RW])
{
    const double pi = 3.14159265358979323846;
    const int N2 = 2 * Bg;
    const double h = pi / Bg;
    const double a0 = ein[3 * s], b0 = ein[3 * s + 1], g0 = ein[3 * s + 2];
    const int ja = (int)llround(a0 / h), jg = (int)llround(g0 / h);
    int kb = (int)llround(b0 / (pi / N2) - 0.5);
    kb = min(max(kb, 2), N2 - 3);
    for (int i = 0; i < RW; ++i) {
        win[0][i] = ((ja + i - 2) % N2 + N2) % N2;
        win[1][i] = kb + i - 2;
        win[2][i] = ((jg + i - 2) % N2 + N2) % N2;
        bet[i] = (2.0 * win[1][i] + 1.0) * pi / (2.0 * N2);
    }
}

// Four blocks per pair, one per quarter of the 8 symmetric variants of every order pair
// 0 <= p <= q < Lc, chosen so that every variant of a quarter has its "row" index +p or
// -p:
//   quarter 0 (X+): v0 (p, q) at beta,  v4 (p, -q) at pi - beta;  rows m = +p
//   quarter 1 (X-): v3 (-p, -q) at beta, v5 (-p, q) at pi - beta; rows m = -p
//   quarter 2 (Y+): v2 (q, p) at beta,  v7 (-q, p) at pi - beta;  rows n = +p
//   quarter 3 (Y-): v1 (-q, -p) at beta, v6 (q, -p) at pi - beta; rows n = -p
// X quarters fold F(beta, m, n) over n with e^{-i n gamma_c} into R[beta][c][p], Y
// quarters over m with e^{-i m alpha_a} into R[beta][a][p]; the remaining sum over the
// row index follows at the end.  Work items are the order pairs in q-major order: the
// lanes of a wave run recurrences of equal length (Lc - q) and write distinct rows p.
// Each item runs the recurrence at the 5 window betas and at pi - beta (10 chains) and
// accumulates F over l in registers.  The quarter stages in LDS (FP64) the coefficient
// set indexed by +-p on its side only (f for X, g for Y) and the other set in full.
// dynamic LDS: R [RW b][RW k][Lc p] | ep [RW k][2Lc-1] fold phases | hs [Lc(Lc+1)/2]
// (the +-p side set) | fs [Lc^2] (the other set); hs / fs hold at least RW Lc entries:
// the row phases of the last phase reuse them
__global__ __launch_bounds__(512) void refine_kernel(const double2 *__restrict__ f, const double2 *__restrict__ g,
                                                     int64_t nb, int n, const double *__restrict__ ein, int Bg, int Lc,
                                                     double2 *__restrict__ part, const double *__restrict__ rca,
                                                     const double *__restrict__ rcx, const double *__restrict__ rcb)
{
    const int s = blockIdx.x >> 2, quarter = blockIdx.x & 3;
    if (s >= n) return;
    const int ytype = quarter >> 1, neg = quarter & 1;
    constexpr int W = RW;
    const int NM = 2 * Lc - 1;
    extern __shared__ double sm[];
    double2 *R = reinterpret_cast<double2 *>(sm);     // [b][k][p]: p fastest, so the lanes' rows p hit distinct banks
    double2 *ep = R + W * Lc * W;                      // [k][j + Lc - 1]: e^{-i j gamma_k} (X) / e^{-i j alpha_k} (Y)
    double2 *hs = ep + W * NM;                         // side set at +-p: [l(l+1)/2 + p]
    double2 *fs = hs + Lc * (Lc + 1) / 2;              // other set: [l(l+1) + j]
    __shared__ double wl[RL_MAX];                      // 2 pi sqrt(2/(2l+1))
    __shared__ double lfac[2 * RL_MAX + 1];            // log k!
    __shared__ double lq[RL_MAX];                      // log sqrt((2q+1)/2)
    __shared__ double lcb[2 * W], lsb[2 * W], cbv[2 * W];  // log cos(b/2), log sin(b/2), cos b; b, pi - b
    __shared__ double bet[W];
    __shared__ int win[3][W];
    __shared__ int qnext;                              // next unclaimed work item (wave queue)
    const int tid = threadIdx.x;
    if (tid == 0) qnext = 0;
    const double pi = 3.14159265358979323846;
    const double h = pi / Bg;
    if (tid == 0) {
        int wi[3][W];
        double be[W];
        refine_window(ein, s, Bg, wi, be);
        for (int i = 0; i < W; ++i) {
            win[0][i] = wi[0][i]; win[1][i] = wi[1][i]; win[2][i] = wi[2][i];
            bet[i] = be[i];
        }
    }
    const double2 *side = (ytype ? g : f) + (int64_t)s * nb, *other = (ytype ? f : g) + (int64_t)s * nb;
    for (int i = tid; i < W * Lc * W; i += blockDim.x) R[i] = make_double2(0.0, 0.0);
    for (int i = tid; i < Lc * Lc; i += blockDim.x) fs[i] = other[i];
    for (int i = tid; i < Lc * (Lc + 1) / 2; i += blockDim.x) {
        int l = (int)((sqrt(8.0 * i + 1.0) - 1.0) * 0.5);
        while (l * (l + 1) / 2 > i) --l;
        while ((l + 1) * (l + 2) / 2 <= i) ++l;
        const int mp = i - l * (l + 1) / 2;
        hs[i] = side[l * (l + 1) + (neg ? -mp : mp)];
    }
    for (int l = tid; l < Lc; l += blockDim.x) {
        wl[l] = 2.0 * pi * sqrt(2.0 / (2.0 * l + 1.0));
        lq[l] = 0.5 * log((2.0 * l + 1.0) * 0.5);
    }
    for (int k = tid; k <= 2 * Lc; k += blockDim.x) lfac[k] = lgamma((double)k + 1.0);
    __syncthreads();
    if (tid < 2 * W) {
        const double b = tid < W ? bet[tid] : pi - bet[tid - W];
        lcb[tid] = log(cos(0.5 * b));
        lsb[tid] = log(sin(0.5 * b));
        cbv[tid] = cos(b);
    }
    for (int i = tid; i < W * NM; i += blockDim.x) {
        const int k = i / NM, j = i % NM - (Lc - 1);
        const double t = (double)j * (ytype ? win[0][k] : win[2][k]) * h;
        ep[i] = make_double2(cos(t), -sin(t));
    }
    __syncthreads();

    // variant at beta (va) and at pi - beta (vb) of this quarter
    const int va = ytype ? (neg ? 1 : 2) : (neg ? 3 : 0);
    const int vb = ytype ? (neg ? 6 : 7) : (neg ? 5 : 4);
    const double cb[2 * W] = {cbv[0], cbv[1], cbv[2], cbv[3], cbv[4], cbv[5], cbv[6], cbv[7], cbv[8], cbv[9]};
    const int npq = Lc * (Lc + 1) / 2;
    // items in q-major order, i.e. by decreasing recurrence length; each wave claims the next
    // 64 from a queue in LDS, so the waves end together (longest items first) instead of the
    // waves holding the long items setting the block's time
    auto item = [&](int P) {
        int q = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
        while (q * (q + 1) / 2 > P) --q;
        while ((q + 1) * (q + 2) / 2 <= P) ++q;
        const int p = P - q * (q + 1) / 2;
        const Variant v0 = variant(va, p, q), v1 = variant(vb, p, q);
        if (!v0.valid && !v1.valid) return;
        // seeds d~^q_{p,q} at beta and pi - beta (closed form, log domain)
        const double lc = 0.5 * (lfac[2 * q] - lfac[q + p] - lfac[q - p]) + lq[q];
        double d[2 * W], dp[2 * W];
#pragma unroll
        for (int bb = 0; bb < 2 * W; ++bb) {
            d[bb] = exp(lc + (double)(q + p) * lcb[bb] + (double)(q - p) * lsb[bb]);
            dp[bb] = 0.0;
        }
        double2 F0[W], F1[W];
#pragma unroll
        for (int bb = 0; bb < W; ++bb) { F0[bb] = make_double2(0.0, 0.0); F1[bb] = make_double2(0.0, 0.0); }
        const double s0 = v0.valid ? (((v0.m + v0.n) & 1) ? -1.0 : 1.0) * v0.sign : 0.0;
        const double s1 = v1.valid ? (((v1.m + v1.n) & 1) ? -1.0 : 1.0) * v1.sign : 0.0;
        // the other set's index of each variant (n for X, m for Y; +-q); the side index is +-p
        const int j0 = v0.valid ? (ytype ? v0.m : v0.n) : 0, j1 = v1.valid ? (ytype ? v1.m : v1.n) : 0;
        // the step constants of (p, q): table row of degree l at rc_row(q, l) + p
        const int64_t rq = rc_base(q);
        struct Ops {
            double2 hv, o0, o1;
            double ca, cx, cb;   // step l -> l+1: a, -a x, -b (degree l < Lc - 1)
        };
        auto load = [&](Ops &o, int l) {
            const int rg = l * (l + 1);
            o.hv = hs[rg / 2 + p];
            o.o0 = fs[rg + j0];
            o.o1 = fs[rg + j1];
            const bool st = l < Lc - 1;   // degree Lc - 1 takes no step (and has no row)
            const int64_t i = st ? rq + (int64_t)(l - q) * (q + 1) + p : 0;
            const double a = rca[i], ax = rcx[i], b = rcb[i];
            o.ca = st ? a : 0.0;
            o.cx = st ? ax : 0.0;
            o.cb = st ? b : 0.0;
        };
        // one degree: F += cc^l d^l, then d^{l+1}
        auto step = [&](const Ops &cu, int l) {
            const double wn = wl[l];
            const double sl0 = wn * s0 * ((v0.lsign && (l & 1)) ? -1.0 : 1.0);
            const double sl1 = wn * s1 * ((v1.lsign && (l & 1)) ? -1.0 : 1.0);
            // f^l_m conj(g^l_n): X (side f) hv conj(o), Y (side g) o conj(hv)
            const double2 z0 = ytype ? cmulc(cu.o0, cu.hv) : cmulc(cu.hv, cu.o0);
            const double2 z1 = ytype ? cmulc(cu.o1, cu.hv) : cmulc(cu.hv, cu.o1);
            const double2 c0 = make_double2(sl0 * z0.x, sl0 * z0.y), c1 = make_double2(sl1 * z1.x, sl1 * z1.y);
#pragma unroll
            for (int bb = 0; bb < W; ++bb) {
                F0[bb].x = fma(c0.x, d[bb], F0[bb].x); F0[bb].y = fma(c0.y, d[bb], F0[bb].y);
                F1[bb].x = fma(c1.x, d[W + bb], F1[bb].x); F1[bb].y = fma(c1.y, d[W + bb], F1[bb].y);
            }
            if (l < Lc - 1) {
#pragma unroll
                for (int bb = 0; bb < 2 * W; ++bb) {
                    const double k = fma(cu.ca, cb[bb], cu.cx);
                    const double dn = fma(k, d[bb], cu.cb * dp[bb]);
                    dp[bb] = d[bb];
                    d[bb] = dn;
                }
            }
        };
        // two operand sets in turn: the loads of degree l+1 are in flight during degree l
        Ops A, B;
        load(A, q);
        int l = q;
        for (; l + 1 < Lc; l += 2) {
            load(B, l + 1);
            step(A, l);
            if (l + 2 < Lc) load(A, l + 2);
            step(B, l + 1);
        }
        if (l < Lc) step(A, l);
        // fold the other index's phase into R[b][k][p] (distinct p across the lanes of one q)
        auto fold = [&](int jj, const double2 (&Fk)[W]) {
#pragma unroll 1
            for (int k = 0; k < W; ++k) {
                const double2 e = ep[k * NM + jj + Lc - 1];
#pragma unroll
                for (int bb = 0; bb < W; ++bb) {
                    const double2 z = Fk[bb];
                    double2 *dst = &R[(bb * W + k) * Lc + p];
                    atomicAdd(&dst->x, z.x * e.x - z.y * e.y);
                    atomicAdd(&dst->y, z.x * e.y + z.y * e.x);
                }
            }
        };
        if (v0.valid) fold(j0, F0);
        if (v1.valid) fold(j1, F1);
    };
    const int lane = tid & 63;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&qnext, 64);
        base = __shfl(base, 0, 64);
        if (base >= npq) break;
        if (base + lane < npq) item(base + lane);
    }
    __syncthreads();
    // the row sums: X quarters sum_p R[b][c][p] e^{-i m alpha_a} (m = +-p), Y quarters
    // sum_p R[b][a][p] e^{-i n gamma_c} (n = +-p); the row phases reuse the hs space
    double2 *er = hs;   // [k'][p]
    const double sgn = neg ? 1.0 : -1.0;
    for (int i = tid; i < W * Lc; i += blockDim.x) {
        const int k = i / Lc, pr = i % Lc;
        double sn, cs;
        sincos((double)pr * (ytype ? win[2][k] : win[0][k]) * h, &sn, &cs);
        er[i] = make_double2(cs, sgn * sn);
    }
    __syncthreads();
    for (int o = tid; o < W * W * W; o += blockDim.x) {
        const int a = o / (W * W), b = (o / W) % W, c = o % W;
        const int kf = ytype ? a : c, kr = ytype ? c : a;   // fold index in R, row-phase index
        double re = 0.0, im = 0.0;
        for (int pr = 0; pr < Lc; ++pr) {
            const double2 z = R[(b * W + kf) * Lc + pr], e = er[kr * Lc + pr];
            re += z.x * e.x - z.y * e.y;
            im += z.x * e.y + z.y * e.x;
        }
        part[((int64_t)s * 4 + quarter) * (W * W * W) + o] = make_double2(re, im);
    }
}

// |C| on the window from the four quarters, argmax (ties: the window centre, then the first
// C-order index), euler_out / score
__global__ void refine_finalize(const double2 *__restrict__ part, int n, const double *__restrict__ ein, int Bg,
                                double *__restrict__ eout, double *__restrict__ score)
{
    constexpr int W = RW;
    const int s = blockIdx.x;
    if (s >= n) return;
    __shared__ double Cabs[W * W * W];
    const double pi = 3.14159265358979323846;
    for (int o = threadIdx.x; o < W * W * W; o += blockDim.x) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < 4; ++k) {
            const double2 u = part[((int64_t)s * 4 + k) * (W * W * W) + o];
            re += u.x;
            im += u.y;
        }
        Cabs[o] = sqrt(re * re + im * im) / (2.0 * pi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int win[3][W];
        double bet[W];
        refine_window(ein, s, Bg, win, bet);
        int best = 62;  // window centre
        for (int o = 0; o < W * W * W; ++o)
            if (Cabs[o] > Cabs[best]) best = o;
        const int a = best / (W * W), b = (best / W) % W, c = best % W;
        const double h = pi / Bg;
        eout[3 * (int64_t)s] = win[0][a] * h;
        eout[3 * (int64_t)s + 1] = bet[b];
        eout[3 * (int64_t)s + 2] = win[2][c] * h;
        if (score) score[s] = Cabs[best];
    }
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
    if (n > (1 << 29)) return SO3FT_EINVAL;
    const int Lc = (int)lmax, NM = 2 * Lc - 1;
    // the alpha table [RW][Lc] reuses the f / g space after the main phase
    const size_t fg = std::max<size_t>((size_t)Lc * (Lc + 1) / 2 + (size_t)Lc * Lc, (size_t)RW * Lc);
    const size_t bytes = sizeof(double2) * ((size_t)RW * Lc * RW + RW * NM + fg);
    static bool attr = false;
    if (!attr) {
        if (hipFuncSetAttribute((const void *)refine_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024) !=
            hipSuccess)
            return SO3FT_EHIP;
        attr = true;
    }
    if (bytes > 150 * 1024) return SO3FT_EINVAL;
    // the four quarters' window sums, per device (grown on demand; calls on one device share it)
    double2 *part = nullptr;
    {
        static std::mutex mu;
        static std::map<int, std::pair<double2 *, int64_t>> cache;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return SO3FT_EHIP;
        std::lock_guard<std::mutex> lk(mu);
        auto &e = cache[dev];
        if (e.second < n) {
            if (e.first) {
                if (hipDeviceSynchronize() != hipSuccess) return SO3FT_EHIP;
                hipFree(e.first);
            }
            e.first = nullptr;
            e.second = 0;
            if (hipMalloc((void **)&e.first, sizeof(double2) * 4 * RW * RW * RW * (size_t)n) != hipSuccess) {
                hipGetLastError();
                return SO3FT_ENOMEM;
            }
            e.second = n;
        }
        part = e.first;
    }
    // the step-constant table of this device, built once
    const double *rca = nullptr, *rcx = nullptr, *rcb = nullptr;
    {
        static std::mutex mu;
        static std::map<int, double *> tabs;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return SO3FT_EHIP;
        std::lock_guard<std::mutex> lk(mu);
        double *&t = tabs[dev];
        const int64_t ne = rc_base(RL_MAX);
        if (!t) {
            std::vector<double> h(3 * ne);
            for (int q = 0; q < RL_MAX; ++q)
                for (int l = q; l < RL_MAX - 1; ++l)
                    for (int p = 0; p <= q; ++p) {
                        const double dl = l, L1 = (double)(l + 1) * (l + 1), L0 = (double)l * l;
                        const double pp = (double)p * p, qq = (double)q * q;
                        const double A = (dl + 1.0) * std::sqrt((2.0 * dl + 1.0) * (2.0 * dl + 3.0));
                        const double B = l == 0 ? 0.0 : (dl + 1.0) * std::sqrt(2.0 * dl + 3.0) / (dl * std::sqrt(2.0 * dl - 1.0));
                        const double X = l == 0 ? 0.0 : (double)p * q / (dl * (dl + 1.0));
                        const double r = 1.0 / std::sqrt((L1 - pp) * (L1 - qq));
                        const double sP = l == q ? 0.0 : std::sqrt((L0 - pp) * (L0 - qq));
                        const int64_t i = rc_base(q) + (int64_t)(l - q) * (q + 1) + p;
                        h[i] = A * r;
                        h[ne + i] = -A * r * X;
                        h[2 * ne + i] = -B * sP * r;
                    }
            if (hipMalloc((void **)&t, sizeof(double) * 3 * ne) != hipSuccess) {
                hipGetLastError();
                t = nullptr;
                return SO3FT_ENOMEM;
            }
            if (hipMemcpy(t, h.data(), sizeof(double) * 3 * ne, hipMemcpyHostToDevice) != hipSuccess) {
                hipFree(t);
                t = nullptr;
                return SO3FT_EHIP;
            }
        }
        rca = t;
        rcx = t + ne;
        rcb = t + 2 * ne;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // one thread per order pair up to 512 (small degree limits leave room for more blocks per CU)
    const int npq = Lc * (Lc + 1) / 2;
    const int threads = std::min(512, std::max(64, (npq + 63) / 64 * 64));
    hipLaunchKernelGGL(refine_kernel, dim3(4 * n), dim3(threads), bytes, st, reinterpret_cast<const double2 *>(f_lms),
                       reinterpret_cast<const double2 *>(g_lms), bw * bw, (int)n, euler_in, (int)grid_bw, Lc, part,
                       rca, rcx, rcb);
    if (hipGetLastError() != hipSuccess) return SO3FT_EHIP;
    hipLaunchKernelGGL(refine_finalize, dim3(n), dim3(128), 0, st, part, (int)n, euler_in, (int)grid_bw, euler_out,
                       score);
    return hipGetLastError() == hipSuccess ? SO3FT_OK : SO3FT_EHIP;
}

}  // extern "C"
