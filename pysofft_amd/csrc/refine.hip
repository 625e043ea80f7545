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
// Two blocks per pair (halves of the symmetric variants; see refine_kernel), then a
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
// (a, -a x, -b) in entries [q][l - q][p] so that the lanes of a wave (consecutive p, one q)
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

// Two blocks per pair, one per half of the 8 symmetric variants of every order pair
// 0 <= p <= q < Lc.  Each half holds the variants whose "row" index is the first (X) or the
// second (Y) index, at +p or -p:
//   X: v0 (p, q), v3 (-p, -q) at beta; v4 (p, -q), v5 (-p, q) at pi - beta; rows m = +-p
//   Y: v2 (q, p), v1 (-q, -p) at beta; v7 (-q, p), v6 (q, -p) at pi - beta; rows n = +-p
// so the four variants of an item share the recurrence chains (5 betas and pi - beta) and
// read the other set at +-q only.  X halves fold F(beta, m, n) over n with e^{-i n gamma_c}
// into R+-[beta][c][p] (rows m = +-p), Y halves over m with e^{-i m alpha_a} into
// R+-[beta][a][p]; the remaining sum over the row index follows at the end.  Work items are
// the order pairs in q-major order: the lanes of a wave run recurrences of about equal length
// (Lc - q) and write distinct rows p.  Each item accumulates F over l in registers, reading
// the coefficient sets from global memory (L2; the pair's two blocks share them).
// dynamic LDS: R+, R- [RW b][RW k][Lc p] (real, imaginary arrays) | ep [RW k][2Lc-1] fold
// phases | er [2][RW k'][Lc p] row phases: 72 KB at Lc = 64, two 256-thread blocks per CU
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void refine_kernel(
    const double2 *__restrict__ f, const double2 *__restrict__ g, int64_t nb, int n, const double *__restrict__ ein,
    int Bg, int Lc, double2 *__restrict__ part, const double *__restrict__ rct, const double *__restrict__ kt)
{
    const int s = blockIdx.x >> 1, ytype = blockIdx.x & 1;
    if (s >= n) return;
    constexpr int W = RW;
    const int NM = 2 * Lc - 1, NR = W * Lc * W;
    extern __shared__ double sm[];
    // R+- [b][k][p] as real and imaginary arrays, p fastest: the lanes' 8-byte atomics on
    // consecutive p fill the banks without conflicts
    double *Rr = sm, *Ri = sm + NR;                    // rows +p; rows -p at + 2 NR
    double2 *ep = reinterpret_cast<double2 *>(sm + 4 * NR);   // [k][j + Lc - 1]: e^{-i j gamma_k} (X) / e^{-i j alpha_k} (Y)
    double2 *er = ep + W * NM;                         // [sign][k'][p]: row phases of the last phase
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
    for (int i = tid; i < 4 * NR; i += blockDim.x) sm[i] = 0.0;
    const double *wlg = kt + 3 * RL_MAX + 1;   // w_l = 2 pi sqrt(2/(2l+1))
    for (int k = tid; k <= 2 * Lc; k += blockDim.x) lfac[k] = kt[k];
    for (int l = tid; l < Lc; l += blockDim.x) lq[l] = kt[2 * RL_MAX + 1 + l];
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

    // the half's variants at beta (a: rows +p, c: rows -p) and at pi - beta (b: +p, d: -p)
    const int va = ytype ? 2 : 0, vc = ytype ? 1 : 3, vb = ytype ? 7 : 4, vd = ytype ? 6 : 5;
    const double cb[2 * W] = {cbv[0], cbv[1], cbv[2], cbv[3], cbv[4], cbv[5], cbv[6], cbv[7], cbv[8], cbv[9]};
    const int npq = Lc * (Lc + 1) / 2;
    // items in q-major order, i.e. by decreasing recurrence length; each wave claims the next
    // 64 from a queue in LDS, so the waves end together (longest items first) instead of the
    // waves holding the long items setting the block's time.  The lanes of a wave run the same
    // number of degree steps, that of its longest chain (lane 0, the smallest q): a lane whose
    // chain has ended takes w_l = 0 and zero step constants, so its sums stay put with no
    // branch in the loop.
    const int lane = tid & 63;
    const uint32_t zi = (uint32_t)rc_base(RL_MAX);   // the zero entry of the step-constant table
    auto item = [&](int P, bool live, int qmin) {
        int q = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
        while (q * (q + 1) / 2 > P) --q;
        while ((q + 1) * (q + 2) / 2 <= P) ++q;
        const int p = P - q * (q + 1) / 2;
        // seeds d~^q_{p,q} at beta and pi - beta (closed form, log domain)
        const double lc = 0.5 * (lfac[2 * q] - lfac[q + p] - lfac[q - p]) + lq[q];
        double d[2 * W], dp[2 * W];
#pragma unroll
        for (int bb = 0; bb < 2 * W; ++bb) {
            d[bb] = exp(lc + (double)(q + p) * lcb[bb] + (double)(q - p) * lsb[bb]);
            dp[bb] = 0.0;
        }
        // F of the variants a, c (at beta) and b, d (at pi - beta)
        double2 Fa[W], Fb[W], Fc[W], Fd[W];
#pragma unroll
        for (int bb = 0; bb < W; ++bb)
            Fa[bb] = Fb[bb] = Fc[bb] = Fd[bb] = make_double2(0.0, 0.0);
        // signs (-1)^(m+n) x variant sign, with the variant's (-1)^l for the even (.0) and the
        // odd (.1) steps; zero for an invalid variant or a padding lane
        auto sgn2 = [&](int v, double (&o)[2]) {
            const Variant vr = variant(v, p, q);
            const double s0 = (live && vr.valid) ? (((vr.m + vr.n) & 1) ? -1.0 : 1.0) * vr.sign : 0.0;
            o[0] = (vr.lsign && (q & 1)) ? -s0 : s0;
            o[1] = vr.lsign ? -o[0] : o[0];
        };
        double sa[2], sb[2], sc[2], sd[2];
        sgn2(va, sa);
        sgn2(vb, sb);
        sgn2(vc, sc);
        sgn2(vd, sd);
        // the other set at +q (variants a, d) and -q (b, c); the side set at +p (a, b), -p (c, d)
        struct Ops {
            double2 hp, hm, oq, om;
            double ca, cx, cb;   // step l -> l+1: a, -a x, -b (zero from degree Lc - 1 on)
        };
        // the step constants of (p, q): degree l at entry rc_base(q) + (l - q)(q + 1) + p
        const uint32_t rq = (uint32_t)(rc_base(q) + p), rs = (uint32_t)(q + 1);
        // operands of step t (degree q + t): the side terms times w_l, both sets read at degree
        // min(l, Lc - 1)
        auto load = [&](Ops &o, int t) {
            const uint32_t l = (uint32_t)(q + t);
            const uint32_t lb = min(l, (uint32_t)(Lc - 1));
            const uint32_t rg = lb * (lb + 1);
            const double2 zp = side[rg + p], zm = side[rg - p];
            const double wn = l < (uint32_t)Lc ? wlg[lb] : 0.0;   // zero once the chain has ended
            o.hp = make_double2(wn * zp.x, wn * zp.y);
            o.hm = make_double2(wn * zm.x, wn * zm.y);
            o.oq = other[rg + q];
            o.om = other[rg - q];
            const uint32_t i = l < (uint32_t)(Lc - 1) ? rq + (uint32_t)t * rs : zi;
            const double *e = rct + 3 * (size_t)i;
            o.ca = e[0];
            o.cx = e[1];
            o.cb = e[2];
        };
        // one degree: F += cc^l d^l (side x conj(other); the Y halves' conjugate is taken in
        // the fold), then d^{l+1}
        auto step = [&](const Ops &cu, int par) {
            const double2 za = cmulc(cu.hp, cu.oq), zb = cmulc(cu.hp, cu.om);
            const double2 zc = cmulc(cu.hm, cu.om), zd = cmulc(cu.hm, cu.oq);
            const double2 ka = make_double2(sa[par] * za.x, sa[par] * za.y);
            const double2 kb = make_double2(sb[par] * zb.x, sb[par] * zb.y);
            const double2 kc = make_double2(sc[par] * zc.x, sc[par] * zc.y);
            const double2 kd = make_double2(sd[par] * zd.x, sd[par] * zd.y);
#pragma unroll
            for (int bb = 0; bb < W; ++bb) {
                Fa[bb].x = fma(ka.x, d[bb], Fa[bb].x); Fa[bb].y = fma(ka.y, d[bb], Fa[bb].y);
                Fc[bb].x = fma(kc.x, d[bb], Fc[bb].x); Fc[bb].y = fma(kc.y, d[bb], Fc[bb].y);
                Fb[bb].x = fma(kb.x, d[W + bb], Fb[bb].x); Fb[bb].y = fma(kb.y, d[W + bb], Fb[bb].y);
                Fd[bb].x = fma(kd.x, d[W + bb], Fd[bb].x); Fd[bb].y = fma(kd.y, d[W + bb], Fd[bb].y);
            }
#pragma unroll
            for (int bb = 0; bb < 2 * W; ++bb) {
                const double k = fma(cu.ca, cb[bb], cu.cx);
                const double dn = fma(k, d[bb], cu.cb * dp[bb]);
                dp[bb] = d[bb];
                d[bb] = dn;
            }
        };
        // the wave's step count, even; two operand sets in turn: the loads of step t+1 are in
        // flight during step t (the barriers keep them there)
        const int nst = (Lc - qmin + 1) & ~1;
        Ops A, B;
        load(A, 0);
#pragma unroll 1
        for (int t = 0; t < nst; t += 2) {
            load(B, t + 1);
            __builtin_amdgcn_sched_barrier(0);
            step(A, 0);
            __builtin_amdgcn_sched_barrier(0);
            load(A, t + 2);
            __builtin_amdgcn_sched_barrier(0);
            step(B, 1);
            __builtin_amdgcn_sched_barrier(0);
        }
        // fold the other index's phase into R+-[b][k][p] (distinct p across the lanes of one
        // q): a and b into the rows +p, c and d into the rows -p
        if (live) {
            const double cy = ytype ? -1.0 : 1.0;   // the Y halves' conjugate
#pragma unroll 1
            for (int k = 0; k < W; ++k) {
                const double2 eq = ep[k * NM + q + Lc - 1], em = ep[k * NM - q + Lc - 1];
#pragma unroll
                for (int bb = 0; bb < W; ++bb) {
                    const double2 za = make_double2(Fa[bb].x, cy * Fa[bb].y);
                    const double2 zb = make_double2(Fb[bb].x, cy * Fb[bb].y);
                    const double2 zc = make_double2(Fc[bb].x, cy * Fc[bb].y);
                    const double2 zd = make_double2(Fd[bb].x, cy * Fd[bb].y);
                    const int o = (bb * W + k) * Lc + p;
                    atomicAdd(&Rr[o], (za.x * eq.x - za.y * eq.y) + (zb.x * em.x - zb.y * em.y));
                    atomicAdd(&Ri[o], (za.x * eq.y + za.y * eq.x) + (zb.x * em.y + zb.y * em.x));
                    atomicAdd(&Rr[2 * NR + o], (zc.x * em.x - zc.y * em.y) + (zd.x * eq.x - zd.y * eq.y));
                    atomicAdd(&Ri[2 * NR + o], (zc.x * em.y + zc.y * em.x) + (zd.x * eq.y + zd.y * eq.x));
                }
            }
        }
    };
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&qnext, 64);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));   // the same in every lane
        if (base >= npq) break;
        // padding lanes past the last item repeat it with zero signs
        const bool live = base + lane < npq;
        const int P = live ? base + lane : npq - 1;
        int q0 = (int)((sqrt(8.0 * base + 1.0) - 1.0) * 0.5);
        while (q0 * (q0 + 1) / 2 > base) --q0;
        while ((q0 + 1) * (q0 + 2) / 2 <= base) ++q0;
        item(P, live, q0);
    }
    __syncthreads();
    // the row sums: X halves sum_p R+-[b][c][p] e^{-i m alpha_a} (m = +-p), Y halves
    // sum_p R+-[b][a][p] e^{-i n gamma_c} (n = +-p)
    for (int i = tid; i < 2 * W * Lc; i += blockDim.x) {
        const int sg = i / (W * Lc), k = (i / Lc) % W, pr = i % Lc;
        double sn, cs;
        sincos((double)pr * (ytype ? win[2][k] : win[0][k]) * h, &sn, &cs);
        er[i] = make_double2(cs, sg ? sn : -sn);
    }
    __syncthreads();
    for (int o = tid; o < W * W * W; o += blockDim.x) {
        const int a = o / (W * W), b = (o / W) % W, c = o % W;
        const int kf = ytype ? a : c, kr = ytype ? c : a;   // fold index in R, row-phase index
        double re = 0.0, im = 0.0;
        for (int sg = 0; sg < 2; ++sg)
            for (int pr = 0; pr < Lc; ++pr) {
                const int ro = 2 * sg * NR + (b * W + kf) * Lc + pr;
                const double2 z = make_double2(Rr[ro], Ri[ro]), e = er[(sg * W + kr) * Lc + pr];
                re += z.x * e.x - z.y * e.y;
                im += z.x * e.y + z.y * e.x;
            }
        part[((int64_t)s * 2 + ytype) * (W * W * W) + o] = make_double2(re, im);
    }
}

// |C| on the window from the two halves, argmax (ties: the window centre, then the first
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
        for (int k = 0; k < 2; ++k) {
            const double2 u = part[((int64_t)s * 2 + k) * (W * W * W) + o];
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
    // R+- (2 x RW Lc RW complex), fold phases RW (2Lc - 1), row phases 2 x RW Lc
    const size_t bytes = sizeof(double2) * (2 * (size_t)RW * Lc * RW + RW * NM + 2 * (size_t)RW * Lc);
    static bool attr = false;
    if (!attr) {
        if (hipFuncSetAttribute((const void *)refine_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024) !=
            hipSuccess)
            return SO3FT_EHIP;
        attr = true;
    }
    if (bytes > 150 * 1024) return SO3FT_EINVAL;
    // the two halves' window sums, per device (grown on demand; calls on one device share it)
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
            if (hipMalloc((void **)&e.first, sizeof(double2) * 2 * RW * RW * RW * (size_t)n) != hipSuccess) {
                hipGetLastError();
                return SO3FT_ENOMEM;
            }
            e.second = n;
        }
        part = e.first;
    }
    // the constant tables of this device, built once: step constants (a, -a x, -b) of every
    // step, then one zero entry (index rc_base(RL_MAX)); then log k! (k <= 2 RL_MAX),
    // log sqrt((2l+1)/2) and w_l = 2 pi sqrt(2/(2l+1)) (l < RL_MAX)
    const double *rct = nullptr, *kt = nullptr;
    {
        static std::mutex mu;
        static std::map<int, double *> tabs;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return SO3FT_EHIP;
        std::lock_guard<std::mutex> lk(mu);
        double *&t = tabs[dev];
        const int64_t ne = rc_base(RL_MAX) + 1;
        const int64_t nt = 3 * ne + 4 * RL_MAX + 1;
        if (!t) {
            std::vector<double> h(nt, 0.0);
            for (int k = 0; k <= 2 * RL_MAX; ++k) h[3 * ne + k] = std::lgamma((double)k + 1.0);
            for (int l = 0; l < RL_MAX; ++l) {
                h[3 * ne + 2 * RL_MAX + 1 + l] = 0.5 * std::log((2.0 * l + 1.0) * 0.5);
                h[3 * ne + 3 * RL_MAX + 1 + l] = 2.0 * M_PI * std::sqrt(2.0 / (2.0 * l + 1.0));
            }
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
                        h[3 * i] = A * r;
                        h[3 * i + 1] = -A * r * X;
                        h[3 * i + 2] = -B * sP * r;
                    }
            if (hipMalloc((void **)&t, sizeof(double) * nt) != hipSuccess) {
                hipGetLastError();
                t = nullptr;
                return SO3FT_ENOMEM;
            }
            if (hipMemcpy(t, h.data(), sizeof(double) * nt, hipMemcpyHostToDevice) != hipSuccess) {
                hipFree(t);
                t = nullptr;
                return SO3FT_EHIP;
            }
        }
        rct = t;
        kt = t + 3 * ne;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // one thread per order pair up to 512 (small degree limits leave room for more blocks per CU)
    const int npq = Lc * (Lc + 1) / 2;
    const int threads = std::min(256, std::max(64, (npq + 63) / 64 * 64));
    hipLaunchKernelGGL(refine_kernel, dim3(2 * n), dim3(threads), bytes, st, reinterpret_cast<const double2 *>(f_lms),
                       reinterpret_cast<const double2 *>(g_lms), bw * bw, (int)n, euler_in, (int)grid_bw, Lc, part,
                       rct, kt);
    if (hipGetLastError() != hipSuccess) return SO3FT_EHIP;
    hipLaunchKernelGGL(refine_finalize, dim3(n), dim3(128), 0, st, part, (int)n, euler_in, (int)grid_bw, euler_out,
                       score);
    return hipGetLastError() == hipSuccess ? SO3FT_OK : SO3FT_EHIP;
}

}  // extern "C"
