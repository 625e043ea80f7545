// C ABI of libso3ft (include/so3ft.h): plan objects and the transform pipelines.
//
// A plan (the counterpart of the reference's so3ft instance, pysofft/soft.f08:16-126)
// holds, on the device: cos(beta_k), the Driscoll-Healy weights, the seeds
// d~^q_{p,q}(beta_k) of the Kostelec recurrence, the mnl slice table, the
// order-pair work list and the FFT twiddles; plus one lazily allocated complex
// scratch array for the DWT <-> FFT intermediate (like fft_c2c_in, soft.f08:30).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace {

using namespace so3;

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(SO3FT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

struct Plan {
    int64_t bw = 0, lmax = 0, N = 0, nc = 0, npair = 0;
    int precompute = 0, init_ffts = 0;
    int64_t recurrence = 0;
    int32_t fftw_flags = 0;
    int device = 0;
    int dwt_engine = SO3FT_DWT_VALU;  // FP64 DWT engine (SO3FT_DWT_*)
    DeviceTables tb;
    FftDesc fft;
    double2 *scratch = nullptr;
    int64_t scratch_T = 0;
    double2 *cstage = nullptr;  // staged mnl coefficients (correlation)
    int64_t cstage_T = 0;
    double *amax_v = nullptr;   // argmax partials
    int64_t *amax_i = nullptr;
    int64_t amax_cap = 0;
    double *red = nullptr;  // integration partials + result
    double *radw = nullptr; // radial shell weights (3-D correlation)
    int64_t radw_cap = 0;
    int lmn = 0;                // coefficients in lmn order (Risbo layout)
    int64_t lcap = 0;           // highest degree summed: lmax for lmn handles, else bw - 1
    int64_t nkeep = 0;          // lmn positions of the degrees l <= lcap: NC(lcap + 1)
    int32_t *lmn_perm = nullptr;
    double2 *pstage = nullptr;  // coefficients reordered to mnl
    int64_t pstage_T = 0;
};

std::mutex g_mu;
std::map<int64_t, Plan *> g_plans;
int64_t g_next = 1;

Plan *get(so3ft_handle h)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_plans.find(h);
    return it == g_plans.end() ? nullptr : it->second;
}

void free_plan_device(Plan *P)
{
    hipFree(P->tb.cosb); hipFree(P->tb.w); hipFree(P->tb.seeds); hipFree(P->tb.slice);
    hipFree(P->tb.pairs); hipFree(P->tb.tw); hipFree(P->tb.rc); hipFree(P->tb.rc_off); hipFree(P->tb.dtab);
    hipFree(P->scratch); hipFree(P->red); hipFree(P->cstage);
    P->tb = DeviceTables();
    P->scratch = nullptr; P->scratch_T = 0; P->red = nullptr;
    P->cstage = nullptr; P->cstage_T = 0;
    hipFree(P->amax_v); hipFree(P->amax_i);
    P->amax_v = nullptr; P->amax_i = nullptr; P->amax_cap = 0;
    hipFree(P->radw);
    P->radw = nullptr; P->radw_cap = 0;
    hipFree(P->lmn_perm); hipFree(P->pstage);
    P->lmn_perm = nullptr; P->pstage = nullptr; P->pstage_T = 0; P->lmn = 0;
}

template <class T>
int upload(T **dst, const std::vector<T> &src)
{
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * (src.empty() ? 1 : src.size())));
    if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return SO3FT_OK;
}

inline int64_t mnidx(int64_t m, int64_t n, int64_t bw) { return (m + bw - 1) * (2 * bw - 1) + (n + bw - 1); }

// mnl layout slice starts by walking the storage order of pysofft/utils.f90:140-201
std::vector<int64_t> slice_table(int64_t bw)
{
    std::vector<int64_t> t((2 * bw - 1) * (2 * bw - 1), 0);
    int64_t pos = 0;
    for (int64_t m1 = 0; m1 < bw; ++m1)
        for (int64_t m2 = m1; m2 < bw; ++m2) {
            const int64_t L = bw - m2;
            t[mnidx(m1, m2, bw)] = pos; pos += L;
            if (m1 == 0 && m2 == 0) continue;
            t[mnidx(-m2, -m1, bw)] = pos; pos += L;
            if (m1 != m2) {
                t[mnidx(m2, m1, bw)] = pos; pos += L;
                t[mnidx(-m1, -m2, bw)] = pos; pos += L;
            }
            if (m1 == 0 || m2 == 0) continue;
            t[mnidx(m1, -m2, bw)] = pos; pos += L;
            t[mnidx(-m1, m2, bw)] = pos; pos += L;
            if (m1 == m2) continue;
            t[mnidx(m2, -m1, bw)] = pos; pos += L;
            t[mnidx(-m2, m1, bw)] = pos; pos += L;
        }
    return t;
}

// lmn order (utils.f90:228-355): degree-major; inside a degree the pairs 0 <= m1 <= m2
// own their distinct variants (m1,m2), (-m2,-m1), (m2,m1), (-m1,-m2), (m1,-m2), (-m1,m2),
// (m2,-m1), (-m2,m1).  perm[i] = mnl position of lmn position i.
std::vector<int32_t> lmn_perm_table(int64_t bw)
{
    const std::vector<int64_t> st = slice_table(bw);
    std::vector<int32_t> perm;
    perm.reserve((4 * bw * bw * bw - bw) / 3);
    for (int64_t l = 0; l < bw; ++l)
        for (int64_t m1 = 0; m1 <= l; ++m1)
            for (int64_t m2 = m1; m2 <= l; ++m2) {
                int64_t v[8][2];
                int nv = 0;
                v[nv][0] = m1; v[nv++][1] = m2;
                if (!(m1 == 0 && m2 == 0)) {
                    v[nv][0] = -m2; v[nv++][1] = -m1;
                    if (m1 != m2) { v[nv][0] = m2; v[nv++][1] = m1; v[nv][0] = -m1; v[nv++][1] = -m2; }
                    if (m1 != 0) {
                        v[nv][0] = m1; v[nv++][1] = -m2; v[nv][0] = -m1; v[nv++][1] = m2;
                        if (m1 != m2) { v[nv][0] = m2; v[nv++][1] = -m1; v[nv][0] = -m2; v[nv++][1] = m1; }
                    }
                }
                for (int k = 0; k < nv; ++k) {
                    const int64_t a = std::max(v[k][0] < 0 ? -v[k][0] : v[k][0], v[k][1] < 0 ? -v[k][1] : v[k][1]);
                    perm.push_back((int32_t)(st[mnidx(v[k][0], v[k][1], bw)] + l - a));
                }
            }
    return perm;
}

// constants (a, x, b) of the normalized Kostelec step from degree l to l + 1 of the pair
// (p, q) (wigner.f90:227-260): d^{l+1} = a (cos(beta) - x) d^l - b d^{l-1}
void rec_consts_host(int64_t l, int64_t p, int64_t q, double &a, double &x, double &b)
{
    const int64_t lp = l + 1;
    const int64_t den = (lp * lp - p * p) * (lp * lp - q * q);
    double cc = (double)lp * std::sqrt(1.0 / (double)den);
    cc = cc * std::sqrt((double)(2 * l + 3));
    a = cc * std::sqrt((double)(2 * l + 1));
    x = (l == 0) ? 0.0 : (double)(p * q) / (double)(l * (l + 1));
    if (l == q) {
        b = 0.0;
    } else {
        double c1 = std::sqrt((double)((l * l - p * p) * (l * l - q * q))) / (double)l;
        c1 = c1 / std::sqrt((double)(2 * l - 1));
        b = cc * c1;
    }
}

int build_plan(Plan *P, int64_t bw, int64_t lmax, int precompute, int init_ffts, int64_t rt, int32_t flags)
{
    if (bw < 1 || bw > 512) return fail(SO3FT_EINVAL, "bw must satisfy 1 <= bw <= 512, got " + std::to_string(bw));
    if (lmax < 0 || lmax >= bw) return fail(SO3FT_EINVAL, "lmax must satisfy 0 <= lmax < bw");
    if (rt != 0 && rt != 1) return fail(SO3FT_EINVAL, "unknown recurrence type " + std::to_string(rt));
    const int64_t N = 2 * bw;
    P->bw = bw; P->lmax = lmax; P->N = N; P->nc = (4 * bw * bw * bw - bw) / 3;
    P->precompute = precompute; P->init_ffts = init_ffts; P->recurrence = rt; P->fftw_flags = flags;
    HIPCHK(hipGetDevice(&P->device));

    const double pi = 3.14159265358979323846;
    std::vector<double> beta(N), cosb(N), sincos(N), cos2(N), w(N);
    for (int64_t k = 0; k < N; ++k) {  // wigner.f90:34-58
        beta[k] = (double)(2 * k + 1) * (pi / (double)(2 * N));
        cosb[k] = std::cos(beta[k]);
        sincos[k] = std::cos(beta[k] / 2) * std::sin(beta[k] / 2);
        cos2[k] = std::pow(std::cos(beta[k] / 2), 2);
    }
    {  // utils.f90:389-404
        const double xi = pi / (double)(4 * bw);
        for (int64_t k = 0; k < N; ++k) {
            const double ko = (double)(2 * k + 1);
            double s = 0.0;
            for (int64_t j = 0; j < bw; ++j) {
                const double jo = (double)(2 * j + 1);
                s += (1.0 / jo) * std::sin(ko * jo * xi);
            }
            w[k] = s * ((2.0 / (double)bw) * std::sin(ko * xi));
        }
    }
    // seeds d~^q_{p,q}(beta_k): stair recursion of wigner.f90:130-187, stored [tri(p,q)][k]
    const int64_t ntri = bw * (bw + 1) / 2;
    std::vector<double> seeds(ntri * N), tmp(bw * N);
    for (int64_t k = 0; k < N; ++k) tmp[k] = std::sqrt(0.5);
    for (int64_t m = 0; m < bw; ++m) {
        const int64_t sw = m + 1, sw2 = (m == 0) ? m : sw;
        const int64_t lim = std::min(bw - m, sw2);
        for (int64_t i = 1; i <= lim; ++i) {
            const int64_t l = m + i - 2;
            const double f = std::sqrt((double)((2 * l + 2) * (2 * l + 3)) / (double)((l + m) * (l + m + 1)));
            for (int64_t k = 0; k < N; ++k) tmp[(i - 1) * N + k] = tmp[(i - 1) * N + k] * f * cos2[k];
        }
        for (int64_t i = sw; i <= bw - m - 1; ++i) {
            const int64_t l = m + i - 1;
            const double f = std::sqrt((double)((2 * l + 2) * (2 * l + 3)) / (double)((l + m + 1) * (l - m + 1)));
            for (int64_t k = 0; k < N; ++k) tmp[i * N + k] = tmp[(i - 1) * N + k] * f * sincos[k];
        }
        const int64_t base = (m * (2 * bw - m + 1)) / 2;
        std::memcpy(&seeds[base * N], tmp.data(), sizeof(double) * (bw - m) * N);
    }
    // recurrence constants per pair (Kostelec recurrence, wigner.f90:227-260, normalized)
    std::vector<int64_t> rc_off(ntri);
    std::vector<double> rct;
    for (int64_t pp = 0; pp < bw; ++pp)
        for (int64_t qq = pp; qq < bw; ++qq) {
            const int64_t tri = pp * (2 * bw - pp + 1) / 2 + (qq - pp), L = bw - qq;
            rc_off[tri] = (int64_t)rct.size() / 3;
            for (int64_t i = 0; i < L - 1 + REC_PAD; ++i) {
                double a = 0.0, x = 0.0, b = 0.0;
                if (i < L - 1) rec_consts_host(qq + i, pp, qq, a, x, b);
                rct.push_back(a); rct.push_back(x); rct.push_back(b);
            }
        }
    // pairs with q <= lmax, most work (largest bw-q) first
    std::vector<int32_t> pairs;
    for (int64_t q = 0; q <= lmax; ++q)
        for (int64_t p = 0; p <= q; ++p) pairs.push_back((int32_t)(p | (q << 16)));
    P->npair = (int64_t)pairs.size();
    // twiddles exp(-2 pi i k/N)
    std::vector<double2> tw(N);
    for (int64_t k = 0; k < N; ++k) {
        const long double th = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
        tw[k] = make_double2((double)cosl(th), -(double)sinl(th));
    }
    FftDesc fd;
    fd.N = (int)N;
    int64_t n = N;
    while (n % 4 == 0) { fd.rad[fd.nrad++] = 4; n /= 4; }
    while (n % 2 == 0) { fd.rad[fd.nrad++] = 2; n /= 2; }
    for (int64_t pr = 3; n > 1; pr += 2)
        while (n % pr == 0) { fd.rad[fd.nrad++] = (int)pr; n /= pr; }
    P->fft = fd;

    int rc;
    if ((rc = upload(&P->tb.cosb, cosb))) return rc;
    if ((rc = upload(&P->tb.w, w))) return rc;
    if ((rc = upload(&P->tb.seeds, seeds))) return rc;
    if ((rc = upload(&P->tb.slice, slice_table(bw)))) return rc;
    if ((rc = upload(&P->tb.pairs, pairs))) return rc;
    if ((rc = upload(&P->tb.tw, tw))) return rc;
    if ((rc = upload(&P->tb.rc, rct))) return rc;
    if ((rc = upload(&P->tb.rc_off, rc_off))) return rc;
    // recurrence_type = Risbo without precomputed Wigners: coefficients in lmn order
    // (soft.py:98-106); the transforms reorder them and run the same d values (Kostelec
    // recurrence, equal to the Risbo values to rounding for bw <= 256)
    // The Risbo transforms sum only the degrees l <= lmax (soft.f08:797,811,1155,1169,1464,
    // 1837,2564,3046): the lmn reordering carries just those (the lmn order is degree-major)
    // and the correlation coefficients of higher degrees are zero.
    P->lcap = bw - 1;
    P->nkeep = P->nc;
    if (rt == 1 && !precompute) {
        P->lmn = 1;
        P->lcap = lmax;
        P->nkeep = (4 * (lmax + 1) * (lmax + 1) * (lmax + 1) - (lmax + 1)) / 3;
        if ((rc = upload(&P->lmn_perm, lmn_perm_table(bw)))) return rc;
    }
    HIPCHK(hipMalloc((void **)&P->red, sizeof(double) * (2 * 1024 + 2)));
    HIPCHK(dwt_set_attributes((int)bw, (int)N));
    // The d table: precompute_wigners (soft.f08:194-219) fills it once with the handle's
    // recurrence (genwig_all_kostelec / genwig_all_risbo); a Risbo handle without it, whose
    // reference transforms rerun the Risbo recurrence on every call (soft.f08:676-824),
    // fills it the same way at creation (the values depend only on bw).  The FP64 DWTs then
    // read d from the table.
    // (A Risbo handle's table is capped at 24 GiB, bw <= 256: beyond it d comes from the
    // Kostelec recurrence, equal to the Risbo values to rounding.)
    const size_t tab_bytes = sizeof(double) * (size_t)dtab_entries((int)bw);
    if (precompute || (rt == 1 && tab_bytes <= (size_t(24) << 30))) {
        const size_t bytes = tab_bytes;
        if (hipMalloc((void **)&P->tb.dtab, bytes) != hipSuccess) {
            hipGetLastError();
            P->tb.dtab = nullptr;
            return fail(SO3FT_ENOMEM, "Wigner d table allocation of " + std::to_string(bytes) + " bytes failed");
        }
        if (rt == 0) {
            HIPCHK(fill_dtab_kostelec(P->tb, (int)bw, P->tb.dtab, nullptr));
        } else {
            double *work = nullptr;
            if (hipMalloc((void **)&work, sizeof(double) * (size_t)dtab_work_entries((int)bw)) != hipSuccess) {
                hipGetLastError();
                return fail(SO3FT_ENOMEM, "Risbo recurrence scratch allocation failed");
            }
            const hipError_t e = fill_dtab_risbo((int)bw, P->tb.dtab, work, nullptr);
            const hipError_t es = hipStreamSynchronize(nullptr);
            hipFree(work);
            HIPCHK(e);
            HIPCHK(es);
        }
        HIPCHK(hipStreamSynchronize(nullptr));
    }
    return SO3FT_OK;
}

// scratch for T transforms (complex N^3 each)
int ensure_scratch(Plan *P, int64_t T)
{
    if (P->scratch_T >= T) return SO3FT_OK;
    hipFree(P->scratch);
    P->scratch = nullptr; P->scratch_T = 0;
    const size_t bytes = sizeof(double2) * (size_t)P->N * P->N * P->N * (size_t)T;
    if (hipMalloc((void **)&P->scratch, bytes) != hipSuccess) {
        hipGetLastError();
        return fail(SO3FT_ENOMEM, "scratch allocation of " + std::to_string(bytes) + " bytes failed");
    }
    P->scratch_T = T;
    return SO3FT_OK;
}

// transforms per launch for the *_many paths: bounded scratch (<= 4 GiB) and at
// most 32768 transforms, which keeps every grid's z extent under 65536
int64_t chunk_of(const Plan *P, int64_t n)
{
    const double per = 16.0 * (double)P->N * P->N * P->N;
    int64_t c = (int64_t)(4.0 * 1024 * 1024 * 1024 / per);
    c = std::max<int64_t>(1, std::min<int64_t>(c, 32768));
    return std::min(c, n);
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

int check_ptrs(const void *a, const void *b)
{
    if (!a || !b) return fail(SO3FT_EINVAL, "null array pointer");
    return SO3FT_OK;
}

// F (i_n,i_m,beta) -> SO(3) function: FFT along m (in place), then along n with 1/(2 pi)
// the inverse transform's two passes: along m (alpha) in place, then along n (gamma)
// with the 1/(2 pi) of the inverse (soft.f08:1896-1906)
static FftAxis inverse_alpha_axis(int64_t N)
{
    const int64_t N2 = N * N, N3 = N2 * N;
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.mask_rows = 1; a.mask_outer = 1; a.sign = -1; a.scale = 1.0;
    return a;
}

static FftAxis inverse_gamma_axis(int64_t N, int out_real)
{
    const int64_t N2 = N * N, N3 = N2 * N;
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.mask_rows = 1; b.mask_outer = 0; b.out_real = out_real; b.sign = -1;
    b.scale = 1.0 / (2.0 * 3.14159265358979323846);
    return b;
}

int inverse_fft_passes(Plan *P, double2 *F, void *out, int out_real, int64_t T, hipStream_t s)
{
    const FftAxis a = inverse_alpha_axis(P->N), b = inverse_gamma_axis(P->N, out_real);
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, F, F, a, (int)T, s));
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, F, out, b, (int)T, s));
    return SO3FT_OK;
}

// SO(3) function -> F: FFT along n (gamma) into F, then along m in place with 2 pi/(2bw)^2
int forward_fft_passes(Plan *P, const void *in, int in_real, double2 *F, int64_t T, hipStream_t s)
{
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.in_real = in_real; b.sign = +1; b.scale = 1.0;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, in, F, b, (int)T, s));
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.mask_outer = 1; a.sign = +1;
    a.scale = 2.0 * 3.14159265358979323846 / ((double)N * (double)N);
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, F, F, a, (int)T, s));
    return SO3FT_OK;
}

int ensure_cstage(Plan *P, int64_t T)
{
    if (P->cstage_T >= T) return SO3FT_OK;
    hipFree(P->cstage);
    P->cstage = nullptr; P->cstage_T = 0;
    const size_t bytes = sizeof(double2) * (size_t)P->nc * (size_t)T;
    if (hipMalloc((void **)&P->cstage, bytes) != hipSuccess) {
        hipGetLastError();
        return fail(SO3FT_ENOMEM, "coefficient staging allocation of " + std::to_string(bytes) + " bytes failed");
    }
    P->cstage_T = T;
    return SO3FT_OK;
}

int ensure_pstage(Plan *P, int64_t T)
{
    if (P->pstage_T >= T) return SO3FT_OK;
    hipFree(P->pstage);
    P->pstage = nullptr; P->pstage_T = 0;
    const size_t bytes = sizeof(double2) * (size_t)P->nc * (size_t)T;
    if (hipMalloc((void **)&P->pstage, bytes) != hipSuccess) {
        hipGetLastError();
        return fail(SO3FT_ENOMEM, "lmn reordering buffer allocation of " + std::to_string(bytes) + " bytes failed");
    }
    P->pstage_T = T;
    return SO3FT_OK;
}

// inverse DWT of T transforms into F on the plan's engine; DWT_CORR takes SH pairs
// (c, aux) with stride c_ts, whose correlation coefficients the kernels form as they load
int dwt_inverse(Plan *P, int mode, const double2 *c, int64_t c_ts, const double2 *aux, double2 *F, int64_t f_ts,
                int64_t T, hipStream_t s)
{
    const int bw = (int)P->bw, N = (int)P->N, np = (int)P->npair;
    // both engines form correlation coefficients as their kernels load them; the MFMA
    // engine stages them first for handles that drop degrees above lmax
    // table plans read d on the default engine's kernels; the MFMA engine's LDS holds the
    // recurrence constants of bw <= 256
    const bool mfma = P->dwt_engine == SO3FT_DWT_MFMA && !P->tb.dtab && P->bw <= 256;
    if (mode == DWT_CORR && mfma && P->lcap < P->bw - 1) {
        if (int rc = ensure_cstage(P, T)) return rc;
        HIPCHK(launch_corr_coeffs(P->tb, bw, c, aux, c_ts, 0, 1, 0, nullptr, P->cstage, P->nc, (int)P->lcap, (int)T,
                                  s));
        c = P->cstage; c_ts = P->nc; mode = DWT_COMPLEX;
    }
    if (mfma) {
        HIPCHK(launch_dwt_inverse(P->tb, bw, N, np, mode, c, c_ts, aux, F, f_ts, (int)T, s));
        return SO3FT_OK;
    }
    HIPCHK(launch_dwt_inverse_valu(P->tb, bw, N, np, mode, c, c_ts, aux, (int)P->lcap, F, f_ts, (int)T, s));
    return SO3FT_OK;
}

int dwt_forward(Plan *P, int mode, const double2 *F, int64_t f_ts, double2 *c, int64_t c_ts, int64_t T, hipStream_t s)
{
    const int bw = (int)P->bw, N = (int)P->N, np = (int)P->npair;
    if (P->dwt_engine == SO3FT_DWT_MFMA && !P->tb.dtab && bw <= 256)
        HIPCHK(launch_dwt_forward(P->tb, bw, N, np, mode, F, f_ts, c, c_ts, (int)T, s));
    else if (mode == DWT_COMPLEX || mode == DWT_REAL)   // lane-beta FP64 MFMA forward (FP64 throughout)
        HIPCHK(launch_dwt_forward_m64(P->tb, bw, N, np, mode, F, f_ts, c, c_ts, (int)T, s));
    else
        HIPCHK(launch_dwt_forward_valu(P->tb, bw, N, np, mode, F, f_ts, c, c_ts, (int)T, s));
    return SO3FT_OK;
}

// perm: the caller's coefficients are in lmn order (plans with P->lmn)
int run_inverse(Plan *P, int mode, const void *coeff, int64_t c_ts, const void *aux, void *out, int out_real,
                int64_t n, hipStream_t s, int perm = 0)
{
    const int64_t N = P->N, N3 = N * N * N;
    const int64_t ch = chunk_of(P, n);
    int rc = ensure_scratch(P, ch);
    if (rc) return rc;
    if (perm && (rc = ensure_pstage(P, ch))) return rc;
    const size_t osz = out_real ? sizeof(double) : sizeof(double2);
    for (int64_t i0 = 0; i0 < n; i0 += ch) {
        const int64_t T = std::min(ch, n - i0);
        const double2 *c = reinterpret_cast<const double2 *>(coeff) + i0 * c_ts;
        if (perm) {
            HIPCHK(launch_permute_lmn(P->lmn_perm, P->nc, P->nkeep, c, c_ts, P->pstage, P->nc, 16, 1, (int)T, s));
            c = P->pstage;
        }
        const double2 *ax = aux ? reinterpret_cast<const double2 *>(aux) + i0 * c_ts : nullptr;
        if ((rc = dwt_inverse(P, mode, c, perm ? P->nc : c_ts, ax, P->scratch, N3, T, s))) return rc;
        void *o = reinterpret_cast<char *>(out) + (size_t)i0 * N3 * osz;
        if ((rc = inverse_fft_passes(P, P->scratch, o, out_real, T, s))) return rc;
    }
    return SO3FT_OK;
}

int run_forward(Plan *P, int mode, const void *in, int in_real, void *coeff, int64_t n, hipStream_t s, int perm = 0)
{
    const int64_t N = P->N, N3 = N * N * N;
    const int64_t ch = chunk_of(P, n);
    int rc = ensure_scratch(P, ch);
    if (rc) return rc;
    if (perm && (rc = ensure_pstage(P, ch))) return rc;
    const size_t isz = in_real ? sizeof(double) : sizeof(double2);
    for (int64_t i0 = 0; i0 < n; i0 += ch) {
        const int64_t T = std::min(ch, n - i0);
        const void *f = reinterpret_cast<const char *>(in) + (size_t)i0 * N3 * isz;
        if ((rc = forward_fft_passes(P, f, in_real, P->scratch, T, s))) return rc;
        double2 *c = reinterpret_cast<double2 *>(coeff) + i0 * P->nc;
        if ((rc = dwt_forward(P, mode, P->scratch, N3, perm ? P->pstage : c, P->nc, T, s))) return rc;
        if (perm) HIPCHK(launch_permute_lmn(P->lmn_perm, P->nc, P->nkeep, P->pstage, P->nc, c, P->nc, 16, 0, (int)T, s));
    }
    return SO3FT_OK;
}

// correlation coefficients of T pairs (layout 0 LMc, bw^2 per set; 1 MLr, bw(bw+1)/2 per
// set) into the plan's staging array; with host weights hw, the weighted sum over nr
// consecutive shells of one pair
int stage_corr(Plan *P, int layout, const void *f, const void *g, int64_t T, int nr, const double *hw, hipStream_t s)
{
    if (int rc = ensure_cstage(P, T)) return rc;
    const int64_t nb = layout == 0 ? P->bw * P->bw : P->bw * (P->bw + 1) / 2;
    const double *w = nullptr;
    if (hw) {
        if (P->radw_cap < nr) {
            HIPCHK(hipStreamSynchronize(s));
            hipFree(P->radw);
            P->radw = nullptr; P->radw_cap = 0;
            HIPCHK(hipMalloc((void **)&P->radw, sizeof(double) * nr));
            P->radw_cap = nr;
        }
        HIPCHK(hipStreamSynchronize(s));  // the previous call's staging has read the weights
        HIPCHK(hipMemcpy(P->radw, hw, sizeof(double) * nr, hipMemcpyHostToDevice));
        w = P->radw;
    }
    HIPCHK(launch_corr_coeffs(P->tb, (int)P->bw, reinterpret_cast<const double2 *>(f),
                              reinterpret_cast<const double2 *>(g), nb, layout, hw ? nr : 1, nb, w, P->cstage, P->nc,
                              (int)P->lcap, (int)T, s));
    return SO3FT_OK;
}

// soft.f08:2696-2697,2737: weights r^2 / (r_hi - r_lo + dr) of the shells r_lo..r_hi
int radial_weights(const double *rad, int64_t n_radial, int64_t r_lo, int64_t r_hi, std::vector<double> &w)
{
    if (!rad) return fail(SO3FT_EINVAL, "null radial_sampling_points");
    if (n_radial < 2 || r_lo < 0 || r_hi >= n_radial || r_lo > r_hi)
        return fail(SO3FT_EINVAL, "radial_limits out of bounds for radial_sampling_points");
    const double step = rad[1] - rad[0];
    const double inv = 1.0 / (rad[r_hi] - rad[r_lo] + step);
    w.clear();
    for (int64_t r = r_lo; r <= r_hi; ++r) w.push_back(rad[r] * rad[r] * inv);
    return SO3FT_OK;
}

// ---- mixed precision: complex64 storage, FP64 recurrence and FFT arithmetic,
// FP32 MFMA contraction with d split into FP32 hi + lo parts
int inverse_fft_passes_c64(Plan *P, float2 *F, float2 *out, int64_t T, hipStream_t s)
{
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.mask_rows = 1; a.mask_outer = 1; a.sign = -1; a.scale = 1.0;
    HIPCHK(launch_fft_reg((int)N, P->tb.tw, (int)P->bw, (int)P->lmax, F, F, a, (int)T, 1, 1, 0, (int)N, s));
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.mask_rows = 1; b.sign = -1; b.scale = 1.0 / (2.0 * 3.14159265358979323846);
    HIPCHK(launch_fft_reg((int)N, P->tb.tw, (int)P->bw, (int)P->lmax, F, out, b, (int)T, 1, 1, 0, (int)N, s));
    return SO3FT_OK;
}

int forward_fft_passes_c64(Plan *P, const float2 *in, float2 *F, int64_t T, hipStream_t s)
{
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.sign = +1; b.scale = 1.0;
    HIPCHK(launch_fft_reg((int)N, P->tb.tw, (int)P->bw, (int)P->lmax, in, F, b, (int)T, 1, 1, 0, (int)N, s));
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.mask_outer = 1; a.sign = +1;
    a.scale = 2.0 * 3.14159265358979323846 / ((double)N * (double)N);
    HIPCHK(launch_fft_reg((int)N, P->tb.tw, (int)P->bw, (int)P->lmax, F, F, a, (int)T, 1, 1, 0, (int)N, s));
    return SO3FT_OK;
}

int check_c64(Plan *P)
{
    (void)P;   // every bandwidth: lengths without a register FFT take the Stockham pass
    return SO3FT_OK;
}

int run_inverse_c64(Plan *P, const float2 *coeff, float2 *out, int64_t n, hipStream_t s)
{
    if (int rc = check_c64(P)) return rc;
    const int64_t N = P->N, N3 = N * N * N;
    const int64_t ch = chunk_of(P, n);
    int rc = ensure_scratch(P, (ch + 1) / 2);  // complex64 scratch: half the bytes
    if (rc) return rc;
    if (P->lmn && (rc = ensure_pstage(P, (ch + 1) / 2))) return rc;
    float2 *F = reinterpret_cast<float2 *>(P->scratch);
    for (int64_t i0 = 0; i0 < n; i0 += ch) {
        const int64_t T = std::min(ch, n - i0);
        const float2 *c = coeff + i0 * P->nc;
        if (P->lmn) {
            float2 *ps = reinterpret_cast<float2 *>(P->pstage);
            HIPCHK(launch_permute_lmn(P->lmn_perm, P->nc, P->nkeep, c, P->nc, ps, P->nc, 8, 1, (int)T, s));
            c = ps;
        }
        HIPCHK(launch_dwt_inverse_mixed(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX, c, P->nc, nullptr, F,
                                        N3, (int)T, s));
        if ((rc = inverse_fft_passes_c64(P, F, out + i0 * N3, T, s))) return rc;
    }
    return SO3FT_OK;
}

// ext: caller-provided scratch (the static array of the zero-allocation entry point)
// holding ext_T transforms' fft arrays followed, for lmn handles, by ext_T coefficient
// sets; nullptr: the plan's own scratch, allocated on demand
int run_forward_c64(Plan *P, const float2 *in, float2 *coeff, int64_t n, hipStream_t s, float2 *ext = nullptr,
                    int64_t ext_T = 0)
{
    if (int rc = check_c64(P)) return rc;
    const int64_t N = P->N, N3 = N * N * N;
    int64_t ch;
    float2 *F, *ps;
    if (ext) {
        ch = std::min(ext_T, n);
        F = ext;
        ps = ext + ext_T * N3;
    } else {
        ch = chunk_of(P, n);
        int rc = ensure_scratch(P, (ch + 1) / 2);
        if (rc) return rc;
        if (P->lmn && (rc = ensure_pstage(P, (ch + 1) / 2))) return rc;
        F = reinterpret_cast<float2 *>(P->scratch);
        ps = reinterpret_cast<float2 *>(P->pstage);
    }
    for (int64_t i0 = 0; i0 < n; i0 += ch) {
        const int64_t T = std::min(ch, n - i0);
        if (int rc = forward_fft_passes_c64(P, in + i0 * N3, F, T, s)) return rc;
        float2 *c = coeff + i0 * P->nc;
        HIPCHK(launch_dwt_forward_mixed(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX, F, N3,
                                        P->lmn ? ps : c, P->nc, (int)T, s));
        if (P->lmn) HIPCHK(launch_permute_lmn(P->lmn_perm, P->nc, P->nkeep, ps, P->nc, c, P->nc, 8, 0, (int)T, s));
    }
    return SO3FT_OK;
}

// ---- scratch of the zero-allocation forward transform (so3ft_soft_many_c64_noalloc).
// The array lives in the library's code object: every GPU the library is loaded on holds
// its own copy until the process ends.  In host code `scratch` names a host-side stand-in;
// only the device address from hipGetSymbolAddress (kept per device) is passed to kernels.
constexpr size_t STATIC_SCRATCH_BYTES = size_t(256) << 20;
__device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;

// device address of `scratch` on device dev, resolved once per device with dev current
int static_scratch(int dev, void **ptr)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *ptr = it->second;
        return SO3FT_OK;
    }
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || p == nullptr)
        return fail(SO3FT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
    *ptr = p;
    return SO3FT_OK;
}

}  // namespace

extern "C" {

int so3ft_get_kostelec_recurrence_val(void) { return 0; }
int so3ft_get_risbo_recurrence_val(void) { return 1; }
const char *so3ft_last_error(void) { return g_err.c_str(); }

int so3ft_init(int64_t bw, int64_t lmax, int precompute_wigners, int init_ffts, int64_t recurrence_type,
               int32_t fftw_flags, so3ft_handle *out)
{
    if (!out) return fail(SO3FT_EINVAL, "null handle pointer");
    Plan *P = new Plan();
    int rc = build_plan(P, bw, lmax, precompute_wigners, init_ffts, recurrence_type, fftw_flags);
    if (rc) {
        free_plan_device(P);
        delete P;
        return rc;
    }
    // resolve the zero-allocation entry point's scratch on this device now (the plan's
    // device is current); a failure is reported by that entry point, which tries again
    void *sp = nullptr;
    if (static_scratch(P->device, &sp) != SO3FT_OK) g_err.clear();
    std::lock_guard<std::mutex> lk(g_mu);
    *out = g_next++;
    g_plans[*out] = P;
    return SO3FT_OK;
}

int so3ft_destroy(so3ft_handle h)
{
    Plan *P = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_plans.find(h);
        if (it == g_plans.end()) return fail(SO3FT_EHANDLE, "unknown handle");
        P = it->second;
        g_plans.erase(it);
    }
    hipDeviceSynchronize();
    free_plan_device(P);
    delete P;
    return SO3FT_OK;
}

int so3ft_reset(so3ft_handle h, int64_t bw, int64_t lmax, int precompute_wigners, int init_ffts,
                int64_t recurrence_type, int32_t fftw_flags)
{
    Plan *P = get(h);
    if (!P) return fail(SO3FT_EHANDLE, "unknown handle");
    Plan fresh;
    int rc = build_plan(&fresh, bw, lmax, precompute_wigners, init_ffts, recurrence_type, fftw_flags);
    if (rc) {
        free_plan_device(&fresh);
        return rc;
    }
    hipDeviceSynchronize();
    free_plan_device(P);
    fresh.dwt_engine = P->dwt_engine;
    *P = fresh;
    return SO3FT_OK;
}

int so3ft_init_fft(so3ft_handle h, int use_real_fft)
{
    (void)use_real_fft;
    return get(h) ? SO3FT_OK : fail(SO3FT_EHANDLE, "unknown handle");
}

int so3ft_get_bw(so3ft_handle h, int64_t *v)
{
    Plan *P = get(h);
    if (!P || !v) return fail(SO3FT_EHANDLE, "unknown handle");
    *v = P->bw;
    return SO3FT_OK;
}
int so3ft_get_lmax(so3ft_handle h, int64_t *v)
{
    Plan *P = get(h);
    if (!P || !v) return fail(SO3FT_EHANDLE, "unknown handle");
    *v = P->lmax;
    return SO3FT_OK;
}
int so3ft_set_lmax(so3ft_handle h, int64_t lmax)
{
    Plan *P = get(h);
    if (!P) return fail(SO3FT_EHANDLE, "unknown handle");
    if (lmax < 0 || lmax >= P->bw) return fail(SO3FT_EINVAL, "lmax must satisfy 0 <= lmax < bw");
    return so3ft_reset(h, P->bw, lmax, P->precompute, P->init_ffts, P->recurrence, P->fftw_flags);
}
int so3ft_get_recurrence_type(so3ft_handle h, int64_t *v)
{
    Plan *P = get(h);
    if (!P || !v) return fail(SO3FT_EHANDLE, "unknown handle");
    *v = P->recurrence;
    return SO3FT_OK;
}
int so3ft_wigners_are_precomputed(so3ft_handle h, int *v)
{
    Plan *P = get(h);
    if (!P || !v) return fail(SO3FT_EHANDLE, "unknown handle");
    *v = P->precompute;
    return SO3FT_OK;
}
int so3ft_get_fftw_flags(so3ft_handle h, int32_t *v)
{
    Plan *P = get(h);
    if (!P || !v) return fail(SO3FT_EHANDLE, "unknown handle");
    *v = P->fftw_flags;
    return SO3FT_OK;
}

int so3ft_set_dwt_engine(so3ft_handle h, int engine)
{
    Plan *P = get(h);
    if (!P) return fail(SO3FT_EHANDLE, "unknown handle");
    if (engine != SO3FT_DWT_VALU && engine != SO3FT_DWT_MFMA) return fail(SO3FT_EINVAL, "unknown DWT engine");
    P->dwt_engine = engine;
    return SO3FT_OK;
}
int so3ft_get_dwt_engine(so3ft_handle h, int *v)
{
    Plan *P = get(h);
    if (!P || !v) return fail(SO3FT_EHANDLE, "unknown handle");
    *v = P->dwt_engine;
    return SO3FT_OK;
}

#define GETP(h)                                                  \
    Plan *P = get(h);                                            \
    if (!P) return fail(SO3FT_EHANDLE, "unknown handle");

int so3ft_isoft_many(so3ft_handle h, const void *coeffs, void *so3funcs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "negative count");
    return run_inverse(P, DWT_COMPLEX, coeffs, P->nc, nullptr, so3funcs, 0, n, S(stream), P->lmn);
}
int so3ft_isoft(so3ft_handle h, const void *coeff, void *so3func, void *stream)
{
    return so3ft_isoft_many(h, coeff, so3func, 1, stream);
}
int so3ft_soft_many(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "negative count");
    return run_forward(P, DWT_COMPLEX, so3funcs, 0, coeffs, n, S(stream), P->lmn);
}
int so3ft_soft(so3ft_handle h, const void *so3func, void *coeff, void *stream)
{
    return so3ft_soft_many(h, so3func, coeff, 1, stream);
}
int so3ft_irsoft_many(so3ft_handle h, const void *coeffs, void *so3funcs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "negative count");
    return run_inverse(P, DWT_REAL, coeffs, P->nc, nullptr, so3funcs, 1, n, S(stream), P->lmn);
}
int so3ft_irsoft(so3ft_handle h, const void *coeff, void *so3func, void *stream)
{
    return so3ft_irsoft_many(h, coeff, so3func, 1, stream);
}
int so3ft_rsoft_many(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "negative count");
    return run_forward(P, DWT_REAL, so3funcs, 1, coeffs, n, S(stream), P->lmn);
}
int so3ft_rsoft(so3ft_handle h, const void *so3func, void *coeff, void *stream)
{
    return so3ft_rsoft_many(h, so3func, coeff, 1, stream);
}

int so3ft_isoft_many_c64(so3ft_handle h, const void *coeffs, void *so3funcs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "n must be >= 0");
    return run_inverse_c64(P, reinterpret_cast<const float2 *>(coeffs), reinterpret_cast<float2 *>(so3funcs), n,
                           S(stream));
}

int so3ft_soft_many_c64(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "n must be >= 0");
    return run_forward_c64(P, reinterpret_cast<const float2 *>(so3funcs), reinterpret_cast<float2 *>(coeffs), n,
                           S(stream));
}

int so3ft_soft_many_c64_noalloc(so3ft_handle h, const void *so3funcs, void *coeffs, int64_t n, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeffs, so3funcs)) return rc;
    if (n < 0) return fail(SO3FT_EINVAL, "n must be >= 0");
    if (int rc = check_c64(P)) return rc;
    const int64_t N3 = P->N * P->N * P->N;
    const size_t per = sizeof(float2) * (size_t)(N3 + (P->lmn ? P->nc : 0));
    if (per > STATIC_SCRATCH_BYTES)
        return fail(SO3FT_EINVAL, "so3ft_soft_many_c64_noalloc: one transform at bw=" + std::to_string(P->bw) +
                                      " needs " + std::to_string(per) + " bytes of scratch, beyond the static " +
                                      std::to_string(STATIC_SCRATCH_BYTES) + " (bw <= 128)");
    void *sp = nullptr;
    if (int rc = static_scratch(P->device, &sp)) return rc;
    return run_forward_c64(P, reinterpret_cast<const float2 *>(so3funcs), reinterpret_cast<float2 *>(coeffs), n,
                           S(stream), reinterpret_cast<float2 *>(sp), (int64_t)(STATIC_SCRATCH_BYTES / per));
}

int so3ft_inverse_wigner_trf_cmplx(so3ft_handle h, const void *coeff, void *fft_array, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeff, fft_array)) return rc;
    const int64_t N3 = P->N * P->N * P->N;
    double2 *F = reinterpret_cast<double2 *>(fft_array);
    const double2 *c = reinterpret_cast<const double2 *>(coeff);
    if (P->lmn) {
        if (int rc = ensure_pstage(P, 1)) return rc;
        HIPCHK(launch_permute_lmn(P->lmn_perm, P->nc, P->nkeep, c, P->nc, P->pstage, P->nc, 16, 1, 1, S(stream)));
        c = P->pstage;
    }
    if (int rc = dwt_inverse(P, DWT_COMPLEX, c, P->nc, nullptr, F, N3, 1, S(stream))) return rc;
    HIPCHK(launch_zero_invalid_columns(F, N3, (int)P->bw, (int)P->N, (int)P->lmax, 1, 0, S(stream)));
    return SO3FT_OK;
}

int so3ft_forward_wigner_trf_cmplx(so3ft_handle h, const void *fft_array, void *coeff, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(coeff, fft_array)) return rc;
    const int64_t N3 = P->N * P->N * P->N;
    double2 *c = reinterpret_cast<double2 *>(coeff);
    if (P->lmn)
        if (int rc = ensure_pstage(P, 1)) return rc;
    if (int rc = dwt_forward(P, DWT_COMPLEX, reinterpret_cast<const double2 *>(fft_array), N3, P->lmn ? P->pstage : c,
                             P->nc, 1, S(stream)))
        return rc;
    if (P->lmn) HIPCHK(launch_permute_lmn(P->lmn_perm, P->nc, P->nkeep, P->pstage, P->nc, c, P->nc, 16, 0, 1, S(stream)));
    return SO3FT_OK;
}

static int fft2_scaled(Plan *P, const void *f1, void *f2, int sign, hipStream_t s)
{
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.sign = sign; b.scale = 1.0;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, f1, f2, b, 1, s));
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.sign = sign; a.scale = 1.0 / (double)N;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, f2, f2, a, 1, s));
    return SO3FT_OK;
}

int so3ft_fft(so3ft_handle h, const void *f1, void *f2, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(f1, f2)) return rc;
    return fft2_scaled(P, f1, f2, -1, S(stream));
}
int so3ft_ifft(so3ft_handle h, const void *f1, void *f2, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(f1, f2)) return rc;
    return fft2_scaled(P, f1, f2, +1, S(stream));
}

// soft.f08:3270-3276 rfft: r2c over (alpha, gamma) of every beta slice, the alpha axis
// halved to bw+1 (f2: C-order (2bw, bw+1, 2bw)), scaled by 1/(2bw)
int so3ft_rfft(so3ft_handle h, const void *f1, void *f2, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(f1, f2)) return rc;
    if (int rc = ensure_scratch(P, 1)) return rc;
    hipStream_t s = S(stream);
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.in_real = 1; b.sign = -1; b.scale = 1.0;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, f1, P->scratch, b, 1, s));
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.sign = -1; a.scale = 1.0 / (double)N;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, P->scratch, P->scratch, a, 1, s));
    HIPCHK(launch_halfcomplex(P->scratch, reinterpret_cast<double2 *>(f2), (int)N, 0, s));
    return SO3FT_OK;
}

// soft.f08:3277-3283 irfft: c2r (FFTW_BACKWARD) of the half-complex f1 into the real f2,
// scaled by 1/(2bw); the redundant half is the Hermitian completion, as FFTW's c2r reads it
int so3ft_irfft(so3ft_handle h, const void *f1, void *f2, void *stream)
{
    GETP(h);
    if (int rc = check_ptrs(f1, f2)) return rc;
    if (int rc = ensure_scratch(P, 1)) return rc;
    hipStream_t s = S(stream);
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    HIPCHK(launch_halfcomplex(reinterpret_cast<const double2 *>(f1), P->scratch, (int)N, 1, s));
    FftAxis a{};
    a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
    a.sign = +1; a.scale = 1.0;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, P->scratch, P->scratch, a, 1, s));
    FftAxis b{};
    b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
    b.out_real = 1; b.sign = +1; b.scale = 1.0 / (double)N;
    HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, P->scratch, f2, b, 1, s));
    return SO3FT_OK;
}

static int integrate(Plan *P, const void *f, int is_real, double *re, double *im, hipStream_t s)
{
    const int nb = 1024;
    HIPCHK(launch_integrate(f, is_real, P->tb.w, (int)P->N, P->red, nb, P->red + 2 * nb, s));
    double res[2];
    HIPCHK(hipMemcpyAsync(res, P->red + 2 * nb, sizeof(res), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double sc = std::pow(3.14159265358979323846 / (double)P->bw, 2);
    *re = res[0] * sc;
    if (im) *im = res[1] * sc;
    return SO3FT_OK;
}

int so3ft_integrate_over_so3_cmplx(so3ft_handle h, const void *f, double *re, double *im, void *stream)
{
    GETP(h);
    if (!f || !re || !im) return fail(SO3FT_EINVAL, "null pointer");
    return integrate(P, f, 0, re, im, S(stream));
}
int so3ft_integrate_over_so3_real(so3ft_handle h, const void *f, double *out, void *stream)
{
    GETP(h);
    if (!f || !out) return fail(SO3FT_EINVAL, "null pointer");
    return integrate(P, f, 1, out, nullptr, S(stream));
}

int so3ft_cross_correlation_ylm_cmplx(so3ft_handle h, const void *f_lm, const void *g_lm, void *cc, void *stream)
{
    GETP(h);
    if (!f_lm || !g_lm || !cc) return fail(SO3FT_EINVAL, "null pointer");
    return run_inverse(P, DWT_CORR, f_lm, P->bw * P->bw, g_lm, cc, 0, 1, S(stream));
}

int so3ft_cross_correlation_ylm_cmplx_many(so3ft_handle h, const void *f_lms, const void *g_lms, void *cc, int64_t n,
                                           void *stream)
{
    GETP(h);
    if (!f_lms || !g_lms || !cc) return fail(SO3FT_EINVAL, "null pointer");
    if (n < 0) return fail(SO3FT_EINVAL, "n must be >= 0");
    return run_inverse(P, DWT_CORR, f_lms, P->bw * P->bw, g_lms, cc, 0, n, S(stream));
}

int so3ft_cross_correlation_argmax_many(so3ft_handle h, const void *f_lms, const void *g_lms, int64_t n, void *idx,
                                        void *val, void *stream)
{
    GETP(h);
    if (!f_lms || !g_lms || !idx) return fail(SO3FT_EINVAL, "null pointer");
    if (n < 0) return fail(SO3FT_EINVAL, "n must be >= 0");
    const int64_t N3 = P->N * P->N * P->N, nb = P->bw * P->bw;
    const int64_t ch = chunk_of(P, n);
    int rc = ensure_scratch(P, ch);
    if (rc) return rc;
    // the gamma pass fused with the peak search where that shape exists: the correlation
    // grids are never written (one grid write and one grid read less per pair)
    const int nfused = fft_reg_supported((int)P->N) ? fft_reg_argmax_partials((int)P->N) : 0;
    const int64_t need = ch * std::max<int64_t>(argmax_partials(N3), nfused);
    if (P->amax_cap < need) {
        hipFree(P->amax_v); hipFree(P->amax_i);
        P->amax_v = nullptr; P->amax_i = nullptr; P->amax_cap = 0;
        if (hipMalloc((void **)&P->amax_v, sizeof(double) * need) != hipSuccess ||
            hipMalloc((void **)&P->amax_i, sizeof(int64_t) * need) != hipSuccess) {
            hipGetLastError();
            return fail(SO3FT_ENOMEM, "argmax partials allocation failed");
        }
        P->amax_cap = need;
    }
    hipStream_t s = S(stream);
    for (int64_t i0 = 0; i0 < n; i0 += ch) {
        const int64_t T = std::min(ch, n - i0);
        const double2 *f = reinterpret_cast<const double2 *>(f_lms) + i0 * nb;
        const double2 *g = reinterpret_cast<const double2 *>(g_lms) + i0 * nb;
        if ((rc = dwt_inverse(P, DWT_CORR, f, nb, g, P->scratch, N3, T, s))) return rc;
        int64_t *ix = reinterpret_cast<int64_t *>(idx) + i0;
        double *vx = val ? reinterpret_cast<double *>(val) + i0 : nullptr;
        if (nfused > 0) {
            const FftAxis a = inverse_alpha_axis(P->N), b = inverse_gamma_axis(P->N, 0);
            HIPCHK(launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, P->scratch, P->scratch, a, (int)T, s));
            int nblk = 0;
            HIPCHK(launch_fft_reg_argmax((int)P->N, P->tb.tw, (int)P->bw, (int)P->lmax, P->scratch, b, (int)T,
                                         P->amax_v, P->amax_i, &nblk, s));
            HIPCHK(launch_argmax_final(P->amax_v, P->amax_i, nblk, (int)T, ix, vx, s));
        } else {
            if ((rc = inverse_fft_passes(P, P->scratch, P->scratch, 0, T, s))) return rc;
            HIPCHK(launch_argmax_abs(P->scratch, N3, (int)T, P->amax_v, P->amax_i, ix, vx, s));
        }
    }
    return SO3FT_OK;
}

int so3ft_cross_correlation_ylm_cmplx_3d(so3ft_handle h, const void *f_lms, const void *g_lms, void *cc,
                                         const double *rad, int64_t n_radial, int64_t r_lo, int64_t r_hi,
                                         void *stream)
{
    GETP(h);
    if (!f_lms || !g_lms || !cc) return fail(SO3FT_EINVAL, "null pointer");
    // soft.f08:2657-2738 sums the shells' correlation grids; the correlation is linear in
    // its coefficients, so the weighted shell sum is taken over the coefficients and one
    // transform follows
    std::vector<double> w;
    if (int rc = radial_weights(rad, n_radial, r_lo, r_hi, w)) return rc;
    const int64_t nb = P->bw * P->bw;
    if (int rc = stage_corr(P, 0, reinterpret_cast<const double2 *>(f_lms) + r_lo * nb,
                            reinterpret_cast<const double2 *>(g_lms) + r_lo * nb, 1, (int)w.size(), w.data(),
                            S(stream)))
        return rc;
    return run_inverse(P, DWT_COMPLEX, P->cstage, P->nc, nullptr, cc, 0, 1, S(stream));
}

int so3ft_cross_correlation_ylm_real(so3ft_handle h, const void *f_ml, const void *g_ml, void *cc, void *stream)
{
    GETP(h);
    if (!f_ml || !g_ml || !cc) return fail(SO3FT_EINVAL, "null pointer");
    // soft.f08:2741-2946,3118-3131: the coefficients of a real pair are Hermitian
    // (cc^l_{-m,-n} = (-1)^{m+n} conj cc^l_{mn}), so the correlation is irsoft of them
    if (int rc = stage_corr(P, 1, f_ml, g_ml, 1, 1, nullptr, S(stream))) return rc;
    return run_inverse(P, DWT_REAL, P->cstage, P->nc, nullptr, cc, 1, 1, S(stream));
}

int so3ft_cross_correlation_ylm_real_3d(so3ft_handle h, const void *f_mls, const void *g_mls, void *cc,
                                        const double *rad, int64_t n_radial, int64_t r_lo, int64_t r_hi, void *stream)
{
    GETP(h);
    if (!f_mls || !g_mls || !cc) return fail(SO3FT_EINVAL, "null pointer");
    std::vector<double> w;
    if (int rc = radial_weights(rad, n_radial, r_lo, r_hi, w)) return rc;
    const int64_t nb = P->bw * (P->bw + 1) / 2;
    if (int rc = stage_corr(P, 1, reinterpret_cast<const double2 *>(f_mls) + r_lo * nb,
                            reinterpret_cast<const double2 *>(g_mls) + r_lo * nb, 1, (int)w.size(), w.data(),
                            S(stream)))
        return rc;
    return run_inverse(P, DWT_REAL, P->cstage, P->nc, nullptr, cc, 1, 1, S(stream));
}

int so3ft_inverse_wigner_trf_corr_cmplx(so3ft_handle h, const void *f_lm, const void *g_lm, void *fft_array,
                                        void *stream)
{
    GETP(h);
    if (!f_lm || !g_lm || !fft_array) return fail(SO3FT_EINVAL, "null pointer");
    if (int rc = stage_corr(P, 0, f_lm, g_lm, 1, 1, nullptr, S(stream))) return rc;
    const int64_t N3 = P->N * P->N * P->N;
    double2 *F = reinterpret_cast<double2 *>(fft_array);
    if (int rc = dwt_inverse(P, DWT_COMPLEX, P->cstage, P->nc, nullptr, F, N3, 1, S(stream))) return rc;
    HIPCHK(launch_zero_invalid_columns(F, N3, (int)P->bw, (int)P->N, (int)P->lmax, 1, 0, S(stream)));
    return SO3FT_OK;
}

int so3ft_inverse_wigner_trf_corr_real(so3ft_handle h, const void *f_ml, const void *g_ml, void *fft_array,
                                       void *stream)
{
    GETP(h);
    if (!f_ml || !g_ml || !fft_array) return fail(SO3FT_EINVAL, "null pointer");
    if (int rc = stage_corr(P, 1, f_ml, g_ml, 1, 1, nullptr, S(stream))) return rc;
    const int64_t N3 = P->N * P->N * P->N;
    double2 *F = reinterpret_cast<double2 *>(fft_array);
    if (int rc = dwt_inverse(P, DWT_REAL, P->cstage, P->nc, nullptr, F, N3, 1, S(stream))) return rc;
    // the half-complex array of soft.f08:2874-2945: columns m >= 0 (and (0, -n)), zero elsewhere
    HIPCHK(launch_zero_invalid_columns(F, N3, (int)P->bw, (int)P->N, (int)P->lmax, 1, 1, S(stream)));
    return SO3FT_OK;
}

int so3ft_bench_kernel(so3ft_handle h, int which, int reps, void *a, void *b, int64_t T, void *stream,
                       float *ms_per_launch)
{
    GETP(h);
    if (!a || !b || !ms_per_launch || reps < 1 || T < 1) return fail(SO3FT_EINVAL, "bad argument");
    hipStream_t s = S(stream);
    const int64_t N = P->N, N2 = N * N, N3 = N2 * N;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto launch = [&]() -> hipError_t {
        switch (which) {
        case 0:
            return launch_dwt_inverse(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                      reinterpret_cast<const double2 *>(a), P->nc, nullptr,
                                      reinterpret_cast<double2 *>(b), N3, (int)T, s);
        case 1:
            return launch_dwt_forward(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                      reinterpret_cast<const double2 *>(a), N3, reinterpret_cast<double2 *>(b),
                                      P->nc, (int)T, s);
        case 4:
            return launch_dwt_inverse_valu(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                           reinterpret_cast<const double2 *>(a), P->nc, nullptr, 0,
                                           reinterpret_cast<double2 *>(b), N3, (int)T, s);
        case 5:   // the default FP64 forward (complex mode)
            return launch_dwt_forward_m64(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                          reinterpret_cast<const double2 *>(a), N3, reinterpret_cast<double2 *>(b),
                                          P->nc, (int)T, s);
        case 12:  // the FP64 forward on the vector ALUs
            return launch_dwt_forward_valu(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                           reinterpret_cast<const double2 *>(a), N3, reinterpret_cast<double2 *>(b),
                                           P->nc, (int)T, s);
        case 2: {
            FftAxis ax{};
            ax.in_t = N3; ax.in_o = N2; ax.in_r = N; ax.out_t = N3; ax.out_o = N2; ax.out_r = N;
            ax.sign = -1; ax.scale = 1.0;
            return launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, a, b, ax, (int)T, s);
        }
        case 3: {
            FftAxis ax{};
            ax.in_t = N3; ax.in_o = N; ax.in_r = N2; ax.out_t = N3; ax.out_o = N; ax.out_r = N2;
            ax.sign = -1; ax.scale = 1.0;
            return launch_fft_axis(P->fft, P->tb.tw, (int)P->bw, (int)P->lmax, a, b, ax, (int)T, s);
        }
        case 6:
            return launch_dwt_inverse_mixed(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                            reinterpret_cast<const float2 *>(a), P->nc, nullptr,
                                            reinterpret_cast<float2 *>(b), N3, (int)T, s);
        case 7:
            return launch_dwt_forward_mixed(P->tb, (int)P->bw, (int)N, (int)P->npair, DWT_COMPLEX,
                                            reinterpret_cast<const float2 *>(a), N3, reinterpret_cast<float2 *>(b),
                                            P->nc, (int)T, s);
        case 8:
        case 9: {
            FftAxis ax{};
            ax.in_t = N3; ax.out_t = N3; ax.sign = -1; ax.scale = 1.0;
            if (which == 8) { ax.in_o = N2; ax.in_r = N; ax.out_o = N2; ax.out_r = N; }
            else { ax.in_o = N; ax.in_r = N2; ax.out_o = N; ax.out_r = N2; }
            return launch_fft_reg((int)N, P->tb.tw, (int)P->bw, (int)P->lmax, a, b, ax, (int)T, 1, 1, 0, (int)N, s);
        }
        case 13:  // copy of the T complex64 grids a -> b: the streaming bandwidth reference
            return launch_copy16(a, b, T * N3 / 2, s);
        default:
            return hipErrorInvalidValue;
        }
    };
    if (which < 0 || (which > 9 && which != 12 && which != 13)) return fail(SO3FT_EINVAL, "unknown kernel id");
    HIPCHK(launch());  // warm-up
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) HIPCHK(launch());
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_launch = ms / (float)reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SO3FT_OK;
}

int so3ft_enforce_real_sym_mnl(int64_t bw, void *coeffs, int64_t n, void *stream)
{
    if (bw < 1 || bw > 512 || !coeffs || n < 0) return fail(SO3FT_EINVAL, "bad argument");
    // slice table for this bw (cached per bw)
    static std::mutex mu;
    static std::map<std::pair<int, int64_t>, int64_t *> cache;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    int64_t *d = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto key = std::make_pair(dev, bw);
        auto it = cache.find(key);
        if (it == cache.end()) {
            int rc = upload(&d, slice_table(bw));
            if (rc) return rc;
            cache[key] = d;
        } else {
            d = it->second;
        }
    }
    const int64_t nc = (4 * bw * bw * bw - bw) / 3;
    HIPCHK(launch_enforce_real_sym(reinterpret_cast<double2 *>(coeffs), nc, d, (int)bw, (int)n, 1, S(stream)));
    return SO3FT_OK;
}

int so3ft_enforce_real_sym_lmn(int64_t bw, void *coeffs, int64_t n, void *stream)
{
    if (bw < 1 || bw > 512 || !coeffs || n < 0) return fail(SO3FT_EINVAL, "bad argument");
    if (n == 0) return SO3FT_OK;
    // reorder to mnl in a work buffer, apply the lmn routine's rule, reorder back
    static std::mutex mu;
    static std::map<std::pair<int, int64_t>, std::pair<int64_t *, int32_t *>> cache;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::pair<int64_t *, int32_t *> tabs;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto key = std::make_pair(dev, bw);
        auto it = cache.find(key);
        if (it == cache.end()) {
            int64_t *sl = nullptr;
            int32_t *pm = nullptr;
            if (int rc = upload(&sl, slice_table(bw))) return rc;
            if (int rc = upload(&pm, lmn_perm_table(bw))) return rc;
            tabs = cache[key] = std::make_pair(sl, pm);
        } else {
            tabs = it->second;
        }
    }
    const int64_t nc = (4 * bw * bw * bw - bw) / 3;
    hipStream_t s = S(stream);
    const int64_t ch = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(1 << 30) / (16 * nc)));
    double2 *w = nullptr;
    HIPCHK(hipMalloc((void **)&w, sizeof(double2) * nc * ch));
    double2 *c = reinterpret_cast<double2 *>(coeffs);
    hipError_t e = hipSuccess;
    for (int64_t i0 = 0; i0 < n && e == hipSuccess; i0 += ch) {
        const int T = (int)std::min(ch, n - i0);
        e = launch_permute_lmn(tabs.second, nc, nc, c + i0 * nc, nc, w, nc, 16, 1, T, s);
        if (e == hipSuccess) e = launch_enforce_real_sym(w, nc, tabs.first, (int)bw, T, 0, s);
        if (e == hipSuccess) e = launch_permute_lmn(tabs.second, nc, nc, w, nc, c + i0 * nc, nc, 16, 0, T, s);
    }
    const hipError_t es = hipStreamSynchronize(s);  // the work buffer is freed below
    hipFree(w);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return fail(SO3FT_EHIP, std::string("enforce_real_sym_lmn: ") + hipGetErrorString(e));
    return SO3FT_OK;
}

}  // extern "C"

namespace so3 {

int set_error(int code, const std::string &msg) { return fail(code, msg); }

int plan_device_tables(int64_t h, DeviceTables *tb, int64_t *bw, int64_t *lmax)
{
    Plan *P = get(h);
    if (!P) return fail(SO3FT_EHANDLE, "unknown handle");
    *tb = P->tb;
    *bw = P->bw;
    *lmax = P->lmax;
    return SO3FT_OK;
}

}  // namespace so3
