// Risbo's recurrence for the Wigner small-d functions on the device, and the plan's
// d table.
//
//  - reduced recurrence (pysofft/wigner.f90:626-748): d^{l-1}_{m,n} (0 <= m <= n <= l-1) ->
//    d^l_{m,n} through the half degree l - 1/2, at every beta of a set closed under
//    beta -> pi - beta (the second half step reads the half-degree values at pi - beta as
//    the reversed array).  Each half step is one launch over (triangle entry, beta);
//    arrays are [tri][nb], betas contiguous, tri(m, n) = m (2 s - m + 1) / 2 + n - m on
//    the triangle of side s.
//  - the d table of a plan: d~^l_{p,q}(beta_k) for 0 <= p <= q <= l < bw at the 2bw beta
//    samples, rows in mnl order (pair-major, then degree; utils.f90:683-695) x 2bw betas,
//    the layout of the reference's precomputed wigner_d before its weights
//    (soft.f08:194-219).  Filled by the reduced Risbo recurrence (genwig_all_risbo,
//    wigner.f90:784-822) or by the Kostelec three-term recurrence (genwig_all_kostelec,
//    wigner.f90:423-465).
//  - full_big_d_risbo (wigner.f90:844-934) and the full-matrix step
//    wigner_mn_recurrence_risbo (wigner.f90:529-585) of the reference's make_wigner module.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace so3 {

// (m, n) of flat index t on the triangle of side s
__device__ inline void tri_mn(int64_t t, int s, int &m, int &n)
{
    // m = largest value with m (2 s - m + 1) / 2 <= t
    const double a = 2.0 * s + 1.0;
    int mm = (int)((a - sqrt(a * a - 8.0 * (double)t)) * 0.5);
    mm = max(0, min(mm, s - 1));
    while (mm > 0 && (int64_t)mm * (2 * s - mm + 1) / 2 > t) --mm;
    while (mm + 1 < s && (int64_t)(mm + 1) * (2 * s - mm) / 2 <= t) ++mm;
    m = mm;
    n = (int)(t - (int64_t)mm * (2 * s - mm + 1) / 2) + mm;
}

__device__ __forceinline__ int64_t tri_s(int m, int n, int s) { return (int64_t)m * (2 * s - m + 1) / 2 + n - m; }

// degrees 0 and 1 (wigner.f90:637-645)
__global__ void risbo_low_kernel(double *d, int l, int nb, const double *__restrict__ c, const double *__restrict__ s,
                                 double sq2, int normalized)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    if (l == 0) {
        d[k] = normalized ? sqrt(0.5) : 1.0;
        return;
    }
    const double nc = normalized ? sqrt(1.5) : 1.0;
    d[k] = (c[k] * c[k] - s[k] * s[k]) * nc;
    d[nb + k] = sq2 * c[k] * s[k] * nc;
    d[2 * nb + k] = c[k] * c[k] * nc;
}

// l - 1 -> l - 1/2 on the side-l triangle (l >= 2)
__global__ void risbo_half1_kernel(const double *__restrict__ d1, double *__restrict__ h, int l, int nb,
                                   const double *__restrict__ c, const double *__restrict__ s, int normalized)
{
    const int64_t nt = (int64_t)l * (l + 1) / 2;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= nt * nb) return;
    const int64_t t = idx / nb;
    const int k = (int)(idx - t * nb);
    int m, n;
    tri_mn(t, l, m, n);
    const double inv = 1.0 / (double)(2 * l - 1);
    const int ll = l - 1;
    const double nc = normalized ? inv * sqrt((double)(2 * l) / (double)(2 * l - 1)) : inv;
    const double ck = c[k], sk = s[k];
    auto D = [&](int a, int b) { return d1[tri_s(a, b, l) * nb + k]; };
    double v;
    if (m == l - 1) {   // m = n = l - 1
        v = D(l - 1, l - 1) * (double)(2 * l - 1) * ck * nc;
    } else if (n == l - 1) {
        v = (D(m, l - 1) * sqrt((double)((l + m) * (2 * l - 1))) * ck +
             D(m + 1, l - 1) * sqrt((double)((ll - m) * (2 * l - 1))) * sk) * nc;
    } else if (n == m) {
        v = (D(m, m) * (double)(l + m) * ck - D(m, m + 1) * sqrt((double)((l + m) * (ll - m))) * sk -
             D(m, m + 1) * sqrt((double)((ll - m) * (l + m))) * sk + D(m + 1, m + 1) * (double)(ll - m) * ck) * nc;
    } else {
        v = (D(m, n) * sqrt((double)((l + m) * (l + n))) * ck - D(m, n + 1) * sqrt((double)((l + m) * (ll - n))) * sk +
             D(m + 1, n) * sqrt((double)((ll - m) * (l + n))) * sk +
             D(m + 1, n + 1) * sqrt((double)((ll - m) * (ll - n))) * ck) * nc;
    }
    h[idx] = v;
}

// l - 1/2 -> l on the side-(l+1) triangle (l >= 2); with tab, every entry is also written
// to its row of the plan's d table (nb = 2 bw betas)
__global__ void risbo_half2_kernel(const double *__restrict__ h, double *__restrict__ d, int l, int nb,
                                   const double *__restrict__ c, const double *__restrict__ s, int normalized,
                                   double *__restrict__ tab, int bw)
{
    const int64_t nt = (int64_t)(l + 1) * (l + 2) / 2;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= nt * nb) return;
    const int64_t t = idx / nb;
    const int k = (int)(idx - t * nb);
    int m, n;
    tri_mn(t, l + 1, m, n);
    const double inv = 1.0 / (double)(2 * l);
    const double nc = normalized ? inv * sqrt((double)(2 * l + 1) / (double)(2 * l)) : inv;
    const double nc2 = normalized ? 0.5 * sqrt((double)(2 * l + 1) / (double)(2 * l)) : 0.5;
    const double lsym = (l & 1) ? -1.0 : 1.0, lpsym = -lsym;
    const double ck = c[k], sk = s[k];
    auto H = [&](int a, int b) { return h[tri_s(a, b, l) * nb + k]; };
    auto HR = [&](int a, int b) { return h[tri_s(a, b, l) * nb + (nb - 1 - k)]; };   // at pi - beta
    double v;
    if (m == 0) {
        if (n == 0) {
            v = (2.0 * H(0, 0) * ck - lpsym * HR(0, 0) * sk + lsym * HR(0, 0) * sk) * nc2;
        } else if (n < l) {
            const double nsym = (n & 1) ? -1.0 : 1.0;
            v = (lsym * nsym * HR(0, n - 1) * sqrt((double)(l * (l + n))) * ck -
                 lpsym * nsym * HR(0, n) * sqrt((double)(l * (l - n))) * sk +
                 H(0, n - 1) * sqrt((double)(l * (l + n))) * sk + H(0, n) * sqrt((double)(l * (l - n))) * ck) * nc;
        } else {
            v = (HR(0, l - 1) * ck + H(0, l - 1) * sk) * nc2 * sqrt(2.0);
        }
    } else if (m == l) {   // m = n = l
        v = H(l - 1, l - 1) * (double)(2 * l) * ck * nc;
    } else if (n == l) {
        v = (H(m - 1, l - 1) * sqrt((double)((l + m) * (2 * l))) * ck +
             H(m, l - 1) * sqrt((double)((l - m) * (2 * l))) * sk) * nc;
    } else if (n == m) {
        v = (H(m - 1, m - 1) * (double)(l + m) * ck - H(m - 1, m) * sqrt((double)((l + m) * (l - m))) * sk -
             H(m - 1, m) * sqrt((double)((l - m) * (l + m))) * sk + H(m, m) * (double)(l - m) * ck) * nc;
    } else {
        v = (H(m - 1, n - 1) * sqrt((double)((l + m) * (l + n))) * ck -
             H(m - 1, n) * sqrt((double)((l + m) * (l - n))) * sk +
             H(m, n - 1) * sqrt((double)((l - m) * (l + n))) * sk + H(m, n) * sqrt((double)((l - m) * (l - n))) * ck) *
            nc;
    }
    d[idx] = v;
    if (tab) tab[(dtab_row(m, n, bw) + (l - n)) * nb + k] = v;
}

// degrees 0 and 1 of the table (the low kernel's values)
__global__ void risbo_low_to_table(const double *__restrict__ d, int l, int nb, double *__restrict__ tab, int bw)
{
    const int64_t nt = (int64_t)(l + 1) * (l + 2) / 2;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= nt * nb) return;
    const int64_t t = idx / nb;
    const int k = (int)(idx - t * nb);
    int m, n;
    tri_mn(t, l + 1, m, n);
    tab[(dtab_row(m, n, bw) + (l - n)) * nb + k] = d[idx];
}

static int nblk(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// one reduced step l - 1 -> l: d1 [l(l+1)/2][nb] -> d [(l+1)(l+2)/2][nb], h scratch
// [l(l+1)/2][nb]; tab (may be null): also write degree l into the plan table (nb = 2 bw)
hipError_t launch_risbo_reduced_step(const double *d1, double *h, double *d, int l, int nb, const double *c,
                                     const double *s, double sq2, int normalized, double *tab, int bw, hipStream_t st)
{
    if (l < 2) {
        hipLaunchKernelGGL(risbo_low_kernel, dim3(nblk(nb, 256)), dim3(256), 0, st, d, l, nb, c, s, sq2, normalized);
        if (tab) {
            const int64_t n = (int64_t)(l + 1) * (l + 2) / 2 * nb;
            hipLaunchKernelGGL(risbo_low_to_table, dim3(nblk(n, 256)), dim3(256), 0, st, d, l, nb, tab, bw);
        }
        return hipGetLastError();
    }
    const int64_t n1 = (int64_t)l * (l + 1) / 2 * nb, n2 = (int64_t)(l + 1) * (l + 2) / 2 * nb;
    hipLaunchKernelGGL(risbo_half1_kernel, dim3(nblk(n1, 256)), dim3(256), 0, st, d1, h, l, nb, c, s, normalized);
    hipLaunchKernelGGL(risbo_half2_kernel, dim3(nblk(n2, 256)), dim3(256), 0, st, h, d, l, nb, c, s, normalized, tab,
                       bw);
    return hipGetLastError();
}

// half-angle cos / sin of the 2bw beta samples (wigner.f90:69-75)
__global__ void risbo_trig_kernel(double *c, double *s, int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double b = (double)(2 * k + 1) * (3.14159265358979323846 / (double)(2 * N));
    c[k] = cos(b / 2);
    s[k] = sin(b / 2);
}

// Kostelec table: one thread per (pair, beta) runs the three-term recurrence from the seed
// (constants rc of the plan: (a, x, b) per step, wigner.f90:227-260)
__global__ void kostelec_table_kernel(double *__restrict__ tab, const double *__restrict__ seeds,
                                      const double *__restrict__ cosb, const double *__restrict__ rc,
                                      const int64_t *__restrict__ rc_off, int bw, int N)
{
    const int64_t ntri = (int64_t)bw * (bw + 1) / 2;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= ntri * N) return;
    const int64_t t = idx / N;
    const int k = (int)(idx - t * N);
    int p, q;
    tri_mn(t, bw, p, q);
    const int L = bw - q;
    const double *r = rc + 3 * rc_off[t];
    const double cb = cosb[k];
    double dc = seeds[t * N + k], dp = 0.0;
    double *row = tab + dtab_row(p, q, bw) * N + k;
    for (int j = 0; j < L; ++j) {
        row[(int64_t)j * N] = dc;
        const double dn = r[3 * j] * (cb - r[3 * j + 1]) * dc - r[3 * j + 2] * dp;
        dp = dc;
        dc = dn;
    }
}

int64_t dtab_entries(int bw) { return (int64_t)bw * (2 + 3 * (int64_t)bw + (int64_t)bw * bw) / 6 * 2 * bw; }

// fills tab (dtab_entries(bw) doubles) with the Risbo recurrence; work: 3 bw(bw+1)/2 x 2bw
// doubles of scratch (two degree triangles and the half step)
hipError_t fill_dtab_risbo(int bw, double *tab, double *work, hipStream_t st)
{
    const int N = 2 * bw;
    const int64_t tsz = (int64_t)bw * (bw + 1) / 2 * N;
    double *a = work, *b = work + tsz, *h = work + 2 * tsz;
    double *c = work + 3 * tsz, *s = c + N;   // half-angle cos / sin of the betas
    hipLaunchKernelGGL(risbo_trig_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, c, s, N);
    for (int l = 0; l < bw; ++l) {
        hipError_t e = launch_risbo_reduced_step(a, h, b, l, N, c, s, sqrt(2.0), 1, tab, bw, st);
        if (e != hipSuccess) return e;
        std::swap(a, b);
    }
    return hipGetLastError();
}

int64_t dtab_work_entries(int bw) { return 3 * (int64_t)bw * (bw + 1) / 2 * 2 * bw + 4 * (int64_t)bw; }

hipError_t fill_dtab_kostelec(const DeviceTables &tb, int bw, double *tab, hipStream_t st)
{
    const int N = 2 * bw;
    const int64_t n = (int64_t)bw * (bw + 1) / 2 * N;
    hipLaunchKernelGGL(kostelec_table_kernel, dim3(nblk(n, 128)), dim3(128), 0, st, tab, tb.seeds, tb.cosb, tb.rc,
                       tb.rc_off, bw, N);
    return hipGetLastError();
}

// full_big_d_risbo: spread degree l of the reduced values at (beta, pi - beta) (d: [tri][2])
// over the eight (m, n) variants of wigner.f90:870-929
__global__ void big_d_fill_kernel(const double *__restrict__ d, int l, int bw, double alpha, double gamma,
                                  double scale, const int64_t *__restrict__ slice, double2 *__restrict__ out)
{
    const int64_t nt = (int64_t)(l + 1) * (l + 2) / 2;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int m, n;
    tri_mn(t, l + 1, m, n);
    const double d1 = d[2 * t], d2 = d[2 * t + 1];
    const double sl = (l & 1) ? -1.0 : 1.0, sm = (m & 1) ? -1.0 : 1.0, sn = (n & 1) ? -1.0 : 1.0;
    auto put = [&](int mm, int nn, double v) {
        const int am = mm < 0 ? -mm : mm, an = nn < 0 ? -nn : nn;
        const int64_t o = slice[(int64_t)(mm + bw - 1) * (2 * bw - 1) + (nn + bw - 1)] + (l - max(am, an));
        // e^{-i mm alpha} e^{-i nn gamma} v
        double sa, ca, sg, cg;
        sincos(-(double)mm * alpha, &sa, &ca);
        sincos(-(double)nn * gamma, &sg, &cg);
        out[o] = make_double2((ca * cg - sa * sg) * v * scale, (ca * sg + sa * cg) * v * scale);
    };
    put(m, n, d1);
    if (m == 0 && n == 0) return;
    put(-n, -m, d1);
    if (m != n) {
        put(n, m, sm * sn * d1);
        put(-m, -n, sm * sn * d1);
    }
    if (m == 0 || n == 0) return;
    put(-m, n, sm * sl * d2);
    put(m, -n, sn * sl * d2);
    if (m == n) return;
    put(-n, m, sm * sl * d2);
    put(n, -m, sn * sl * d2);
}

// full-matrix Risbo step (wigner.f90:529-585), both half steps in gather form: output
// (I, J) of degree l (side 2l+1) from the (2l)^2 half-degree values, each of those from
// the (2l-1)^2 inputs.  Matrices row-major per beta; sq[k] = sqrt(k).
__device__ inline double risbo_full_half(const double *__restrict__ x, int deg, int I, int J, double c, double s,
                                         const double *__restrict__ sq)
{
    // target (I, J) of the (deg+1)^2 result: sources (i, j) of the deg^2 input with
    // i in {I-1, I}, j in {J-1, J}, in the order the scattering loop adds them
    const double inv = 1.0 / (double)deg;
    double acc = 0.0;
    if (I >= 1 && J >= 1) acc += inv * sq[I] * sq[J] * x[(I - 1) * deg + (J - 1)] * c;
    if (I >= 1 && J < deg) acc -= inv * sq[I] * sq[deg - J] * x[(I - 1) * deg + J] * s;
    if (I < deg && J >= 1) acc += inv * sq[deg - I] * sq[J] * x[I * deg + (J - 1)] * s;
    if (I < deg && J < deg) acc += inv * sq[deg - I] * sq[deg - J] * x[I * deg + J] * c;
    return acc;
}

__global__ void risbo_full_kernel(const double *__restrict__ dprev, int l, int nb, const double *__restrict__ cb,
                                  const double *__restrict__ sb, const double *__restrict__ sq,
                                  double *__restrict__ dout)
{
    const int w = 2 * l + 1;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nb * w * w) return;
    const int k = (int)(idx / ((int64_t)w * w));
    const int e = (int)(idx - (int64_t)k * w * w);
    const int I = e / w, J = e - I * w;
    const double c = cb[k], s = sb[k];
    double v;
    if (l == 0) {
        v = 1.0;
    } else if (l == 1) {
        const double d0 = c * c, d1 = sq[2] * c * s, d2 = s * s;
        const double tbl[9] = {d0, d1, d2, -d1, d0 - d2, d1, d2, -d1, d0};
        v = tbl[e];
    } else {
        const int deg1 = 2 * l - 1, deg2 = 2 * l;
        const double *x = dprev + (int64_t)k * deg1 * deg1;
        const double inv = 1.0 / (double)deg2;
        double t11 = 0.0, t10 = 0.0, t01 = 0.0, t00 = 0.0;
        if (I >= 1 && J >= 1) t11 = risbo_full_half(x, deg1, I - 1, J - 1, c, s, sq);
        if (I >= 1 && J < deg2) t10 = risbo_full_half(x, deg1, I - 1, J, c, s, sq);
        if (I < deg2 && J >= 1) t01 = risbo_full_half(x, deg1, I, J - 1, c, s, sq);
        if (I < deg2 && J < deg2) t00 = risbo_full_half(x, deg1, I, J, c, s, sq);
        v = 0.0;
        if (I >= 1 && J >= 1) v += inv * sq[I] * sq[J] * t11 * c;
        if (I >= 1 && J < deg2) v -= inv * sq[I] * sq[deg2 - J] * t10 * s;
        if (I < deg2 && J >= 1) v += inv * sq[deg2 - I] * sq[J] * t01 * s;
        if (I < deg2 && J < deg2) v += inv * sq[deg2 - I] * sq[deg2 - J] * t00 * c;
    }
    dout[idx] = v;
}

}  // namespace so3

extern "C" {

using namespace so3;

static hipStream_t STR(void *s) { return reinterpret_cast<hipStream_t>(s); }

int so3ft_wigner_mn_recurrence_risbo(const void *d_prev, int64_t l, const void *cos_b, const void *sin_b,
                                     const void *sqrts, int64_t nsqrt, int64_t nbeta, void *d_out, void *stream)
{
    if (l < 0 || nbeta < 0 || !cos_b || !sin_b || !sqrts || !d_out || (l >= 2 && !d_prev))
        return set_error(SO3FT_EINVAL, "so3ft_wigner_mn_recurrence_risbo: bad argument");
    if (nsqrt < std::max<int64_t>(3, 2 * l + 1))
        return set_error(SO3FT_EINVAL, "so3ft_wigner_mn_recurrence_risbo: sqrts needs at least max(3, 2l+1) entries");
    if (l > 4096) return set_error(SO3FT_EINVAL, "so3ft_wigner_mn_recurrence_risbo: l > 4096");
    if (nbeta == 0) return SO3FT_OK;
    const int w = 2 * (int)l + 1;
    const int64_t n = nbeta * w * w;
    hipLaunchKernelGGL(risbo_full_kernel, dim3(nblk(n, 256)), dim3(256), 0, STR(stream),
                       reinterpret_cast<const double *>(d_prev), (int)l, (int)nbeta,
                       reinterpret_cast<const double *>(cos_b), reinterpret_cast<const double *>(sin_b),
                       reinterpret_cast<const double *>(sqrts), reinterpret_cast<double *>(d_out));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SO3FT_EHIP, std::string("risbo_full_kernel: ") + hipGetErrorString(e));
    return SO3FT_OK;
}

int so3ft_wigner_recurrence_risbo_reduced(const void *dl_prev, int64_t l, const void *cos_b, const void *sin_b,
                                          const void *sqrts, int64_t nsqrt, int64_t nbeta, int normalized,
                                          void *dl_out, void *stream)
{
    if (l < 0 || nbeta < 0 || !cos_b || !sin_b || !dl_out || (l >= 2 && !dl_prev) || (l == 1 && !sqrts))
        return set_error(SO3FT_EINVAL, "so3ft_wigner_recurrence_risbo_reduced: bad argument");
    if (l == 1 && nsqrt < 3)
        return set_error(SO3FT_EINVAL, "so3ft_wigner_recurrence_risbo_reduced: sqrts needs at least 3 entries");
    if (l > 4096) return set_error(SO3FT_EINVAL, "so3ft_wigner_recurrence_risbo_reduced: l > 4096");
    if (nbeta == 0) return SO3FT_OK;
    hipStream_t st = STR(stream);
    double sq2 = std::sqrt(2.0);
    if (l == 1) {   // the reference reads sqrts(3) (= sqrt 2) from the caller's table
        hipError_t e = hipMemcpyAsync(&sq2, reinterpret_cast<const double *>(sqrts) + 2, sizeof(double),
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return set_error(SO3FT_EHIP, std::string("sqrts read: ") + hipGetErrorString(e));
    }
    double *h = nullptr;
    if (l >= 2) {
        const size_t bytes = sizeof(double) * (size_t)(l * (l + 1) / 2) * (size_t)nbeta;
        if (hipMallocAsync((void **)&h, bytes, st) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(SO3FT_ENOMEM, "so3ft_wigner_recurrence_risbo_reduced: scratch allocation failed");
        }
    }
    hipError_t e = launch_risbo_reduced_step(reinterpret_cast<const double *>(dl_prev), h,
                                             reinterpret_cast<double *>(dl_out), (int)l, (int)nbeta,
                                             reinterpret_cast<const double *>(cos_b),
                                             reinterpret_cast<const double *>(sin_b), sq2, normalized, nullptr, 0, st);
    if (h) (void)hipFreeAsync(h, st);
    if (e != hipSuccess) return set_error(SO3FT_EHIP, std::string("risbo step: ") + hipGetErrorString(e));
    return SO3FT_OK;
}

int so3ft_genwig_all(int64_t bw, int64_t recurrence_type, void *out, void *stream)
{
    if (bw < 1 || bw > 256 || !out || (recurrence_type != 0 && recurrence_type != 1))
        return set_error(SO3FT_EINVAL, "so3ft_genwig_all: bad argument");
    // the Kostelec fill needs a plan's seeds and constants: build a temporary plan
    so3ft_handle h = 0;
    if (recurrence_type == 0) {
        if (int rc = so3ft_init(bw, bw - 1, 0, 0, 0, 0, &h)) return rc;
        DeviceTables tb;
        int64_t b, lm;
        int rc = plan_device_tables(h, &tb, &b, &lm);
        hipError_t e = rc ? hipSuccess : fill_dtab_kostelec(tb, (int)bw, reinterpret_cast<double *>(out), STR(stream));
        if (!rc && e == hipSuccess) e = hipStreamSynchronize(STR(stream));   // before the plan's tables go
        so3ft_destroy(h);
        if (rc) return rc;
        if (e != hipSuccess) return set_error(SO3FT_EHIP, std::string("kostelec table: ") + hipGetErrorString(e));
        return SO3FT_OK;
    }
    double *work = nullptr;
    hipStream_t st = STR(stream);
    if (hipMallocAsync((void **)&work, sizeof(double) * (size_t)dtab_work_entries((int)bw), st) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(SO3FT_ENOMEM, "so3ft_genwig_all: scratch allocation failed");
    }
    hipError_t e = fill_dtab_risbo((int)bw, reinterpret_cast<double *>(out), work, st);
    (void)hipFreeAsync(work, st);
    if (e != hipSuccess) return set_error(SO3FT_EHIP, std::string("risbo table: ") + hipGetErrorString(e));
    return SO3FT_OK;
}

int so3ft_full_big_d_risbo(int64_t bw, double alpha, double beta, double gamma, int normalized, void *out,
                           void *stream)
{
    if (bw < 1 || bw > 1024 || !out) return set_error(SO3FT_EINVAL, "so3ft_full_big_d_risbo: bad argument");
    hipStream_t st = STR(stream);
    const int64_t nt = bw * (bw + 1) / 2;
    // work: two degree triangles and the half step at 2 betas, the trig values, the slice table
    const int64_t nsl = (2 * bw - 1) * (2 * bw - 1);
    std::vector<int64_t> sl((size_t)nsl, 0);
    {
        int64_t pos = 0;   // mnl slice starts (utils.f90:140-201)
        auto at = [&](int64_t m, int64_t n) -> int64_t & { return sl[(size_t)((m + bw - 1) * (2 * bw - 1) + n + bw - 1)]; };
        for (int64_t m1 = 0; m1 < bw; ++m1)
            for (int64_t m2 = m1; m2 < bw; ++m2) {
                const int64_t L = bw - m2;
                at(m1, m2) = pos; pos += L;
                if (m1 == 0 && m2 == 0) continue;
                at(-m2, -m1) = pos; pos += L;
                if (m1 != m2) { at(m2, m1) = pos; pos += L; at(-m1, -m2) = pos; pos += L; }
                if (m1 == 0 || m2 == 0) continue;
                at(m1, -m2) = pos; pos += L; at(-m1, m2) = pos; pos += L;
                if (m1 == m2) continue;
                at(m2, -m1) = pos; pos += L; at(-m2, m1) = pos; pos += L;
            }
    }
    const size_t wbytes = sizeof(double) * (size_t)(3 * 2 * nt + 4) + sizeof(int64_t) * (size_t)nsl;
    char *work = nullptr;
    if (hipMallocAsync((void **)&work, wbytes, st) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(SO3FT_ENOMEM, "so3ft_full_big_d_risbo: scratch allocation failed");
    }
    double *a = reinterpret_cast<double *>(work), *b = a + 2 * nt, *h = b + 2 * nt, *cs = h + 2 * nt;
    int64_t *slice = reinterpret_cast<int64_t *>(cs + 4);
    const double bb[2] = {beta, 3.14159265358979323846 - beta};
    const double trig[4] = {std::cos(bb[0] / 2), std::cos(bb[1] / 2), std::sin(bb[0] / 2), std::sin(bb[1] / 2)};
    hipError_t e = hipMemcpyAsync(cs, trig, sizeof(trig), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(slice, sl.data(), sizeof(int64_t) * (size_t)nsl, hipMemcpyHostToDevice, st);
    const double scale = normalized ? 1.0 / (2.0 * 3.14159265358979323846) : 1.0;
    for (int l = 0; l < bw && e == hipSuccess; ++l) {
        e = launch_risbo_reduced_step(a, h, b, l, 2, cs, cs + 2, std::sqrt(2.0), normalized, nullptr, 0, st);
        if (e != hipSuccess) break;
        const int64_t n = (int64_t)(l + 1) * (l + 2) / 2;
        hipLaunchKernelGGL(big_d_fill_kernel, dim3(nblk(n, 128)), dim3(128), 0, st, b, l, (int)bw, alpha, gamma, scale,
                           slice, reinterpret_cast<double2 *>(out));
        e = hipGetLastError();
        std::swap(a, b);
    }
    // the host arrays above are read by the copies: wait before they go out of scope
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipFreeAsync(work, st);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return set_error(SO3FT_EHIP, std::string("full_big_d_risbo: ") + hipGetErrorString(e));
    return SO3FT_OK;
}

}  // extern "C"
