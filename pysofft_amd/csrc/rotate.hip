// Wigner-D rotation of spherical-harmonic coefficient sets (pysofft/soft.py:713-764):
//   y'^l_n = e^{-i n gamma} sum_m y^l_m e^{-i m alpha} d^l(-beta)[m][n]
// with the full small-d matrices built degree by degree by Risbo's recurrence
// (pysofft/wigner.f90:529-585), evaluated here in gather form: each output entry
// sums its (up to) four sources instead of scattering every input.
// One workgroup per rotation; the two (2bw-1)^2 work matrices live in a per-rotation
// global scratch slice (L2 resident for the sizes used).
#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace so3 {

static constexpr int ROT_THREADS = 256;

// one half step deg-1 -> deg (matrices row-major, sizes deg x deg -> (deg+1) x (deg+1))
__device__ inline void risbo_half(const double *__restrict__ src, double *__restrict__ dst, int deg, double c,
                                  double s)
{
    const int pd = deg + 1;
    const double inv = 1.0 / (double)deg;
    for (int idx = threadIdx.x; idx < pd * pd; idx += ROT_THREADS) {
        const int I = idx / pd, J = idx - I * pd;
        double acc = 0.0;
        // sources x(i,j) with i in {I, I-1}, j in {J, J-1}, 0 <= i,j < deg
        if (I < deg && J < deg) acc += inv * sqrt((double)(deg - I)) * sqrt((double)(deg - J)) * src[I * deg + J] * c;
        if (I >= 1 && J < deg) acc -= inv * sqrt((double)I) * sqrt((double)(deg - J)) * src[(I - 1) * deg + J] * s;
        if (I < deg && J >= 1) acc += inv * sqrt((double)(deg - I)) * sqrt((double)J) * src[I * deg + J - 1] * s;
        if (I >= 1 && J >= 1) acc += inv * sqrt((double)I) * sqrt((double)J) * src[(I - 1) * deg + J - 1] * c;
        dst[idx] = acc;
    }
}

__global__ __launch_bounds__(ROT_THREADS) void rotate_ylm_kernel(const double2 *__restrict__ y, int nset,
                                                                 const double *__restrict__ euler, int bw,
                                                                 double2 *__restrict__ out, double *__restrict__ work,
                                                                 int64_t work_stride, int per_rot_sets)
{
    const int r = blockIdx.x;
    const double al = euler[3 * r], be = euler[3 * r + 1], ga = euler[3 * r + 2];
    const double c = cos(be / 2), s = sin(-be / 2);
    double *A = work + (int64_t)r * work_stride;
    double *B = A + (int64_t)(2 * bw - 1) * (2 * bw - 1);
    const int nb = bw * bw;
    // per_rot_sets: 1 -> set r uses rotation r only (paired mode); 0 -> every set x every rotation
    const int s0 = per_rot_sets ? r : 0, s1 = per_rot_sets ? r + 1 : nset;
    for (int l = 0; l < bw; ++l) {
        const int w = 2 * l + 1;
        if (l == 0) {
            if (threadIdx.x == 0) A[0] = 1.0;
        } else if (l == 1) {
            if (threadIdx.x == 0) {
                const double d0 = c * c, d1 = sqrt(2.0) * c * s, d2 = s * s;
                A[0] = d0; A[1] = d1; A[2] = d2;
                A[3] = -d1; A[4] = d0 - d2; A[5] = d1;
                A[6] = d2; A[7] = -d1; A[8] = d0;
            }
        } else {
            risbo_half(A, B, 2 * l - 1, c, s);
            __syncthreads();
            risbo_half(B, A, 2 * l, c, s);
        }
        __syncthreads();
        // apply: out[j] = e^{-i (j-l) gamma} sum_i y[i] e^{-i (i-l) alpha} A[i][j]
        for (int idx = threadIdx.x; idx < (s1 - s0) * w; idx += ROT_THREADS) {
            const int st = s0 + idx / w, j = idx % w;
            const double2 *ys = y + (int64_t)st * nb + l * l;
            double re = 0.0, im = 0.0;
            for (int i = 0; i < w; ++i) {
                const double2 v = ys[i];
                double sa, ca;
                sincos(-(double)(i - l) * al, &sa, &ca);
                const double d = A[i * w + j];
                re += d * (v.x * ca - v.y * sa);
                im += d * (v.x * sa + v.y * ca);
            }
            double sg, cg;
            sincos(-(double)(j - l) * ga, &sg, &cg);
            const int64_t oi = per_rot_sets ? ((int64_t)r * nb) : (((int64_t)r * nset + st) * nb);
            out[oi + l * l + j] = make_double2(re * cg - im * sg, re * sg + im * cg);
        }
        __syncthreads();
    }
}

}  // namespace so3

extern "C" {

// Rotate nset coefficient sets (LMc layout, bw^2 each) by nrot ZYZ rotations.
// paired = 0: out is (nrot, nset, bw^2), every set by every rotation (soft.py:713-764);
// paired = 1: nset == nrot and set i is rotated by rotation i only, out is (nrot, bw^2).
// euler: device array (nrot, 3).  No counterpart entry point in the reference (there
// this is pure numpy); declared in include/so3ft.h.
int so3ft_rotate_ylm_cmplx(int64_t bw, const void *ylms, int64_t nset, const void *euler, int64_t nrot,
                           int paired, void *out, void *stream)
{
    using namespace so3;
    if (bw < 1 || bw > 512 || !ylms || !euler || !out || nset < 0 || nrot < 0) return SO3FT_EINVAL;
    if (paired && nset != nrot) return SO3FT_EINVAL;
    if (nrot == 0 || nset == 0) return SO3FT_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t stride = 2 * (2 * bw - 1) * (2 * bw - 1);
    // bounded scratch: rotations in chunks
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nrot, (int64_t)(512ll * 1024 * 1024 / (8 * stride))));
    double *work = nullptr;
    if (hipMallocAsync((void **)&work, sizeof(double) * stride * chunk, s) != hipSuccess) return SO3FT_ENOMEM;
    for (int64_t r0 = 0; r0 < nrot; r0 += chunk) {
        const int64_t nr = std::min(chunk, nrot - r0);
        const double *e = reinterpret_cast<const double *>(euler) + 3 * r0;
        const double2 *y = reinterpret_cast<const double2 *>(ylms) + (paired ? r0 * bw * bw : 0);
        double2 *o = reinterpret_cast<double2 *>(out) + r0 * (paired ? 1 : nset) * bw * bw;
        hipLaunchKernelGGL(rotate_ylm_kernel, dim3((unsigned)nr), dim3(ROT_THREADS), 0, s, y,
                           (int)(paired ? nr : nset), e, (int)bw, o, work, stride, paired);
        if (hipGetLastError() != hipSuccess) {
            (void)hipFreeAsync(work, s);
            return SO3FT_EHIP;
        }
    }
    (void)hipFreeAsync(work, s);
    return SO3FT_OK;
}

}  // extern "C"
