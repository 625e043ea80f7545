// Peak search over correlation grids: per transform, the flat index of the maximum of
// |C| over the C-order (gamma, alpha, beta) grid, the way the reference's tests locate
// the relative rotation (tests/test_fortran_backend.py:718, np.argmax(np.abs(corr)));
// ties resolve to the smallest C-order index.
#include <algorithm>
#include <cstdint>

#include "so3ft_internal.h"

namespace so3 {

// stage 1: grid (nblk, T); block b of transform t scans [b*chunk, (b+1)*chunk)
__global__ __launch_bounds__(256) void argmax_partial_kernel(const double2 *__restrict__ x, int64_t n_per, int64_t chunk,
                                                             double *__restrict__ pv, int64_t *__restrict__ pi)
{
    __shared__ double sv[4];
    __shared__ int64_t si[4];
    const int64_t t = blockIdx.y;
    const double2 *xt = x + t * n_per;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(lo + chunk, n_per);
    double v = -1.0;
    int64_t i = INT64_MAX;
    for (int64_t k = lo + threadIdx.x; k < hi; k += blockDim.x) {
        const double2 z = xt[k];
        amax_better(v, i, z.x * z.x + z.y * z.y, k);
    }
    amax_block_reduce(v, i, sv, si);
    if (threadIdx.x == 0) {
        pv[t * gridDim.x + blockIdx.x] = v;
        pi[t * gridDim.x + blockIdx.x] = i;
    }
}

// stage 2: one block per transform over its nblk partials
__global__ __launch_bounds__(256) void argmax_final_kernel(const double *__restrict__ pv, const int64_t *__restrict__ pi,
                                                           int nblk, int64_t *__restrict__ idx, double *__restrict__ val)
{
    __shared__ double sv[4];
    __shared__ int64_t si[4];
    const int64_t t = blockIdx.x;
    double v = -1.0;
    int64_t i = INT64_MAX;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) amax_better(v, i, pv[t * nblk + k], pi[t * nblk + k]);
    amax_block_reduce(v, i, sv, si);
    if (threadIdx.x == 0) {
        idx[t] = i;
        if (val) val[t] = sqrt(v);
    }
}

int argmax_partials(int64_t n_per) { return (int)std::min<int64_t>(256, (n_per + 8191) / 8192); }

hipError_t launch_argmax_abs(const double2 *x, int64_t n_per, int T, double *pv, int64_t *pi, int64_t *idx,
                             double *val, hipStream_t s)
{
    const int nblk = argmax_partials(n_per);
    const int64_t chunk = (n_per + nblk - 1) / nblk;
    hipLaunchKernelGGL(argmax_partial_kernel, dim3(nblk, T), dim3(256), 0, s, x, n_per, chunk, pv, pi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_argmax_final(pv, pi, nblk, T, idx, val, s);
}

hipError_t launch_argmax_final(const double *pv, const int64_t *pi, int nblk, int T, int64_t *idx, double *val,
                               hipStream_t s)
{
    hipLaunchKernelGGL(argmax_final_kernel, dim3(T), dim3(256), 0, s, pv, pi, nblk, idx, val);
    return hipGetLastError();
}

}  // namespace so3
