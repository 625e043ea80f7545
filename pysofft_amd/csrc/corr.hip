// Peak search over correlation grids: per transform, the flat index of the maximum of
// |C| over the C-order (gamma, alpha, beta) grid, the way the reference's tests locate
// the relative rotation (tests/test_fortran_backend.py:718, np.argmax(np.abs(corr)));
// ties resolve to the smallest C-order index.
#include <algorithm>
#include <cstdint>

#include "so3ft_internal.h"

namespace so3 {

// (value, index) with the larger value winning and the smaller index breaking ties
__device__ __forceinline__ void better(double &v, int64_t &i, double v2, int64_t i2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__device__ __forceinline__ void block_reduce(double &v, int64_t &i, double *sv, int64_t *si)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_down(v, o, 64);
        const int64_t i2 = __shfl_down(i, o, 64);
        better(v, i, v2, i2);
    }
    if (lane == 0) { sv[wid] = v; si[wid] = i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) better(v, i, sv[w], si[w]);
    }
}

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
        better(v, i, z.x * z.x + z.y * z.y, k);
    }
    block_reduce(v, i, sv, si);
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
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) better(v, i, pv[t * nblk + k], pi[t * nblk + k]);
    block_reduce(v, i, sv, si);
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
    hipLaunchKernelGGL(argmax_final_kernel, dim3(T), dim3(256), 0, s, pv, pi, nblk, idx, val);
    return hipGetLastError();
}

}  // namespace so3
