// FP64 reductions and scalar-ratio updates for least-squares migration by
// conjugate gradients. The vectors stay FP32 in HBM; inner products are
// accumulated in FP64 (psktm_dot_f64_ws in a fixed order: per-workgroup
// partials in a caller workspace, then one workgroup adds them, so equal
// inputs give equal scalars on every run), and step lengths alpha = num/den
// are formed on the device from those FP64 scalars, so an iteration never
// syncs with the host.
#include "internal.hpp"

#include <algorithm>

namespace psktm {
namespace {

constexpr int kDotThreads = 256;

// Sum of a thread's value over the workgroup in a fixed order (wave
// butterflies, then the waves in index order); the total is valid in thread 0.
__device__ __forceinline__ double block_sum_f64(double v, double *s)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s[w];
    return t;
}

__global__ void zero_f64_kernel(double *r) { r[0] = 0.0; }

// psktm_dot_f64: each workgroup's FP64 sum is added to r[0] with an atomic,
// so the last bits depend on the order in which workgroups finish.
__global__ __launch_bounds__(kDotThreads) void dot_atomic_kernel(const float *__restrict__ a,
                                                                 const float *__restrict__ b,
                                                                 size_t n, double *__restrict__ r)
{
    __shared__ double s[kDotThreads / 64];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    const double t = block_sum_f64(acc, s);
    if (threadIdx.x == 0) atomicAdd(r, t);
}

// Pass 1: block b adds a[i]*b[i] over its grid-stride share in FP64 and
// writes partial[b]. The share of every block depends only on n and the grid,
// so the partials are the same numbers on every run.
__global__ __launch_bounds__(kDotThreads) void dot_partial_kernel(const float *__restrict__ a,
                                                                  const float *__restrict__ b,
                                                                  size_t n,
                                                                  double *__restrict__ partial)
{
    __shared__ double s[kDotThreads / 64];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    const double t = block_sum_f64(acc, s);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Pass 2: one workgroup adds the nb partials in a fixed order and writes
// result[0] (no clearing of the result beforehand, so no memset in a graph).
__global__ __launch_bounds__(kDotThreads) void dot_final_kernel(const double *__restrict__ partial,
                                                                int nb, double *__restrict__ r)
{
    __shared__ double s[kDotThreads / 64];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) acc += partial[q];
    const double t = block_sum_f64(acc, s);
    if (threadIdx.x == 0) r[0] = t;
}

__device__ __forceinline__ float ratio(const double *num, const double *den)
{
    const double d = den[0];
    return (d != 0.0) ? (float)(num[0] / d) : 0.0f;
}

__global__ void axpy_ratio_kernel(float *__restrict__ y, const float *__restrict__ x, size_t n,
                                  const double *__restrict__ num, const double *__restrict__ den,
                                  float sign)
{
    const float alpha = sign * ratio(num, den);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = fmaf(alpha, x[i], y[i]);
}

__global__ void xpby_ratio_kernel(float *__restrict__ p, const float *__restrict__ r, size_t n,
                                  const double *__restrict__ num, const double *__restrict__ den)
{
    const float beta = ratio(num, den);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = fmaf(beta, p[i], r[i]);
}

unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 4096);
}

} // namespace

// Both passes on `st`, partials in `partial` (kDotPartials doubles).
psktm_status launch_dot(const float *a, const float *b, size_t n, double *result,
                        double *partial, hipStream_t st)
{
    const unsigned blocks =
        (unsigned)std::max<size_t>(1, std::min<size_t>((n + kDotThreads - 1) / kDotThreads, kDotPartials));
    if (n > 0) {
        hipLaunchKernelGGL(dot_partial_kernel, dim3(blocks), dim3(kDotThreads), 0, st, a, b, n,
                           partial);
        psktm_status s;
        if ((s = check_launch("dot_partial_kernel")) != PSKTM_OK) return s;
    }
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotThreads), 0, st, (const double *)partial,
                       n > 0 ? (int)blocks : 0, result);
    return check_launch("dot_final_kernel");
}

} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_dot_f64(const float *a, const float *b, size_t n, double *result, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(result && (n == 0 || (a && b)), PSKTM_E_INVAL, "NULL pointer");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(1), 0, st, result);
    if (n == 0) return check_launch("zero_f64_kernel");
    const unsigned blocks =
        (unsigned)std::min<size_t>((n + kDotThreads - 1) / kDotThreads, kDotPartials);
    hipLaunchKernelGGL(dot_atomic_kernel, dim3(blocks), dim3(kDotThreads), 0, st, a, b, n, result);
    return check_launch("dot_atomic_kernel");
}

psktm_status psktm_dot_f64_ws(const float *a, const float *b, size_t n, double *result,
                              double *workspace, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(result && workspace && (n == 0 || (a && b)), PSKTM_E_INVAL, "NULL pointer");
    return launch_dot(a, b, n, result, workspace, as_stream(stream));
}

psktm_status psktm_axpy_ratio(float *y, const float *x, size_t n, const double *num,
                              const double *den, float sign, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(num && den && (n == 0 || (x && y)), PSKTM_E_INVAL, "NULL pointer");
    if (n == 0) return PSKTM_OK;
    hipLaunchKernelGGL(axpy_ratio_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), y, x,
                       n, num, den, sign);
    return check_launch("axpy_ratio_kernel");
}

psktm_status psktm_xpby_ratio(float *p, const float *r, size_t n, const double *num,
                              const double *den, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(num && den && (n == 0 || (p && r)), PSKTM_E_INVAL, "NULL pointer");
    if (n == 0) return PSKTM_OK;
    hipLaunchKernelGGL(xpby_ratio_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), p, r,
                       n, num, den);
    return check_launch("xpby_ratio_kernel");
}

} // extern "C"
