// Trace preparation on the device: the steps the reference's MATLAB writers
// perform before a block is migrated.
//
//   psktm_offset_sort     sort by half-offset magnitude
//                         (scripts/ktmGetSortedOffTraces.m:17-22)
//   psktm_offset_classes  full-offset classes of a fixed width (the batched
//                         gathers' offset classes)
//   psktm_select_offset   constant-offset selection by round(|h|) within
//                         (h - tol, h + tol) (scripts/ktmGetConstOffTraces.m:18-28)
//   psktm_gather_pad      gather rows and zero-pad both dimensions to a
//                         multiple of sze (scripts/ktmWriteTraceBlockConstOffset.m:100-118)
//   psktm_upsample_traces linear interpolation onto a 2^-N sample grid
//                         (scripts/ktmInterpTraceBlock.m:134-154)
//
// Sorting and selection are stable (hipCUB radix sort / flagged selection),
// so equal keys keep trace order, as MATLAB's sort and find do. These entry
// points allocate temporary device memory and return counts to the host;
// they are set-up steps, not graph-capturable.
#include "internal.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

namespace psktm {
namespace {

struct Tmp {
    void *p = nullptr;
    ~Tmp() { if (p) (void)hipFree(p); }
};

__global__ void magnitude_key_kernel(const float *__restrict__ hx, const float *__restrict__ hy,
                                     int n, unsigned *__restrict__ key, int *__restrict__ idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float m = sqrtf(fmaf(hx[k], hx[k], hy[k] * hy[k]));
    key[k] = __float_as_uint(m);          // non-negative floats order as integers
    idx[k] = k;
}

__global__ void key_to_float_kernel(const unsigned *__restrict__ key, int n, float *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = __uint_as_float(key[k]);
}

__global__ void class_key_kernel(const float *__restrict__ hx, const float *__restrict__ hy, int n,
                                 double width, int n_classes, unsigned *__restrict__ key,
                                 int *__restrict__ idx, int *__restrict__ counts)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double full = 2.0 * sqrt((double)hx[k] * hx[k] + (double)hy[k] * hy[k]);
    const double c = floor(full / width);
    const int cls = (c >= 0.0 && c < (double)n_classes) ? (int)c : n_classes;
    key[k] = (unsigned)cls;
    idx[k] = k;
    if (cls < n_classes) atomicAdd(counts + cls, 1);
}

__global__ void select_flag_kernel(const float *__restrict__ hx, const float *__restrict__ hy,
                                   int n, double h, double tol, char *__restrict__ flag)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    // MATLAB round(): halves away from zero; the magnitude is non-negative
    const double r = floor(sqrt((double)hx[k] * hx[k] + (double)hy[k] * hy[k]) + 0.5);
    flag[k] = (r > h - tol) && (r < h + tol);
}

__global__ void gather_pad_kernel(const float *__restrict__ in, int m, const int *__restrict__ rows,
                                  int n_rows, int out_rows, int out_cols, float *__restrict__ out)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)out_rows * out_cols) return;
    const int r = (int)(q / out_cols), s = (int)(q % out_cols);
    float v = 0.f;
    if (r < n_rows && s < m) {
        const int src = rows ? rows[r] : r;
        v = in[(size_t)src * m + s];
    }
    out[q] = v;
}

__global__ void upsample_kernel(const float *__restrict__ in, int n, int m, int N,
                                float *__restrict__ out, int mo)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)n * mo) return;
    const int k = (int)(q / mo), p = (int)(q % mo);
    const int i = p >> N;
    const float *row = in + (size_t)k * m;
    if (i >= m - 1) {
        out[q] = row[m - 1];
        return;
    }
    const float f = (float)(p & ((1 << N) - 1)) * __uint_as_float((unsigned)(127 - N) << 23);
    const float d0 = row[i], d1 = row[i + 1];
    out[q] = fmaf(f, d1 - d0, d0);
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

} // namespace
} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_offset_sort(const float *off_x, const float *off_y, int n, int *perm,
                               float *key, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(n >= 0, PSKTM_E_INVAL, "negative trace count");
    if (n == 0) return PSKTM_OK;
    PSKTM_CHECK(off_x && off_y && perm, PSKTM_E_INVAL, "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipStream_t s = as_stream(stream);
    Tmp k0, k1, i0, tmp;
    PSKTM_HIP(hipMalloc(&k0.p, sizeof(unsigned) * n));
    PSKTM_HIP(hipMalloc(&k1.p, sizeof(unsigned) * n));
    PSKTM_HIP(hipMalloc(&i0.p, sizeof(int) * n));
    hipLaunchKernelGGL(magnitude_key_kernel, dim3(blocks_for(n)), dim3(256), 0, s, off_x, off_y, n,
                       (unsigned *)k0.p, (int *)i0.p);
    size_t bytes = 0;
    PSKTM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (unsigned *)k0.p, (unsigned *)k1.p,
                                                 (int *)i0.p, perm, n, 0, 32, s));
    PSKTM_HIP(hipMalloc(&tmp.p, bytes > 0 ? bytes : 1));
    PSKTM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, (unsigned *)k0.p, (unsigned *)k1.p,
                                                 (int *)i0.p, perm, n, 0, 32, s));
    if (key)
        hipLaunchKernelGGL(key_to_float_kernel, dim3(blocks_for(n)), dim3(256), 0, s,
                           (const unsigned *)k1.p, n, key);
    if ((st = check_launch("offset sort")) != PSKTM_OK) return st;
    PSKTM_HIP(hipStreamSynchronize(s));       // temporaries are freed on return
    return PSKTM_OK;
}

psktm_status psktm_offset_classes(const float *off_x, const float *off_y, int n, float width,
                                  int n_classes, int *perm, int *class_start, int *n_kept,
                                  void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(n >= 0 && n_classes > 0 && width > 0.0f, PSKTM_E_INVAL,
                "bad arguments (n=%d classes=%d width=%g)", n, n_classes, width);
    PSKTM_CHECK(class_start && n_kept && (n == 0 || (off_x && off_y && perm)), PSKTM_E_INVAL,
                "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipStream_t s = as_stream(stream);
    std::vector<int> h(n_classes + 1, 0);
    if (n > 0) {
        Tmp k0, k1, i0, cnt, tmp;
        PSKTM_HIP(hipMalloc(&k0.p, sizeof(unsigned) * n));
        PSKTM_HIP(hipMalloc(&k1.p, sizeof(unsigned) * n));
        PSKTM_HIP(hipMalloc(&i0.p, sizeof(int) * n));
        PSKTM_HIP(hipMalloc(&cnt.p, sizeof(int) * n_classes));
        PSKTM_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * n_classes, s));
        hipLaunchKernelGGL(class_key_kernel, dim3(blocks_for(n)), dim3(256), 0, s, off_x, off_y, n,
                           (double)width, n_classes, (unsigned *)k0.p, (int *)i0.p, (int *)cnt.p);
        int end_bit = 1;
        while ((1ll << end_bit) <= n_classes) ++end_bit;
        size_t bytes = 0;
        PSKTM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (unsigned *)k0.p,
                                                     (unsigned *)k1.p, (int *)i0.p, perm, n, 0,
                                                     end_bit, s));
        PSKTM_HIP(hipMalloc(&tmp.p, bytes > 0 ? bytes : 1));
        PSKTM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, (unsigned *)k0.p,
                                                     (unsigned *)k1.p, (int *)i0.p, perm, n, 0,
                                                     end_bit, s));
        if ((st = check_launch("offset classes")) != PSKTM_OK) return st;
        PSKTM_HIP(hipMemcpyAsync(h.data() + 1, cnt.p, sizeof(int) * n_classes,
                                 hipMemcpyDeviceToHost, s));
        PSKTM_HIP(hipStreamSynchronize(s));
    }
    for (int c = 0; c < n_classes; ++c) h[c + 1] += h[c];
    *n_kept = h[n_classes];
    PSKTM_HIP(hipMemcpyAsync(class_start, h.data(), sizeof(int) * (n_classes + 1),
                             hipMemcpyHostToDevice, s));
    PSKTM_HIP(hipStreamSynchronize(s));
    return PSKTM_OK;
}

psktm_status psktm_select_offset(const float *off_x, const float *off_y, int n, float h, float tol,
                                 int *selected, int *n_kept, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(n >= 0 && n_kept && (n == 0 || (off_x && off_y && selected)), PSKTM_E_INVAL,
                "bad arguments");
    *n_kept = 0;
    if (n == 0) return PSKTM_OK;
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipStream_t s = as_stream(stream);
    Tmp flag, cnt, tmp;
    PSKTM_HIP(hipMalloc(&flag.p, n));
    PSKTM_HIP(hipMalloc(&cnt.p, sizeof(int)));
    hipLaunchKernelGGL(select_flag_kernel, dim3(blocks_for(n)), dim3(256), 0, s, off_x, off_y, n,
                       (double)h, (double)tol, (char *)flag.p);
    hipcub::CountingInputIterator<int> ids(0);
    size_t bytes = 0;
    PSKTM_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, ids, (const char *)flag.p, selected,
                                            (int *)cnt.p, n, s));
    PSKTM_HIP(hipMalloc(&tmp.p, bytes > 0 ? bytes : 1));
    PSKTM_HIP(hipcub::DeviceSelect::Flagged(tmp.p, bytes, ids, (const char *)flag.p, selected,
                                            (int *)cnt.p, n, s));
    if ((st = check_launch("offset selection")) != PSKTM_OK) return st;
    PSKTM_HIP(hipMemcpyAsync(n_kept, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PSKTM_HIP(hipStreamSynchronize(s));
    return PSKTM_OK;
}

psktm_status psktm_gather_pad(const float *in, int m, const int *rows, int n_rows, int sze,
                              float *out, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(m > 0 && n_rows >= 0 && sze > 0, PSKTM_E_INVAL, "bad sizes");
    const long long R = ((long long)n_rows + sze - 1) / sze * sze;
    const long long C = ((long long)m + sze - 1) / sze * sze;
    PSKTM_CHECK(R < (1ll << 31) && C < (1ll << 31) && R * C < (1ll << 40), PSKTM_E_INVAL,
                "padded block too large");
    if (R * C == 0) return PSKTM_OK;
    PSKTM_CHECK(in && out, PSKTM_E_INVAL, "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipLaunchKernelGGL(gather_pad_kernel, dim3(blocks_for((size_t)R * C)), dim3(256), 0,
                       as_stream(stream), in, m, rows, n_rows, (int)R, (int)C, out);
    return check_launch("gather_pad_kernel");
}

psktm_status psktm_upsample_traces(const float *in, int n, int m, int N, float *out, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(n >= 0 && m >= 1 && N >= 0 && N <= 16, PSKTM_E_INVAL, "bad sizes or N");
    const long long mo = (long long)(m - 1) * (1ll << N) + 1;
    PSKTM_CHECK(mo < (1ll << 31) && (long long)n * mo < (1ll << 40), PSKTM_E_INVAL,
                "upsampled block too large");
    if (n == 0) return PSKTM_OK;
    PSKTM_CHECK(in && out, PSKTM_E_INVAL, "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipLaunchKernelGGL(upsample_kernel, dim3(blocks_for((size_t)n * mo)), dim3(256), 0,
                       as_stream(stream), in, n, m, N, out, (int)mo);
    return check_launch("upsample_kernel");
}

} // extern "C"
