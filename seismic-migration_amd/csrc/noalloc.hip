// psktm_dot_f64_noalloc: the fixed-order FP64 dot product of
// psktm_dot_f64_ws for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc) and
// have no workspace to pass. It runs the same two kernels in the same order
// (launch_dot, cg.hip), so its result has the same bits as psktm_dot_f64_ws's,
// with the partials in `scratch`, a file-scope device array, instead of the
// caller's workspace.
//
// `scratch` holds the kDotPartials doubles that any n needs (the first pass
// uses at most kDotPartials workgroups and strides over n), so no size is
// refused. Each GPU holds its own copy of the array until the process ends.
// Every call on one GPU shares it: calls that overlap on one GPU (several
// streams, several host threads) race. Nothing in it needs clearing: the
// second kernel reads only the partials the first one wrote in the same call.
#include "internal.hpp"

#include <mutex>

namespace psktm {
namespace {

static __device__ __align__(256) double scratch[kDotPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call there (with that device current) and kept. The host-side name of the
// array is only a stand-in: kernels and copies take this address.
psktm_status scratch_address(double **out)
{
    int dev = 0;
    PSKTM_HIP(hipGetDevice(&dev));
    PSKTM_CHECK(dev >= 0 && dev < kMaxDevices, PSKTM_E_UNSUPPORTED, "device index %d", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        PSKTM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return PSKTM_OK;
}

} // namespace
} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_dot_f64_noalloc(const float *a, const float *b, size_t n, double *result,
                                   void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(result && (n == 0 || (a && b)), PSKTM_E_INVAL, "NULL pointer");
    double *partial = nullptr;
    psktm_status s;
    if ((s = scratch_address(&partial)) != PSKTM_OK) return s;
    return launch_dot(a, b, n, result, partial, as_stream(stream));
}

} // extern "C"
