// Rho filter (scripts/ktmApplyRhoFilter.m:22-58) and per-trace quantisation.
//
// One workgroup filters two traces at once as the real and imaginary parts of
// one complex signal z = a + i b of length M = 2^ceil(log2 m), zero padded
// (:26, :37). The reference's filter H(w) is Hermitian (H(-w) = conj H(w)),
// DC maps to 0 and at the Nyquist bin only Re H survives its real() (:43,
// :50, :57), so with H_nyq replaced by Re H_nyq the product Y = H Z inverts to
// filtered a in the real part and filtered b in the imaginary part: one
// forward and one inverse FFT serve two traces, with no spectrum separation.
// The FFT is a radix-2 Stockham transform ping-ponging between two LDS
// buffers, twiddles from a table computed on the host in double precision.
//
// Frequencies follow the reference exactly: w_k = 2 pi k' / Dr with
// k' = k (k < M/2), k - M (k > M/2), -M/2 at Nyquist, Dr = dw * 1e-6 (:31-33),
// with no 1/M (so outputs carry the reference's factor M, or sqrt(M)).
//
// Outputs: FP32, or FP16/BF16 with a per-trace scale s_k = max|y_k| / q
// (q = 32768 for FP16, keeping values well inside its range; q = 1 for BF16),
// stored value = y / s_k rounded to nearest even. With a gain, y at sample n
// is first multiplied by sqrtf(n * gain_dt): the sqrt(t) part of the
// migration weight, applied once per trace sample instead of once per update.
#include "internal.hpp"

#include <hip/hip_fp16.h>

#include <cmath>
#include <mutex>
#include <vector>

namespace psktm {
namespace {

constexpr int kRhoThreads = 256;
constexpr int kRhoMaxM = 4096;

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-LDS radix-2 Stockham FFT of length M (power of two); sign -1 forward,
// +1 inverse (unnormalised). Returns the buffer holding the result.
__device__ float2 *stockham(float2 *b0, float2 *b1, const float2 *tw, int M, int logM,
                            bool inverse)
{
    float2 *src = b0, *dst = b1;
    const int half = M >> 1;
    for (int s = 0, Ns = 1; s < logM; ++s, Ns <<= 1) {
        for (int j = threadIdx.x; j < half; j += blockDim.x) {
            const int k = j & (Ns - 1);
            float2 v0 = src[j], v1 = src[j + half];
            float2 w = tw[k * (half / Ns)];      // exp(-2 pi i k / (2 Ns))
            if (inverse) w.y = -w.y;
            v1 = cmul(v1, w);
            const int o = ((j - k) << 1) + k;    // (j / Ns) * 2 Ns + k
            dst[o] = make_float2(v0.x + v1.x, v0.y + v1.y);
            dst[o + Ns] = make_float2(v0.x - v1.x, v0.y - v1.y);
        }
        __syncthreads();
        float2 *t = src; src = dst; dst = t;
    }
    return src;
}

__device__ __forceinline__ unsigned short f2bf16_rne(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float block_max(float v, float *s_red)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    float r = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmaxf(r, s_red[w]);
    return r;
}

__device__ __forceinline__ void store_q(void *out, int dtype, size_t idx, float y, float inv)
{
    if (dtype == PSKTM_F32) static_cast<float *>(out)[idx] = y;
    else if (dtype == PSKTM_F16) static_cast<__half *>(out)[idx] = __float2half_rn(y * inv);
    else static_cast<unsigned short *>(out)[idx] = f2bf16_rne(y * inv);
}

// sqrtf(n * gain_dt), or 1 without a gain (gain_dt == 0); one rounding for the
// product, correctly rounded sqrt.
__device__ __forceinline__ float row_gain(int n, float gain_dt)
{
#pragma clang fp contract(off)
    if (gain_dt == 0.0f) return 1.0f;
    const float t = (float)n * gain_dt;
    return sqrtf(t);
}

__global__ __launch_bounds__(kRhoThreads) void rho_kernel(
    const float *__restrict__ in, int n, int m, int M, int logM, float wscale, int opt,
    int dtype, float qmax, void *__restrict__ out, float *__restrict__ scale_out,
    int shift, const float2 *__restrict__ twg, float gain_dt)
{
    extern __shared__ float2 smem[];
    __shared__ float s_red[kRhoThreads / 64];
    float2 *b0 = smem, *b1 = smem + M, *tw = smem + 2 * M;
    const int half = M >> 1;
    const int ra = 2 * blockIdx.x, rb = ra + 1;        // output rows
    const bool has_b = rb < n;
    // the reference's 2-D fftshift pair rotates odd trace counts by one row
    const int rot = (shift && (n & 1)) ? 1 : 0;
    const int ia = (ra + rot) % n, ib = has_b ? (rb + rot) % n : 0;
    for (int q = threadIdx.x; q < half; q += blockDim.x) tw[q] = twg[q];
    for (int q = threadIdx.x; q < M; q += blockDim.x) {
        float va = 0.f, vb = 0.f;
        if (q < m) {
            va = in[(size_t)ia * m + q];
            if (has_b) vb = in[(size_t)ib * m + q];
        }
        b0[q] = make_float2(va, vb);
    }
    __syncthreads();
    float2 *Z = stockham(b0, b1, tw, M, logM, false);
    float2 *other = (Z == b0) ? b1 : b0;
    // Y_k = H_k Z_k (natural order), written into the other buffer's partner
    const float c45 = 0.70710678118654752f;
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        float2 z = Z[k], y;
        if (k == 0) {
            y = make_float2(0.f, 0.f);
        } else {
            const int kp = (k < half) ? k : k - M;        // Nyquist -> -M/2
            const float w = wscale * (float)kp;           // 2 pi k' / Dr
            const float aw = fabsf(w);
            float2 h;
            if (opt == 1) h = make_float2(aw, 0.f);
            else if (opt == 2) {
                const float r = sqrtf(aw) * c45;
                h = make_float2(r, w < 0.f ? -r : r);
            } else h = make_float2(0.f, w);
            if (k == half) h.y = 0.f;                     // real() keeps Re H only
            y = cmul(h, z);
        }
        other[k] = y;
    }
    __syncthreads();
    float2 *buf_a = other, *buf_b = Z;
    float2 *Y = stockham(buf_a, buf_b, tw, M, logM, true);
    const float invM = 1.0f / (float)M;
    if (gain_dt != 0.0f) {
        __syncthreads();
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            const float gq = row_gain(q, gain_dt);
            const float ya = Y[q].x * invM, yb = Y[q].y * invM;
            // stored back pre-multiplied by M, so the common path below
            // divides by M exactly as without a gain (M is a power of two)
            Y[q] = make_float2((ya * gq) * (float)M, (yb * gq) * (float)M);
        }
        __syncthreads();
    }
    float inv_a = 1.f, inv_b = 1.f;
    if (dtype != PSKTM_F32) {
        float ma = 0.f, mb = 0.f;
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            ma = fmaxf(ma, fabsf(Y[q].x * invM));
            mb = fmaxf(mb, fabsf(Y[q].y * invM));
        }
        ma = block_max(ma, s_red);
        mb = block_max(mb, s_red);
        const float sa = ma / qmax, sb = mb / qmax;
        inv_a = (sa > 0.f) ? 1.0f / sa : 0.f;
        inv_b = (sb > 0.f) ? 1.0f / sb : 0.f;
        if (threadIdx.x == 0) {
            scale_out[ra] = sa;
            if (has_b) scale_out[rb] = sb;
        }
    }
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        store_q(out, dtype, (size_t)ra * m + q, Y[q].x * invM, inv_a);
        if (has_b) store_q(out, dtype, (size_t)rb * m + q, Y[q].y * invM, inv_b);
    }
}

// One trace per workgroup: optional gain, then F32 out or per-trace max-abs
// quantisation.
__global__ __launch_bounds__(kRhoThreads) void quantize_kernel(
    const float *__restrict__ in, int m, int dtype, float qmax, void *__restrict__ out,
    float *__restrict__ scale_out, float gain_dt)
{
#pragma clang fp contract(off)
    __shared__ float s_red[kRhoThreads / 64];
    const size_t row = (size_t)blockIdx.x * m;
    auto value = [&](int q) { return in[row + q] * row_gain(q, gain_dt); };
    if (dtype == PSKTM_F32) {
        for (int q = threadIdx.x; q < m; q += blockDim.x)
            static_cast<float *>(out)[row + q] = value(q);
        return;
    }
    float mx = 0.f;
    for (int q = threadIdx.x; q < m; q += blockDim.x) mx = fmaxf(mx, fabsf(value(q)));
    mx = block_max(mx, s_red);
    const float s = mx / qmax;
    const float inv = (s > 0.f) ? 1.0f / s : 0.f;
    if (threadIdx.x == 0) scale_out[blockIdx.x] = s;
    for (int q = threadIdx.x; q < m; q += blockDim.x) store_q(out, dtype, row + q, value(q), inv);
}

__global__ void ibm_kernel(const unsigned *__restrict__ in, float *__restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned w = in[i];
    const int e = (int)((w >> 24) & 0x7fu) - 64;
    const float v = ldexpf((float)(w & 0x00ffffffu), 4 * e - 24);
    out[i] = (w & 0x80000000u) ? -v : v;
}

// Twiddle tables per (device, log2 M); kept for the life of the process so
// that a kernel still in flight never reads a freed table.
constexpr int kMaxDevices = 64;
std::mutex g_tw_mutex;
float2 *g_tw[kMaxDevices][13] = {};

psktm_status twiddles(int dev, int logM, const float2 **out)
{
    PSKTM_CHECK(dev >= 0 && dev < kMaxDevices, PSKTM_E_UNSUPPORTED, "device index %d", dev);
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    float2 *&slot = g_tw[dev][logM];
    if (!slot) {
        const int M = 1 << logM;
        std::vector<float2> h(M / 2);
        for (int q = 0; q < M / 2; ++q) {
            const double a = -2.0 * M_PI * q / M;
            h[q] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        float2 *d = nullptr;
        PSKTM_HIP(hipMalloc(&d, sizeof(float2) * (M / 2)));
        hipError_t e = hipMemcpy(d, h.data(), sizeof(float2) * (M / 2), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return hip_fail(e, "hipMemcpy(twiddles)", __FILE__, __LINE__);
        }
        slot = d;
    }
    *out = slot;
    return PSKTM_OK;
}

float qmax_of(int dtype) { return dtype == PSKTM_F16 ? 32768.0f : 1.0f; }

} // namespace
} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_rho_filter(const float *in, int n_traces, int n_samples, float dw_us,
                              int opt, int out_dtype, void *out, float *scale_out,
                              int matlab_trace_shift, void *stream)
{
    return psktm_rho_filter_gain(in, n_traces, n_samples, dw_us, opt, out_dtype, out, scale_out,
                                 matlab_trace_shift, 0.0f, stream);
}

psktm_status psktm_rho_filter_gain(const float *in, int n_traces, int n_samples, float dw_us,
                                   int opt, int out_dtype, void *out, float *scale_out,
                                   int matlab_trace_shift, float gain_dt, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(std::isfinite(gain_dt) && gain_dt >= 0.0f, PSKTM_E_INVAL,
                "gain_dt must be finite and non-negative");
    PSKTM_CHECK(n_traces >= 0 && n_samples > 0, PSKTM_E_INVAL, "bad sizes (%d x %d)", n_traces,
                n_samples);
    PSKTM_CHECK(opt == 1 || opt == 2 || opt == 3, PSKTM_E_INVAL, "rho option must be 1, 2 or 3");
    PSKTM_CHECK(dw_us > 0.0f, PSKTM_E_INVAL, "sample interval must be positive");
    PSKTM_CHECK(out_dtype == PSKTM_F32 || out_dtype == PSKTM_F16 || out_dtype == PSKTM_BF16,
                PSKTM_E_INVAL, "unknown output dtype %d", out_dtype);
    if (n_traces == 0) return PSKTM_OK;
    PSKTM_CHECK(in && out, PSKTM_E_INVAL, "NULL trace pointer");
    PSKTM_CHECK(out_dtype == PSKTM_F32 || scale_out, PSKTM_E_INVAL, "scale_out is NULL");
    // M = 2^ceil(log(m)/log(2)) as the reference computes it (:26)
    const double lg = std::ceil(std::log((double)n_samples) / std::log(2.0));
    const int logM = (int)lg;
    PSKTM_CHECK(logM >= 1 && (1 << logM) <= kRhoMaxM, PSKTM_E_UNSUPPORTED,
                "rho filter supports 2 <= M <= %d (m = %d)", kRhoMaxM, n_samples);
    const int M = 1 << logM;
    PSKTM_CHECK(M >= n_samples, PSKTM_E_INVAL, "FFT length %d below trace length %d", M, n_samples);
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    int dev = 0;
    PSKTM_HIP(hipGetDevice(&dev));
    const float2 *tw = nullptr;
    if ((st = twiddles(dev, logM, &tw)) != PSKTM_OK) return st;
    const double Dr = (double)dw_us * 1e-6;
    const float wscale = (float)(2.0 * M_PI / Dr);
    const size_t lds = sizeof(float2) * (2 * (size_t)M + M / 2);
    { psktm_status s_; if ((s_ = ensure_dynamic_lds((const void *)rho_kernel, lds)) != PSKTM_OK) return s_; }
    const unsigned blocks = (unsigned)((n_traces + 1) / 2);
    hipLaunchKernelGGL(rho_kernel, dim3(blocks), dim3(kRhoThreads), lds, as_stream(stream), in,
                       n_traces, n_samples, M, logM, wscale, opt, out_dtype,
                       qmax_of(out_dtype), out, scale_out, matlab_trace_shift ? 1 : 0, tw,
                       gain_dt);
    return check_launch("rho_kernel");
}

psktm_status psktm_trace_gain(const float *in, int n_traces, int n_samples, float gain_dt,
                              int out_dtype, void *out, float *scale_out, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(n_traces >= 0 && n_samples > 0, PSKTM_E_INVAL, "bad sizes");
    PSKTM_CHECK(std::isfinite(gain_dt) && gain_dt > 0.0f, PSKTM_E_INVAL,
                "gain_dt must be positive");
    PSKTM_CHECK(out_dtype == PSKTM_F32 || out_dtype == PSKTM_F16 || out_dtype == PSKTM_BF16,
                PSKTM_E_INVAL, "unknown output dtype %d", out_dtype);
    if (n_traces == 0) return PSKTM_OK;
    PSKTM_CHECK(in && out && (out_dtype == PSKTM_F32 || scale_out), PSKTM_E_INVAL,
                "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipLaunchKernelGGL(quantize_kernel, dim3(n_traces), dim3(kRhoThreads), 0, as_stream(stream),
                       in, n_samples, out_dtype, qmax_of(out_dtype), out, scale_out, gain_dt);
    return check_launch("quantize_kernel");
}

psktm_status psktm_quantize(const float *in, int n_traces, int n_samples, int out_dtype,
                            void *out, float *scale_out, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(n_traces >= 0 && n_samples > 0, PSKTM_E_INVAL, "bad sizes");
    PSKTM_CHECK(out_dtype == PSKTM_F16 || out_dtype == PSKTM_BF16, PSKTM_E_INVAL,
                "quantize writes F16 or BF16");
    if (n_traces == 0) return PSKTM_OK;
    PSKTM_CHECK(in && out && scale_out, PSKTM_E_INVAL, "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipLaunchKernelGGL(quantize_kernel, dim3(n_traces), dim3(kRhoThreads), 0, as_stream(stream),
                       in, n_samples, out_dtype, qmax_of(out_dtype), out, scale_out, 0.0f);
    return check_launch("quantize_kernel");
}

psktm_status psktm_ibm32_to_ieee_device(const unsigned int *input, float *output, size_t n,
                                        void *stream)
{
    psktm__clear_error();
    if (n == 0) return PSKTM_OK;
    PSKTM_CHECK(input && output, PSKTM_E_INVAL, "NULL pointer");
    psktm_status st;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipLaunchKernelGGL(ibm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), input, output, n);
    return check_launch("ibm_kernel");
}

} // extern "C"
