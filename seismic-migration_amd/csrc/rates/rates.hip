// gfx950 rate microbenchmarks behind the roofline figures: VALU FMA and
// v_sqrt_f32 issue rates, LDS gather rate, and LDS atomic-add rates for the
// access patterns of the adjoint scatter. Built as librates.so next to
// libpsktm.so and driven by `bench.py --workload rates`; not part of the
// migration library.
//
// Every kernel runs a fixed number of iterations per lane on values that
// depend on the lane and the loop, and stores a per-lane checksum, so no
// work is removed by the compiler. LDS addresses are masked into the
// workgroup's array.
#include <hip/hip_runtime.h>

namespace {

constexpr int kThreads = 256;
constexpr int kLdsSlots = 4096;      // 16 KiB of 32-bit / 32 KiB of 64-bit slots

// 8 independent FMA chains per lane: 8 * iters FMAs.
__global__ __launch_bounds__(kThreads) void fma_kernel(int iters, float *out)
{
    float a[8];
    const float m = 1.0f + 1e-7f * (float)threadIdx.x, c = 1e-6f;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = (float)(threadIdx.x + q);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = fmaf(a[q], m, c);
    }
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// 8 independent chains per lane of one VALU op with every operand in a
// VGPR (op 0: v_fma_f32, 1: v_add_f32, 2: v_mul_f32): 8 * iters ops.
template <int OP>
__global__ __launch_bounds__(kThreads) void valu_kernel(int iters, float *out, float seed)
{
    float a[8];
    // lane-dependent operands, so that they live in VGPRs, not as constants
    const float m = 1.0f + 1e-7f * (float)threadIdx.x * seed;
    const float c = 1e-6f * (float)(threadIdx.x & 7) * seed;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = (float)(threadIdx.x + q);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (OP == 0) a[q] = fmaf(a[q], m, c);
            else if (OP == 1) a[q] = a[q] + c;
            else a[q] = a[q] * m;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// Issue-rate probes of the migration loop's instruction types: 8 independent
// chains per lane, every operand a VGPR, iters iterations. OP:
//   0 v_add_f32  1 v_fma_f32  2 v_cvt_i32_f32 + v_cvt_f32_i32  3 v_fract_f32
//   4 v_lshl_add_u32  5 v_fma_mix_f32 (FP16 operands)  6 v_fma_f32 with an
//   SGPR operand  7 v_add_f32 and v_sqrt_f32 alternating chains
// Dynamic LDS per workgroup sets the resident waves per SIMD.
template <int OP>
__global__ __launch_bounds__(kThreads) void issue_kernel(int iters, float *out, float seed, float s)
{
    extern __shared__ float pad[];
    float a[8];
    unsigned u[8];
    const float m = 1.0f + 1e-7f * (float)threadIdx.x * seed;
    const float c = 1e-6f * (float)(threadIdx.x & 7) * seed;
    const unsigned cu = threadIdx.x * (unsigned)seed;
    const unsigned hpair = 0x3c003800u ^ (threadIdx.x & 1u);   // (0.5, 1.0) as FP16
#pragma unroll
    for (int q = 0; q < 8; ++q) { a[q] = (float)(threadIdx.x + q) * 0.01f; u[q] = threadIdx.x + q; }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (OP == 0) a[q] = a[q] + c;
            else if (OP == 1) a[q] = fmaf(a[q], m, c);
            else if (OP == 2) a[q] = (float)(int)a[q];
            else if (OP == 3) a[q] = __builtin_amdgcn_fractf(a[q]);
            else if (OP == 4) u[q] = (u[q] << 2) + cu;
            else if (OP == 5) {
                // the loop's interpolation: a * hi(h) + lo(h), FP16 halves
                float r;
                asm volatile("v_fma_mix_f32 %0, %1, %2, %2 op_sel:[0,1,0] op_sel_hi:[0,1,1]"
                             : "=v"(r) : "v"(a[q]), "v"(hpair));
                a[q] = r;
            } else if (OP == 6) a[q] = fmaf(a[q], s, c);
            else a[q] = (q & 1) ? __builtin_amdgcn_sqrtf(a[q]) : a[q] + c;   // 4 + 4
        }
    }
    float t = pad[0] * 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) t += a[q] + (float)u[q];
    out[blockIdx.x * kThreads + threadIdx.x] = t;
}

// 8 independent v_sqrt_f32 chains per lane, plus `fmas` independent FMA
// chains interleaved (0 = sqrt alone): 8 * iters sqrts.
template <int FMAS>
__global__ __launch_bounds__(kThreads) void sqrt_kernel(int iters, float *out)
{
    float a[8], b[FMAS > 0 ? FMAS : 1];
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = 2.0f + (float)(threadIdx.x + q);
#pragma unroll
    for (int q = 0; q < (FMAS > 0 ? FMAS : 1); ++q) b[q] = (float)q;
    const float m = 1.0f + 1e-7f * (float)threadIdx.x;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = __builtin_amdgcn_sqrtf(a[q]) + 1.5f;
        if (FMAS > 0) {
#pragma unroll
            for (int q = 0; q < FMAS; ++q) b[q] = fmaf(b[q], m, 1e-6f);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += a[q];
    if (FMAS > 0) {
#pragma unroll
        for (int q = 0; q < FMAS; ++q) s += b[q];
    }
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// iters x 8 ds_read_b64 per lane at pseudo-random slots of a 32 KiB table.
__global__ __launch_bounds__(kThreads) void lds_read_kernel(int iters, float *out)
{
    __shared__ float2 tab[kLdsSlots];
    for (int q = threadIdx.x; q < kLdsSlots; q += kThreads) tab[q] = make_float2((float)q, 1.f);
    __syncthreads();
    unsigned h = threadIdx.x * 2654435761u + blockIdx.x;
    float2 s = make_float2(0.f, 0.f);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            h = h * 1664525u + 1013904223u;
            const float2 v = tab[(h >> 16) & (kLdsSlots - 1)];
            s.x += v.x;
            s.y += v.y;
        }
    }
    out[blockIdx.x * kThreads + threadIdx.x] = s.x + s.y;
}

// iters x 8 LDS atomic adds per lane. pattern 0: lane-consecutive slots
// (one wave covers 64 neighbouring slots), 1: slots 4 apart per lane (the
// adjoint's lanes each own 4 consecutive tau samples), 2: pseudo-random.
template <typename T>
__global__ __launch_bounds__(kThreads) void lds_atomic_kernel(int iters, int pattern, float *out)
{
    __shared__ T acc[kLdsSlots];
    for (int q = threadIdx.x; q < kLdsSlots; q += kThreads) acc[q] = T(0);
    __syncthreads();
    unsigned h = threadIdx.x * 2654435761u + blockIdx.x;
    const T v = T(1);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            unsigned slot;
            if (pattern == 0) slot = threadIdx.x + 256u * (unsigned)q + 7u * (unsigned)it;
            else if (pattern == 1) slot = 4u * threadIdx.x + (unsigned)q + 3u * (unsigned)it;
            else {
                h = h * 1664525u + 1013904223u;
                slot = h >> 16;
            }
            atomicAdd(&acc[slot & (kLdsSlots - 1)], v);
        }
    }
    __syncthreads();
    out[blockIdx.x * kThreads + threadIdx.x] = (float)acc[threadIdx.x];
}

template <typename F>
int timed(F launch, float *ms)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
    launch();                                   // warm-up
    hipEventRecord(a, 0);
    launch();
    hipEventRecord(b, 0);
    if (hipEventSynchronize(b) != hipSuccess) return 2;
    hipEventElapsedTime(ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

} // namespace

extern "C" {

// Runs microbenchmark `kind` with `blocks` workgroups of 256 lanes and
// `iters` iterations; *ms = kernel time of the second of two launches.
//   0 fma (8*iters FMA per lane)
//   1 sqrt alone, 2 sqrt + 8 FMA per 8 sqrt, 3 sqrt + 64 FMA per 8 sqrt
//     (8*iters sqrt per lane)
//   4 LDS ds_read_b64 random (8*iters per lane)
//   5 v_fma_f32, 6 v_add_f32, 7 v_mul_f32, every operand a VGPR (8*iters per lane)
//   10+p LDS ds_add_f32, 20+p ds_add_f64, 30+p ds_add_u32, 40+p ds_add_u64
//     with pattern p
//     (8*iters per lane)
// Returns 0 on success.
int psktm_rates_issue(int op, int blocks, int iters, int lds_bytes, float *ms)
{
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)blocks * kThreads) != hipSuccess) return 10;
    const dim3 g(blocks), t(kThreads);
    const size_t lds = (size_t)(lds_bytes < 4 ? 4 : lds_bytes);
    int rc = 0;
#define ISSUE(K)                                                                              \
    case K:                                                                                   \
        if (hipFuncSetAttribute((const void *)issue_kernel<K>,                                \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) { \
            rc = 12; break;                                                                   \
        }                                                                                     \
        rc = timed([&] { hipLaunchKernelGGL(issue_kernel<K>, g, t, lds, 0, iters, out, 1.0f, 1.0000001f); }, ms); \
        break;
    switch (op) {
    ISSUE(0) ISSUE(1) ISSUE(2) ISSUE(3) ISSUE(4) ISSUE(5) ISSUE(6) ISSUE(7)
    default: rc = 11;
    }
#undef ISSUE
    (void)hipFree(out);
    return rc;
}

int psktm_rates_run(int kind, int blocks, int iters, float *ms)
{
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)blocks * kThreads) != hipSuccess) return 10;
    const dim3 g(blocks), t(kThreads);
    int rc = 0;
    switch (kind) {
    case 0: rc = timed([&] { hipLaunchKernelGGL(fma_kernel, g, t, 0, 0, iters, out); }, ms); break;
    case 1: rc = timed([&] { hipLaunchKernelGGL(sqrt_kernel<0>, g, t, 0, 0, iters, out); }, ms); break;
    case 2: rc = timed([&] { hipLaunchKernelGGL(sqrt_kernel<8>, g, t, 0, 0, iters, out); }, ms); break;
    case 3: rc = timed([&] { hipLaunchKernelGGL(sqrt_kernel<64>, g, t, 0, 0, iters, out); }, ms); break;
    case 4: rc = timed([&] { hipLaunchKernelGGL(lds_read_kernel, g, t, 0, 0, iters, out); }, ms); break;
    case 5: rc = timed([&] { hipLaunchKernelGGL(valu_kernel<0>, g, t, 0, 0, iters, out, 1.0f); }, ms); break;
    case 6: rc = timed([&] { hipLaunchKernelGGL(valu_kernel<1>, g, t, 0, 0, iters, out, 1.0f); }, ms); break;
    case 7: rc = timed([&] { hipLaunchKernelGGL(valu_kernel<2>, g, t, 0, 0, iters, out, 1.0f); }, ms); break;
    default: {
        const int p = kind % 10, ty = kind / 10;
        if (p > 2 || ty < 1 || ty > 4) { rc = 11; break; }
        if (ty == 1)
            rc = timed([&] { hipLaunchKernelGGL(lds_atomic_kernel<float>, g, t, 0, 0, iters, p, out); }, ms);
        else if (ty == 2)
            rc = timed([&] { hipLaunchKernelGGL(lds_atomic_kernel<double>, g, t, 0, 0, iters, p, out); }, ms);
        else if (ty == 3)
            rc = timed([&] { hipLaunchKernelGGL(lds_atomic_kernel<unsigned>, g, t, 0, 0, iters, p, out); }, ms);
        else
            rc = timed([&] { hipLaunchKernelGGL(lds_atomic_kernel<unsigned long long>, g, t, 0, 0, iters, p, out); }, ms);
    }
    }
    (void)hipFree(out);
    return rc;
}

} // extern "C"
