// Reference-exact PSTM migration on gfx950.
//
// Behaviour of ktmMigrationCPU (src/psktmCPU.c:42-87) and of the CUDA kernel
// ktmCUDA (src/psktmCUDA.cu:64-82,127-200), written for CDNA4:
//   - every operation is binary32 with its own rounding: contraction is off
//     for this file, sqrtf and '/' are the correctly rounded forms;
//   - each image sample is owned by one lane and receives the trace
//     contributions in increasing k order, so the sum is bit-identical to the
//     single-thread reference loop (the reference's own OpenMP loop races on
//     the image, src/psktmCPU.c:56,84);
//   - no trace or depth is dropped (the CUDA form drops the remainders of
//     THREADS_X/THREADS_Y, src/psktmCUDA.cu:95-97) and no grid limit applies
//     (src/psktmCUDA.cu:298-302).
//
// Mapping: workgroup (x, y) owns kBins consecutive image bins (b = i*ny + j)
// from x and kThreads consecutive tau samples from y; lane = tau sample, kBins accumulators in
// registers. Per stage of kStage traces the workgroup tabulates the distance
// terms A = |s-x|^2, B = |r-x|^2 (src/psktmCPU.c:65-72) in LDS; trace samples
// are gathered from global memory (neighbouring lanes read neighbouring
// samples of one trace, so the gather stays within a few cache lines).
//
// Images too small to fill the GPU that way (a few hundred workgroups or
// fewer, e.g. a 64-bin line) take the split form: workgroups (x, y, z) compute
// the contributions of trace stage z into a scratch array, without summing,
// and a second kernel adds them to each image sample one trace at a time in
// increasing k order. The sums are the direct form's: a skipped contribution
// is added as +0.0, which leaves every sum unchanged (sums start at +0.0 and
// never become -0.0 under round-to-nearest).
#include "internal.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace psktm {
namespace {

constexpr int kThreads = 256;
constexpr int kBins = 8;
constexpr int kStage = 32;

// initThreeAxes (src/psktmFunctions.c:28-34): start + (float)i * step.
__device__ __forceinline__ float axis_value(float start, float step, int i)
{
    float fi = (float)i;
    float p = fi * step;
    return start + p;
}

// TERMS: the split form's first kernel. Workgroup z takes traces
// kbase + z kStage .. (at most kStage, below kend) and writes each
// contribution (+0.0 where the reference skips the sample) to
// terms[(k - kbase) * nbins * nt + b * nt + l].
template <bool TERMS>
__global__ __launch_bounds__(kThreads) void pstm_exact_kernel(
    int nx, float x0, float dx, int ny, float y0, float dy, int nt, float tau0,
    float dtau, int ntr, int ns, float dt, const float *__restrict__ midX,
    const float *__restrict__ midY, const float *__restrict__ offX,
    const float *__restrict__ offY, const float *__restrict__ traces,
    const float *__restrict__ vel, float *__restrict__ image, int kbase, int kend,
    float *__restrict__ terms)
{
    __shared__ float sA[kStage][kBins];
    __shared__ float sB[kStage][kBins];

    const int l = blockIdx.y * kThreads + threadIdx.x;
    const long nbins = (long)nx * ny;
    const long b0 = (long)blockIdx.x * kBins;

    // tauS[l] = zz*zz; slownessS[l] = 1.0f/(v*v) (src/psktmCPU.c:46-49).
    float tauS = 0.0f, sS = 0.0f;
    if (l < nt) {
        float z = axis_value(tau0, dtau, l);
        tauS = z * z;
        float v = vel[l];
        float v2 = v * v;
        sS = 1.0f / v2;
    }
    const float nsf = (float)ns;

    float acc[kBins];
#pragma unroll
    for (int bb = 0; bb < kBins; ++bb) acc[bb] = 0.0f;

    const int kfirst = TERMS ? kbase + (int)blockIdx.z * kStage : 0;
    const int klast = TERMS ? min(kend, kfirst + kStage) : ntr;
    const size_t nsamp = (size_t)nbins * (size_t)nt;
    for (int k0 = kfirst; k0 < klast; k0 += kStage) {
        const int kn = min(kStage, klast - k0);
        for (int q = threadIdx.x; q < kStage * kBins; q += kThreads) {
            const int g = q / kBins, bb = q % kBins;
            const long b = b0 + bb;
            float A = 0.0f, B = 0.0f;
            if (g < kn && b < nbins) {
                const int i = (int)(b / ny), j = (int)(b % ny);
                const float x = axis_value(x0, dx, i);
                const float y = axis_value(y0, dy, j);
                const int k = k0 + g;
                const float mx = midX[k], my = midY[k], hx = offX[k], hy = offY[k];
                float ax = (mx - x) - hx, bx = (mx - x) + hx;     // :65-66
                float tempC = ax * ax, tempD = bx * bx;
                float ay = (my - y) - hy, by = (my - y) + hy;     // :71-72
                float ay2 = ay * ay, by2 = by * by;
                A = tempC + ay2;
                B = tempD + by2;
            }
            sA[g][bb] = A;
            sB[g][bb] = B;
        }
        __syncthreads();
        if (l < nt) {
            for (int g = 0; g < kn; ++g) {
                const float *tr = traces + (size_t)(k0 + g) * (size_t)ns;
#pragma unroll
                for (int bb = 0; bb < kBins; ++bb) {
                    const float pa = sA[g][bb] * sS;
                    const float pb = sB[g][bb] * sS;
                    float t = sqrtf(tauS + pa);                    // :78
                    t = t + sqrtf(tauS + pb);                      // :79
                    const float q = t / dt;                        // :80
                    // (int)q lies in (0, ns) exactly when 1 <= q < ns
                    // (truncation of a non-negative value; NaN fails both).
                    float c = 0.0f;
                    if (q >= 1.0f && q < nsf) {                    // :82
                        const int n = (int)q;
                        const float dtm = tr[n] * t;               // :84-85
                        const float r = tauS / t;
                        c = dtm * sqrtf(r);
                        if (!TERMS) acc[bb] = acc[bb] + c;
                    }
                    if (TERMS && b0 + bb < nbins)
                        terms[(size_t)(k0 + g - kbase) * nsamp + (size_t)(b0 + bb) * nt + l] = c;
                }
            }
        }
        __syncthreads();
    }
    if (!TERMS && l < nt) {
#pragma unroll
        for (int bb = 0; bb < kBins; ++bb)
            if (b0 + bb < nbins) image[(size_t)(b0 + bb) * (size_t)nt + l] = acc[bb];
    }
}

// The split form's second kernel: image[s] += terms[k][s] for k = 0 .. kn-1
// in order (one binary32 add each), one lane per image sample.
__global__ __launch_bounds__(kThreads) void pstm_exact_sum_kernel(
    size_t nsamp, int kn, const float *__restrict__ terms, float *__restrict__ image)
{
    const size_t s = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= nsamp) return;
    float acc = image[s];
#pragma unroll 8
    for (int k = 0; k < kn; ++k) acc = acc + terms[(size_t)k * nsamp + s];
    image[s] = acc;
}

__global__ void zero_image_kernel(float *__restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Scratch of the split form: at most this many bytes of contributions.
constexpr size_t kTermsBytes = size_t(1) << 30;

} // namespace

psktm_status launch_exact(const psktm_grid &g, const psktm_survey &s,
                          const float *traces, const float *vel, float *image,
                          hipStream_t stream)
{
    const long nbins = (long)g.nx * g.ny;
    const long bx = (nbins + kBins - 1) / kBins;
    const long ty = (g.nt + kThreads - 1) / kThreads;
    PSKTM_CHECK(bx <= 2147483647L && ty <= 65535, PSKTM_E_INVAL,
                "image too large for one launch (%ld bins, %d samples)", nbins, g.nt);
    dim3 grid((unsigned)bx, (unsigned)ty);
    auto direct = [&]() {
        hipLaunchKernelGGL(pstm_exact_kernel<false>, grid, dim3(kThreads), 0, stream, g.nx,
                           g.x0, g.dx, g.ny, g.y0, g.dy, g.nt, g.tau0, g.dtau,
                           s.n_traces, s.n_samples, s.dt, s.mid_x, s.mid_y, s.off_x,
                           s.off_y, traces, vel, image, 0, 0, (float *)nullptr);
        return check_launch("pstm_exact_kernel");
    };
    // the split form where the direct grid leaves most CUs idle and at least
    // one stage of contributions fits the scratch budget
    int dev = 0, cus = 0;
    PSKTM_HIP(hipGetDevice(&dev));
    PSKTM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t nsamp = (size_t)nbins * (size_t)g.nt;
    const size_t per_trace = nsamp * sizeof(float);
    if (bx * ty >= 2L * cus || s.n_traces < 2 * kStage || per_trace * kStage > kTermsBytes)
        return direct();
    // traces per pass: as many whole stages as the budget holds (grid z <= 65535)
    long ks = (long)std::min<size_t>(kTermsBytes / per_trace, (size_t)65535 * kStage);
    ks = std::min<long>(ks / kStage * kStage, s.n_traces);
    float *terms = nullptr;
    if (hipMallocAsync((void **)&terms, (size_t)ks * per_trace, stream) != hipSuccess) {
        (void)hipGetLastError();
        return direct();                      // no scratch: the direct form
    }
    hipLaunchKernelGGL(zero_image_kernel, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0,
                       stream, image, nsamp);
    psktm_status st = check_launch("zero_image_kernel");
    for (long k0 = 0; st == PSKTM_OK && k0 < s.n_traces; k0 += ks) {
        const int kn = (int)std::min<long>(ks, s.n_traces - k0);
        const dim3 gz((unsigned)bx, (unsigned)ty, (unsigned)((kn + kStage - 1) / kStage));
        hipLaunchKernelGGL(pstm_exact_kernel<true>, gz, dim3(kThreads), 0, stream, g.nx,
                           g.x0, g.dx, g.ny, g.y0, g.dy, g.nt, g.tau0, g.dtau,
                           s.n_traces, s.n_samples, s.dt, s.mid_x, s.mid_y, s.off_x,
                           s.off_y, traces, vel, image, (int)k0, (int)k0 + kn, terms);
        if ((st = check_launch("pstm_exact_kernel")) != PSKTM_OK) break;
        hipLaunchKernelGGL(pstm_exact_sum_kernel, dim3((unsigned)((nsamp + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, stream, nsamp, kn, (const float *)terms, image);
        st = check_launch("pstm_exact_sum_kernel");
    }
    (void)hipFreeAsync(terms, stream);
    return st;
}

} // namespace psktm
