// C ABI core of libpsktm.so: error state, device checks, the drop-ins
// ktmMigrationGPU (src/psktmCUDA.cu:211-332) and ktmMigrationCPU
// (src/psktmCPU.c:13-101), and psktm_migrate_exact.
#include "internal.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace {
thread_local char g_last_error[512] = "";
std::mutex g_legacy_mutex;
hipStream_t g_legacy_stream = nullptr;
} // namespace

extern "C" {

void psktm__set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

void psktm__clear_error(void) { g_last_error[0] = '\0'; }

const char *psktm_last_error(void) { return g_last_error; }

const char *psktm_version(void) { return "psktm-mi355x 0.1 (gfx950)"; }

int psktm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

} // extern "C"

namespace psktm {

psktm_status hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    (void)hipGetLastError();
    const char *base = strrchr(file, '/');
    psktm__set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e),
                     base ? base + 1 : file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PSKTM_E_ALLOC;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PSKTM_E_NODEV;
    return PSKTM_E_HIP;
}

psktm_status require_device()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        psktm__set_error("no HIP device available (%s)",
                         e != hipSuccess ? hipGetErrorString(e) : "device count 0");
        return PSKTM_E_NODEV;
    }
    int dev = 0;
    PSKTM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PSKTM_HIP(hipGetDeviceProperties(&p, dev));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        psktm__set_error("device %d is %s; libpsktm is built for gfx950 only", dev,
                         p.gcnArchName);
        return PSKTM_E_NODEV;
    }
    return PSKTM_OK;
}

psktm_status check_launch(const char *kernel)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        psktm__set_error("launch of %s failed: %s", kernel, hipGetErrorString(e));
        return PSKTM_E_LAUNCH;
    }
    return PSKTM_OK;
}

psktm_status ensure_dynamic_lds(const void *kernel, size_t bytes)
{
    if (bytes <= 64 * 1024) return PSKTM_OK;
    static std::mutex m;
    static std::vector<std::pair<std::pair<const void *, int>, size_t>> done;
    int dev = 0;
    PSKTM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(m);
    for (auto &d : done)
        if (d.first.first == kernel && d.first.second == dev && d.second >= bytes) return PSKTM_OK;
    PSKTM_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    for (auto &d : done)
        if (d.first.first == kernel && d.first.second == dev) {
            d.second = bytes;
            return PSKTM_OK;
        }
    done.push_back({{kernel, dev}, bytes});
    return PSKTM_OK;
}

psktm_status validate_grid(const psktm_grid *g)
{
    PSKTM_CHECK(g, PSKTM_E_INVAL, "grid is NULL");
    PSKTM_CHECK(g->nx > 0 && g->ny > 0 && g->nt > 0, PSKTM_E_INVAL,
                "image sizes must be positive (nx=%d ny=%d nt=%d)", g->nx, g->ny, g->nt);
    return PSKTM_OK;
}

psktm_status validate_survey(const psktm_survey *s)
{
    PSKTM_CHECK(s, PSKTM_E_INVAL, "survey is NULL");
    PSKTM_CHECK(s->n_traces >= 0 && s->n_samples > 0, PSKTM_E_INVAL,
                "trace count must be >= 0 and samples > 0 (traces=%d samples=%d)",
                s->n_traces, s->n_samples);
    PSKTM_CHECK(s->n_samples < (1 << 24), PSKTM_E_INVAL,
                "samples per trace must be < 2^24 (got %d)", s->n_samples);
    PSKTM_CHECK(s->dt > 0.0f, PSKTM_E_INVAL, "sample interval must be positive");
    PSKTM_CHECK(s->n_traces == 0 ||
                    (s->mid_x && s->mid_y && s->off_x && s->off_y),
                PSKTM_E_INVAL, "geometry pointer is NULL");
    return PSKTM_OK;
}

} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_migrate_exact(const psktm_grid *grid, const psktm_survey *survey,
                                 const float *traces, const float *vel, float *image,
                                 void *stream)
{
    psktm__clear_error();
    psktm_status st;
    if ((st = validate_grid(grid)) != PSKTM_OK) return st;
    if ((st = validate_survey(survey)) != PSKTM_OK) return st;
    PSKTM_CHECK(image && vel, PSKTM_E_INVAL, "image or velocity pointer is NULL");
    PSKTM_CHECK(survey->n_traces == 0 || traces, PSKTM_E_INVAL, "traces pointer is NULL");
    if ((st = require_device()) != PSKTM_OK) return st;
    return launch_exact(*grid, *survey, traces, vel, image, as_stream(stream));
}

} // extern "C"

namespace {

// The legacy 11-argument migration on host arrays, shared by ktmMigrationGPU
// (src/psktmCUDA.cu:211-332) and ktmMigrationCPU (src/psktmCPU.c:13-101):
// synchronous, image zeroed first (src/psktmCPU.c:52-53, src/psktmCUDA.cu:
// 241-242) and left zeroed on any failure, then the reference-exact HIP
// kernel, which is bit-identical to the single-thread CPU loop. Unlike the
// CUDA form it keeps the device context alive (:327), never prints (:299,312)
// and has no block limit (:298-302). `who` names the entry in error texts.
void legacy_migrate(const char *who, struct imageGrid *imageX, struct imageGrid *imageY,
                    struct imageGrid *imageZ, struct jobParams *config, float *midX,
                    float *midY, float *offX, float *offY, float *traces, float *slowness,
                    float *image)
{
    psktm__clear_error();
    if (!imageX || !imageY || !imageZ || !config) {
        psktm__set_error("%s: NULL grid or job struct", who);
        return;
    }
    const int nx = imageX->imagePnts, ny = imageY->imagePnts, nt = imageZ->imagePnts;
    if (nx <= 0 || ny <= 0 || nt <= 0) {
        psktm__set_error("%s: image sizes must be positive", who);
        return;
    }
    const size_t nimg = (size_t)nx * ny * nt;
    if (!image) {
        psktm__set_error("%s: image is NULL", who);
        return;
    }
    memset(image, 0, nimg * sizeof(float));

    psktm_grid g{nx, ny, nt, imageX->imageStart, imageX->imageStep,
                 imageY->imageStart, imageY->imageStep, imageZ->imageStart,
                 imageZ->imageStep};
    const int ntr = config->traceNo, ns = config->tracePts;
    if (ntr < 0 || ns <= 0 || !(config->traceSampleRate > 0.0f) || ns >= (1 << 24)) {
        psktm__set_error("%s: invalid job (traceNo=%d tracePts=%d)", who, ntr, ns);
        return;
    }
    if (!slowness || (ntr > 0 && (!midX || !midY || !offX || !offY || !traces))) {
        psktm__set_error("%s: NULL input array", who);
        return;
    }
    if (ntr == 0) return;   // nothing to add: the zeroed image is the answer

    std::lock_guard<std::mutex> lock(g_legacy_mutex);
    if (require_device() != PSKTM_OK) return;
    auto fail = [&](hipError_t e, const char *what) {
        hip_fail(e, what, __FILE__, __LINE__);
    };
    if (!g_legacy_stream) {
        hipError_t e = hipStreamCreateWithFlags(&g_legacy_stream, hipStreamNonBlocking);
        if (e != hipSuccess) { fail(e, "hipStreamCreate"); g_legacy_stream = nullptr; return; }
    }
    hipStream_t st = g_legacy_stream;

    const size_t ntraceb = (size_t)ntr * ns * sizeof(float);
    const size_t ngeob = (size_t)ntr * sizeof(float);
    float *dImg = nullptr, *dTr = nullptr, *dGeo = nullptr, *dVel = nullptr;
    hipError_t e;
    bool ok = false;
    do {
        if ((e = hipMalloc(&dImg, nimg * sizeof(float))) != hipSuccess) { fail(e, "hipMalloc(image)"); break; }
        if ((e = hipMalloc(&dTr, ntraceb)) != hipSuccess) { fail(e, "hipMalloc(traces)"); break; }
        if ((e = hipMalloc(&dGeo, 4 * ngeob)) != hipSuccess) { fail(e, "hipMalloc(geometry)"); break; }
        if ((e = hipMalloc(&dVel, (size_t)nt * sizeof(float))) != hipSuccess) { fail(e, "hipMalloc(velocity)"); break; }
        if ((e = hipMemcpyAsync(dTr, traces, ntraceb, hipMemcpyHostToDevice, st)) != hipSuccess) { fail(e, "H2D traces"); break; }
        if ((e = hipMemcpyAsync(dGeo, midX, ngeob, hipMemcpyHostToDevice, st)) != hipSuccess) { fail(e, "H2D midX"); break; }
        if ((e = hipMemcpyAsync(dGeo + ntr, midY, ngeob, hipMemcpyHostToDevice, st)) != hipSuccess) { fail(e, "H2D midY"); break; }
        if ((e = hipMemcpyAsync(dGeo + 2 * (size_t)ntr, offX, ngeob, hipMemcpyHostToDevice, st)) != hipSuccess) { fail(e, "H2D offX"); break; }
        if ((e = hipMemcpyAsync(dGeo + 3 * (size_t)ntr, offY, ngeob, hipMemcpyHostToDevice, st)) != hipSuccess) { fail(e, "H2D offY"); break; }
        if ((e = hipMemcpyAsync(dVel, slowness, (size_t)nt * sizeof(float), hipMemcpyHostToDevice, st)) != hipSuccess) { fail(e, "H2D velocity"); break; }
        psktm_survey s{ntr, ns, config->traceSampleRate, dGeo, dGeo + ntr,
                       dGeo + 2 * (size_t)ntr, dGeo + 3 * (size_t)ntr};
        if (launch_exact(g, s, dTr, dVel, dImg, st) != PSKTM_OK) break;
        if ((e = hipMemcpyAsync(image, dImg, nimg * sizeof(float), hipMemcpyDeviceToHost, st)) != hipSuccess) { fail(e, "D2H image"); break; }
        if ((e = hipStreamSynchronize(st)) != hipSuccess) { fail(e, "hipStreamSynchronize"); break; }
        ok = true;
    } while (0);
    if (!ok) {
        (void)hipStreamSynchronize(st);
        memset(image, 0, nimg * sizeof(float));
    }
    if (dImg) (void)hipFree(dImg);
    if (dTr) (void)hipFree(dTr);
    if (dGeo) (void)hipFree(dGeo);
    if (dVel) (void)hipFree(dVel);
}

} // namespace

extern "C" {

// src/psktmCUDA.cu:211-221
void ktmMigrationGPU(struct imageGrid *imageX, struct imageGrid *imageY,
                     struct imageGrid *imageZ, struct jobParams *config,
                     float *midX, float *midY, float *offX, float *offY,
                     float *traces, float *slowness, float *image)
{
    legacy_migrate("ktmMigrationGPU", imageX, imageY, imageZ, config, midX, midY, offX, offY,
                   traces, slowness, image);
}

// src/psktmCPU.c:13-23: the same call, the same image as the reference's
// loop run on one thread (its OpenMP loop races on the image,
// src/psktmCPU.c:56,84), computed by the reference-exact HIP kernel.
void ktmMigrationCPU(struct imageGrid *imageX, struct imageGrid *imageY,
                     struct imageGrid *imageZ, struct jobParams *config,
                     float *midX, float *midY, float *offX, float *offY,
                     float *traces, float *slowness, float *image)
{
    legacy_migrate("ktmMigrationCPU", imageX, imageY, imageZ, config, midX, midY, offX, offY,
                   traces, slowness, image);
}

} // extern "C"
