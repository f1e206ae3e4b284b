// Internal declarations shared by the HIP translation units of libpsktm.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/psktm.h"

extern "C" {
// Thread-local error text behind psktm_last_error(); printf-style.
void psktm__set_error(const char *fmt, ...);
void psktm__clear_error(void);
}

namespace psktm {

// Status from a HIP call, with the failing expression recorded.
psktm_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define PSKTM_HIP(expr)                                                        \
    do {                                                                       \
        hipError_t e__ = (expr);                                               \
        if (e__ != hipSuccess)                                                 \
            return ::psktm::hip_fail(e__, #expr, __FILE__, __LINE__);          \
    } while (0)

#define PSKTM_CHECK(cond, status, ...)                                         \
    do {                                                                       \
        if (!(cond)) {                                                         \
            psktm__set_error(__VA_ARGS__);                                     \
            return (status);                                                   \
        }                                                                      \
    } while (0)

// Fails with PSKTM_E_NODEV unless the current device is a gfx950.
psktm_status require_device();

// Launch-error check after a kernel launch.
psktm_status check_launch(const char *kernel);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Validates a grid and survey (sizes positive, pointers present).
psktm_status validate_grid(const psktm_grid *g);
psktm_status validate_survey(const psktm_survey *s);

// Reference-exact migration launcher (exact.hip).
psktm_status launch_exact(const psktm_grid &g, const psktm_survey &s,
                          const float *traces, const float *vel, float *image,
                          hipStream_t stream);

} // namespace psktm
