// Internal declarations shared by the HIP translation units of libpsktm.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

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

// Raises a kernel's dynamic-LDS limit to at least `bytes` (above 64 KiB).
// The attribute is set once per kernel and size, before any graph capture
// that replays the launch.
psktm_status ensure_dynamic_lds(const void *kernel, size_t bytes);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Validates a grid and survey (sizes positive, pointers present).
psktm_status validate_grid(const psktm_grid *g);
psktm_status validate_survey(const psktm_survey *s);

// Reference-exact migration launcher (exact.hip).
psktm_status launch_exact(const psktm_grid &g, const psktm_survey &s,
                          const float *traces, const float *vel, float *image,
                          hipStream_t stream);

// Binary32 aperture rule shared with the oracle: d2 = (mx-x)^2 + (my-y)^2
// with every operation rounded separately, inside iff d2 <= r2.
__device__ __forceinline__ bool in_aperture_f32(float mx, float my, float x,
                                                float y, float r2)
{
#pragma clang fp contract(off)
    float ex = mx - x;
    float ey = my - y;
    float ex2 = ex * ex;
    float ey2 = ey * ey;
    float d2 = ex2 + ey2;
    return d2 <= r2;
}

// Binary32 axis value start + (float)i * step without contraction
// (src/psktmFunctions.c:28-34).
__device__ __forceinline__ float axis_f32(float start, float step, int i)
{
#pragma clang fp contract(off)
    float fi = (float)i;
    float p = fi * step;
    return start + p;
}

} // namespace psktm

// A migration plan: geometry binned into output tiles (see psktm_plan_create).
// 64-bit slots per adjoint accumulator row for traces of ns samples: samples
// 0 .. ns-1, then two slots that take discarded contributions (ns+1, ns+2),
// rounded up to an odd count (rows of different traces start on different
// LDS banks).
inline int adj_row(int ns) { return (ns + 3) | 1; }
// Traces per adjoint work item (upper bound): one. Eight 256-lane workgroups
// per CU on 2,000-sample traces; on the 512-CMP line 2, 3 and 4 traces per
// item were 3%, 11% and 40% slower.
constexpr int kAdjMax = 1;

struct psktm_plan {
    psktm_grid grid;
    psktm_options opt;
    int n_traces = 0, n_samples = 0;
    float dt = 0.0f;
    int device = 0;

    // output tiles of tile_x * tile_y bins, numbered T = TI * tiles_y + TJ
    int tile_x = 0, tile_y = 0, tiles_x = 0, tiles_y = 0, n_tiles = 0;
    bool all_traces = false;        // no aperture: every trace in every tile

    // midpoint cells (same size as tiles, extended by a margin of mx/my cells)
    int cell_mx = 0, cell_my = 0, cells_x = 0, cells_y = 0, n_cells = 0;
    int win_x = 0, win_y = 0;       // tile/cell search half-width in cells

    // device buffers
    float4 *geo = nullptr;          // [n_traces] (mid_x, mid_y, off_x, off_y)
    int *sorted_ids = nullptr;      // trace ids sorted by cell (stable)
    int *cell_start = nullptr;      // [n_cells + 1] offsets into sorted_ids
    long long *tile_off = nullptr;  // [n_tiles + 1] offsets into list
    int *list = nullptr;            // per-tile trace ids, tile after tile
    int *tile_order = nullptr;      // tiles by decreasing work
    float *reach = nullptr;         // [1] bound on |source - bin| and
                                    // |receiver - bin| over the plan's
                                    // (trace, in-aperture bin) pairs
    int *adj_items = nullptr;       // [n_adj_items * 2] (cell, first sorted pos)
    int n_adj_items = 0;
    int adj_batch = 0;              // traces per adjoint work item
    long long list_len = 0;
    double pairs = 0.0;             // (trace, in-aperture bin) pairs
    std::vector<double> tile_pairs; // per tile, host copy (work estimates)

    // per-call workspace: scaled slowness (1D or 3D) and tau^2/dt^2
    float *slow = nullptr;
    size_t slow_cap = 0;
    float *tsq = nullptr;
    size_t tsq_cap = 0;
    float *partial = nullptr;       // split-K partial images [splits][image]
    size_t partial_cap = 0;
    unsigned *adj_wmax = nullptr;   // [1] adjoint: max |image * tau * sqrt(dt)| (bits)

    size_t bytes = 0;
};

namespace psktm {

// Forward and adjoint launchers of the fast kernels (migrate.hip).
psktm_status launch_migrate(psktm_plan *p, const psktm_traces &tr,
                            const psktm_velocity &vel, float *image,
                            int accumulate, hipStream_t stream);
psktm_status launch_demigrate(psktm_plan *p, const float *image,
                              const psktm_velocity &vel, float *traces_out, int accumulate,
                              int gained, hipStream_t st);

// FP64 dot product in a fixed order (cg.hip): at most kDotPartials
// workgroups write one partial each into `partial`, then one workgroup adds
// them into result[0].
constexpr int kDotPartials = PSKTM_DOT_WORKSPACE;
psktm_status launch_dot(const float *a, const float *b, size_t n, double *result,
                        double *partial, hipStream_t stream);

// Per-call slowness/tau tables for a plan (migrate.hip).
psktm_status prepare_tables(psktm_plan *p, const psktm_velocity &vel,
                            hipStream_t stream);

} // namespace psktm
