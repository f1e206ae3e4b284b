// Migration plan: bins the survey's midpoints into output tiles.
//
// The reference migrates every trace into every image bin
// (src/psktmCPU.c:56-76); with an aperture most (trace, bin) pairs are empty,
// so each output tile needs the list of traces inside its aperture. Built once
// per geometry and reused by every migration and demigration:
//
//   1. assign each midpoint to a tile-sized cell on the bin lattice, extended
//      by the aperture; pack (mid, off) into one float4 per trace;
//   2. stable radix sort of trace ids by cell (hipCUB);
//   3. per tile, walk the cells within reach and keep the traces whose
//      midpoint lies within the aperture of the tile's bin rectangle - an
//      ordered block compaction, so lists are deterministic (cell order, then
//      trace id);
//   4. tiles ordered by decreasing work for the launch; adjoint work items
//      are (cell, batch of traces).
//
// The tile-level test is conservative (slightly larger radius); the exact
// per-bin binary32 test is applied inside the kernels.
#include "internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <vector>

namespace psktm {
namespace {

constexpr int kListThreads = 256;
constexpr int kListU = 4;   // candidates per lane per pass of the list kernels

struct Lattice {
    int nx, ny;
    float x0, dx, y0, dy;
    int tile_x, tile_y, tiles_x, tiles_y;
    int cell_mx, cell_my, cells_x, cells_y;
    int win_x, win_y;
    double rm2;      // conservative squared radius for tile-level tests
    float r2;        // binary32 squared aperture for the per-bin rule
};

// Cell index of coordinate m along one axis (-1 outside the extended range).
__device__ __forceinline__ int cell_axis(float m, float o, float d, int tile,
                                         int margin, int ncells, int nb)
{
    if (nb <= 1 || ncells <= 1) return 0;
    double u = ((double)m - (double)o) / ((double)d * tile) + 0.5 / tile;
    double c = floor(u) + margin;
    if (!(c >= 0.0) || c >= (double)ncells) return -1;
    return (int)c;
}

__global__ void assign_cells_kernel(int n, const float *__restrict__ mx,
                                    const float *__restrict__ my,
                                    const float *__restrict__ hx,
                                    const float *__restrict__ hy, Lattice L,
                                    float4 *__restrict__ geo,
                                    int *__restrict__ keys, int *__restrict__ vals,
                                    int drop_key)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float4 g = make_float4(mx[k], my[k], hx[k], hy[k]);
    geo[k] = g;
    int cx = cell_axis(g.x, L.x0, L.dx, L.tile_x, L.cell_mx, L.cells_x, L.nx);
    int cy = cell_axis(g.y, L.y0, L.dy, L.tile_y, L.cell_my, L.cells_y, L.ny);
    keys[k] = (cx < 0 || cy < 0) ? drop_key : cx * L.cells_y + cy;
    vals[k] = k;
}

// reach = max over traces of min(R, distance from the midpoint to the farthest
// image corner) + |h|: no source or receiver of a trace is farther than that
// from any bin the trace is migrated into (|s - x| <= |m - x| + |h|, and
// |m - x| <= R inside the aperture). Evaluated in FP64 and rounded up, then
// maxed over the bit patterns of the non-negative floats (a NaN coordinate
// gives a NaN pattern above every finite one, which no bound test passes).
__global__ void reach_kernel(int n, const float4 *__restrict__ geo, double cx0, double cx1,
                             double cy0, double cy1, double R, unsigned *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 g = geo[k];
    const double fx = fmax(fabs((double)g.x - cx0), fabs((double)g.x - cx1));
    const double fy = fmax(fabs((double)g.y - cy0), fabs((double)g.y - cy1));
    const double far = sqrt(fx * fx + fy * fy);
    const double h = sqrt((double)g.z * g.z + (double)g.w * g.w);
    const float r = (float)((fmin(R, far) + h) * (1.0 + 1e-6));
    atomicMax(out, __float_as_uint(r));
}

__global__ void fill_int_kernel(int *p, long long n, int v)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void iota_kernel(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// cell_start[c] = first sorted position with key c; cell_start[n_cells] = end
// of the binned traces. Empty cells get the start of the next non-empty one
// (filled by a backward max-scan on the host side of the plan).
__global__ void cell_bounds_kernel(int n, const int *__restrict__ skeys,
                                   int n_cells, int *__restrict__ first,
                                   int *__restrict__ last)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int key = skeys[p];
    if (key >= n_cells) return;
    if (p == 0 || skeys[p - 1] != key) first[key] = p;
    if (p == n - 1 || skeys[p + 1] != key) last[key] = p + 1;
}

__device__ __forceinline__ double axis_d(float start, float step, int i)
{
    return (double)axis_f32(start, step, i);
}

// Squared distance from a point to an axis-aligned rectangle.
__device__ __forceinline__ double point_rect_d2(double px, double py, double x0,
                                                double x1, double y0, double y1)
{
    double ex = fmax(fmax(x0 - px, px - x1), 0.0);
    double ey = fmax(fmax(y0 - py, py - y1), 0.0);
    return ex * ex + ey * ey;
}

// Conservative cell test: the rectangle of midpoints that cell (ci, cj) can
// hold (midpoints of cell c along x lie in [x0 + ((c-m)*tile - 1/2) dx,
// x0 + ((c-m+1)*tile - 1/2) dx)) comes within the aperture of the tile's bin
// rectangle [rx0, rx1] x [ry0, ry1]. Along cj at fixed ci the distance is
// convex, so the passing cells of a column form one run of cells.
__device__ __forceinline__ bool cell_in_reach(const Lattice &L, int ci, int cj, double rx0,
                                              double rx1, double ry0, double ry1)
{
    double ex = 0.0, ey = 0.0;
    if (L.cells_x > 1) {
        const double cx0 = L.x0 + ((double)(ci - L.cell_mx) * L.tile_x - 0.5) * L.dx;
        const double cx1 = cx0 + (double)L.tile_x * L.dx;
        ex = fmax(fmax(cx0 - rx1, rx0 - cx1), 0.0);
    }
    if (L.cells_y > 1) {
        const double cy0 = L.y0 + ((double)(cj - L.cell_my) * L.tile_y - 0.5) * L.dy;
        const double cy1 = cy0 + (double)L.tile_y * L.dy;
        ey = fmax(fmax(cy0 - ry1, ry0 - cy1), 0.0);
    }
    return !(ex * ex + ey * ey > L.rm2);
}

// One workgroup per output tile. COUNT: number of listed traces and of
// (trace, bin) pairs passing the binary32 rule. FILL: writes the list.
//
// The tile walks the cell columns within reach; in each column the cells
// that pass the cell test are one run of sorted positions, read as one
// stream of midpoints (smid: sorted position -> (mid x, mid y)), kListU
// candidates per lane per pass. The list keeps sorted-position order (cell
// order, then trace id): FILL ranks each pass's survivors by (slice, lane)
// with one exchange of wave counts; COUNT only needs totals and runs with no
// barrier per pass. Should rounding ever break a column's run, its passing
// cells are read one by one (same list).
template <bool FILL>
__global__ __launch_bounds__(kListThreads) void tile_list_kernel(
    Lattice L, const float2 *__restrict__ smid, const int *__restrict__ sorted_ids,
    const int *__restrict__ cell_start, long long *__restrict__ tile_count,
    unsigned long long *__restrict__ tile_pairs,
    const long long *__restrict__ tile_off, int *__restrict__ list)
{
    constexpr int kWaves = kListThreads / 64;
    __shared__ int s_cnt[kListU * kWaves];
    __shared__ int s_lo, s_hi, s_n;
    __shared__ unsigned long long s_pairs;
    __shared__ long long s_count;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int T = blockIdx.x;
    const int TI = T / L.tiles_y, TJ = T % L.tiles_y;
    const int i_lo = TI * L.tile_x, i_hi = min(L.nx, i_lo + L.tile_x) - 1;
    const int j_lo = TJ * L.tile_y, j_hi = min(L.ny, j_lo + L.tile_y) - 1;
    const double rx0 = axis_d(L.x0, L.dx, i_lo), rx1 = axis_d(L.x0, L.dx, i_hi);
    const double ry0 = axis_d(L.y0, L.dy, j_lo), ry1 = axis_d(L.y0, L.dy, j_hi);
    if (!FILL && tid == 0) {
        s_pairs = 0ull;
        s_count = 0;
    }
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;

    long long out = FILL ? tile_off[T] : 0;
    int my_n = 0;
    unsigned long long my_pairs = 0;
    auto scan = [&](int p0, int p1) {
        for (int base = p0; base < p1; base += kListThreads * kListU) {
            bool pass[kListU];
            int kk[kListU];
            float2 m[kListU];
            // clamped loads (p1 > base): all slices in flight at once
#pragma unroll
            for (int u = 0; u < kListU; ++u) {
                const int pc = min(base + u * kListThreads + tid, p1 - 1);
                m[u] = smid[pc];
                kk[u] = FILL ? sorted_ids[pc] : 0;
            }
#pragma unroll
            for (int u = 0; u < kListU; ++u)
                pass[u] = base + u * kListThreads + tid < p1 &&
                          point_rect_d2(m[u].x, m[u].y, rx0, rx1, ry0, ry1) <= L.rm2;
            if (FILL) {
                unsigned long long bal[kListU];
#pragma unroll
                for (int u = 0; u < kListU; ++u) bal[u] = __ballot(pass[u]);
                __syncthreads();   // the previous pass's counts have been read
                if (lane == 0) {
#pragma unroll
                    for (int u = 0; u < kListU; ++u) s_cnt[u * kWaves + wave] = __popcll(bal[u]);
                }
                __syncthreads();
                int run = 0;
#pragma unroll
                for (int u = 0; u < kListU; ++u) {
                    int before = run;
#pragma unroll
                    for (int w = 0; w < kWaves; ++w) {
                        const int c = s_cnt[u * kWaves + w];
                        before += (w < wave) ? c : 0;
                        run += c;
                    }
                    if (pass[u]) list[out + before + __popcll(bal[u] & below)] = kk[u];
                }
                out += run;
            } else {
#pragma unroll
                for (int u = 0; u < kListU; ++u) {
                    if (!pass[u]) continue;
                    ++my_n;
                    for (int i = i_lo; i <= i_hi; ++i) {
                        const float x = axis_f32(L.x0, L.dx, i);
                        for (int j = j_lo; j <= j_hi; ++j)
                            my_pairs += in_aperture_f32(m[u].x, m[u].y, x, axis_f32(L.y0, L.dy, j), L.r2);
                    }
                }
            }
        }
    };

    const int ci0 = max(0, TI + L.cell_mx - L.win_x), ci1 = min(L.cells_x - 1, TI + L.cell_mx + L.win_x);
    const int cj0 = max(0, TJ + L.cell_my - L.win_y), cj1 = min(L.cells_y - 1, TJ + L.cell_my + L.win_y);
    for (int ci = ci0; ci <= ci1; ++ci) {
        // first, last and number of passing cells of the column
        if (tid == 0) {
            s_lo = INT_MAX;
            s_hi = INT_MIN;
            s_n = 0;
        }
        __syncthreads();
        {
            int lo = INT_MAX, hi = INT_MIN, n = 0;
            for (int cj = cj0 + tid; cj <= cj1; cj += kListThreads)
                if (cell_in_reach(L, ci, cj, rx0, rx1, ry0, ry1)) {
                    lo = min(lo, cj);
                    hi = max(hi, cj);
                    ++n;
                }
            if (n) {
                atomicMin(&s_lo, lo);
                atomicMax(&s_hi, hi);
                atomicAdd(&s_n, n);
            }
        }
        __syncthreads();
        const int lo = s_lo, hi = s_hi, n = s_n;
        __syncthreads();   // read by every thread before the next column resets them
        if (n == 0) continue;
        const int cb = ci * L.cells_y;
        if (n == hi - lo + 1) {
            scan(cell_start[cb + lo], cell_start[cb + hi + 1]);
        } else {
            for (int cj = lo; cj <= hi; ++cj)
                if (cell_in_reach(L, ci, cj, rx0, rx1, ry0, ry1))
                    scan(cell_start[cb + cj], cell_start[cb + cj + 1]);
        }
    }
    if (!FILL) {
        if (my_n) atomicAdd((unsigned long long *)&s_count, (unsigned long long)my_n);
        if (my_pairs) atomicAdd(&s_pairs, my_pairs);
        __syncthreads();
        if (tid == 0) {
            tile_count[T] = s_count;
            tile_pairs[T] = s_pairs;
        }
    }
}

// Sorted position -> midpoint, so that the list passes read candidates as
// one coalesced stream instead of gathering geo[sorted_ids[p]].
__global__ void gather_mid_kernel(int n, const int *__restrict__ sorted_ids,
                                  const float4 *__restrict__ geo, float2 *__restrict__ smid)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 g = geo[sorted_ids[p]];
    smid[p] = make_float2(g.x, g.y);
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

} // namespace

static void plan_free(psktm_plan *p)
{
    if (!p) return;
    void *bufs[] = {p->geo, p->sorted_ids, p->cell_start, p->tile_off, p->list,
                    p->tile_order, p->adj_items, p->slow, p->tsq, p->partial, p->reach,
                    p->adj_wmax};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

#define PLAN_HIP(expr)                                                         \
    do {                                                                       \
        hipError_t e__ = (expr);                                               \
        if (e__ != hipSuccess) {                                               \
            psktm_status s__ = hip_fail(e__, #expr, __FILE__, __LINE__);       \
            plan_free(p);                                                      \
            return s__;                                                        \
        }                                                                      \
    } while (0)

template <typename T>
static hipError_t dalloc(psktm_plan *p, T **ptr, size_t n)
{
    hipError_t e = hipMalloc((void **)ptr, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) p->bytes += std::max<size_t>(n, 1) * sizeof(T);
    return e;
}

static psktm_status plan_build(psktm_plan *p, const psktm_survey &s, hipStream_t st)
{
    const psktm_grid &g = p->grid;
    const float R = p->opt.aperture;
    const int ntr = s.n_traces;
    Lattice L{};
    L.nx = g.nx; L.ny = g.ny; L.x0 = g.x0; L.dx = g.dx; L.y0 = g.y0; L.dy = g.dy;
    L.tile_x = p->tile_x; L.tile_y = p->tile_y;
    L.tiles_x = p->tiles_x; L.tiles_y = p->tiles_y;
    if (!p->all_traces) {
        const double Rd = R;
        L.cell_mx = (g.nx > 1) ? (int)std::ceil(Rd / ((double)p->tile_x * g.dx)) + 1 : 0;
        L.cell_my = (g.ny > 1) ? (int)std::ceil(Rd / ((double)p->tile_y * g.dy)) + 1 : 0;
        L.cells_x = (g.nx > 1) ? p->tiles_x + 2 * L.cell_mx : 1;
        L.cells_y = (g.ny > 1) ? p->tiles_y + 2 * L.cell_my : 1;
        L.win_x = L.cell_mx;
        L.win_y = L.cell_my;
        const double rm = Rd * (1.0 + 1e-5) + 1e-2;
        L.rm2 = rm * rm;
        L.r2 = R * R;
    } else {
        L.cells_x = L.cells_y = 1;
        L.win_x = p->tiles_x;
        L.win_y = p->tiles_y;
        L.rm2 = INFINITY;
        L.r2 = INFINITY;
    }
    p->cell_mx = L.cell_mx; p->cell_my = L.cell_my;
    p->cells_x = L.cells_x; p->cells_y = L.cells_y;
    p->n_cells = L.cells_x * L.cells_y;
    p->win_x = L.win_x; p->win_y = L.win_y;

    PLAN_HIP(dalloc(p, &p->geo, (size_t)ntr));
    PLAN_HIP(dalloc(p, &p->sorted_ids, (size_t)ntr));
    PLAN_HIP(dalloc(p, &p->cell_start, (size_t)p->n_cells + 1));
    PLAN_HIP(dalloc(p, &p->reach, 1));
    PLAN_HIP(dalloc(p, &p->adj_wmax, 1));
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, st, (int *)p->reach, 1ll, 0);
    const int nb = 256;
    const int gtr = (ntr + nb - 1) / nb;

    std::vector<int> h_start(p->n_cells + 1, 0);
    DevBuf smid;   // sorted position -> midpoint, for the list passes only
    if (ntr > 0) {
        DevBuf keys, vals, skeys, tmp;
        PLAN_HIP(hipMalloc(&keys.p, sizeof(int) * ntr));
        PLAN_HIP(hipMalloc(&vals.p, sizeof(int) * ntr));
        PLAN_HIP(hipMalloc(&skeys.p, sizeof(int) * ntr));
        const int drop_key = p->n_cells;
        hipLaunchKernelGGL(assign_cells_kernel, dim3(gtr), dim3(nb), 0, st, ntr,
                           s.mid_x, s.mid_y, s.off_x, s.off_y, L, p->geo,
                           (int *)keys.p, (int *)vals.p, drop_key);
        if (check_launch("assign_cells_kernel") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
        const double cx1 = (double)g.x0 + (double)(g.nx - 1) * g.dx;
        const double cy1 = (double)g.y0 + (double)(g.ny - 1) * g.dy;
        hipLaunchKernelGGL(reach_kernel, dim3(gtr), dim3(nb), 0, st, ntr, (const float4 *)p->geo,
                           (double)g.x0, cx1, (double)g.y0, cy1,
                           p->all_traces ? (double)INFINITY : (double)R, (unsigned *)p->reach);
        if (check_launch("reach_kernel") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
        int end_bit = 1;
        while ((1ll << end_bit) <= drop_key) ++end_bit;
        size_t tmp_bytes = 0;
        PLAN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (int *)keys.p,
                                                    (int *)skeys.p, (int *)vals.p,
                                                    p->sorted_ids, ntr, 0, end_bit, st));
        PLAN_HIP(hipMalloc(&tmp.p, std::max<size_t>(tmp_bytes, 1)));
        PLAN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, (int *)keys.p,
                                                    (int *)skeys.p, (int *)vals.p,
                                                    p->sorted_ids, ntr, 0, end_bit, st));
        if (!p->all_traces) {
            PLAN_HIP(hipMalloc(&smid.p, sizeof(float2) * ntr));
            hipLaunchKernelGGL(gather_mid_kernel, dim3(gtr), dim3(nb), 0, st, ntr,
                               (const int *)p->sorted_ids, (const float4 *)p->geo, (float2 *)smid.p);
            if (check_launch("gather_mid_kernel") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
        }
        // first/last sorted position of every non-empty cell
        DevBuf first, last;
        PLAN_HIP(hipMalloc(&first.p, sizeof(int) * p->n_cells));
        PLAN_HIP(hipMalloc(&last.p, sizeof(int) * p->n_cells));
        const int gc = (p->n_cells + nb - 1) / nb;
        hipLaunchKernelGGL(fill_int_kernel, dim3(gc), dim3(nb), 0, st, (int *)first.p,
                           (long long)p->n_cells, -1);
        hipLaunchKernelGGL(fill_int_kernel, dim3(gc), dim3(nb), 0, st, (int *)last.p,
                           (long long)p->n_cells, -1);
        hipLaunchKernelGGL(cell_bounds_kernel, dim3(gtr), dim3(nb), 0, st, ntr,
                           (const int *)skeys.p, p->n_cells, (int *)first.p, (int *)last.p);
        if (check_launch("cell_bounds_kernel") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
        std::vector<int> h_first(p->n_cells), h_last(p->n_cells);
        PLAN_HIP(hipMemcpyAsync(h_first.data(), first.p, sizeof(int) * p->n_cells, hipMemcpyDeviceToHost, st));
        PLAN_HIP(hipMemcpyAsync(h_last.data(), last.p, sizeof(int) * p->n_cells, hipMemcpyDeviceToHost, st));
        PLAN_HIP(hipStreamSynchronize(st));
        // cell_start: non-decreasing, empty cells start where the next begins
        int binned = 0;
        for (int c = 0; c < p->n_cells; ++c)
            if (h_last[c] >= 0) binned = std::max(binned, h_last[c]);
        h_start[p->n_cells] = binned;
        for (int c = p->n_cells - 1; c >= 0; --c)
            h_start[c] = (h_first[c] >= 0) ? h_first[c] : h_start[c + 1];
    }
    PLAN_HIP(hipMemcpyAsync(p->cell_start, h_start.data(), sizeof(int) * (p->n_cells + 1),
                            hipMemcpyHostToDevice, st));

    // ---- per-tile lists --------------------------------------------------
    const int nT = p->n_tiles;
    std::vector<long long> h_count(nT, 0);
    std::vector<unsigned long long> h_pairs(nT, 0);
    if (!p->all_traces && ntr > 0) {
        DevBuf dcount, dpairs;
        PLAN_HIP(hipMalloc(&dcount.p, sizeof(long long) * nT));
        PLAN_HIP(hipMalloc(&dpairs.p, sizeof(unsigned long long) * nT));
        hipLaunchKernelGGL(tile_list_kernel<false>, dim3(nT), dim3(kListThreads), 0, st, L,
                           (const float2 *)smid.p, p->sorted_ids, p->cell_start, (long long *)dcount.p,
                           (unsigned long long *)dpairs.p, (const long long *)nullptr,
                           (int *)nullptr);
        if (check_launch("tile_list_kernel<count>") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
        PLAN_HIP(hipMemcpyAsync(h_count.data(), dcount.p, sizeof(long long) * nT, hipMemcpyDeviceToHost, st));
        PLAN_HIP(hipMemcpyAsync(h_pairs.data(), dpairs.p, sizeof(unsigned long long) * nT, hipMemcpyDeviceToHost, st));
        PLAN_HIP(hipStreamSynchronize(st));
    } else if (ntr > 0) {
        for (int T = 0; T < nT; ++T) {
            const int TI = T / p->tiles_y, TJ = T % p->tiles_y;
            const long long bx = std::min(g.nx, (TI + 1) * p->tile_x) - TI * p->tile_x;
            const long long by = std::min(g.ny, (TJ + 1) * p->tile_y) - TJ * p->tile_y;
            h_count[T] = ntr;
            h_pairs[T] = (unsigned long long)ntr * bx * by;
        }
    }
    std::vector<long long> h_off(nT + 1, 0);
    double pairs = 0.0;
    for (int T = 0; T < nT; ++T) {
        h_off[T + 1] = h_off[T] + h_count[T];
        pairs += (double)h_pairs[T];
    }
    p->pairs = pairs;
    p->tile_pairs.assign(h_pairs.begin(), h_pairs.end());
    p->list_len = p->all_traces ? 0 : h_off[nT];
    // launch order: decreasing (traces x bins) so that long tiles start first
    std::vector<int> order(nT);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h_pairs[a] > h_pairs[b]; });
    PLAN_HIP(dalloc(p, &p->tile_off, (size_t)nT + 1));
    PLAN_HIP(dalloc(p, &p->tile_order, (size_t)nT));
    PLAN_HIP(hipMemcpyAsync(p->tile_off, h_off.data(), sizeof(long long) * (nT + 1), hipMemcpyHostToDevice, st));
    PLAN_HIP(hipMemcpyAsync(p->tile_order, order.data(), sizeof(int) * nT, hipMemcpyHostToDevice, st));
    if (!p->all_traces && p->list_len > 0) {
        PLAN_HIP(dalloc(p, &p->list, (size_t)p->list_len));
        hipLaunchKernelGGL(tile_list_kernel<true>, dim3(nT), dim3(kListThreads), 0, st, L,
                           (const float2 *)smid.p, p->sorted_ids, p->cell_start, (long long *)nullptr,
                           (unsigned long long *)nullptr, (const long long *)p->tile_off,
                           p->list);
        if (check_launch("tile_list_kernel<fill>") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
    }
    if (p->all_traces && ntr > 0) {
        hipLaunchKernelGGL(iota_kernel, dim3(gtr), dim3(nb), 0, st, p->sorted_ids, ntr);
        if (check_launch("iota_kernel") != PSKTM_OK) { plan_free(p); return PSKTM_E_LAUNCH; }
    }

    // ---- adjoint work items: (cell, first sorted position) ----------------
    // traces per adjoint work item: kAdjMax, as long as their 64-bit
    // accumulator rows fit in 150 KB of LDS
    p->adj_batch = (int)std::max(1ll, std::min((long long)kAdjMax, (150ll * 1024) /
                                                        (8ll * adj_row(s.n_samples))));
    std::vector<int> items;
    for (int c = 0; c < p->n_cells; ++c)
        for (int q = h_start[c]; q < h_start[c + 1]; q += p->adj_batch) {
            items.push_back(c);
            items.push_back(q);
        }
    p->n_adj_items = (int)(items.size() / 2);
    PLAN_HIP(dalloc(p, &p->adj_items, std::max<size_t>(items.size(), 2)));
    if (!items.empty())
        PLAN_HIP(hipMemcpyAsync(p->adj_items, items.data(), sizeof(int) * items.size(),
                                hipMemcpyHostToDevice, st));
    PLAN_HIP(hipStreamSynchronize(st));
    return PSKTM_OK;
}

} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_plan_create(psktm_plan **plan, const psktm_grid *grid,
                               const psktm_survey *survey, const psktm_options *opt,
                               void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(plan, PSKTM_E_INVAL, "plan output pointer is NULL");
    *plan = nullptr;
    psktm_status st;
    if ((st = validate_grid(grid)) != PSKTM_OK) return st;
    if ((st = validate_survey(survey)) != PSKTM_OK) return st;
    PSKTM_CHECK(opt, PSKTM_E_INVAL, "options are NULL");
    PSKTM_CHECK(opt->interp == PSKTM_INTERP_TRUNC || opt->interp == PSKTM_INTERP_LINEAR,
                PSKTM_E_INVAL, "unknown interpolation %d", opt->interp);
    PSKTM_CHECK(opt->math == PSKTM_MATH_FAST || opt->math == PSKTM_MATH_IEEE, PSKTM_E_INVAL,
                "unknown math mode %d", opt->math);
    PSKTM_CHECK(opt->math == PSKTM_MATH_FAST, PSKTM_E_UNSUPPORTED,
                "plans run the FAST form; use psktm_migrate_exact for IEEE arithmetic");
    PSKTM_CHECK(std::isfinite(opt->aperture), PSKTM_E_INVAL, "aperture must be finite");
    PSKTM_CHECK(opt->splits >= 0 && opt->splits <= 1024, PSKTM_E_INVAL,
                "splits must be in [0, 1024]");
    PSKTM_CHECK(opt->row_pad >= -1 && opt->row_pad <= (1 << 16), PSKTM_E_INVAL,
                "row_pad must be in [-1, 65536]");
    PSKTM_CHECK((grid->nx == 1 || grid->dx > 0.0f) && (grid->ny == 1 || grid->dy > 0.0f),
                PSKTM_E_INVAL, "bin spacings must be positive");
    // Default tiles of 8 bins: 4x2 on a 3-D grid, whose forward kernel (24
    // FP32 accumulators and slowness samples per lane, 79 VGPRs) keeps three
    // 512-lane workgroups per CU where 16-bin tiles keep two (3% faster on the
    // 3-D workload with 1-D and with 3-D velocity); 8x1 on a line, where the
    // precise-mode forward drops from 165 to 108 VGPRs (two workgroups per CU
    // instead of one: 12% faster on the least-squares line, whose adjoint
    // walks the smaller tiles 2% slower).
    int tx = opt->tile_x, ty = opt->tile_y;
    if (tx <= 0 || ty <= 0) {
        if (grid->ny == 1) { tx = 8; ty = 1; }
        else if (grid->nx == 1) { tx = 1; ty = 8; }
        else { tx = 4; ty = 2; }
    }
    PSKTM_CHECK(tx * ty == 8 || tx * ty == 16, PSKTM_E_INVAL,
                "tile must hold 8 or 16 bins (got %dx%d)", tx, ty);
    if ((st = require_device()) != PSKTM_OK) return st;

    psktm_plan *p = new psktm_plan();
    p->grid = *grid;
    p->opt = *opt;
    p->n_traces = survey->n_traces;
    p->n_samples = survey->n_samples;
    p->dt = survey->dt;
    (void)hipGetDevice(&p->device);
    p->tile_x = tx;
    p->tile_y = ty;
    p->tiles_x = (grid->nx + tx - 1) / tx;
    p->tiles_y = (grid->ny + ty - 1) / ty;
    p->n_tiles = p->tiles_x * p->tiles_y;
    p->all_traces = !(opt->aperture > 0.0f);
    st = plan_build(p, *survey, as_stream(stream));
    if (st != PSKTM_OK) return st;   // plan_build freed p
    *plan = p;
    return PSKTM_OK;
}

psktm_status psktm_plan_destroy(psktm_plan *plan)
{
    psktm__clear_error();
    if (plan) {
        (void)hipDeviceSynchronize();
        plan_free(plan);
    }
    return PSKTM_OK;
}

psktm_status psktm_plan_updates(const psktm_plan *plan, double *updates)
{
    psktm__clear_error();
    PSKTM_CHECK(plan && updates, PSKTM_E_INVAL, "NULL argument");
    *updates = plan->pairs * (double)plan->grid.nt;
    return PSKTM_OK;
}

psktm_status psktm_plan_bytes(const psktm_plan *plan, size_t *bytes)
{
    psktm__clear_error();
    PSKTM_CHECK(plan && bytes, PSKTM_E_INVAL, "NULL argument");
    *bytes = plan->bytes;
    return PSKTM_OK;
}

} // extern "C"
