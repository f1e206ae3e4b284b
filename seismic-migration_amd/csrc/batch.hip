// Batched gathers in one persistent launch.
//
// Offset-class gathers are independent migrations of very different size
// (thousands to tens of thousands of traces, patches of 32-128 bins a side).
// Launching them one by one leaves the GPU idle at every tail; here every
// gather gets its own plan (binning + per-tile lists), all (gather, tile,
// tau pass) items of the batch are sorted by decreasing cost, and one
// persistent launch of exactly as many workgroups as fit on the device pulls
// items from a device counter until the queue is empty. Each item runs the
// same tile body as the single-plan kernel (fwd_body.hpp), reading its
// parameters from a per-gather descriptor. The launch sequence - tau tables,
// queue reset (a kernel, so graph replays never contain a memset node),
// persistent kernel - can be captured in a HIP graph and replayed.
#include "fwd_body.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace psktm {
namespace {

template <int NTH, int TB, int TL, bool KH, bool GN, bool HP>
__global__ __launch_bounds__(NTH) void migrate_batch_kernel(const FwdArgs *__restrict__ descs,
                                                            const int2 *__restrict__ items,
                                                            int n_items,
                                                            unsigned *__restrict__ counter)
{
    extern __shared__ __align__(16) float2 s_pairs[];
    __shared__ float2 s_ab[kGMax][TB];
    __shared__ unsigned s_mask[2 * kGMax];   // aperture bits, then trace scales (HP)
    __shared__ int s_item;
    for (;;) {
        __syncthreads();                 // LDS of the previous item is free
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= n_items) break;        // every workgroup reaches this exit
        const int g = __builtin_amdgcn_readfirstlane(items[it].x);
        const int tc = __builtin_amdgcn_readfirstlane(items[it].y);
        const FwdArgs a = descs[g];
        fwd_tile<NTH, TB, TL, false, KH, GN, HP>(a, tc / a.n_chunks, tc % a.n_chunks, 0, s_pairs,
                                             s_ab, s_mask);
    }
}

__global__ void reset_counter_kernel(unsigned *c) { *c = 0u; }

__global__ void batch_tables_kernel(int nt, float tau0, float dtau, float inv_dt, float inv_dt2,
                                    const float *__restrict__ v, float *__restrict__ slow,
                                    float *__restrict__ tsq)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nt) return;
    const float vv = v[l];
    slow[l] = inv_dt2 / (vv * vv);
    const float t = axis_f32(tau0, dtau, l) * inv_dt;
    tsq[l] = t * t;
}

using BatchFn = void (*)(const FwdArgs *, const int2 *, int, unsigned *);

template <int NTH, int TB, int TL, bool KH, bool GN, bool HP>
BatchFn batch_pick()
{
    if constexpr (acc_regs(TB, TL, false, KH) <= reg_budget(KH) && !(KH && GN) && (GN || !HP))
        return migrate_batch_kernel<NTH, TB, TL, KH, GN, HP>;
    else
        return nullptr;
}

template <int NTH, int TB, bool KH, bool GN, bool HP = false>
BatchFn batch_by_tl(int tl)
{
    switch (tl) {
    case 2: return batch_pick<NTH, TB, 2, KH, GN, HP>();
    case 3: return batch_pick<NTH, TB, 3, KH, GN, HP>();
    case 4: return batch_pick<NTH, TB, 4, KH, GN, HP>();
    case 6: return batch_pick<NTH, TB, 6, KH, GN, HP>();
    case 8: return batch_pick<NTH, TB, 8, KH, GN, HP>();
    default: return nullptr;
    }
}

// hp: gained FP16 traces staged as half pairs
template <int NTH, int TB>
BatchFn batch_by_mode(int tl, bool kh, bool gn, bool hp)
{
    if (kh) return batch_by_tl<NTH, TB, true, false>(tl);
    if (gn && hp) return batch_by_tl<NTH, TB, false, true, true>(tl);
    return gn ? batch_by_tl<NTH, TB, false, true>(tl) : batch_by_tl<NTH, TB, false, false>(tl);
}

BatchFn batch_kernel(int nth, int tb, int tl, bool kh, bool gn, bool hp = false)
{
    if (nth == 256) return tb == 8 ? batch_by_mode<256, 8>(tl, kh, gn, hp) : batch_by_mode<256, 16>(tl, kh, gn, hp);
    return tb == 8 ? batch_by_mode<512, 8>(tl, kh, gn, hp) : batch_by_mode<512, 16>(tl, kh, gn, hp);
}

} // namespace
} // namespace psktm

struct psktm_batch {
    std::vector<psktm_plan *> plans;
    std::vector<psktm_gather> gathers;
    std::vector<size_t> image_off;
    psktm_options opt{};
    int ns = 0, nt = 0, tb = 0;
    float dt = 0.f, tau0 = 0.f, dtau = 0.f;
    psktm::FwdShape shape{};
    psktm::BatchFn fn = nullptr;        // raw traces
    psktm::BatchFn fn_gained = nullptr; // gained traces (same shape), NULL in precise mode
    psktm::BatchFn fn_half = nullptr;   // gained FP16 traces as half pairs (same shape)
    int G = 4, nsp = 0, lim = 0;
    int G_h = 4, nsp_h = 0, lim_h = 0;  // the half-pair kernel's stage, rows and limit
    int nth_h = 0;                      // and its workgroup size
    size_t lds = 0, lds_h = 0;
    int blocks = 0, blocks_h = 0;
    int2 *d_items = nullptr;
    int n_items = 0;
    unsigned *d_counter = nullptr;
    psktm::FwdArgs *d_args = nullptr;
    std::vector<psktm::FwdArgs> h_args;
    float *slow = nullptr, *tsq = nullptr;
    double pairs = 0.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;       // private stream used only to record the graph
    const void *key[5] = {};
};

namespace psktm {
namespace {

void batch_free(psktm_batch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (psktm_plan *p : b->plans)
        if (p) psktm_plan_destroy(p);
    void *bufs[] = {b->d_items, b->d_counter, b->d_args, b->slow, b->tsq};
    for (void *q : bufs)
        if (q) (void)hipFree(q);
    delete b;
}

psktm_status enqueue(psktm_batch *b, const float *v, bool gained, bool hp, hipStream_t st)
{
    const double inv_dt = 1.0 / (double)b->dt;
    hipLaunchKernelGGL(batch_tables_kernel, dim3((b->nt + 255) / 256), dim3(256), 0, st, b->nt,
                       b->tau0, b->dtau, (float)inv_dt, (float)(inv_dt * inv_dt), v, b->slow,
                       b->tsq);
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, st, b->d_counter);
    if (b->n_items > 0)
        hipLaunchKernelGGL(hp ? b->fn_half : (gained ? b->fn_gained : b->fn),
                           dim3(hp ? b->blocks_h : b->blocks), dim3(hp ? b->nth_h : b->shape.nth),
                           hp ? b->lds_h : b->lds, st,
                           (const FwdArgs *)b->d_args, (const int2 *)b->d_items, b->n_items,
                           b->d_counter);
    return check_launch("migrate_batch_kernel");
}

} // namespace
} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_batch_create(psktm_batch **batch, const psktm_gather *gathers, int n_gathers,
                                const psktm_survey *survey, const psktm_options *opt,
                                void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(batch && gathers && survey && opt, PSKTM_E_INVAL, "NULL argument");
    *batch = nullptr;
    PSKTM_CHECK(n_gathers > 0, PSKTM_E_INVAL, "a batch needs at least one gather");
    psktm_status st;
    if ((st = validate_survey(survey)) != PSKTM_OK) return st;
    const psktm_grid &g0 = gathers[0].grid;
    for (int g = 0; g < n_gathers; ++g) {
        const psktm_gather &q = gathers[g];
        if ((st = validate_grid(&q.grid)) != PSKTM_OK) return st;
        PSKTM_CHECK(q.grid.nt == g0.nt && q.grid.tau0 == g0.tau0 && q.grid.dtau == g0.dtau,
                    PSKTM_E_INVAL, "gather %d: all patches must share the tau axis", g);
        PSKTM_CHECK(q.first_trace >= 0 && q.n_traces >= 0 &&
                        (long long)q.first_trace + q.n_traces <= survey->n_traces,
                    PSKTM_E_INVAL, "gather %d: trace range outside the survey", g);
    }
    psktm_batch *b = new psktm_batch();
    b->opt = *opt;
    b->ns = survey->n_samples;
    b->dt = survey->dt;
    b->nt = g0.nt;
    b->tau0 = g0.tau0;
    b->dtau = g0.dtau;
    b->gathers.assign(gathers, gathers + n_gathers);
    size_t off = 0;
    for (int g = 0; g < n_gathers; ++g) {
        const psktm_gather &q = gathers[g];
        b->image_off.push_back(off);
        off += (size_t)q.grid.nx * q.grid.ny * q.grid.nt;
        psktm_survey s = *survey;
        s.n_traces = q.n_traces;
        s.mid_x += q.first_trace;
        s.mid_y += q.first_trace;
        s.off_x += q.first_trace;
        s.off_y += q.first_trace;
        psktm_plan *p = nullptr;
        st = psktm_plan_create(&p, &q.grid, &s, opt, stream);
        if (st != PSKTM_OK) {
            batch_free(b);
            return st;
        }
        b->plans.push_back(p);
        b->pairs += p->pairs;
    }
    // all plans share the tile size (same options)
    b->tb = b->plans[0]->tile_x * b->plans[0]->tile_y;
    const bool kh = opt->precise != 0;
    psktm::FwdLaunch L{};
    st = choose_launch(b->nt, b->ns, b->tb, false, kh, false, opt->traces_per_stage, opt->row_pad,
                       [&](int nth, int tl) { return (const void *)batch_kernel(nth, b->tb, tl, kh, false); },
                       &L);
    if (st != PSKTM_OK) {
        batch_free(b);
        return st;
    }
    b->shape = L.shape;
    b->fn = (BatchFn)L.fn;
    b->fn_gained = kh ? nullptr : batch_kernel(L.shape.nth, b->tb, L.shape.tl, false, true);
    b->nsp = L.nsp;
    b->lim = opt->row_pad < 0 ? 0 : pair_limit(b->nsp);
    const int G = L.G;
    b->G = G;
    b->lds = L.lds;
    if ((st = ensure_dynamic_lds((const void *)b->fn, b->lds)) != PSKTM_OK ||
        (b->fn_gained && (st = ensure_dynamic_lds((const void *)b->fn_gained, b->lds)) != PSKTM_OK)) {
        batch_free(b);
        return st;
    }
    if (!kh) {
        // the half-pair kernel: its own shape when that covers tau in as many
        // passes (the work items are per pass), else the same shape; its own
        // stage size either way
        psktm::FwdLaunch H{};
        const psktm::FwdShape sh = L.shape;
        st = choose_launch(b->nt, b->ns, b->tb, false, false, true, opt->traces_per_stage,
                           opt->row_pad,
                           [&](int nth, int tl) {
                               return (const void *)batch_kernel(nth, b->tb, tl, false, true, true);
                           },
                           &H);
        if (st == PSKTM_OK && H.shape.chunks != sh.chunks)
            st = choose_launch(b->nt, b->ns, b->tb, false, false, true, opt->traces_per_stage,
                               opt->row_pad,
                               [&](int nth, int tl) {
                                   return (nth == sh.nth && tl == sh.tl)
                                       ? (const void *)batch_kernel(nth, b->tb, tl, false, true, true)
                                       : nullptr;
                               },
                               &H);
        if (st == PSKTM_OK && H.shape.chunks == sh.chunks) {
            b->nth_h = H.shape.nth;
            b->fn_half = (BatchFn)H.fn;
            b->G_h = H.G;
            b->nsp_h = H.nsp;
            b->lim_h = opt->row_pad < 0 ? 0 : pair_limit(H.nsp);
            b->lds_h = H.lds;
            st = ensure_dynamic_lds((const void *)b->fn_half, b->lds_h);
        }
        if (st != PSKTM_OK) {
            batch_free(b);
            return st;
        }
    }
    // work items, longest first: cost = pairs of the tile x its tau pass
    std::vector<std::pair<double, int2>> items;
    for (int g = 0; g < n_gathers; ++g) {
        const psktm_plan *p = b->plans[g];
        for (int T = 0; T < p->n_tiles; ++T) {
            if (p->tile_pairs[T] <= 0.0 && !p->all_traces) {
                // still written (zeros): keep the item, it is cheap
            }
            for (int c = 0; c < b->shape.chunks; ++c)
                items.push_back({p->tile_pairs[T], make_int2(g, T * b->shape.chunks + c)});
        }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<double, int2> &x, const std::pair<double, int2> &y) {
                         return x.first > y.first;
                     });
    b->n_items = (int)items.size();
    std::vector<int2> h_items(items.size());
    for (size_t q = 0; q < items.size(); ++q) h_items[q] = items[q].second;
    hipStream_t s = as_stream(stream);
    auto fail = [&](hipError_t e, const char *what) {
        psktm_status r = hip_fail(e, what, __FILE__, __LINE__);
        batch_free(b);
        return r;
    };
    hipError_t e;
    if ((e = hipMalloc(&b->d_items, sizeof(int2) * std::max<size_t>(h_items.size(), 1))) != hipSuccess) return fail(e, "hipMalloc(items)");
    if ((e = hipMalloc(&b->d_counter, sizeof(unsigned))) != hipSuccess) return fail(e, "hipMalloc(counter)");
    if ((e = hipMalloc(&b->d_args, sizeof(FwdArgs) * n_gathers)) != hipSuccess) return fail(e, "hipMalloc(descriptors)");
    if ((e = hipMalloc(&b->slow, sizeof(float) * b->nt)) != hipSuccess) return fail(e, "hipMalloc(slowness)");
    if ((e = hipMalloc(&b->tsq, sizeof(float) * b->nt)) != hipSuccess) return fail(e, "hipMalloc(tau table)");
    if (!h_items.empty() &&
        (e = hipMemcpyAsync(b->d_items, h_items.data(), sizeof(int2) * h_items.size(),
                            hipMemcpyHostToDevice, s)) != hipSuccess)
        return fail(e, "H2D items");
    // persistent grid: as many workgroups as can be resident at once
    int dev = 0, cus = 0, per_cu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return fail(e, "hipGetDevice");
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return fail(e, "hipDeviceGetAttribute");
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)b->fn,
                                                          b->shape.nth, b->lds)) != hipSuccess)
        return fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    b->blocks = std::max(1, std::min(b->n_items, std::max(1, per_cu) * cus));
    if (b->fn_half) {
        if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)b->fn_half,
                                                              b->nth_h, b->lds_h)) != hipSuccess)
            return fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        b->blocks_h = std::max(1, std::min(b->n_items, std::max(1, per_cu) * cus));
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    *batch = b;
    return PSKTM_OK;
}

psktm_status psktm_batch_destroy(psktm_batch *batch)
{
    psktm__clear_error();
    if (batch) {
        (void)hipDeviceSynchronize();
        batch_free(batch);
    }
    return PSKTM_OK;
}

psktm_status psktm_batch_updates(const psktm_batch *batch, double *updates)
{
    psktm__clear_error();
    PSKTM_CHECK(batch && updates, PSKTM_E_INVAL, "NULL argument");
    *updates = batch->pairs * (double)batch->nt;
    return PSKTM_OK;
}

psktm_status psktm_batch_migrate(psktm_batch *b, const psktm_traces *tr, const psktm_velocity *vel,
                                 float *images, int use_graph, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(b && tr && vel && images, PSKTM_E_INVAL, "NULL argument");
    PSKTM_CHECK(vel->kind == PSKTM_VEL_1D && vel->v, PSKTM_E_INVAL,
                "batched gathers take a shared 1-D velocity v[nt]");
    PSKTM_CHECK(tr->dtype == PSKTM_F32 || tr->dtype == PSKTM_F16 || tr->dtype == PSKTM_BF16,
                PSKTM_E_INVAL, "unknown trace dtype %d", tr->dtype);
    PSKTM_CHECK(tr->data && (tr->dtype == PSKTM_F32 || tr->scale), PSKTM_E_INVAL,
                "trace data or scales missing");
    PSKTM_CHECK(tr->gained == 0 || tr->gained == 1, PSKTM_E_INVAL, "gained must be 0 or 1");
    PSKTM_CHECK(!tr->gained || b->fn_gained, PSKTM_E_UNSUPPORTED,
                "precise mode takes raw traces (gained = 0)");
    const bool hp = tr->gained && tr->dtype == PSKTM_F16 && b->fn_half;
    PSKTM_CHECK(!hp || ((uintptr_t)tr->data & 1u) == 0, PSKTM_E_INVAL,
                "FP16 trace data must be 2-byte aligned");
    hipStream_t st = as_stream(stream);
    const void *key[5] = {tr->data, tr->scale, vel->v, images,
                          (const void *)(intptr_t)(tr->dtype + 1 + 16 * tr->gained)};
    const bool same = b->exec && std::equal(key, key + 5, b->key);
    if (use_graph && same) {
        PSKTM_HIP(hipGraphLaunch(b->exec, st));
        return PSKTM_OK;
    }
    // descriptors for these pointers
    const size_t esz = tr->dtype == PSKTM_F32 ? 4 : 2;
    b->h_args.resize(b->plans.size());
    for (size_t g = 0; g < b->plans.size(); ++g) {
        FwdArgs a{};
        fwd_args_from_plan(b->plans[g], a);
        a.n_chunks = b->shape.chunks;
        a.dtype = tr->dtype;
        a.gained = tr->gained;
        if (tr->gained) a.wscale = 1.0f;
        a.G = hp ? b->G_h : b->G;
        a.nsp = hp ? b->nsp_h : b->nsp;
        a.lim = hp ? b->lim_h : b->lim;
        a.accumulate = 0;
        a.n_split = 1;
        a.traces = (const char *)tr->data + (size_t)b->gathers[g].first_trace * b->ns * esz;
        a.scale = tr->dtype == PSKTM_F32 ? nullptr : tr->scale + b->gathers[g].first_trace;
        a.slow = b->slow;
        a.tsq = b->tsq;
        a.image = images + b->image_off[g];
        b->h_args[g] = a;
    }
    if (!use_graph) {
        PSKTM_HIP(hipMemcpyAsync(b->d_args, b->h_args.data(), sizeof(FwdArgs) * b->h_args.size(),
                                 hipMemcpyHostToDevice, st));
        // h_args must outlive the copy: wait for it before returning
        PSKTM_HIP(hipStreamSynchronize(st));
        return enqueue(b, vel->v, tr->gained != 0, hp, st);
    }
    // (re)capture for these pointers. Recording happens on a private
    // non-blocking stream (the caller's stream may be the legacy default
    // stream, which cannot be captured); the graph is launched on the
    // caller's stream, so it stays ordered with the caller's work.
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    if (!b->cap) PSKTM_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    PSKTM_HIP(hipStreamSynchronize(st));
    PSKTM_HIP(hipMemcpy(b->d_args, b->h_args.data(), sizeof(FwdArgs) * b->h_args.size(),
                        hipMemcpyHostToDevice));
    PSKTM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    psktm_status s = enqueue(b, vel->v, tr->gained != 0, hp, b->cap);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &graph);
    if (s != PSKTM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return s;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    b->graph = graph;
    PSKTM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    std::copy(key, key + 5, b->key);
    PSKTM_HIP(hipGraphLaunch(b->exec, st));
    return PSKTM_OK;
}

} // extern "C"
