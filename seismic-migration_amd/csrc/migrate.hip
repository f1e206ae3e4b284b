// Fast PSTM kernels for gfx950: forward migration (traces -> image) and its
// adjoint (image -> traces), both driven by a psktm_plan.
//
// Forward (gather, output-stationary). A workgroup owns one output tile of
// TB bins and NTH*TL tau samples; every lane keeps TB*TL image samples in
// FP32 registers for the whole tile, so the image is written once. The tile's
// trace list is streamed in stages of G traces:
//   fill    - each trace is expanded into LDS as (d[n], d[n+1]-d[n]) pairs in
//             FP32 (16-bit traces are decoded and scaled here), and lanes
//             (g, b) tabulate the distance terms A = |s-x|^2, B = |r-x|^2 and
//             the binary32 aperture bit of every (trace, bin);
//   compute - per (trace, bin) that passed (a wave-uniform scalar branch),
//             each lane evaluates its TL samples:
//               x  = sqrt(tau'^2 + A s') + sqrt(tau'^2 + B s')   (= t/dt)
//               d  = fma(fract(x), dd[n], d[n]),  n = (int)min(x, ns)
//               acc += d * sqrt(x)
//             with s' = 1/(v dt)^2 and tau' = tau/dt; the weight
//             t*sqrt(tau^2/t) = tau*sqrt(dt)*sqrt(x) (src/psktmCPU.c:84-85)
//             is completed by one multiply per sample at the end. Gained
//             traces (sample n already times sqrt(n dt)) accumulate d alone
//             (GN instantiation: one v_sqrt_f32 fewer) and end with tau.
//             Gained FP16 traces (HP) stage half pairs instead, one dword
//             (h[n]/2, fp16(h[n+1]/2-h[n]/2)) per entry of the stored values
//             (halved so that no difference of finite FP16 values
//             overflows): half the LDS per row, twice the traces per stage,
//             interpolation by v_fma_mix_f32 and twice the trace's scale
//             applied in the accumulate.
// Out-of-range samples need no test: the pair table's entries outside the
// valid range of the interpolation rule are zero (TRUNC: 0 < n < ns,
// src/psktmCPU.c:82; LINEAR: 0 <= n, n+1 < ns), and x is clamped to ns unless
// the tile's traveltimes are bounded inside the zero-padded rows.
// The path is bound by VALU issue (two or three v_sqrt_f32 and ~8 other VALU
// ops per update); HBM traffic is the trace rows, read once per tile.
//
// Adjoint (scatter, trace-stationary). A workgroup owns a batch of traces
// from one midpoint cell, visits every tile within reach of the cell and
// writes each trace once. Lanes merge their runs of equal trace samples in
// FP32 registers and add them to LDS as exact 64-bit fixed-point integers
// (ds_add_u64; no FP64, and the sums do not depend on the atomics' order). It
// uses the same traveltime, validity rule, interpolation weights and aperture
// bits as the forward, so the pair is an exact adjoint up to FP32 rounding.
#include "fwd_body.hpp"

#include <algorithm>
#include <cmath>

namespace psktm {
namespace {


// An adjoint accumulator row holds one 64-bit slot per trace sample,
// n = 0 .. ns-1, and two more (ns+1, ns+2) that take the contributions of
// invalid samples, which are never read (adj_row(ns) >= ns + 3).

struct AdjArgs {
    int nx, ny, nt, ns;
    float x0, dx, y0, dy, tau0, dtau;
    int tile_x, tile_y, tiles_x, tiles_y, n_chunks;
    int cells_y, cell_mx, cell_my, win_x, win_y;
    float nsf, r2, wscale;
    float bound;                  // bound on (sum of |contributions| to one slot) / max|image*w|
    float cbound;                 // bound on |a lane's merged contribution| / max|image*w|
    float gain_dt;                // gained operator: output sample n times sqrtf(n gain_dt); 0: none
    const unsigned *wmax;         // max |image * tau * sqrt(dt)| (bits), per launch
    int use_ap, interp, nsa, batch;
    const float4 *geo;
    const int *sorted_ids;
    const int *cell_start;
    const int *items;
    const float *slow;
    const float *tsq;
    const float *image;
    float *traces;
};

// Resident waves per SIMD asked of the forward kernel: three 512-lane
// workgroups per CU (at most 80 VGPRs) for the half-pair 3-D tiles of three
// tau samples per lane, whose fill alone would take it past 80; the other
// forms fit as compiled (asking them too changes their schedule: FP32 pairs
// ran 5% slower).
constexpr int fwd_waves_per_eu(int nth, int tl, bool hp)
{
    return nth == 512 && tl <= 3 && hp ? 6 : 1;
}

template <int NTH, int TB, int TL, bool V3, bool KH, bool GN, bool HP>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(fwd_waves_per_eu(NTH, TL, HP))))
void migrate_fwd_kernel(const FwdArgs a)
{
    extern __shared__ __align__(16) float2 s_pairs[];
    __shared__ float2 s_ab[kGMax][TB];
    __shared__ unsigned s_mask[2 * kGMax];   // aperture bits, then trace scales (HP)
    // block -> (tile in work order, tau pass, split); splits of a tile adjacent
    const int split = blockIdx.x % a.n_split;
    const int rest = blockIdx.x / a.n_split;
    fwd_tile<NTH, TB, TL, V3, KH, GN, HP, HP>(a, a.tile_order[rest / a.n_chunks], rest % a.n_chunks,
                                              split, s_pairs, s_ab, s_mask);
}

// Exact fixed-point accumulation of the adjoint. A lane's merged
// contribution y, already multiplied by 2^S, is rounded to the nearest
// integer (ties to even) and added as a 64-bit two's complement integer.
// With h = rint(y / 2^22) and r = y - h 2^22 (|r| <= 2^21):
//   t1 = fma(y, 2^-22, 1.5 2^23) = 1.5 2^23 + h,
//   t2 = fma(h - 3, -2^22, y)   = 1.5 2^23 + rint(r),
// each one rounding into [2^23, 2^24), where the floats are the integers
// (so |y| < 2^43), leaving h and rint(r) in the low mantissa bits. As
// bits(1.5 2^23) = 301 2^22, the integer is
//   h 2^22 + rint(r) = (bits(t1) - 301 2^22 - 301) 2^22 + bits(t2),
// one 64-bit multiply-add. S is chosen per launch (from adj_max_kernel) so
// that every merged contribution has |y| < 2^43 and no slot's sum can reach
// 2^62 in magnitude: the integer sums never overflow, and integer addition
// is associative, so the adjoint's output does not depend on the order of
// the LDS atomics. No FP64 is used.
__device__ __forceinline__ unsigned long long to_fixed(float y)
{
#pragma clang fp contract(off)
    const float t1 = fmaf(y, 2.384185791015625e-07f, 12582912.0f);
    const float hm3 = t1 - 12582915.0f;                         // h - 3, exact
    const float t2 = fmaf(hm3, -4194304.0f, y);
#if defined(__HIP_DEVICE_COMPILE__)
    const int hi = (int)(__float_as_uint(t1) - 0x4B40012Du);
    unsigned long long u;
    asm("v_mad_i64_i32 %0, vcc, %1, %2, %3"
        : "=v"(u) : "v"(hi), "s"(4194304), "v"((unsigned long long)__float_as_uint(t2))
        : "vcc");
    return u;
#else
    return (unsigned long long)((long long)(t1 - 12582912.0f) * 4194304ll +
                                (long long)(t2 - 12582912.0f));
#endif
}

// The float nearest (within an ulp) to a fixed-point sum u (|u| < 2^62).
__device__ __forceinline__ float from_fixed(unsigned long long u)
{
    const int hi = (int)(unsigned)(u >> 32);
    return fmaf((float)hi, 4294967296.0f, (float)(unsigned)u);
}

// Per launch: m = max over the image of |image * tau * sqrt(dt)| (order-free
// atomic max of the bit patterns; any non-finite sample gives a NaN pattern
// above every finite one).
__global__ void adj_max_kernel(const float *__restrict__ image, long long n, int nt, float tau0,
                               float dtau, float wscale, unsigned *__restrict__ out)
{
    float m = 0.f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int l = (int)(i % nt);
        const float v = fabsf(image[i] * (axis_f32(tau0, dtau, l) * wscale));
        m = (v > m || v != v) ? v : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float u = __shfl_xor(m, o);
        m = (u > m || u != u) ? u : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m != m ? 0x7fc00000u : __float_as_uint(m));
}

template <int NTH, int TB, int TL, bool V3, bool LIN, bool GN>
__global__ __launch_bounds__(NTH) void demigrate_kernel(const AdjArgs a)
{
    // Products and sums round one at a time (explicit fmaf calls stay
    // fused), so the scatter weights do not depend on how the compiler
    // schedules the merged accumulation below.
#pragma clang fp contract(off)
    static_assert(TL % 4 == 0, "a lane's tau samples are loaded 4 at a time");
    extern __shared__ __align__(8) unsigned char s_raw[];
    unsigned long long *s_acc = reinterpret_cast<unsigned long long *>(s_raw);
    __shared__ float2 s_ab[kAdjMax][TB];
    __shared__ unsigned s_mask[kAdjMax];
    __shared__ int s_k[kAdjMax];

    const int tid = threadIdx.x;
    const int c = a.items[2 * blockIdx.x], p0 = a.items[2 * blockIdx.x + 1];
    const int p1 = min(p0 + a.batch, a.cell_start[c + 1]);
    const int ng = p1 - p0;
    for (int q = tid; q < a.batch * a.nsa; q += NTH) s_acc[q] = 0ull;
    if (tid < kAdjMax) s_k[tid] = (tid < ng) ? a.sorted_ids[p0 + tid] : -1;

    // fixed-point scale 2^S (see to_fixed): mx * cbound < 2^(e1 + e3) bounds
    // a lane's merged contribution (which must stay below 2^43), mx * bound <
    // 2^(e1 + e2) the sum of |contributions| to a slot
    const float mx = __uint_as_float(*a.wmax);
    const bool bad = !(mx <= 3.0e38f);           // NaN or inf in the image
    int e1 = 0, e2 = 0, e3 = 0;
    (void)frexpf(bad ? 0.f : mx, &e1);
    (void)frexpf(a.bound, &e2);
    (void)frexpf(a.cbound, &e3);
    // S reaches 250 for images near the FP32 denormal range (and is negative
    // for images near FLT_MAX): applied by ldexp, which rounds once
    const int S = (bad || !(mx > 0.f)) ? 0
                : min(250, max(-120, min(43 - (e1 + e3), 62 - (e1 + e2))));
    // 2^S folded into the weight w = tau sqrt(dt) (one instruction fewer per
    // update) where w 2^S stays finite and well above the subnormals at both
    // ends of the tau axis: image * (w 2^S) is then the same product as
    // image * w times 2^S, rounded once
    const float wend = fmaxf(fabsf(axis_f32(a.tau0, a.dtau, 0)),
                             fabsf(axis_f32(a.tau0, a.dtau, a.nt - 1))) * a.wscale;
    const bool fold = S >= -60 && S <= 100 && ldexpf(wend, S) < 1.0e37f;
    __syncthreads();

    const int CI = c / a.cells_y, CJ = c % a.cells_y;
    const int ti0 = max(0, CI - a.cell_mx - a.win_x), ti1 = min(a.tiles_x - 1, CI - a.cell_mx + a.win_x);
    const int tj0 = max(0, CJ - a.cell_my - a.win_y), tj1 = min(a.tiles_y - 1, CJ - a.cell_my + a.win_y);
    // valid sample index: LINEAR 0 <= n <= ns-2, TRUNC 1 <= n <= ns-1
    // (src/psktmCPU.c:82)
    const float xhi = a.nsf - 1.f;
    const float xmid = 0.5f * (a.nsf + 1.f), xrad = 0.5f * (a.nsf - 1.f);
    // the discard slot in a VGPR (the validity select then reads no SGPR)
    int disc = a.ns + 1;
    asm volatile("" : "+v"(disc));

    for (int TI = ti0; TI <= ti1; ++TI) {
        for (int TJ = tj0; TJ <= tj1; ++TJ) {
            const int i0 = TI * a.tile_x, j0 = TJ * a.tile_y;
            if (tid < kAdjMax * TB) {
                const int g = tid / TB, b = tid % TB;
                float2 ab = make_float2(0.f, 0.f);
                bool m = false;
                const int bi = b / a.tile_y, bj = b % a.tile_y;
                const int i = i0 + bi, jj = j0 + bj;
                if (g < ng && bi < a.tile_x && i < a.nx && jj < a.ny)
                    m = pair_terms(a.geo[s_k[g]], i, jj, a.x0, a.dx, a.y0, a.dy, a.use_ap, a.r2, &ab);
                s_ab[g][b] = ab;
                const unsigned long long bal = __ballot(m);
                if ((tid % TB) == 0) {
                    const int shift = (tid & 63) - ((tid & 63) % TB);
                    s_mask[g] = (unsigned)((bal >> shift) & ((1ull << TB) - 1ull));
                }
            }
            __syncthreads();
            unsigned any = 0;
            for (int g = 0; g < kAdjMax; ++g) any |= s_mask[g];
            any = __builtin_amdgcn_readfirstlane(any);
            if (any) {
                for (int ch = 0; ch < a.n_chunks; ++ch) {
                    // each lane takes TL consecutive tau samples, so that the
                    // lanes of a wave scatter to samples further apart
                    const int lbase = ch * (NTH * TL) + tid * TL;
                    float tq[TL], wt[TL], sl1[TL];
#pragma unroll
                    for (int j = 0; j < TL; ++j) {
                        const int l = lbase + j;
                        const bool ok = l < a.nt;
                        tq[j] = ok ? a.tsq[l] : 0.f;
                        // weight tau*sqrt(dt) (the forward's epilogue)
                        wt[j] = ok ? (axis_f32(a.tau0, a.dtau, l) * a.wscale) : 0.f;
                        if (fold) wt[j] = ldexpf(wt[j], S);
                        sl1[j] = (!V3 && ok) ? a.slow[l] : 0.f;
                    }
                    // bins and traces in rolled loops: the kernel's code
                    // stays small (TL steps per trace and bin)
#pragma unroll 1
                    for (int b = 0; b < TB; ++b) {
                        if (!((any >> b) & 1u)) continue;
                        const int i = i0 + b / a.tile_y, jj = j0 + b % a.tile_y;
                        // the bin's image (and 3-D slowness) column through
                        // buffer descriptors whose range is the column: the
                        // lane's TL samples in 16-byte loads, zeros past nt
                        const size_t col = ((size_t)i * a.ny + jj) * a.nt;
                        const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc(
                            (void *)(a.image + col), (short)0, a.nt * 4, 0x00020000);
                        const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
                            (void *)(a.slow + (V3 ? col : 0)), (short)0, a.nt * 4, 0x00020000);
                        float im[TL], sl[TL];
#pragma unroll
                        for (int q = 0; q < TL; q += 4) {
                            const auto vi = __builtin_amdgcn_raw_buffer_load_b128(ri, 4 * (lbase + q), 0, 0);
                            const auto vv = V3 ? __builtin_amdgcn_raw_buffer_load_b128(rv, 4 * (lbase + q), 0, 0)
                                               : vi;
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                im[q + e] = __uint_as_float(vi[e]) * wt[q + e];
                                sl[q + e] = V3 ? __uint_as_float(vv[e]) : sl1[q + e];
                            }
                        }
                        if (!fold) {
#pragma unroll
                            for (int j = 0; j < TL; ++j) im[j] = ldexpf(im[j], S);
                        }
#pragma unroll 1
                        for (int g = 0; g < kAdjMax; ++g) {
                            const unsigned m = __builtin_amdgcn_readfirstlane(s_mask[g]);
                            if (!((m >> b) & 1u)) continue;
                            const float2 ab = s_ab[g][b];
                            unsigned long long *acc = s_acc + g * a.nsa;
                            // A lane's TL consecutive tau samples reach few
                            // distinct trace samples (at most one more per tau
                            // step where the time map is flat or slow): hold
                            // the running samples nc, nc+1 in registers and add
                            // them to LDS when the index moves on.
                            int nc = 0;
                            float c0 = 0.f, c1 = 0.f;
#pragma unroll
                            for (int j = 0; j < TL; ++j) {
                                const float ua = fmaf(ab.x, sl[j], tq[j]);
                                const float ub = fmaf(ab.y, sl[j], tq[j]);
                                const float x = __builtin_amdgcn_sqrtf(ua) + __builtin_amdgcn_sqrtf(ub);
                                // LINEAR: 0 <= x < ns-1 (x >= 0 holds); TRUNC:
                                // 1 <= x < ns, as |x - (ns+1)/2| < (ns-1)/2. A
                                // NaN time fails both. Invalid samples go to
                                // the discarded slots ns+1, ns+2.
                                const bool ok = LIN ? x < xhi : fabsf(x - xmid) < xrad;
                                const int n = ok ? (int)x : disc;
                                const float f = __builtin_amdgcn_fractf(x);
                                // gained operator: weight tau alone (in im)
                                const float w = GN ? im[j] : __builtin_amdgcn_sqrtf(x) * im[j];
                                const float w1 = LIN ? w * f : 0.f;
                                const float v0 = w - w1, v1 = w1;
                                if (j == 0) {
                                    nc = n;
                                    c0 = v0;
                                    c1 = v1;
                                    continue;
                                }
                                // moved on: the running sample nc is complete;
                                // one sample on, nc + 1 becomes the running
                                // sample, else (a jump) it is complete too
                                const bool moved = n != nc, next = n == nc + 1;
                                if (moved) {
                                    atomicAdd(acc + nc, to_fixed(c0));
                                    if (LIN && !next) atomicAdd(acc + nc + 1, to_fixed(c1));
                                }
                                c0 = (moved ? (next ? c1 : 0.f) : c0) + v0;
                                c1 = (moved ? 0.f : c1) + v1;
                                nc = n;
                            }
                            atomicAdd(acc + nc, to_fixed(c0));
                            if (LIN) atomicAdd(acc + nc + 1, to_fixed(c1));
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int g = 0; g < ng; ++g) {
        float *out = a.traces + (size_t)s_k[g] * a.ns;
        const unsigned long long *acc = s_acc + g * a.nsa;
        for (int n = tid; n < a.ns; n += NTH) {
            float v = bad ? __int_as_float(0x7fc00000) : ldexpf(from_fixed(acc[n]), -S);
            // the gained operator's trace gain (psktm_trace_gain's sqrtf(n dt)),
            // transposed: applied to the output sample
            if (GN) v = v * sqrtf((float)n * a.gain_dt);
            out[n] = out[n] + v;        // zeroed beforehand unless accumulating
        }
    }
}

__global__ void tables_kernel(int nt, float tau0, float dtau, float inv_dt, float inv_dt2,
                              const float *__restrict__ v, size_t nv, float *__restrict__ slow,
                              float *__restrict__ tsq)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nv) {
        const float vv = v[q];
        slow[q] = inv_dt2 / (vv * vv);
    }
    if (q < (size_t)nt) {
        const float t = axis_f32(tau0, dtau, (int)q) * inv_dt;
        tsq[q] = t * t;
    }
}

__global__ void zero_kernel(float *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// image[i] (+)= sum over splits of partial[s][i], in split order. Precise
// mode (compensated != 0): partial also holds n_split low-order planes after
// the n_split high-order ones, and everything is added by compensated FP32
// summation (each addition's rounding error recovered exactly by TwoSum and
// summed separately, then added once: as accurate as summing in twice the
// precision and rounding, with no FP64).
__global__ void reduce_partials_kernel(const float *__restrict__ partial, int n_split,
                                       long long n, int accumulate, int compensated,
                                       float *__restrict__ image)
{
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (compensated) {
        float t = accumulate ? image[i] : 0.f, e = 0.f;
        for (int s = 0; s < 2 * n_split; ++s) {
            const float x = partial[(size_t)s * n + i];
            const float u = t + x, bv = u - t;
            e += (t - (u - bv)) + (x - bv);
            t = u;
        }
        image[i] = t + e;
        return;
    }
    float t = partial[i];
    for (int s = 1; s < n_split; ++s) t += partial[(size_t)s * n + i];
    image[i] = accumulate ? image[i] + t : t;
}

// ---- kernel selection ------------------------------------------------------

using FwdFn = void (*)(FwdArgs);
using AdjFn = void (*)(AdjArgs);

template <int NTH, int TB, int TL, bool V3, bool KH, bool GN, bool HP>
FwdFn fwd_pick()
{
    if constexpr (acc_regs(TB, TL, V3, KH) <= reg_budget(KH) && !(KH && HP) && (GN || !HP))
        return migrate_fwd_kernel<NTH, TB, TL, V3, KH, GN, HP>;
    else
        return nullptr;
}

template <int NTH, int TB, bool V3, bool KH, bool GN, bool HP = false>
FwdFn fwd_by_tl(int tl)
{
    switch (tl) {
    case 2: return fwd_pick<NTH, TB, 2, V3, KH, GN, HP>();
    case 3: return fwd_pick<NTH, TB, 3, V3, KH, GN, HP>();
    case 4: return fwd_pick<NTH, TB, 4, V3, KH, GN, HP>();
    case 6: return fwd_pick<NTH, TB, 6, V3, KH, GN, HP>();
    case 8: return fwd_pick<NTH, TB, 8, V3, KH, GN, HP>();
    default: return nullptr;
    }
}

// mode: 0 raw traces, 1 precise (compensated), 2 gained traces, 3 gained
// FP16 traces staged as half pairs, 4 precise with gained traces
template <int NTH, int TB>
FwdFn fwd_by_flags(int tl, bool v3, int mode)
{
    if (v3) {
        if (mode == 4) return fwd_by_tl<NTH, TB, true, true, true>(tl);
        if (mode == 1) return fwd_by_tl<NTH, TB, true, true, false>(tl);
        if (mode == 2) return fwd_by_tl<NTH, TB, true, false, true>(tl);
        if (mode == 3) return fwd_by_tl<NTH, TB, true, false, true, true>(tl);
        return fwd_by_tl<NTH, TB, true, false, false>(tl);
    }
    if (mode == 4) return fwd_by_tl<NTH, TB, false, true, true>(tl);
    if (mode == 1) return fwd_by_tl<NTH, TB, false, true, false>(tl);
    if (mode == 2) return fwd_by_tl<NTH, TB, false, false, true>(tl);
    if (mode == 3) return fwd_by_tl<NTH, TB, false, false, true, true>(tl);
    return fwd_by_tl<NTH, TB, false, false, false>(tl);
}

FwdFn fwd_kernel(int nth, int tb, int tl, bool v3, bool kh, bool gn, bool hp)
{
    const int mode = kh ? (gn ? 4 : 1) : (gn ? (hp ? 3 : 2) : 0);
    if (nth == 256) return tb == 8 ? fwd_by_flags<256, 8>(tl, v3, mode) : fwd_by_flags<256, 16>(tl, v3, mode);
    return tb == 8 ? fwd_by_flags<512, 8>(tl, v3, mode) : fwd_by_flags<512, 16>(tl, v3, mode);
}

} // namespace

psktm_status prepare_tables(psktm_plan *p, const psktm_velocity &vel, hipStream_t st)
{
    const psktm_grid &g = p->grid;
    const size_t nv = (vel.kind == PSKTM_VEL_3D) ? (size_t)g.nx * g.ny * g.nt : (size_t)g.nt;
    if (p->slow_cap < nv) {
        if (p->slow) { (void)hipFree(p->slow); p->bytes -= p->slow_cap * sizeof(float); p->slow = nullptr; }
        PSKTM_HIP(hipMalloc(&p->slow, nv * sizeof(float)));
        p->slow_cap = nv;
        p->bytes += nv * sizeof(float);
    }
    if (p->tsq_cap < (size_t)g.nt) {
        if (p->tsq) { (void)hipFree(p->tsq); p->bytes -= p->tsq_cap * sizeof(float); p->tsq = nullptr; }
        PSKTM_HIP(hipMalloc(&p->tsq, (size_t)g.nt * sizeof(float)));
        p->tsq_cap = g.nt;
        p->bytes += (size_t)g.nt * sizeof(float);
    }
    const double inv_dt = 1.0 / (double)p->dt;
    const size_t n = std::max(nv, (size_t)g.nt);
    hipLaunchKernelGGL(tables_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g.nt,
                       g.tau0, g.dtau, (float)inv_dt, (float)(inv_dt * inv_dt), vel.v, nv,
                       p->slow, p->tsq);
    return check_launch("tables_kernel");
}

static psktm_status validate_velocity(const psktm_velocity &vel)
{
    PSKTM_CHECK(vel.v, PSKTM_E_INVAL, "velocity pointer is NULL");
    PSKTM_CHECK(vel.kind == PSKTM_VEL_1D || vel.kind == PSKTM_VEL_3D, PSKTM_E_INVAL,
                "unknown velocity kind %d", vel.kind);
    return PSKTM_OK;
}

psktm_status launch_migrate(psktm_plan *p, const psktm_traces &tr, const psktm_velocity &vel,
                            float *image, int accumulate, hipStream_t st)
{
    psktm_status s;
    if ((s = validate_velocity(vel)) != PSKTM_OK) return s;
    PSKTM_CHECK(tr.dtype == PSKTM_F32 || tr.dtype == PSKTM_F16 || tr.dtype == PSKTM_BF16,
                PSKTM_E_INVAL, "unknown trace dtype %d", tr.dtype);
    PSKTM_CHECK(p->n_traces == 0 || tr.data, PSKTM_E_INVAL, "trace data is NULL");
    PSKTM_CHECK(tr.dtype == PSKTM_F32 || p->n_traces == 0 || tr.scale, PSKTM_E_INVAL,
                "16-bit traces need per-trace scales");
    PSKTM_CHECK(tr.gained == 0 || tr.gained == 1, PSKTM_E_INVAL, "gained must be 0 or 1");
    const psktm_grid &g = p->grid;
    const int tb = p->tile_x * p->tile_y;
    const bool v3 = vel.kind == PSKTM_VEL_3D;
    const bool kh = p->opt.precise != 0;
    const bool gn = tr.gained != 0;
    // gained FP16 traces: half pairs in LDS (half the bytes per staged row);
    // precise mode stages FP32 pairs
    const bool hp = gn && !kh && tr.dtype == PSKTM_F16;
    PSKTM_CHECK(!hp || ((uintptr_t)tr.data & 1u) == 0, PSKTM_E_INVAL,
                "FP16 trace data must be 2-byte aligned");
    FwdLaunch L{};
    if ((s = choose_launch(g.nt, p->n_samples, tb, v3, kh, hp, p->opt.traces_per_stage,
                           p->opt.row_pad,
                           [&](int nth, int tl) { return (const void *)fwd_kernel(nth, tb, tl, v3, kh, gn, hp); },
                           &L)) != PSKTM_OK)
        return s;
    const FwdShape sh = L.shape;
    const FwdFn fn = (FwdFn)L.fn;
    const int G = L.G, nsp = L.nsp;
    if ((s = prepare_tables(p, vel, st)) != PSKTM_OK) return s;

    FwdArgs a{};
    fwd_args_from_plan(p, a);
    a.n_chunks = sh.chunks;
    a.dtype = tr.dtype;
    a.gained = tr.gained;
    if (tr.gained) a.wscale = 1.0f;
    a.G = G;
    a.nsp = nsp;
    a.lim = p->opt.row_pad < 0 ? 0 : pair_limit(nsp);
    a.accumulate = accumulate;
    a.traces = tr.data; a.scale = tr.scale; a.slow = p->slow; a.tsq = p->tsq; a.image = image;
    const size_t lds = L.lds;
    { psktm_status s_; if ((s_ = ensure_dynamic_lds((const void *)fn, lds)) != PSKTM_OK) return s_; }

    // Split-K: several workgroups share each tile's trace list and write
    // partial images that reduce_partials_kernel adds in split order. When
    // the image has too few (tile, tau pass) work units to fill the GPU twice
    // over, enough splits to do so; when it has fewer than ~40 rounds of
    // resident workgroups, up to 8 splits so that the last round (the
    // smallest tiles, launched last) is short: the 1e6-trace 3-D workload
    // (5,000 tiles, 768 resident workgroups) ran 2.6% faster with 6-7 splits.
    const long long base = (long long)p->n_tiles * sh.chunks;
    int n_split = p->opt.splits;
    if (n_split == 0) {
        int cus = 0, per_cu = 0;
        PSKTM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
        PSKTM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn,
                                                               sh.nth, lds));
        const long long slots = (long long)cus * std::max(per_cu, 1);
        const double per_tile = p->all_traces ? (double)p->n_traces
                                              : (double)p->list_len / std::max(p->n_tiles, 1);
        const long long by_len = (long long)(per_tile / (4.0 * G));   // >= 4 stages each
        n_split = 1;
        if (base > 0 && base < 2 * slots) {
            const long long want = (2 * slots + base - 1) / base;
            n_split = (int)std::max(1ll, std::min({want, by_len, 64ll}));
        } else if (base > 0 && base < 40 * slots) {
            const long long want = (40 * slots + base - 1) / base;
            n_split = (int)std::max(1ll, std::min({want, by_len, 8ll}));
        }
    }
    a.n_split = n_split;
    a.image_elems = (long long)g.nx * g.ny * g.nt;
    if (n_split > 1) {
        const size_t need = (size_t)n_split * (size_t)a.image_elems * (kh ? 2 : 1);
        if (p->partial_cap < need) {
            if (p->partial) {
                PSKTM_HIP(hipStreamSynchronize(st));
                (void)hipFree(p->partial);
                p->bytes -= p->partial_cap * sizeof(float);
                p->partial = nullptr;
                p->partial_cap = 0;
            }
            PSKTM_HIP(hipMalloc(&p->partial, need * sizeof(float)));
            p->partial_cap = need;
            p->bytes += need * sizeof(float);
        }
        a.partial = p->partial;
    }
    const long long blocks = base * n_split;
    PSKTM_CHECK(blocks < (1ll << 31), PSKTM_E_UNSUPPORTED, "too many workgroups (%lld)", blocks);
    if (blocks > 0) {
        hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(sh.nth), lds, st, a);
        if ((s = check_launch("migrate_fwd_kernel")) != PSKTM_OK) return s;
    }
    if (n_split > 1) {
        const long long n = a.image_elems;
        hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           st, (const float *)p->partial, n_split, n, accumulate, kh ? 1 : 0, image);
        return check_launch("reduce_partials_kernel");
    }
    return PSKTM_OK;
}

psktm_status launch_demigrate(psktm_plan *p, const float *image, const psktm_velocity &vel,
                              float *traces_out, int accumulate, int gained, hipStream_t st)
{
    psktm_status s;
    if ((s = validate_velocity(vel)) != PSKTM_OK) return s;
    const psktm_grid &g = p->grid;
    const int tb = p->tile_x * p->tile_y;
    const bool v3 = vel.kind == PSKTM_VEL_3D;
    const int nsa = adj_row(p->n_samples);
    // 64-bit fixed-point accumulators: on gfx950 ds_add_u64 runs this
    // scatter's patterns at 2.3-2.8e12/s, ds_add_f64 at 1.2-2.7e12/s and
    // ds_add_f32 at 2.0e11/s (bench.py --workload rates)
    const size_t lds = (size_t)p->adj_batch * nsa * sizeof(unsigned long long);
    PSKTM_CHECK(lds <= 150 * 1024, PSKTM_E_UNSUPPORTED,
                "traces of %d samples do not fit in LDS for the adjoint", p->n_samples);
    if ((s = prepare_tables(p, vel, st)) != PSKTM_OK) return s;
    const size_t ntot = (size_t)p->n_traces * p->n_samples;
    if (!accumulate && ntot)
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st,
                           traces_out, ntot);
    if (p->n_adj_items == 0) return check_launch("zero_kernel");
    // the fixed-point scale: max |image * tau * sqrt(dt)|, stream-ordered
    const long long nimg = (long long)g.nx * g.ny * g.nt;
    // weight tau*sqrt(dt) per update (times sqrt(x)), or tau alone for the
    // gained operator (its sqrt(n dt) acts on the output samples)
    const float wscale = gained ? 1.0f : (float)std::sqrt((double)p->dt);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, (float *)p->adj_wmax, (size_t)1);
    hipLaunchKernelGGL(adj_max_kernel,
                       dim3((unsigned)std::min<long long>((nimg + 255) / 256, 2048)), dim3(256),
                       0, st, image, nimg, g.nt, g.tau0, g.dtau, wscale, p->adj_wmax);
    if ((s = check_launch("adj_max_kernel")) != PSKTM_OK) return s;

    // 256 lanes x 8 consecutive tau samples, one trace per work item: eight
    // workgroups (32 waves) per CU on 2,000-sample traces. On the 512-CMP
    // line: 31.0 ms; two traces per item 31.8-32.1 ms, 512 x 4 with four
    // traces 36.2 ms, 128 x 16 with one trace 32.6 ms.
    constexpr int TL = 8, NTH = 256;
    // Without an aperture every work item visits every tile, so the adjoint
    // may take tiles of its own: 8-bin plan tiles are doubled along their
    // long axis (half the tile visits and barriers; 48.8 vs 49.7 ms on a
    // 512-CMP line, same output bits there).
    int atx = p->tile_x, aty = p->tile_y;
    if (p->all_traces && tb == 8) {
        if (atx >= aty) atx *= 2; else aty *= 2;
    }
    const int atb = atx * aty;
    const bool lin = p->opt.interp == PSKTM_INTERP_LINEAR;
    AdjFn fn;
#define PSKTM_ADJ_PICK(TBV)                                                                  \
    (gained ? (v3 ? (lin ? demigrate_kernel<NTH, TBV, TL, true, true, true>                  \
                         : demigrate_kernel<NTH, TBV, TL, true, false, true>)                \
                  : (lin ? demigrate_kernel<NTH, TBV, TL, false, true, true>                 \
                         : demigrate_kernel<NTH, TBV, TL, false, false, true>))              \
            : (v3 ? (lin ? demigrate_kernel<NTH, TBV, TL, true, true, false>                 \
                         : demigrate_kernel<NTH, TBV, TL, true, false, false>)               \
                  : (lin ? demigrate_kernel<NTH, TBV, TL, false, true, false>                \
                         : demigrate_kernel<NTH, TBV, TL, false, false, false>)))
    fn = atb == 8 ? PSKTM_ADJ_PICK(8) : PSKTM_ADJ_PICK(16);
#undef PSKTM_ADJ_PICK
    AdjArgs a{};
    a.nx = g.nx; a.ny = g.ny; a.nt = g.nt; a.ns = p->n_samples;
    a.x0 = g.x0; a.dx = g.dx; a.y0 = g.y0; a.dy = g.dy; a.tau0 = g.tau0; a.dtau = g.dtau;
    a.tile_x = atx; a.tile_y = aty;
    a.tiles_x = (g.nx + atx - 1) / atx; a.tiles_y = (g.ny + aty - 1) / aty;
    a.n_chunks = (g.nt + NTH * TL - 1) / (NTH * TL);
    a.cells_y = p->cells_y; a.cell_mx = p->cell_mx; a.cell_my = p->cell_my;
    a.win_x = p->win_x; a.win_y = p->win_y;
    a.nsf = (float)p->n_samples;
    a.r2 = p->opt.aperture * p->opt.aperture;
    a.wscale = wscale;
    // |contribution| <= sqrt(x) * max|image * w| with x < ns, and a slot sums
    // at most every update of its trace (<= nx * ny * nt)
    a.bound = (float)(1.01 * std::sqrt((double)p->n_samples) * (double)g.nx * g.ny * g.nt);
    // a merged contribution adds at most TL of them
    a.cbound = (float)(1.01 * std::sqrt((double)p->n_samples) * TL);
    a.gain_dt = gained ? p->dt : 0.0f;
    a.wmax = p->adj_wmax;
    a.use_ap = !p->all_traces;
    a.interp = p->opt.interp;
    a.nsa = nsa;
    a.batch = p->adj_batch;
    a.geo = p->geo; a.sorted_ids = p->sorted_ids; a.cell_start = p->cell_start;
    a.items = p->adj_items; a.slow = p->slow; a.tsq = p->tsq; a.image = image; a.traces = traces_out;
    { psktm_status s_; if ((s_ = ensure_dynamic_lds((const void *)fn, lds)) != PSKTM_OK) return s_; }
    hipLaunchKernelGGL(fn, dim3(p->n_adj_items), dim3(NTH), lds, st, a);
    return check_launch("demigrate_kernel");
}

} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_migrate(psktm_plan *plan, const psktm_traces *traces,
                           const psktm_velocity *vel, float *image, int accumulate,
                           void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(plan && traces && vel && image, PSKTM_E_INVAL, "NULL argument");
    return launch_migrate(plan, *traces, *vel, image, accumulate, as_stream(stream));
}

psktm_status psktm_demigrate(psktm_plan *plan, const float *image, const psktm_velocity *vel,
                             float *traces_out, int accumulate, void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(plan && image && vel && (traces_out || plan->n_traces == 0), PSKTM_E_INVAL,
                "NULL argument");
    return launch_demigrate(plan, image, *vel, traces_out, accumulate, 0, as_stream(stream));
}

psktm_status psktm_demigrate_gained(psktm_plan *plan, const float *image,
                                    const psktm_velocity *vel, float *traces_out, int accumulate,
                                    void *stream)
{
    psktm__clear_error();
    PSKTM_CHECK(plan && image && vel && (traces_out || plan->n_traces == 0), PSKTM_E_INVAL,
                "NULL argument");
    return launch_demigrate(plan, image, *vel, traces_out, accumulate, 1, as_stream(stream));
}

} // extern "C"
