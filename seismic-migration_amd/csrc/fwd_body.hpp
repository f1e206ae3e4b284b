// Forward migration tile body shared by migrate.hip (one plan per launch)
// and batch.hip (persistent launch over many gathers).
#pragma once

#include "internal.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

namespace psktm {

constexpr int kGMax = 8;      // traces per forward stage (upper bound)
constexpr int kTBMax = 16;

struct FwdArgs {
    int nx, ny, nt, ns;
    float x0, dx, y0, dy, tau0, dtau;
    int tile_x, tile_y, tiles_y, n_chunks;
    float nsf, r2, wscale;
    int use_ap, all_traces, ntr, interp, dtype, G, nsp, accumulate;
    int n_split;                  // workgroups per (tile, tau pass); > 1 writes partials
    float *partial;               // [n_split][image] when n_split > 1
    long long image_elems;
    int gained;                   // traces carry sqrt(n dt): no per-update weight
    int lim;                      // pair-table entries from n = 0 in every staged row
    const float *reach;           // plan's bound on |s - x| and |r - x| (device scalar)
    const float4 *geo;
    const int *list;
    const long long *tile_off;
    const int *tile_order;
    const void *traces;
    const float *scale;
    const float *slow;
    const float *tsq;
    float *image;
};


// LDS pair table of one staged trace: entry n holds (d[n], d[n+1]-d[n]) for
// LINEAR (valid n in [0, ns-2]) or (d[n], 0) for TRUNC (valid n in [1, ns-1]),
// and (0, 0) elsewhere; d = stored sample x per-trace scale. A row keeps
// kPairFront entries in front of n = 0 and enough behind n = ns that the
// 16-byte chunk writes of the fill never leave the row (the fill writes
// entries -7 .. ns+7), plus `pad` zero entries that no fill overwrites.
constexpr int kPairFront = 8;
__host__ __device__ inline int pair_stride(int ns, int pad = 0)
{
    return (ns + pad + 2 * kPairFront + 1) & ~1;
}

// Row stride of the half pair table (FP16 traces, gained form): entry n is
// one dword, (h[n]/2, h[n+1]/2 - h[n]/2) as two FP16 values (low, high half),
// with twice the per-trace scale applied in the accumulate. Halving keeps the
// difference of any two finite FP16 values finite (|h| <= 65504, so the
// halved difference is at most 65504). Rows are a multiple of 4 entries so
// that every row starts 16-byte aligned.
__host__ __device__ inline int half_stride(int ns, int pad = 0)
{
    return (ns + pad + 2 * kPairFront + 3) & ~3;
}

__host__ __device__ inline int row_stride(int ns, int pad, bool hp)
{
    return hp ? half_stride(ns, pad) : pair_stride(ns, pad);
}

__host__ __device__ inline int entry_bytes(bool hp) { return hp ? 4 : 8; }

// Entries from n = 0 to the end of a row of stride nsp.
__host__ __device__ inline int pair_limit(int nsp) { return nsp - kPairFront; }

__device__ __forceinline__ float decode16(unsigned bits, int dtype, float sc)
{
    if (dtype == PSKTM_F16) return __half2float(__ushort_as_half((unsigned short)bits)) * sc;
    return __uint_as_float(bits << 16) * sc;    // BF16
}

// Stages the trace rows s .. s+G-1 of the tile list into the LDS pair table.
// Work is split into (trace, group of 64 chunks) items, one item per wave at
// a time, so the row address is wave-uniform and each row is read through a
// buffer descriptor whose range ends at the dword holding the row's last byte.
// Each lane loads one 16-byte chunk (8 FP16/BF16 or 4 FP32 samples, from a
// 16-byte aligned base) and the first sample of the next chunk, then writes one
// pair per sample. Dwords beyond the range read as zeros; a chunk whose start
// lies inside it stays inside one 16-byte block that holds a byte of the row,
// so no access leaves the pages of the trace buffer. Samples
// outside [0, ns] land in the row's padding and are never read.
template <int NTH>
__device__ __forceinline__ void fill_stage(const FwdArgs &a, long long s, long long s1,
                                           float2 *__restrict__ s_pairs)
{
    // the pair table is (fl(h * sc), fl(d[n+1] - d[n])) exactly, whatever the
    // surrounding code: no fused multiply-subtract
#pragma clang fp contract(off)
    constexpr int W = NTH / 64;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int E = a.dtype == PSKTM_F32 ? 4 : 2;
    const int per = 16 / E;
    const int G = (int)((s1 - s) < (long long)a.G ? (s1 - s) : (long long)a.G);
    const int groups = ((14 + a.ns * E) / 16 + 1 + 63) / 64;
    const bool lin = a.interp == PSKTM_INTERP_LINEAR;
    for (int it = wave; it < G * groups; it += W) {
        const int g = it / groups, cg = it - g * groups;
        const long long q = s + g;
        const int k = a.all_traces ? (int)q : a.list[q];
        const uintptr_t ra = (uintptr_t)a.traces + (size_t)k * a.ns * E;
        const int sb = (int)(ra & 15u);
        const int n_chunks = (sb + a.ns * E) / 16 + 1;
        const int c = cg * 64 + lane;
        if (c >= n_chunks) continue;
        const float sc = a.dtype == PSKTM_F32 ? 1.f : a.scale[k];
        // range rounded up to whole dwords: the bounds check is per dword, and
        // a row of 16-bit samples may end in the middle of one
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(ra & ~(uintptr_t)15), (short)0, (sb + a.ns * E + 3) & ~3, 0x00020000);
        const int n0 = (16 * c - sb) / E;           // first sample of the chunk
        float2 *P = s_pairs + g * a.nsp + kPairFront + n0;
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, 16 * c, 0, 0);
        const unsigned nx = E == 4 ? __builtin_amdgcn_raw_buffer_load_b32(rs, 16 * c + 16, 0, 0)
                                   : __builtin_amdgcn_raw_buffer_load_b16(rs, 16 * c + 16, 0, 0);
        // sample i of the chunk (i = per is the next chunk's first sample)
        auto sample = [&](int i) -> float {
            if (E == 4) return __uint_as_float(i < 4 ? v[i] : nx);
            const unsigned w = i < 8 ? ((v[i >> 1] >> (16 * (i & 1))) & 0xffffu) : nx;
            return decode16(w, a.dtype, sc);
        };
        float cur = sample(0);
        if ((n0 & 1) == 0) {
            // entry n0 starts a 16-byte boundary: two pairs per ds_write_b128
            // (half as many bank-conflicted writes as one pair per write)
#pragma unroll
            for (int i = 0; i < 8; i += 2) {
                if (i >= per) break;
                const float n1 = sample(i + 1), n2 = sample(i + 2);
                const int n = n0 + i;
                const bool ok0 = lin ? (n >= 0 && n <= a.ns - 2) : (n >= 1 && n <= a.ns - 1);
                const bool ok1 = lin ? (n + 1 >= 0 && n + 1 <= a.ns - 2)
                                     : (n + 1 >= 1 && n + 1 <= a.ns - 1);
                const float2 p0 = ok0 ? make_float2(cur, lin ? n1 - cur : 0.f) : make_float2(0.f, 0.f);
                const float2 p1 = ok1 ? make_float2(n1, lin ? n2 - n1 : 0.f) : make_float2(0.f, 0.f);
                reinterpret_cast<float4 *>(P)[i >> 1] = make_float4(p0.x, p0.y, p1.x, p1.y);
                cur = n2;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i >= per) break;
                const float nxt = sample(i + 1);
                const int n = n0 + i;
                const bool ok = lin ? (n >= 0 && n <= a.ns - 2) : (n >= 1 && n <= a.ns - 1);
                P[i] = ok ? make_float2(cur, lin ? nxt - cur : 0.f) : make_float2(0.f, 0.f);
                cur = nxt;
            }
        }
    }
}

// Stages rows s .. s+G-1 of FP16 traces as half pairs, one trace per wave at
// a time, with the chunks and buffer ranges of fill_stage: each lane writes
// the 8 entries of its chunk as dwords (u[n] | (u[n+1] - u[n]) << 16 with
// u = h/2, exact for every normal FP16 value, and one FP16 rounding of the
// difference), zero outside the valid range. Without the halving, neighbours such as +32768 and
// -32752 (both valid quantised samples) would difference to -inf. Rows start
// 16-byte aligned, so a chunk whose first entry is a multiple of 4 writes two
// 16-byte stores.
template <int NTH, bool PF>
__device__ __forceinline__ void fill_half_rows(const FwdArgs &a, long long s, int G,
                                               unsigned *__restrict__ s_half, int k_first)
{
    constexpr int W = NTH / 64;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const bool lin = a.interp == PSKTM_INTERP_LINEAR;
    const int lo = lin ? 0 : 1, hi = lin ? a.ns - 2 : a.ns - 1;
    typedef _Float16 h2v __attribute__((ext_vector_type(2)));
    const h2v half2 = {(_Float16)0.5f, (_Float16)0.5f};
    const unsigned lm = lin ? 0xffffu : 0u;     // TRUNC: no difference term
    // one trace per wave at a time: its row set-up once, then its chunks
    for (int g = wave; g < G; g += W) {
        const long long q = s + g;
        // PF: the wave's first trace of the stage comes loaded one stage ahead
        const int k = (PF && g == wave) ? k_first : (a.all_traces ? (int)q : a.list[q]);
        const uintptr_t ra = (uintptr_t)a.traces + (size_t)k * a.ns * 2;
        const int sb = (int)(ra & 15u);
        const int n_chunks = (sb + a.ns * 2) / 16 + 1;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(ra & ~(uintptr_t)15), (short)0, (sb + a.ns * 2 + 3) & ~3, 0x00020000);
        unsigned *row = s_half + g * a.nsp + kPairFront;
        // 16-byte aligned entries (the same for every chunk of the row)
        const bool fast = lin && (sb & 6) == 0;
        for (int c = lane; c < n_chunks; c += 64) {
            const int n0 = (16 * c - sb) >> 1;      // rows start on a whole sample
            unsigned *P = row + n0;
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, 16 * c, 0, 0);
            const unsigned nx = __builtin_amdgcn_raw_buffer_load_b16(rs, 16 * c + 16, 0, 0);
            // entry i from the sample pair (h[n0+i], h[n0+i+1]) as one dword
            // (a funnel shift for odd i): both halved by one packed FP16
            // multiply with an inline constant, then (u0, u1 - u0)
            auto entry = [&](int i) -> unsigned {
                const unsigned pr = (i & 1) ? __builtin_amdgcn_alignbit(i < 7 ? v[(i + 1) >> 1] : nx,
                                                                        v[i >> 1], 16)
                                            : v[i >> 1];
                const h2v uu = __builtin_bit_cast(h2v, pr) * half2;
                const unsigned dd = (unsigned)__builtin_bit_cast(unsigned short, (_Float16)(uu.y - uu.x));
                return (unsigned)__builtin_bit_cast(unsigned short, uu.x) | ((dd & lm) << 16);
            };
            // whole waves of interior chunks: no range tests
            const bool inner = n0 >= lo && n0 + 7 <= hi;
            // LINEAR on aligned rows: halve the chunk (and the next chunk's
            // first sample) by packed FP16 multiplies, then make each entry by
            // one packed FMA of halved samples a = u[n+1], c = u[n]:
            //   lo = a * 0 + c = u[n],  hi = a * 1 - c = fp16(u[n+1] - u[n])
            // (the same roundings as entry(): an exact halving, one rounding
            // of the difference)
            auto halve = [&](int k) -> unsigned {
                return __builtin_bit_cast(unsigned, __builtin_bit_cast(h2v, k < 4 ? v[k] : nx) * half2);
            };
#if defined(__HIP_DEVICE_COMPILE__)
            const unsigned zero_one = 0x3c000000u;          // (0.0, 1.0) as FP16 (lo, hi)
            // even entry: (u.lo, u.hi - u.lo)
            auto even = [&](unsigned u) -> unsigned {
                unsigned r;
                asm("v_pk_fma_f16 %0, %1, %2, %1 op_sel_hi:[1,1,0] neg_hi:[0,0,1]"
                    : "=v"(r) : "v"(u), "s"(zero_one));
                return r;
            };
            // odd entry: (u.hi, w.lo - u.hi), w the next dword
            auto odd = [&](unsigned u, unsigned w) -> unsigned {
                unsigned r;
                asm("v_pk_fma_f16 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,1,1] neg_hi:[0,0,1]"
                    : "=v"(r) : "v"(w), "s"(zero_one), "v"(u));
                return r;
            };
#else
            auto even = [&](unsigned) -> unsigned { return 0u; };
            auto odd = [&](unsigned, unsigned) -> unsigned { return 0u; };
#endif
            if (fast && __builtin_amdgcn_ballot_w64(!inner) == 0) {
#if defined(__HIP_DEVICE_COMPILE__)
                // the first four entries are stored before the rest are made
                const unsigned u0 = halve(0), u1 = halve(1), u2 = halve(2);
                reinterpret_cast<uint4 *>(P)[0] = make_uint4(even(u0), odd(u0, u1), even(u1), odd(u1, u2));
                const unsigned u3 = halve(3), u4 = halve(4);
                reinterpret_cast<uint4 *>(P)[1] = make_uint4(even(u2), odd(u2, u3), even(u3), odd(u3, u4));
#else
                reinterpret_cast<uint4 *>(P)[0] = make_uint4(entry(0), entry(1), entry(2), entry(3));
                reinterpret_cast<uint4 *>(P)[1] = make_uint4(entry(4), entry(5), entry(6), entry(7));
#endif
            } else if (fast) {
                // waves holding a row end: the same entries one at a time,
                // zero outside [lo, hi]
                auto put = [&](int i, unsigned e) {
                    P[i] = (unsigned)(n0 + i - lo) <= (unsigned)(hi - lo) ? e : 0u;
                };
#if defined(__HIP_DEVICE_COMPILE__)
                unsigned u = halve(0);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const unsigned w = halve(k + 1);
                    put(2 * k, even(u));
                    put(2 * k + 1, odd(u, w));
                    u = w;
                }
#else
#pragma unroll
                for (int i = 0; i < 8; ++i) put(i, entry(i));
#endif
            } else {
                // unaligned rows and TRUNC: entry by entry
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const bool ok = (unsigned)(n0 + i - lo) <= (unsigned)(hi - lo);
                    P[i] = ok ? entry(i) : 0u;
                }
            }
        }
    }
}

// Distance terms and aperture bit of (trace geometry ge, bin (i, j)).
__device__ __forceinline__ bool pair_terms(float4 ge, int i, int j, float x0, float dx,
                                           float y0, float dy, int use_ap, float r2,
                                           float2 *ab)
{
    const float x = axis_f32(x0, dx, i), y = axis_f32(y0, dy, j);
    const bool m = !use_ap || in_aperture_f32(ge.x, ge.y, x, y, r2);
    const float ex = ge.x - x, ey = ge.y - y;
    const float ax = ex - ge.z, ay = ey - ge.w, bx = ex + ge.z, by = ey + ge.w;
    *ab = make_float2(fmaf(ax, ax, ay * ay), fmaf(bx, bx, by * by));
    return m;
}

// One output tile (T, tau pass `chunk`) of a migration: the whole forward
// algorithm described in migrate.hip. GN: gained traces (sample n carries
// sqrt(n dt)), accumulated without the per-update sqrt(t) weight; a template
// parameter so that each form keeps its own register allocation. PF (half
// pairs only): each stage's list entries are loaded one stage ahead. Shared
// by the per-plan kernel and the persistent batch kernel; the caller
// provides the LDS arrays.
template <int NTH, int TB, int TL, bool V3, bool KH, bool GN, bool HP = false, bool PF = false>
__device__ __forceinline__ void fwd_tile(const FwdArgs &a, const int T, const int chunk,
                                         const int split,
                                         float2 *__restrict__ s_pairs,
                                         float2 (*__restrict__ s_ab)[TB],
                                         unsigned *__restrict__ s_mask)
{
    const int tid = threadIdx.x;
    const int TI = T / a.tiles_y, TJ = T % a.tiles_y;
    const int i0 = TI * a.tile_x, j0 = TJ * a.tile_y;
    const int lbase = chunk * (NTH * TL) + tid;

    float tq[TL];
    float sv[V3 ? TB : 1][TL];
    float acc[TB][TL];
    float cmp[KH ? TB : 1][KH ? TL : 1];   // Kahan compensation (precise mode)
#pragma unroll
    for (int j = 0; j < TL; ++j) {
        const int l = lbase + j * NTH;
        tq[j] = (l < a.nt) ? a.tsq[l] : 0.f;
        if (!V3) sv[0][j] = (l < a.nt) ? a.slow[l] : 0.f;
    }
#pragma unroll
    for (int b = 0; b < TB; ++b) {
        const int i = i0 + b / a.tile_y, jj = j0 + b % a.tile_y;
        const bool bin_ok = (b / a.tile_y) < a.tile_x && i < a.nx && jj < a.ny;
#pragma unroll
        for (int j = 0; j < TL; ++j) {
            acc[b][j] = 0.f;
            if (KH) cmp[KH ? b : 0][KH ? j : 0] = 0.f;
            if (V3) {
                const int l = lbase + j * NTH;
                sv[V3 ? b : 0][j] = (bin_ok && l < a.nt)
                    ? a.slow[((size_t)i * a.ny + jj) * a.nt + l] : 0.f;
            }
        }
    }

    // The pair table is cleared once per tile: entries past what the fill
    // writes (-7 .. ns+7 of each row) then read as zeros for the whole tile.
    if (HP) {
        unsigned *h = reinterpret_cast<unsigned *>(s_pairs);
        for (int q = tid; q < a.G * a.nsp; q += NTH) h[q] = 0u;
    } else {
        for (int q = tid; q < a.G * a.nsp; q += NTH) s_pairs[q] = make_float2(0.f, 0.f);
    }
    // Traveltime bound of the tile: every (trace, in-aperture bin) pair has
    // A, B <= reach^2, so x <= 2 sqrt(tau'^2 + reach^2 s') at each tau. When
    // that stays inside the padded rows (with a margin far above the FP32
    // rounding of A, B and v_sqrt_f32), the index needs no clamp: samples
    // at and past ns read zero pairs either way, so the image is the same.
    // Every comparison is written so that a NaN (from a NaN coordinate, a
    // NaN velocity, or inf * 0) fails it, so the unclamped loop only ever
    // sees finite times inside the rows (fmaxf would drop a NaN operand).
    bool unclamped = false;
    {
        float xb = 0.f;
        bool finite = true;
        if (a.lim > a.ns) {
            const float r = *a.reach;
            const float r2 = r * r;
#pragma unroll
            for (int j = 0; j < TL; ++j) {
                float sm = sv[0][j];
                if (V3) {
#pragma unroll
                    for (int b = 1; b < TB; ++b) {
                        const float v = sv[V3 ? b : 0][j];
                        finite = finite && (v == v);
                        sm = v > sm ? v : sm;
                    }
                }
                const float v = sqrtf(fmaf(r2, sm, tq[j]));
                finite = finite && (v == v);
                xb = v > xb ? v : xb;
            }
        }
        const bool ok = finite && a.lim > a.ns && 2.f * xb * 1.0001f + 2.f < (float)a.lim;
        unclamped = !__syncthreads_or(!ok);
    }

    // this workgroup's part of the tile's trace list (split `split` of n_split)
    const long long l0 = a.all_traces ? 0 : a.tile_off[T];
    const long long l1 = a.all_traces ? a.ntr : a.tile_off[T + 1];
    const long long s0 = l0 + (l1 - l0) * split / a.n_split;
    const long long s1 = l0 + (l1 - l0) * (split + 1) / a.n_split;
    const int G = a.G;
    // PF: list entries are loaded one stage ahead, during the previous
    // stage's compute: the distance-term lane's trace (kp, lane tid < G*TB:
    // trace tid / TB) and the fill's first trace of each wave (kf, wave w:
    // trace w). The stage's set-up then waits for one load (geometry, rows)
    // instead of two in a row. Only the per-plan half-pair kernel, which
    // caps its registers, takes it: the live entries cost the FP32-pair
    // forms a wave per SIMD, and the uncapped batch kernel 7-10 VGPRs.
    const int wave_u = __builtin_amdgcn_readfirstlane(tid >> 6);
    auto list_at = [&](long long q) -> int {
        return q < s1 ? (a.all_traces ? (int)q : a.list[q]) : 0;
    };
    static_assert(!PF || HP, "the stage look-ahead is for half pairs");
    int kp = (PF && tid < G * TB) ? list_at(s0 + tid / TB) : 0;
    int kf = PF ? list_at(s0 + wave_u) : 0;
    auto stages = [&](auto clamp_tag) {
    constexpr bool CL = decltype(clamp_tag)::value;
    for (long long s = s0; s < s1; s += G) {
        // ---- fill: distance terms + aperture bits, then trace pairs -------
        if (tid < G * TB) {
            const int g = tid / TB, b = tid % TB;
            const long long q = s + g;
            float2 ab = make_float2(0.f, 0.f);
            bool m = false;
            if (q < s1) {
                const int k = PF ? kp : (a.all_traces ? (int)q : a.list[q]);
                const int bi = b / a.tile_y, bj = b % a.tile_y;
                const int i = i0 + bi, jj = j0 + bj;
                if (bi < a.tile_x && i < a.nx && jj < a.ny)
                    m = pair_terms(a.geo[k], i, jj, a.x0, a.dx, a.y0, a.dy, a.use_ap, a.r2, &ab);
            }
            s_ab[g][b] = ab;
            const unsigned long long bal = __ballot(m);
            if ((tid % TB) == 0) {
                const int shift = (tid & 63) - ((tid & 63) % TB);
                s_mask[g] = (unsigned)((bal >> shift) & ((1ull << TB) - 1ull));
                // half pairs hold half the stored values: twice the trace's
                // scale (exact) restores them in the accumulate
                if constexpr (PF) {
                    s_mask[kGMax + g] = __float_as_uint(q < s1 ? 2.f * a.scale[kp] : 0.f);
                } else if (HP) {
                    const int k = q < s1 ? (a.all_traces ? (int)q : a.list[q]) : 0;
                    s_mask[kGMax + g] = __float_as_uint(q < s1 ? 2.f * a.scale[k] : 0.f);
                }
            }
        }
        if (HP)
            fill_half_rows<NTH, PF>(a, s, (int)((s1 - s) < (long long)a.G ? (s1 - s) : (long long)a.G),
                                reinterpret_cast<unsigned *>(s_pairs), kf);
        else
            fill_stage<NTH>(a, s, s1, s_pairs);
        __syncthreads();
        if (PF && tid < G * TB) kp = list_at(s + G + tid / TB);
        if (PF) kf = list_at(s + G + wave_u);
        // ---- compute --------------------------------------------------------
        // traces staged in this pass (a 32-bit scalar bound: the 64-bit test
        // s + g < s1 would cost two VALU instructions per trace)
        const int gn = (int)((s1 - s) < (long long)G ? (s1 - s) : (long long)G);
        for (int g = 0; g < gn; ++g) {
            // the trace's scale (read with its mask) stays in a VGPR: every
            // lane reads the same LDS word, and an accumulate FMA with no
            // SGPR operand can dual-issue
            unsigned mv = s_mask[g];
            float scv = HP ? __uint_as_float(s_mask[kGMax + g]) : 0.f;
#if defined(__HIP_DEVICE_COMPILE__)
            asm volatile("" : "+v"(mv), "+v"(scv));
#endif
            const unsigned m = __builtin_amdgcn_readfirstlane(mv);
            if (!m) continue;
            const float2 *P = s_pairs + g * a.nsp + kPairFront;
            const unsigned *PH = reinterpret_cast<const unsigned *>(s_pairs) + g * a.nsp + kPairFront;
#if defined(__HIP_DEVICE_COMPILE__)
            // LDS address of this trace's distance terms, moved into a VGPR
            // once per trace by an opaque move: the compiler then reads every
            // bin's (A, B) at an immediate offset from it instead of
            // re-materialising the address from the scalar base per bin (one
            // v_mov per (trace, bin), 1.5% of the loop; same image). The host
            // pass never runs kernel bodies and keeps the plain read.
            unsigned abase;
            asm volatile("v_mov_b32 %0, %1" : "=v"(abase)
                         : "s"((unsigned)(uintptr_t)(const void *)&s_ab[g][0]));
            typedef __attribute__((address_space(3))) const float2 lds_float2;
#endif
            if constexpr (KH) {
                // Precise mode with every bin live in this trace and the next
                // (a plan without an aperture cut): both traces in one block
                // with no branch per bin, each Kahan accumulator updated twice
                // in place. After one update in a block that a skipped bin
                // would bypass, the sum must be moved back into its register
                // (one move per update). Same operations, same order.
                constexpr unsigned full = TB >= 32 ? 0xffffffffu : (1u << TB) - 1u;
                if (m == full && g + 1 < gn &&
                    __builtin_amdgcn_readfirstlane(s_mask[g + 1]) == full) {
                    auto kahan = [&](float2 ab, const float2 *Pt, int b, int j) {
                        const float sl = sv[V3 ? b : 0][j];
                        const float ua = fmaf(ab.x, sl, tq[j]);
                        const float ub = fmaf(ab.y, sl, tq[j]);
                        float x = __builtin_amdgcn_sqrtf(ua) + __builtin_amdgcn_sqrtf(ub);
                        if (CL) x = fminf(x, a.nsf);
                        const int n = (int)x;
                        const float f = __builtin_amdgcn_fractf(x);
                        const float2 pr = Pt[n];
                        const float d = fmaf(f, pr.y, pr.x);
                        float &c = cmp[KH ? b : 0][KH ? j : 0];
                        const float y = GN ? d - c : fmaf(d, __builtin_amdgcn_sqrtf(x), -c);
                        const float t = acc[b][j] + y;
                        c = (t - acc[b][j]) - y;
                        acc[b][j] = t;
                    };
                    const float2 *P2 = P + a.nsp;
#pragma unroll
                    for (int b = 0; b < TB; ++b) {
#if defined(__HIP_DEVICE_COMPILE__)
                        const float2 ab0 = *(lds_float2 *)(uintptr_t)(abase + 8u * b);
                        const float2 ab1 = *(lds_float2 *)(uintptr_t)(abase + 8u * (TB + b));
#else
                        const float2 ab0 = s_ab[g][b], ab1 = s_ab[g + 1][b];
#endif
#pragma unroll
                        for (int j = 0; j < TL; ++j) {
                            kahan(ab0, P, b, j);
                            kahan(ab1, P2, b, j);
                        }
                    }
                    ++g;
                    continue;
                }
            }
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                if (!((m >> b) & 1u)) continue;
#if defined(__HIP_DEVICE_COMPILE__)
                const float2 ab = *(lds_float2 *)(uintptr_t)(abase + 8u * b);
#else
                const float2 ab = s_ab[g][b];
#endif
#pragma unroll
                for (int j = 0; j < TL; ++j) {
                    const float sl = sv[V3 ? b : 0][j];
                    const float ua = fmaf(ab.x, sl, tq[j]);
                    const float ub = fmaf(ab.y, sl, tq[j]);
                    float x = __builtin_amdgcn_sqrtf(ua) + __builtin_amdgcn_sqrtf(ub);
                    if (CL) x = fminf(x, a.nsf);
                    const int n = (int)x;
                    const float f = __builtin_amdgcn_fractf(x);
                    float d;
                    if (HP) {
                        const unsigned e = PH[n];
                        d = fmaf(f, __half2float(__ushort_as_half((unsigned short)(e >> 16))),
                                 __half2float(__ushort_as_half((unsigned short)(e & 0xffffu))));
                    } else {
                        const float2 pr = P[n];
                        d = fmaf(f, pr.y, pr.x);
                    }
                    if (HP) {
                        acc[b][j] = fmaf(d, scv, acc[b][j]);
                    } else if (KH) {
                        // gained traces: no per-update weight
                        float &c = cmp[KH ? b : 0][KH ? j : 0];
                        const float y = GN ? d - c : fmaf(d, __builtin_amdgcn_sqrtf(x), -c);
                        const float t = acc[b][j] + y;
                        c = (t - acc[b][j]) - y;
                        acc[b][j] = t;
                    } else if (GN) {
                        acc[b][j] = acc[b][j] + d;
                    } else {
                        acc[b][j] = fmaf(d, __builtin_amdgcn_sqrtf(x), acc[b][j]);
                    }
                }
            }
        }
        __syncthreads();
    }
    };
    if (unclamped)
        stages(std::false_type{});
    else
        stages(std::true_type{});
    // ---- epilogue: weight tau*sqrt(dt) (tau for gained traces), one write ---
#pragma unroll
    for (int b = 0; b < TB; ++b) {
        const int bi = b / a.tile_y, bj = b % a.tile_y;
        const int i = i0 + bi, jj = j0 + bj;
        if (bi >= a.tile_x || i >= a.nx || jj >= a.ny) continue;
#pragma unroll
        for (int j = 0; j < TL; ++j) {
            const int l = lbase + j * NTH;
            if (l >= a.nt) continue;
            const float w = axis_f32(a.tau0, a.dtau, l) * a.wscale;
            const size_t idx = ((size_t)i * a.ny + jj) * a.nt + l;
            const float sum = KH ? acc[b][j] - cmp[KH ? b : 0][KH ? j : 0] : acc[b][j];
            const float v = sum * w;
            if (a.n_split > 1) {
                a.partial[(size_t)split * a.image_elems + idx] = v;
                if (KH) {
                    // low-order part of (acc - cmp) * w: Fast2Sum residual of
                    // the difference (|cmp| <= |acc|) and the exact product
                    // error, so that the compensated reduction of the splits
                    // keeps what each split's compensated sum carried
                    const float lo = (acc[b][j] - sum) - cmp[KH ? b : 0][KH ? j : 0];
                    a.partial[(size_t)(a.n_split + split) * a.image_elems + idx] =
                        fmaf(sum, w, -v) + lo * w;
                }
            } else {
                a.image[idx] = a.accumulate ? a.image[idx] + v : v;
            }
        }
    }
}


// Per-lane register budget for the accumulators, the Kahan compensations and
// the 3-D slowness samples; shapes beyond it are not instantiated.
constexpr int acc_regs(int tb, int tl, bool v3, bool kh)
{
    return tb * tl * ((kh ? 2 : 1) + (v3 ? 1 : 0));
}
constexpr int reg_budget(bool kh) { return kh ? 128 : 96; }

constexpr int kTLSet[] = {2, 3, 4, 6, 8};

struct FwdShape {
    int nth, tl, chunks;
};

// Register-only shape choice (no device query): the first 256- or 512-lane
// shape that covers nt in one pass, else the widest that fits, in passes.
inline FwdShape choose_shape(int nt, int tb, bool v3, bool kh)
{
    auto fits = [&](int tl) { return acc_regs(tb, tl, v3, kh) <= reg_budget(kh); };
    for (int nth : {256, 512}) {
        const int need = (nt + nth - 1) / nth;
        for (int tl : kTLSet)
            if (tl >= need && fits(tl)) return {nth, tl, 1};
    }
    int tl = 2;
    for (int t : kTLSet)
        if (fits(t)) tl = t;
    return {512, tl, (nt + 512 * tl - 1) / (512 * tl)};
}

// Zero entries appended to every staged row (psktm_options.row_pad): -1 none,
// > 0 that many, 0 the largest of ns/4, ns/8, ... (at least 16) whose LDS
// footprint keeps the kernel's resident workgroups per CU.
inline psktm_status choose_row_pad(int opt_pad, int ns, int G, bool hp, const void *fn, int nth,
                                   int *pad)
{
    *pad = 0;
    if (opt_pad < 0) return PSKTM_OK;
    if (opt_pad > 0) { *pad = opt_pad; return PSKTM_OK; }
    const size_t eb = entry_bytes(hp);
    const size_t base_lds = (size_t)G * row_stride(ns, 0, hp) * eb;
    int base = 0;
    PSKTM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&base, fn, nth, base_lds));
    for (int cand = ns / 4; cand >= 16; cand /= 2) {
        const size_t lds = (size_t)G * row_stride(ns, cand, hp) * eb;
        if (lds > 150 * 1024) continue;
        int occ = 0;
        PSKTM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, nth, lds));
        if (occ >= base) { *pad = cand; break; }
    }
    return PSKTM_OK;
}

// Launch configuration of the forward kernel for traces of ns samples.
struct FwdLaunch {
    FwdShape shape;
    const void *fn;
    int G, pad, nsp;
    size_t lds;
};

// Chooses the shape, the traces per stage and the row padding from the
// device's occupancy rules. The kernel hides the latency of its LDS round
// trips and trace loads with other resident waves, so occupancy decides:
//  * shape: among the shapes that cover nt in one pass and fit the register
//    budget, most resident waves per CU, then fewest idle lanes, then more
//    samples per lane (each (trace, bin) costs a distance-term read and a
//    mask test, amortised over the lane's samples: 256 x 4 beat 512 x 2 by
//    11% on the batched gathers). For 8-bin tiles with a 3-D velocity,
//    512 lanes x 3 samples at 79 VGPRs keep three workgroups per CU where
//    256 x 6 keeps one;
//  * G (traces per stage): the largest of 5 .. 1 whose pair table keeps the
//    register-limited workgroups per CU (measured on the 3-D workload: 5 for
//    16-bin tiles at two workgroups per CU, 4 for 8-bin tiles at three);
//  * pad: choose_row_pad at that G.
// `lookup(nth, tl)` returns the kernel of that shape or nullptr. `opt_G` > 0
// and `opt_pad` != 0 override the choices.
template <typename Lookup>
inline psktm_status choose_launch(int nt, int ns, int tb, bool v3, bool kh, bool hp, int opt_G,
                                  int opt_pad, Lookup lookup, FwdLaunch *out)
{
    const size_t eb = entry_bytes(hp);
    const size_t slot = (size_t)row_stride(ns, 0, hp) * eb;
    FwdShape best{0, 0, 0};
    const void *best_fn = nullptr;
    long long best_waves = -1, best_waste = 0;
    for (int nth : {512, 256}) {
        for (int tl : kTLSet) {
            if ((long long)nth * tl < nt || acc_regs(tb, tl, v3, kh) > reg_budget(kh)) continue;
            const void *fn = lookup(nth, tl);
            if (!fn) continue;
            int blocks = 0;
            PSKTM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, nth, slot));
            const long long waves = (long long)blocks * (nth / 64);
            const long long waste = (long long)nth * tl - nt;
            if (waves > best_waves || (waves == best_waves && (waste < best_waste ||
                                       (waste == best_waste && tl > best.tl)))) {
                best = {nth, tl, 1};
                best_fn = fn;
                best_waves = waves;
                best_waste = waste;
            }
        }
    }
    if (!best_fn) {
        best = choose_shape(nt, tb, v3, kh);
        best_fn = lookup(best.nth, best.tl);
    }
    PSKTM_CHECK(best_fn, PSKTM_E_UNSUPPORTED, "no forward kernel for tile %d", tb);
    // occupancy queries and launches below may use more than 64 KiB of LDS
    psktm_status s;
    if ((s = ensure_dynamic_lds(best_fn, 150 * 1024)) != PSKTM_OK) return s;
    int G = opt_G > 0 ? std::min(opt_G, kGMax) : 0;
    if (G == 0) {
        // rows as long as an explicit pad makes them
        const size_t gslot = (size_t)row_stride(ns, std::max(opt_pad, 0), hp) * eb;
        int reg_blocks = 0;
        PSKTM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&reg_blocks, best_fn, best.nth, slot));
        G = 1;
        // half pairs take half the LDS per row: up to kGMax rows per stage
        for (int g = hp ? kGMax : 5; g >= 1; --g) {
            int b = 0;
            if ((size_t)g * gslot > 150 * 1024) continue;
            PSKTM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, best_fn, best.nth, g * gslot));
            if (b >= reg_blocks) { G = g; break; }
        }
    }
    PSKTM_CHECK((size_t)G * slot <= 150 * 1024, PSKTM_E_UNSUPPORTED,
                "traces of %d samples do not fit in LDS", ns);
    int pad = 0;
    if ((s = choose_row_pad(opt_pad, ns, G, hp, best_fn, best.nth, &pad)) != PSKTM_OK) return s;
    const int nsp = row_stride(ns, pad, hp);
    PSKTM_CHECK((size_t)G * nsp * eb <= 150 * 1024, PSKTM_E_UNSUPPORTED,
                "row_pad %d does not fit in LDS", pad);
    *out = FwdLaunch{best, best_fn, G, pad, nsp, (size_t)G * nsp * eb};
    return PSKTM_OK;
}

// Fills the FwdArgs fields that depend only on the plan.
inline void fwd_args_from_plan(const psktm_plan *p, FwdArgs &a)
{
    const psktm_grid &g = p->grid;
    a.nx = g.nx; a.ny = g.ny; a.nt = g.nt; a.ns = p->n_samples;
    a.x0 = g.x0; a.dx = g.dx; a.y0 = g.y0; a.dy = g.dy; a.tau0 = g.tau0; a.dtau = g.dtau;
    a.tile_x = p->tile_x; a.tile_y = p->tile_y; a.tiles_y = p->tiles_y;
    a.nsf = (float)p->n_samples;
    a.r2 = p->opt.aperture * p->opt.aperture;
    a.wscale = (float)std::sqrt((double)p->dt);
    a.use_ap = !p->all_traces;
    a.all_traces = p->all_traces;
    a.ntr = p->n_traces;
    a.interp = p->opt.interp;
    a.geo = p->geo; a.list = p->list; a.tile_off = p->tile_off; a.tile_order = p->tile_order;
    a.reach = p->reach;
}

} // namespace psktm
