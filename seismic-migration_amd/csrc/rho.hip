// Rho filter (scripts/ktmApplyRhoFilter.m:22-58) and per-trace quantisation.
//
// One workgroup filters two traces at once as the real and imaginary parts of
// one complex signal z = a + i b of length M = 2^ceil(log2 m), zero padded
// (:26, :37). The reference's filter H(w) is Hermitian (H(-w) = conj H(w)),
// DC maps to 0 and at the Nyquist bin only Re H survives its real() (:43,
// :50, :57), so with H_nyq replaced by Re H_nyq the product Y = H Z inverts to
// filtered a in the real part and filtered b in the imaginary part: one
// forward and one inverse FFT serve two traces, with no spectrum separation.
// For M = 1024, 2048 and 4096 (traces of 513-4096 samples) the FFT is a
// register-blocked Stockham transform of radix 16, 16 and M/256 (described
// at rho_fast_kernel); other M use a radix-2 Stockham transform ping-ponging
// between two LDS buffers. Twiddles come from a table computed on the host
// in double precision.
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

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <mutex>
#include <vector>

namespace psktm {
namespace {

constexpr int kRhoThreads = 256;
constexpr int kRhoMaxM = 4096;
// resident waves per SIMD asked of the radix-16 kernel (at most 256 VGPRs):
// two 256-lane workgroups per CU, which their LDS (43-68 KB each) allows
constexpr int kRhoWaves = 2;

// Complex product with its two fused multiply-adds written out: the radix-16
// kernel is compiled without contraction so that every output type computes
// the same bits (left to the compiler, contraction differs between the
// instantiations).
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
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

// y * inv rounded to binary32, then to the output type (nearest even). Left
// alone, the compiler fuses the product and the FP16 conversion into one
// mixed-precision FMA (v_fma_mixlo_f16) that rounds once and so can differ on
// ties; an empty asm statement holds the binary32 product in a register.
__device__ __forceinline__ void store_q(void *out, int dtype, size_t idx, float y, float inv)
{
    if (dtype == PSKTM_F32) static_cast<float *>(out)[idx] = y;
    else if (dtype == PSKTM_F16) {
        float p = y * inv;
        asm volatile("" : "+v"(p));
        static_cast<__half *>(out)[idx] = __float2half_rn(p);
    }
    else static_cast<unsigned short *>(out)[idx] = f2bf16_rne(y * inv);
}

// The 16-bit code store_q writes for y (FP16 or BF16), for paired stores.
template <int DT>
__device__ __forceinline__ unsigned q16(float y, float inv)
{
    if (DT == PSKTM_F16) {
        float p = y * inv;
        asm volatile("" : "+v"(p));
        return __half_as_ushort(__float2half_rn(p));
    }
    return f2bf16_rne(y * inv);
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

// The filter at bin k of an M-point transform (scripts/ktmApplyRhoFilter.m:
// 31-58): w = 2 pi k' / Dr with k' = k (k < M/2), k - M (k >= M/2, so the
// Nyquist bin is -M/2); opt 1 |w|, opt 2 sqrt|w| e^{i pi/4 sgn w}, opt 3 i w;
// DC -> 0; at the Nyquist bin the real() keeps Re H only.
__device__ __forceinline__ float2 rho_response(int k, int M, float wscale, int opt)
{
    if (k == 0) return make_float2(0.f, 0.f);
    const int kp = (k < (M >> 1)) ? k : k - M;
    const float w = wscale * (float)kp;
    const float aw = fabsf(w);
    float2 h;
    if (opt == 1) h = make_float2(aw, 0.f);
    else if (opt == 2) {
        const float r = sqrtf(aw) * 0.70710678118654752f;
        h = make_float2(r, w < 0.f ? -r : r);
    } else h = make_float2(0.f, w);
    if (k == (M >> 1)) h.y = 0.f;
    return h;
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
    for (int k = threadIdx.x; k < M; k += blockDim.x)
        other[k] = k == 0 ? make_float2(0.f, 0.f) : cmul(rho_response(k, M, wscale, opt), Z[k]);
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

// ---- register-blocked Stockham FFT for M = 1024, 2048, 4096 -----------------
//
// M = 16 x 16 x R3 (R3 = M/256): three Stockham passes of radix 16, 16 and R3
// (pass p reads x[jj + r M/R], twiddles by W_M^(r k M/(Ns R)) with k = jj mod
// Ns, and writes the R outputs to (jj/Ns) Ns R + k + s Ns; Ns = 1, 16, 256).
// L = M/16 lanes serve one complex FFT (two real traces): one butterfly per
// lane in passes 1-2, 16/R3 in pass 3, so that after the forward transform
// lane j holds X[j + L r] for r = 0..15 in registers. That is exactly the
// input of the first inverse pass, so the filter is applied in registers and
// a forward-inverse pair needs four LDS exchanges (the radix-2 form: 22
// passes, each through LDS behind a barrier). The inverse leaves lane j with
// y[j + L r], stored with consecutive lanes on consecutive samples.

// W_16^e = exp(-2 pi i e / 16) (forward) for e = 0..7; INV conjugates.
template <int E, bool INV>
__device__ __forceinline__ float2 rot16(float2 v)
{
#pragma clang fp contract(off)
    constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f;
    constexpr float c2 = 0.70710678118654752f;
    constexpr float sg = INV ? -1.f : 1.f;      // forward multiplies by cos - i sin
    if constexpr (E == 0) return v;
    else if constexpr (E == 4) return INV ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
    else if constexpr (E == 2 || E == 6) {
        // (cos, -sg sin) with cos = +-c2, sin = c2
        const float c = E == 2 ? c2 : -c2;
        const float sn = sg * c2;
        return make_float2(fmaf(v.x, c, v.y * sn), fmaf(v.y, c, -(v.x * sn)));
    } else {
        const float c = (E == 1) ? c1 : (E == 3) ? s1 : (E == 5) ? -s1 : -c1;
        const float sn = sg * ((E == 1 || E == 7) ? s1 : c1);
        return make_float2(fmaf(v.x, c, v.y * sn), fmaf(v.y, c, -(v.x * sn)));
    }
}

// In-register DFT of R = 4, 8 or 16 points (radix-2 decimation in
// frequency, outputs put back in natural order at compile time).
template <int R, bool INV>
__device__ __forceinline__ void dft_regs(float2 (&v)[16])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int span = R / 2; span >= 1; span >>= 1) {
#pragma unroll
        for (int st = 0; st < R; st += 2 * span) {
#pragma unroll
            for (int i = 0; i < span; ++i) {
                const float2 a = v[st + i], b = v[st + i + span];
                v[st + i] = make_float2(a.x + b.x, a.y + b.y);
                const float2 d = make_float2(a.x - b.x, a.y - b.y);
                // W_{2 span}^i = W_16^(i * 8 / span)
                const int e = i * (8 / span);
                float2 r;
                switch (e) {
                case 0: r = rot16<0, INV>(d); break;
                case 1: r = rot16<1, INV>(d); break;
                case 2: r = rot16<2, INV>(d); break;
                case 3: r = rot16<3, INV>(d); break;
                case 4: r = rot16<4, INV>(d); break;
                case 5: r = rot16<5, INV>(d); break;
                case 6: r = rot16<6, INV>(d); break;
                default: r = rot16<7, INV>(d); break;
                }
                v[st + i + span] = r;
            }
        }
    }
    // bit-reversed -> natural order
    float2 t[16];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int b = 0;
#pragma unroll
        for (int q = 1, r = R >> 1; q < R; q <<= 1, r >>= 1)
            if (k & q) b |= r;
        t[k] = v[b];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = t[k];
}

// Transforms of up to 2048 points keep the filter H[k] (M float2) and the
// per-sample gain (M floats) in LDS tables instead of evaluating them per
// transform (two 256-lane workgroups per CU still fit: 75 KB at M = 2048).
__host__ __device__ constexpr bool rho_tables(int M) { return M <= 2048; }
__host__ __device__ constexpr size_t rho_table_bytes(int M)
{
    return rho_tables(M) ? (size_t)M * (sizeof(float2) + sizeof(float)) : 0;
}

// Twiddles of the radix-16 kernel in the order its lanes read them: for the
// Ns = 16 pass W_M^(r k M/256) at (r - 1) 16 + k (0 < r < 16, k < 16), then
// for the Ns = 256 pass W_M^(r k) at 240 + (r - 1) 256 + k (0 < r < M/256,
// k < 256). Lanes on consecutive k read consecutive entries, free of bank
// conflicts (as W_M^e the Ns = 16 pass's reads fell 16 k dwords apart, on
// two banks).
__host__ __device__ constexpr bool rho_fast_length(int M)
{
    return M == 1024 || M == 2048 || M == 4096;
}
__host__ __device__ constexpr int rho_fast_tw_count(int M) { return 240 + (M / 256 - 1) * 256; }
__host__ __device__ constexpr int rho_fast_tw_index(int Ns, int r, int k)
{
    return (Ns == 16 ? 0 : 240) + (r - 1) * Ns + k;
}

// LDS index with one pad slot per 16 entries (the 16-apart writes of the
// radix-16 passes would otherwise fall on one bank).
__device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }

// One Stockham pass on L lanes (lane j): radix R, span Ns, Q = 16/R
// butterflies per lane (jj = j + L q). Inputs come from regs (FROM_REGS:
// input r is regs[r]) or from LDS; outputs go to LDS, or (TO_REGS) output s
// of butterfly q to regs[q + Q s]. In place (src == dst) every lane reads all
// its inputs, then the workgroup meets at a barrier, then writes.
template <int M, int R, int Ns, bool INV, bool FROM_REGS, bool TO_REGS>
__device__ __forceinline__ void stockham_pass(float2 (&regs)[16], const float2 *src,
                                              float2 *dst, const float2 *__restrict__ tw, int j)
{
#pragma clang fp contract(off)
    constexpr int L = M / 16, Q = 16 / R, STRIDE = M / R;
    static_assert((M / R) == L * Q, "one butterfly per lane and input slot");
    float2 vin[16];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            vin[q * R + r] = FROM_REGS ? regs[r] : src[pidx(j + L * q + r * STRIDE)];
    }
    if (!FROM_REGS && !TO_REGS) __syncthreads();          // in place: reads before writes
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int jj = j + L * q;
        const int k = jj & (Ns - 1);
        float2 v[16];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = vin[q * R + r];
        if constexpr (Ns > 1) {
            static_assert(Ns == 16 || Ns == 256, "twiddle layout covers the Ns = 16, 256 passes");
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[rho_fast_tw_index(Ns, r, k)];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_regs<R, INV>(v);
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (TO_REGS) regs[q + Q * s] = v[s];
            else dst[pidx((jj / Ns) * Ns * R + k + s * Ns)] = v[s];
        }
    }
}

// One workgroup of 256 lanes runs 256/L transforms at a time (two traces
// each) and loops over trace groups; the pass-ordered twiddle table, the
// filter and gain tables (M <= 2048) and one padded exchange buffer per
// transform live in LDS.
template <int M, int DT>
__global__ __launch_bounds__(kRhoThreads) __attribute__((amdgpu_waves_per_eu(kRhoWaves)))
void rho_fast_kernel(
    const float *__restrict__ in, int n, int m, float wscale, int opt, float qmax,
    void *__restrict__ out, float *__restrict__ scale_out, int shift,
    const float2 *__restrict__ twg, float gain_dt)
{
#pragma clang fp contract(off)
    constexpr int L = M / 16, FPW = kRhoThreads / L, R3 = M / 256;
    constexpr int BUF = M + M / 16;
    constexpr bool TAB = rho_tables(M);
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    __shared__ float s_red[2][kRhoThreads / 64];
    const int tid = threadIdx.x, slot = tid / L, j = tid % L;
    float2 *buf = smem + M + slot * BUF;
    // TAB: the filter H[k] and the gain of every sample, once per workgroup
    float2 *hk = smem + M + FPW * BUF;
    float *gtab = reinterpret_cast<float *>(hk + M);
    static_assert(rho_fast_tw_count(M) <= M, "twiddles fit the table slot");
    for (int q = tid; q < M; q += kRhoThreads) {
        if (q < rho_fast_tw_count(M)) tw[q] = twg[q];
        if (TAB) {
            hk[q] = rho_response(q, M, wscale, opt);
            gtab[q] = row_gain(q, gain_dt);
        }
    }
    const int rot = (shift && (n & 1)) ? 1 : 0;
    const float invM = 1.0f / (float)M;
    const int pairs = (n + 1) / 2;
    // the two rows of transform `slot` of the group at p, into registers
    // (zero past m and for rows past n); lane j takes samples j + L r
    auto load_group = [&](int p, float2 (&regs)[16]) {
        const int pr = p + slot;
        const bool live = pr < pairs;
        const int ra = 2 * pr, rb = ra + 1;
        const bool has_b = live && rb < n;
        // a transform spans whole waves (L >= 64): its rows are wave-uniform,
        // so the row bases live in scalar registers and each load needs only
        // a 32-bit lane offset
        const int ia = __builtin_amdgcn_readfirstlane(live ? (ra + rot) % n : 0);
        const int ib = __builtin_amdgcn_readfirstlane(has_b ? (rb + rot) % n : 0);
        const float *rowa = in + (size_t)ia * m, *rowb = in + (size_t)ib * m;
        // buffer descriptors whose range is the row (none for a row past n):
        // loads past it return zeros, so the zero padding needs no branches
        const int na = __builtin_amdgcn_readfirstlane(live ? m * 4 : 0);
        const int nb = __builtin_amdgcn_readfirstlane(has_b ? m * 4 : 0);
        const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc(
            (void *)rowa, (short)0, na, 0x00020000);
        const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc(
            (void *)rowb, (short)0, nb, 0x00020000);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int q = j + L * r;
            regs[r] = make_float2(
                __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsa, 4 * q, 0, 0)),
                __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsb, 4 * q, 0, 0)));
        }
    };
    const int stride = gridDim.x * FPW;
    const bool paired = (m & 1) == 0 && ((size_t)out & 7) == 0;
    float2 regs[16];
    if (blockIdx.x * FPW < pairs) load_group(blockIdx.x * FPW, regs);
    for (int p0 = blockIdx.x * FPW; p0 < pairs; p0 += stride) {
        const int pr = p0 + slot;
        const bool live = pr < pairs;
        const int ra = 2 * pr, rb = ra + 1;
        const bool has_b = live && rb < n;
        __syncthreads();                       // previous group done with buf (and tw loaded)
        stockham_pass<M, 16, 1, false, true, false>(regs, nullptr, buf, tw, j);
        __syncthreads();
        stockham_pass<M, 16, 16, false, false, false>(regs, buf, buf, tw, j);
        __syncthreads();
        stockham_pass<M, R3, 256, false, false, true>(regs, buf, nullptr, tw, j);
        // filter in registers: lane j holds X[j + L r]
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int k = j + L * r;
            regs[r] = k == 0 ? make_float2(0.f, 0.f)
                             : cmul(TAB ? hk[k] : rho_response(k, M, wscale, opt), regs[r]);
        }
        __syncthreads();                       // pass-3 reads done before buf is rewritten
        stockham_pass<M, 16, 1, true, true, false>(regs, nullptr, buf, tw, j);
        __syncthreads();
        stockham_pass<M, 16, 16, true, false, false>(regs, buf, buf, tw, j);
        __syncthreads();
        stockham_pass<M, R3, 256, true, false, true>(regs, buf, nullptr, tw, j);
        // lane j holds y[j + L r] (times M): gain, per-trace max-abs, then the
        // values go through the exchange buffer in natural order, so that
        // the stores are a plain loop with consecutive lanes on consecutive
        // samples
        float ma = 0.f, mb = 0.f;
        __syncthreads();                       // pass-3 reads done before buf is rewritten
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int q = j + L * r;
            const float gq = TAB ? gtab[q] : row_gain(q, gain_dt);
            const float ya = (regs[r].x * invM) * gq, yb = (regs[r].y * invM) * gq;
            buf[pidx(q)] = make_float2(ya, yb);
            if (q < m) {
                ma = fmaxf(ma, fabsf(ya));
                mb = fmaxf(mb, fabsf(yb));
            }
        }
        // the registers are free: the next group's loads go out now and land
        // while this group's scale reduction and stores run
        if (p0 + stride < pairs) load_group(p0 + stride, regs);
        float inv_a = 1.f, inv_b = 1.f;
        if (DT != PSKTM_F32) {
            // max over the L lanes of this transform: waves, then LDS
            for (int o = 32; o > 0; o >>= 1) {
                ma = fmaxf(ma, __shfl_xor(ma, o));
                mb = fmaxf(mb, __shfl_xor(mb, o));
            }
            if (L > 64) {
                if ((tid & 63) == 0) {
                    s_red[0][tid >> 6] = ma;
                    s_red[1][tid >> 6] = mb;
                }
                __syncthreads();
                ma = 0.f;
                mb = 0.f;
                const int w0 = slot * (L / 64);
                for (int w = 0; w < L / 64; ++w) {
                    ma = fmaxf(ma, s_red[0][w0 + w]);
                    mb = fmaxf(mb, s_red[1][w0 + w]);
                }
            }
            const float sa = ma / qmax, sb = mb / qmax;
            inv_a = (sa > 0.f) ? 1.0f / sa : 0.f;
            inv_b = (sb > 0.f) ? 1.0f / sb : 0.f;
            if (live && j == 0) {
                scale_out[ra] = sa;
                if (has_b) scale_out[rb] = sb;
            }
        }
        __syncthreads();
        if (live) {
            const size_t oa = (size_t)__builtin_amdgcn_readfirstlane(ra) * m;
            const size_t ob = oa + m;
            if (paired) {
                // even m (and an 8-byte aligned output): every row starts on
                // an even sample, so each lane stores samples q and q + 1 of
                // both traces as one 4- (FP16/BF16) or 8-byte (FP32) word;
                // pidx keeps an even q and q + 1 adjacent in buf
                for (int q = 2 * j; q < m; q += 2 * L) {
                    const int e = pidx(q);
                    const float2 y0 = buf[e], y1 = buf[e + 1];
                    if (DT == PSKTM_F32) {
                        float *o = static_cast<float *>(out);
                        *reinterpret_cast<float2 *>(o + oa + q) = make_float2(y0.x, y1.x);
                        if (has_b) *reinterpret_cast<float2 *>(o + ob + q) = make_float2(y0.y, y1.y);
                    } else {
                        unsigned short *o = static_cast<unsigned short *>(out);
                        *reinterpret_cast<unsigned *>(o + oa + q) =
                            q16<DT>(y0.x, inv_a) | (q16<DT>(y1.x, inv_a) << 16);
                        if (has_b)
                            *reinterpret_cast<unsigned *>(o + ob + q) =
                                q16<DT>(y0.y, inv_b) | (q16<DT>(y1.y, inv_b) << 16);
                    }
                }
            } else {
                for (int q = j; q < m; q += L) {
                    const float2 y = buf[pidx(q)];
                    store_q(out, DT, oa + q, y.x, inv_a);
                    if (has_b) store_q(out, DT, ob + q, y.y, inv_b);
                }
            }
        }
    }
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
        // W_M^e = exp(-2 pi i e / M), computed in double
        auto w = [M](int e) {
            const double a = -2.0 * M_PI * e / M;
            return make_float2((float)std::cos(a), (float)std::sin(a));
        };
        // the radix-16 kernel's pass-ordered table, else W_M^q for q < M
        // (the radix-2 passes read the first half)
        std::vector<float2> h(M);
        if (rho_fast_length(M)) {
            for (int r = 1; r < 16; ++r)
                for (int k = 0; k < 16; ++k)
                    h[rho_fast_tw_index(16, r, k)] = w((r * k * (M / 256)) & (M - 1));
            for (int r = 1; r < M / 256; ++r)
                for (int k = 0; k < 256; ++k)
                    h[rho_fast_tw_index(256, r, k)] = w((r * k) & (M - 1));
        } else {
            for (int q = 0; q < M; ++q) h[q] = w(q);
        }
        float2 *d = nullptr;
        PSKTM_HIP(hipMalloc(&d, sizeof(float2) * M));
        hipError_t e = hipMemcpy(d, h.data(), sizeof(float2) * M, hipMemcpyHostToDevice);
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
    if (rho_fast_length(M)) {
        // register-blocked radix-16 form; each workgroup loops over trace
        // pairs (one twiddle-table load and table build per workgroup)
        auto pick = [&](auto mtag) -> const void * {
            constexpr int MM = decltype(mtag)::value;
            return out_dtype == PSKTM_F32 ? (const void *)rho_fast_kernel<MM, PSKTM_F32>
                 : out_dtype == PSKTM_F16 ? (const void *)rho_fast_kernel<MM, PSKTM_F16>
                                          : (const void *)rho_fast_kernel<MM, PSKTM_BF16>;
        };
        const void *fn = M == 1024 ? pick(std::integral_constant<int, 1024>{})
                       : M == 2048 ? pick(std::integral_constant<int, 2048>{})
                                   : pick(std::integral_constant<int, 4096>{});
        const int L = M / 16, fpw = kRhoThreads / L;
        const size_t lds = sizeof(float2) * ((size_t)M + (size_t)fpw * (M + M / 16)) +
                           rho_table_bytes(M);
        { psktm_status s_; if ((s_ = ensure_dynamic_lds(fn, lds)) != PSKTM_OK) return s_; }
        int cus = 0;
        PSKTM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long long groups = ((long long)(n_traces + 1) / 2 + fpw - 1) / fpw;
        // about 16 workgroups in turn per resident slot (two per CU): on
        // 1e6 x 1500 traces 32 per CU ran 4.01 ms, 8 per CU 4.06 ms (slots
        // stay in lockstep longer), 128 per CU 4.20 ms (table set-up)
        const unsigned blocks = (unsigned)std::max(1ll, std::min(groups, 32ll * cus));
        const float qmax = qmax_of(out_dtype);
        const int shift = matlab_trace_shift ? 1 : 0;
        void *args[] = {(void *)&in, (void *)&n_traces, (void *)&n_samples, (void *)&wscale,
                        (void *)&opt, (void *)&qmax, (void *)&out,
                        (void *)&scale_out, (void *)&shift, (void *)&tw, (void *)&gain_dt};
        PSKTM_HIP(hipLaunchKernel(fn, dim3(blocks), dim3(kRhoThreads), args, lds,
                                  as_stream(stream)));
        return check_launch("rho_fast_kernel");
    }
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
