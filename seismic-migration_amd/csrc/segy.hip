// SEG-Y (rev 0) / SU ingest: the reference's MATLAB readers restated in C++
// on the host, with the sample and trace-header decode on the device.
//
//   psktm_segy_text_header   scripts/ktmGetInputMeta.m:30-56   (opt 1)
//   psktm_segy_reel_header   scripts/ktmGetInputMeta.m:58-76   (opt 2)
//   psktm_segy_open          scripts/ktmGetTotalTraceNo.m:18-32 (+ SU layout)
//   psktm_segy_trace_headers scripts/ktmGetInputMeta.m:78-153  (opt 3)
//   psktm_segy_decode        scripts/ktmGetInputTrace.m:27-75 + src/ioFunctions.c:4-40
//   psktm_segy_load          file -> pinned chunks -> device -> decode
//                            (chunk size: PSKTM_SEGY_CHUNK_BYTES, default 64 MiB)
//
// The trace header is read as 120 16-bit words in the file's byte order and
// mapped to 101 columns (segy_column); 32-bit values are word pairs, first
// word high for big-endian files (ktmGetInputMeta.m:116-149).
#include "internal.hpp"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>

namespace psktm {
namespace {

// Column c of the 101-column trace-header layout: first 16-bit word and
// whether the value spans two words.
__host__ __device__ inline int segy_column(int c, bool *is32)
{
    if (c < 7) { *is32 = true; return 2 * c; }
    if (c < 11) { *is32 = false; return 14 + (c - 7); }
    if (c < 19) { *is32 = true; return 18 + 2 * (c - 11); }
    if (c < 21) { *is32 = false; return 34 + (c - 19); }
    if (c < 25) { *is32 = true; return 36 + 2 * (c - 21); }
    *is32 = false;
    return 44 + (c - 25);
}

__host__ __device__ inline unsigned word16(const unsigned char *p, bool be)
{
    return be ? ((unsigned)p[0] << 8) | p[1] : p[0] | ((unsigned)p[1] << 8);
}

// Header column as two's complement, or clamped like MATLAB's int32()/int16()
// of the unsigned value when `saturate` is set.
__host__ __device__ inline int header_value(const unsigned char *hdr, int c, bool be,
                                            bool saturate)
{
    bool is32;
    const int w = segy_column(c, &is32);
    const unsigned a = word16(hdr + 2 * w, be);
    if (is32) {
        const unsigned b = word16(hdr + 2 * w + 2, be);
        const unsigned v = be ? (a << 16) | b : a | (b << 16);
        if (saturate && v > 0x7fffffffu) return 0x7fffffff;
        return (int)v;
    }
    if (saturate && a > 0x7fffu) return 0x7fff;
    return (int)(short)(unsigned short)a;
}

__device__ __forceinline__ unsigned word32(const unsigned char *p, bool be)
{
    const unsigned w = *reinterpret_cast<const unsigned *>(p);     // 4-byte aligned
    return be ? __builtin_bswap32(w) : w;
}

// IBM System/360 single -> float. Reference arithmetic: sign * (fraction /
// 0xFFFFFF) * 16^(e-64), each step rounded as in src/ioFunctions.c:18-38;
// otherwise the exact value fraction * 2^(4(e-64)-24).
__device__ __forceinline__ float ibm_to_float(unsigned w, bool reference)
{
    const int e = (int)((w >> 24) & 0x7fu) - 64;
    if (reference) {
        const float sign = (w & 0x80000000u) ? -1.0f : 1.0f;
        const float q = (float)(w & 0x00ffffffu) / (float)0x00ffffffu;
        const float p = ldexpf(1.0f, 4 * e);
        const float sq = sign * q;
        return sq * p;
    }
    const float v = ldexpf((float)(w & 0x00ffffffu), 4 * e - 24);
    return (w & 0x80000000u) ? -v : v;
}

__global__ void segy_samples_kernel(const unsigned char *__restrict__ rec, long long count,
                                    int ns, long long record_bytes, int format, int be,
                                    int ibm_reference, float *__restrict__ out)
{
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count * ns) return;
    const long long r = q / ns;
    const int s = (int)(q - r * ns);
    const unsigned char *p = rec + r * record_bytes + 240;
    float v;
    if (format == 3) {
        v = (float)(short)(unsigned short)word16(p + 2 * s, be);
    } else {
        const unsigned w = word32(p + 4 * (size_t)s, be);
        if (format == 1) v = ibm_to_float(w, ibm_reference != 0);
        else if (format == 2) v = (float)(int)w;
        else v = __uint_as_float(w);
    }
    out[q] = v;
}

__global__ void segy_headers_kernel(const unsigned char *__restrict__ rec, long long count,
                                    long long record_bytes, int be, int *__restrict__ out)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count * PSKTM_SEGY_HEADER_COLUMNS) return;
    const long long r = q / PSKTM_SEGY_HEADER_COLUMNS;
    const int c = (int)(q - r * PSKTM_SEGY_HEADER_COLUMNS);
    out[q] = header_value(rec + r * record_bytes, c, be != 0, false);
}

bool is_su(const char *path)
{
    const size_t n = std::strlen(path);
    return n >= 3 && std::strcmp(path + n - 3, ".su") == 0;
}

struct File {
    FILE *f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

psktm_status read_at(const char *path, long long off, void *buf, size_t n)
{
    File fh;
    fh.f = std::fopen(path, "rb");
    PSKTM_CHECK(fh.f, PSKTM_E_IO, "cannot open %s: %s", path, std::strerror(errno));
    PSKTM_CHECK(fseeko(fh.f, (off_t)off, SEEK_SET) == 0, PSKTM_E_IO, "seek failed in %s", path);
    PSKTM_CHECK(std::fread(buf, 1, n, fh.f) == n, PSKTM_E_IO,
                "%s: short read of %zu bytes at offset %lld", path, n, off);
    return PSKTM_OK;
}

psktm_status check_layout(const psktm_segy_layout *L)
{
    PSKTM_CHECK(L, PSKTM_E_INVAL, "NULL layout");
    PSKTM_CHECK(L->n_samples > 0 && L->record_bytes == 240 + (long long)L->n_samples *
                L->bytes_per_sample, PSKTM_E_INVAL, "inconsistent layout");
    PSKTM_CHECK(L->format == 1 || L->format == 2 || L->format == 3 || L->format == 5,
                PSKTM_E_UNSUPPORTED, "sample format %d is not supported (1, 2, 3, 5)",
                L->format);
    return PSKTM_OK;
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

} // namespace
} // namespace psktm

using namespace psktm;

extern "C" {

psktm_status psktm_segy_text_header(const char *path, char *out)
{
    psktm__clear_error();
    PSKTM_CHECK(path && out, PSKTM_E_INVAL, "NULL argument");
    // the script's table: index = byte value, unlisted codes -> '?'
    char tbl[256];
    std::memset(tbl, '?', sizeof tbl);
    auto put = [&](int lo, const char *s) { for (int i = 0; s[i]; ++i) tbl[lo + i] = s[i]; };
    put(129, "abcdefghi");
    put(145, "jklmnopqr");
    put(161, "~stuvwxyz");
    put(192, "{ABCDEFGHI");
    put(208, "}JKLMNOPQR");
    tbl[224] = '\\';
    tbl[225] = '\0';
    put(226, "STUVWXYZ");
    put(240, "0123456789");
    put(96, "-/");
    tbl[64] = ' ';
    tbl[122] = ':';
    tbl[123] = '#';
    tbl[124] = '@';
    unsigned char raw[3200];
    psktm_status st;
    if ((st = read_at(path, 0, raw, sizeof raw)) != PSKTM_OK) return st;
    char *o = out;
    for (int line = 0; line < 40; ++line) {
        for (int i = 0; i < 80; ++i) *o++ = tbl[raw[80 * line + i]];
        *o++ = '\n';
    }
    *o = '\0';
    return PSKTM_OK;
}

psktm_status psktm_segy_reel_header(const char *path, long long values[27])
{
    psktm__clear_error();
    PSKTM_CHECK(path && values, PSKTM_E_INVAL, "NULL argument");
    unsigned char b[400];
    psktm_status st;
    if ((st = read_at(path, 3200, b, sizeof b)) != PSKTM_OK) return st;
    for (int i = 0; i < 3; ++i)
        values[i] = ((long long)b[4 * i] << 24) | ((long long)b[4 * i + 1] << 16) |
                    ((long long)b[4 * i + 2] << 8) | b[4 * i + 3];
    for (int i = 0; i < 24; ++i)
        values[3 + i] = ((long long)b[12 + 2 * i] << 8) | b[13 + 2 * i];
    return PSKTM_OK;
}

psktm_status psktm_segy_open(const char *path, int big_endian, psktm_segy_layout *L)
{
    psktm__clear_error();
    PSKTM_CHECK(path && L, PSKTM_E_INVAL, "NULL argument");
    std::memset(L, 0, sizeof *L);
    struct stat sb;
    PSKTM_CHECK(stat(path, &sb) == 0, PSKTM_E_IO, "cannot stat %s: %s", path, std::strerror(errno));
    const long long size = (long long)sb.st_size;
    L->big_endian = big_endian ? 1 : 0;
    psktm_status st;
    if (is_su(path)) {
        // SU: records only, IEEE samples; sample count and interval from the
        // first trace header (bytes 115-118)
        unsigned char h[240];
        PSKTM_CHECK(size >= 240, PSKTM_E_IO, "%s is shorter than one trace header", path);
        if ((st = read_at(path, 0, h, sizeof h)) != PSKTM_OK) return st;
        L->data_offset = 0;
        L->n_samples = (int)word16(h + 114, L->big_endian);
        L->sample_interval_us = (int)word16(h + 116, L->big_endian);
        L->format = 5;
    } else {
        long long v[27];
        PSKTM_CHECK(size >= 3600, PSKTM_E_IO, "%s is shorter than the SEG-Y reel headers", path);
        if ((st = psktm_segy_reel_header(path, v)) != PSKTM_OK) return st;
        L->data_offset = 3600;
        L->n_samples = (int)v[7];
        L->sample_interval_us = (int)v[5];
        L->format = (int)v[9];
    }
    L->bytes_per_sample = L->format == 3 ? 2 : 4;
    L->record_bytes = 240 + (long long)L->n_samples * L->bytes_per_sample;
    PSKTM_CHECK(L->n_samples > 0, PSKTM_E_IO, "%s: zero samples per trace", path);
    const long long body = size - L->data_offset;
    L->n_traces = body / L->record_bytes;
    L->ragged = (body % L->record_bytes) != 0;
    return check_layout(L);
}

psktm_status psktm_segy_trace_headers(const psktm_segy_layout *L, const char *path,
                                      long long first, long long count, int saturate,
                                      int *out)
{
    psktm__clear_error();
    psktm_status st;
    if ((st = check_layout(L)) != PSKTM_OK) return st;
    PSKTM_CHECK(path && (count == 0 || out), PSKTM_E_INVAL, "NULL argument");
    PSKTM_CHECK(first >= 0 && count >= 0 && first + count <= L->n_traces, PSKTM_E_INVAL,
                "traces [%lld, %lld) outside the file's %lld", first, first + count,
                L->n_traces);
    File fh;
    fh.f = std::fopen(path, "rb");
    PSKTM_CHECK(fh.f, PSKTM_E_IO, "cannot open %s: %s", path, std::strerror(errno));
    unsigned char h[240];
    for (long long r = 0; r < count; ++r) {
        const long long off = L->data_offset + (first + r) * L->record_bytes;
        PSKTM_CHECK(fseeko(fh.f, (off_t)off, SEEK_SET) == 0 && std::fread(h, 1, 240, fh.f) == 240,
                    PSKTM_E_IO, "%s: cannot read trace header %lld", path, first + r);
        for (int c = 0; c < PSKTM_SEGY_HEADER_COLUMNS; ++c)
            out[r * PSKTM_SEGY_HEADER_COLUMNS + c] =
                header_value(h, c, L->big_endian != 0, saturate != 0);
    }
    return PSKTM_OK;
}

psktm_status psktm_segy_decode(const psktm_segy_layout *L, const void *records,
                               long long count, int ibm_reference, float *traces,
                               int *headers, void *stream)
{
    psktm__clear_error();
    psktm_status st;
    if ((st = check_layout(L)) != PSKTM_OK) return st;
    PSKTM_CHECK(count >= 0, PSKTM_E_INVAL, "negative count");
    if (count == 0) return PSKTM_OK;
    PSKTM_CHECK(records && traces, PSKTM_E_INVAL, "NULL pointer");
    PSKTM_CHECK(L->bytes_per_sample == 2 || L->record_bytes % 4 == 0, PSKTM_E_INVAL,
                "4-byte samples need 4-byte aligned records");
    PSKTM_CHECK(((uintptr_t)records & 3u) == 0, PSKTM_E_INVAL, "records must be 4-byte aligned");
    PSKTM_CHECK(count * L->n_samples < (1ll << 40), PSKTM_E_INVAL, "too many samples");
    if ((st = require_device()) != PSKTM_OK) return st;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(segy_samples_kernel, dim3(blocks_for(count * L->n_samples)), dim3(256),
                       0, s, (const unsigned char *)records, count, L->n_samples,
                       L->record_bytes, L->format, L->big_endian, ibm_reference, traces);
    if ((st = check_launch("segy_samples_kernel")) != PSKTM_OK) return st;
    if (headers) {
        hipLaunchKernelGGL(segy_headers_kernel,
                           dim3(blocks_for(count * PSKTM_SEGY_HEADER_COLUMNS)), dim3(256), 0, s,
                           (const unsigned char *)records, count, L->record_bytes,
                           L->big_endian, headers);
        if ((st = check_launch("segy_headers_kernel")) != PSKTM_OK) return st;
    }
    return PSKTM_OK;
}

psktm_status psktm_segy_load(const psktm_segy_layout *L, const char *path, long long first,
                             long long count, int ibm_reference, float *traces, int *headers,
                             void *stream)
{
    psktm__clear_error();
    psktm_status st;
    if ((st = check_layout(L)) != PSKTM_OK) return st;
    PSKTM_CHECK(path && (count == 0 || traces), PSKTM_E_INVAL, "NULL argument");
    PSKTM_CHECK(first >= 0 && count >= 0 && first + count <= L->n_traces, PSKTM_E_INVAL,
                "traces [%lld, %lld) outside the file's %lld", first, first + count,
                L->n_traces);
    if (count == 0) return PSKTM_OK;
    if ((st = require_device()) != PSKTM_OK) return st;
    hipStream_t s = as_stream(stream);
    // two pinned host buffers and two device buffers of ~64 MiB: chunk i is
    // read into host buffer i%2 while chunk i-1 is copied and decoded
    long long chunk_bytes = 64ll << 20;
    if (const char *env = std::getenv("PSKTM_SEGY_CHUNK_BYTES"))
        chunk_bytes = std::max(1ll, std::atoll(env));
    const long long chunk = std::max(1ll, chunk_bytes / L->record_bytes);
    const size_t bytes = (size_t)(chunk * L->record_bytes);
    void *host[2] = {nullptr, nullptr}, *dev[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    File fh;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        for (int i = 0; i < 2; ++i) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (host[i]) (void)hipHostFree(host[i]);
            if (dev[i]) (void)hipFree(dev[i]);
        }
    };
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipHostMalloc(&host[i], bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&dev[i], bytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "segy staging buffers", __FILE__, __LINE__);
    }
    fh.f = std::fopen(path, "rb");
    if (!fh.f || fseeko(fh.f, (off_t)(L->data_offset + first * L->record_bytes), SEEK_SET) != 0) {
        cleanup();
        psktm__set_error("cannot open or seek %s", path);
        return PSKTM_E_IO;
    }
    st = PSKTM_OK;
    for (long long r = 0, i = 0; r < count && st == PSKTM_OK; r += chunk, ++i) {
        const int b = (int)(i & 1);
        const long long n = std::min(chunk, count - r);
        // host buffer b is free once the copy issued from it two chunks ago ended
        if ((e = hipEventSynchronize(done[b])) != hipSuccess) {
            st = hip_fail(e, "hipEventSynchronize", __FILE__, __LINE__);
            break;
        }
        if (std::fread(host[b], 1, (size_t)(n * L->record_bytes), fh.f) !=
            (size_t)(n * L->record_bytes)) {
            psktm__set_error("%s: short read at trace %lld", path, first + r);
            st = PSKTM_E_IO;
            break;
        }
        if ((e = hipMemcpyAsync(dev[b], host[b], (size_t)(n * L->record_bytes),
                                hipMemcpyHostToDevice, s)) != hipSuccess ||
            (e = hipEventRecord(done[b], s)) != hipSuccess) {
            st = hip_fail(e, "segy chunk copy", __FILE__, __LINE__);
            break;
        }
        st = psktm_segy_decode(L, dev[b], n, ibm_reference, traces + r * L->n_samples,
                               headers ? headers + r * PSKTM_SEGY_HEADER_COLUMNS : nullptr,
                               stream);
    }
    cleanup();
    return st;
}

} // extern "C"
