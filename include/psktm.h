/*
 * psktm.h - C ABI of the MI355X-native prestack Kirchhoff time migration (PSTM).
 *
 * Two layers live behind this header, both with C linkage in libpsktm.so:
 *
 *  1. The reference's own C API for the migration path, same names and
 *     argument meaning, so an existing caller relinks against libpsktm.so:
 *       initTwoAxes / initThreeAxes          src/psktmFunctions.c:1,19
 *       ktmIBM32toIEEE, fileSizeFourBytes, binReadUints, binReadFloats,
 *       fileWrite, errorMessage, loadInputData, load*Config
 *                                            src/ioFunctions.c:4-175
 *       ktmMigrationGPU                      src/psktmCUDA.cu:211-221
 *       ktmMigrationCPU                      src/psktmCPU.c:13-23
 *     The reference never defines its structs or IBM masks; the layouts
 *     below are this library's definition (fields inferred from their uses).
 *
 *  2. The extended psktm_* API: device pointers (HBM-resident inputs), an
 *     explicit HIP stream (passed as void*), status codes, and the options the
 *     reference lacks (linear interpolation, aperture, 3-D velocity, FP16/BF16
 *     trace storage, accumulate mode, the adjoint, the rho filter, FP64 CG
 *     reductions and batched gathers).
 *
 * Units: seconds and metres. The image time axis is one-way vertical time tau
 * (the DSR time is 2*tau at zero offset, src/psktmCPU.c:78-79); offsets are
 * half-offset vectors (off = (rec - src)/2, scripts/ktmConvertXYtoOffMid.m:12);
 * the reference's "slowness" argument holds velocity in m/s
 * (src/psktmCPU.c:48). Layouts: traces[k*tracePts + n] (trace-major),
 * image[(i*Lj + j)*Ll + l] (x slowest, tau fastest) (src/psktmCPU.c:84-85).
 *
 * Errors: legacy entries return void like the reference, leave the image
 * zeroed on failure and record the reason for psktm_last_error(); extended
 * entries return psktm_status. The library never prints and never tears down
 * the device context.
 */
#ifndef PSKTM_H
#define PSKTM_H

#include <stdio.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Reference data structures (undefined in the reference).                  */
/* ------------------------------------------------------------------------ */

/* One image axis: imagePnts samples at imageStart + i*imageStep
 * (src/psktmFunctions.c:8-12, src/psktmCUDA.cu:226-237). */
struct imageGrid {
    int   imagePnts;
    float imageStart;
    float imageStep;
};

/* Job scalars: trace count, samples per trace, sample interval in seconds
 * (src/psktmCPU.c:29,59,80,82; src/psktmCUDA.cu:245-247). */
struct jobParams {
    int   traceNo;
    int   tracePts;
    float traceSampleRate;
};

/* Open input files of the on-disk job (src/ioFunctions.c:94-153). */
struct fileInputPtrs {
    FILE *inputData;
    FILE *inputSrcX;
    FILE *inputSrcY;
    FILE *inputRecX;
    FILE *inputRecY;
    int   flag;          /* 1 if any file failed to open */
};

/* IBM System/360 single: bit 31 sign, bits 30:24 base-16 exponent (bias 64),
 * bits 23:0 fraction (src/ioFunctions.c:7-9). */
#define SIGN_MASK     0x80000000u
#define EXPONENT_MASK 0x7F000000u
#define FRACTION_MASK 0x00FFFFFFu

/* ------------------------------------------------------------------------ */
/* Reference API (legacy entry points).                                     */
/* ------------------------------------------------------------------------ */

/* axis[i] = start + (float)i * step, binary32, no contraction.
 * src/psktmFunctions.c:1-17 */
void initTwoAxes(struct imageGrid *imageY, struct imageGrid *imageZ,
                 float *imageYY, float *imageZZ);
/* src/psktmFunctions.c:19-39 */
void initThreeAxes(struct imageGrid *imageX, struct imageGrid *imageY,
                   struct imageGrid *imageZ, float *imageXX, float *imageYY,
                   float *imageZZ);

/* IBM -> IEEE with the reference's arithmetic: sign * (fraction / 0xFFFFFF)
 * * 16^(exp-64) (src/ioFunctions.c:4-40). Divides by 2^24-1, so results are
 * about 1 ulp high; psktm_ibm32_to_ieee_exact gives the true value. */
void ktmIBM32toIEEE(unsigned int *input, float *output, int n);
/* Number of 4-byte words in an open file; rewinds it (src/ioFunctions.c:44-56). */
int fileSizeFourBytes(FILE *fin);
/* malloc'd buffer of n words read from fin; rewinds it; caller frees
 * (src/ioFunctions.c:60-68, 70-78). */
unsigned int *binReadUints(FILE *fin, int n);
float *binReadFloats(FILE *fin, int n);
/* Writes t floats (src/ioFunctions.c:80-84). */
void fileWrite(FILE *fout, float *ptr, int t);
/* Records "Can't open <str>." as the last error; the reference printf's it
 * (src/ioFunctions.c:88-91). */
void errorMessage(char *str);
/* Opens ../data/{inputData,inputSrcX,inputSrcY,inputRecX,inputRecY}.bin;
 * flag = 1 if any fails (src/ioFunctions.c:94-153). Files stay open. */
void loadInputData(struct fileInputPtrs *input);
/* Empty in the reference (src/ioFunctions.c:155-175); kept as no-ops.
 * psktm_load_ktm_config reads inp.txt for real. */
void loadRTM2dConfig(void);
void loadKTMConfig(void);
void loadKDMConfig(void);

/* Drop-in for src/psktmCUDA.cu:211-221: host arrays in, image zeroed then
 * overwritten, synchronous. Runs the reference-exact HIP form (truncating
 * index, IEEE sqrt/div, no contraction, traces summed in k order per image
 * sample), which is bit-identical to the single-thread reference loop. No
 * aperture (the reference has none). On any failure the image stays zeroed
 * and psktm_last_error() says why. */
void ktmMigrationGPU(struct imageGrid *imageX, struct imageGrid *imageY,
                     struct imageGrid *imageZ, struct jobParams *config,
                     float *midX, float *midY, float *offX, float *offY,
                     float *traces, float *slowness, float *image);
/* Drop-in for src/psktmCPU.c:13-23 (the identical argument list): the same
 * semantics and the same image bits as ktmMigrationGPU, which equal the
 * reference CPU loop run on one thread (its OpenMP loop races on the image,
 * src/psktmCPU.c:56,84). Computed on the GPU by the same kernel; fails like
 * ktmMigrationGPU (zeroed image, psktm_last_error()) without a gfx950. */
void ktmMigrationCPU(struct imageGrid *imageX, struct imageGrid *imageY,
                     struct imageGrid *imageZ, struct jobParams *config,
                     float *midX, float *midY, float *offX, float *offY,
                     float *traces, float *slowness, float *image);

/* ------------------------------------------------------------------------ */
/* Extended API.                                                            */
/* ------------------------------------------------------------------------ */

typedef enum {
    PSKTM_OK = 0,
    PSKTM_E_INVAL = 1,        /* bad argument or size */
    PSKTM_E_NODEV = 2,        /* no usable gfx950 device */
    PSKTM_E_HIP = 3,          /* HIP runtime error */
    PSKTM_E_ALLOC = 4,        /* device allocation failed */
    PSKTM_E_LAUNCH = 5,       /* kernel launch failed */
    PSKTM_E_IO = 6,           /* file I/O */
    PSKTM_E_UNSUPPORTED = 7   /* valid but not implemented combination */
} psktm_status;

typedef enum { PSKTM_INTERP_TRUNC = 0, PSKTM_INTERP_LINEAR = 1 } psktm_interp;
/* IEEE: reference arithmetic (correctly rounded sqrt and divide, no FMA);
 * FAST: hardware v_sqrt_f32, reciprocal multiply, FMA. */
typedef enum { PSKTM_MATH_IEEE = 0, PSKTM_MATH_FAST = 1 } psktm_math;
typedef enum { PSKTM_F32 = 0, PSKTM_F16 = 1, PSKTM_BF16 = 2 } psktm_dtype;
typedef enum { PSKTM_VEL_1D = 0, PSKTM_VEL_3D = 1 } psktm_vel_kind;

/* Output image axes (x slowest, tau fastest). tau is one-way time. */
typedef struct psktm_grid {
    int   nx, ny, nt;
    float x0, dx, y0, dy, tau0, dtau;
} psktm_grid;

/* Acquisition: device arrays of length n_traces (half offsets). */
typedef struct psktm_survey {
    int          n_traces;
    int          n_samples;   /* samples per trace */
    float        dt;          /* sample interval, s */
    const float *mid_x, *mid_y, *off_x, *off_y;
} psktm_survey;

/* Trace samples on the device, trace-major [n_traces][n_samples]. For F16 and
 * BF16 the sample value is stored[k][n] * scale[k]. */
typedef struct psktm_traces {
    int          dtype;       /* psktm_dtype */
    const void  *data;
    const float *scale;       /* [n_traces]; NULL for F32 */
    int          gained;      /* 0: raw samples; the forward applies the weight
                                 t*sqrt(tau^2/t) = tau*sqrt(t) per update
                                 (src/psktmCPU.c:84-85). 1: sample n already
                                 carries sqrt(n*dt) (psktm_rho_filter_gain,
                                 psktm_trace_gain), and the forward applies
                                 tau only: one v_sqrt_f32 fewer per update.
                                 The interpolated gained sample differs from
                                 the interpolated sample times sqrt(t) by
                                 f(1-f)(g[n+1]-g[n]) terms, ~1e-4 relative on
                                 30 Hz data at 4 ms. psktm_demigrate is the
                                 adjoint of the raw form,
                                 psktm_demigrate_gained of the gained form;
                                 precise mode takes either. */
} psktm_traces;

/* RMS velocity in m/s on the device: VEL_1D -> v[nt] (laterally invariant, as
 * in the reference); VEL_3D -> v[(i*ny + j)*nt + l] on the image grid. */
typedef struct psktm_velocity {
    int          kind;        /* psktm_vel_kind */
    const float *v;
} psktm_velocity;

typedef struct psktm_options {
    int   interp;             /* psktm_interp */
    int   math;               /* psktm_math */
    float aperture;           /* metres, midpoint-to-bin; <= 0 means none */
    int   tile_x, tile_y;     /* bins per tile (8 or 16 in all); 0 = library
                                 choice: 4x2 on a 3-D grid, 8x1 on a line */
    int   traces_per_stage;   /* traces staged in LDS at once (1-8); 0 = the
                                 largest up to 5 (8 for gained FP16 traces,
                                 whose rows take half the LDS) that keeps the
                                 kernel's workgroups per CU */
    int   precise;            /* 1: compensated (Kahan) FP32 accumulation in
                                 the forward, for least squares and adjoint
                                 tests. 0: plain FP32 accumulation. The
                                 adjoint's sums are exact 64-bit fixed-point
                                 integers either way. */
    int   splits;             /* forward: workgroups sharing one tile's trace
                                 list, each summing a contiguous part into a
                                 partial image that a second kernel adds in
                                 fixed order. 0 = library choice (more than
                                 one only when the image has too few tiles to
                                 fill the GPU), 1 = one workgroup per tile. */
    int   row_pad;            /* forward: zero entries appended to every trace
                                 row staged in LDS. A tile whose traveltimes
                                 are bounded (from the plan's geometry and the
                                 tile's slowness) below the padded row's end
                                 reads the row without clamping the sample
                                 index; images are bit-identical either way.
                                 0 = library choice (up to half a row, as
                                 long as the kernel's occupancy is kept),
                                 -1 = no padding (the index is always
                                 clamped), > 0 = that many entries. */
} psktm_options;

/* Last error text of the calling thread ("" if none). */
const char *psktm_last_error(void);
/* Library version string. */
const char *psktm_version(void);
/* Number of visible HIP devices (0 when none; never fails). */
int psktm_device_count(void);

/* A plan holds what depends only on the survey geometry, the image grid and
 * the options: midpoints binned into tiles and, per output tile, the ordered
 * list of traces inside its aperture. Build once, migrate many times
 * (FP32/FP16 forms, every CG iteration). Geometry arrays are read during
 * creation only. */
typedef struct psktm_plan psktm_plan;
psktm_status psktm_plan_create(psktm_plan **plan, const psktm_grid *grid,
                               const psktm_survey *survey,
                               const psktm_options *opt, void *stream);
psktm_status psktm_plan_destroy(psktm_plan *plan);
/* Total updates (trace, in-aperture bin, tau) one migration performs. */
psktm_status psktm_plan_updates(const psktm_plan *plan, double *updates);
/* Bytes of device memory the plan holds. */
psktm_status psktm_plan_bytes(const psktm_plan *plan, size_t *bytes);

/* Migration (traces -> image), stream-ordered, device pointers. The image is
 * overwritten unless accumulate != 0. */
psktm_status psktm_migrate(psktm_plan *plan, const psktm_traces *traces,
                           const psktm_velocity *vel, float *image,
                           int accumulate, void *stream);
/* Adjoint of psktm_migrate (image -> traces, Kirchhoff demigration) with the
 * same discrete traveltime, interpolation, aperture and weight. traces_out is
 * FP32 [n_traces][n_samples], overwritten unless accumulate != 0. */
psktm_status psktm_demigrate(psktm_plan *plan, const float *image,
                             const psktm_velocity *vel, float *traces_out,
                             int accumulate, void *stream);
/* Adjoint of the gained operator: psktm_trace_gain (sample n times
 * sqrtf(n*dt)) followed by psktm_migrate with gained = 1. Demigration with
 * the weight tau per update (no sqrt(t)), each output sample then multiplied
 * by sqrtf(n*dt): one v_sqrt_f32 fewer per update than psktm_demigrate. With
 * it, least squares runs on the operator of the gained forward form. No
 * counterpart in the reference (it has no adjoint). */
psktm_status psktm_demigrate_gained(psktm_plan *plan, const float *image,
                                    const psktm_velocity *vel, float *traces_out,
                                    int accumulate, void *stream);

/* Batched gathers: several independent migrations (e.g. offset classes),
 * each with its own traces and output patch, run as ONE persistent launch
 * whose workgroups pull (gather, tile) items from a device work queue in
 * order of decreasing cost. All gathers share the survey's sample count and
 * interval, the tau axis (nt, tau0, dtau) of their patches, the options and
 * a 1-D velocity v[nt]. Traces of gather g are rows first_trace ..
 * first_trace + n_traces - 1 of the batch survey and trace arrays; images are
 * the patches' images concatenated in gather order. No counterpart in the
 * reference (its migration handles one trace block per call). */
typedef struct psktm_gather {
    psktm_grid grid;          /* output patch of this gather */
    int        first_trace;
    int        n_traces;
} psktm_gather;
typedef struct psktm_batch psktm_batch;
psktm_status psktm_batch_create(psktm_batch **batch, const psktm_gather *gathers,
                                int n_gathers, const psktm_survey *survey,
                                const psktm_options *opt, void *stream);
psktm_status psktm_batch_destroy(psktm_batch *batch);
psktm_status psktm_batch_updates(const psktm_batch *batch, double *updates);
/* use_graph != 0: the launch sequence (tables, queue reset, persistent
 * kernel) is recorded into a HIP graph on the first call with a given set of
 * pointers and replayed, on the caller's stream, by later calls with the same
 * ones. A batch serves one stream at a time. */
psktm_status psktm_batch_migrate(psktm_batch *batch, const psktm_traces *traces,
                                 const psktm_velocity *vel, float *images,
                                 int use_graph, void *stream);

/* Reference-exact migration on device pointers (the kernel behind
 * ktmMigrationGPU): TRUNC + IEEE, 1-D velocity v[nt], no aperture, FP32.
 * Images too small to fill the GPU with one lane per sample (under two
 * workgroups of 8 bins x 256 tau per CU) are computed in trace stages into
 * stream-ordered scratch (hipMallocAsync/hipFreeAsync on `stream`, at most
 * 1 GiB) and summed per sample in trace order; the image bits are the same.
 * Without that memory the call falls back to the one-lane-per-sample form. */
psktm_status psktm_migrate_exact(const psktm_grid *grid,
                                 const psktm_survey *survey,
                                 const float *traces, const float *vel,
                                 float *image, void *stream);

/* Rho filter (scripts/ktmApplyRhoFilter.m:1-64) on device traces
 * [n_traces][n_samples]: zero-pad to M = 2^ceil(log2 m), multiply the
 * spectrum by |w| (opt 1), sqrt|w| e^{i pi/4 sgn w} (opt 2) or i w (opt 3)
 * with w_k = 2 pi k / (dw_us * 1e-6) (no 1/M, as in the reference), keep the
 * first m samples. out_dtype F32 writes float; F16/BF16 write the quantised
 * value and scale_out[k] = max|y_k| / qmax (per-trace max-abs scale).
 * matlab_trace_shift != 0 reproduces the reference's 2-D fftshift, which for
 * an odd trace count returns trace (c+1) mod n in row c. */
psktm_status psktm_rho_filter(const float *in, int n_traces, int n_samples,
                              float dw_us, int opt, int out_dtype, void *out,
                              float *scale_out, int matlab_trace_shift,
                              void *stream);
/* psktm_rho_filter followed, in the same kernel, by the trace-domain part of
 * the migration weight: output sample n is the filtered value times
 * sqrtf(n * gain_dt) (gain_dt = the sample interval in seconds; 0 = no gain,
 * the same output as psktm_rho_filter), quantised after the gain. Traces
 * written with a gain are migrated with psktm_traces.gained = 1. */
psktm_status psktm_rho_filter_gain(const float *in, int n_traces, int n_samples,
                                   float dw_us, int opt, int out_dtype, void *out,
                                   float *scale_out, int matlab_trace_shift,
                                   float gain_dt, void *stream);
/* The gain alone: out[k][n] = in[k][n] * sqrtf(n * gain_dt) (each product
 * rounded once), as F32, or F16/BF16 with per-trace max-abs scales. */
psktm_status psktm_trace_gain(const float *in, int n_traces, int n_samples,
                              float gain_dt, int out_dtype, void *out,
                              float *scale_out, void *stream);
/* FP32 -> F16/BF16 with per-trace max-abs scales, no filtering. */
psktm_status psktm_quantize(const float *in, int n_traces, int n_samples,
                            int out_dtype, void *out, float *scale_out,
                            void *stream);

/* True IBM decode (fraction / 2^24), bit-exact where the value is a normal
 * float; flushes below FLT_MIN to signed zero and saturates to +-inf above. */
void psktm_ibm32_to_ieee_exact(const unsigned int *input, float *output, int n);
/* Same decode on the device (n words, stream-ordered). */
psktm_status psktm_ibm32_to_ieee_device(const unsigned int *input,
                                        float *output, size_t n, void *stream);

/* Reads inp.txt (12 floats: tracePts, dt, traceNo, xStart, xStep, xPnts,
 * yStart, yStep, yPnts, zStart, zStep, zPnts; written by
 * scripts/ktmWriteTraceBlockConstOffset.m:154-167, read by
 * scripts/ktmVisualizeOutputFile.m:22-48) into the structs. */
psktm_status psktm_load_ktm_config(const char *path, struct imageGrid *imageX,
                                   struct imageGrid *imageY,
                                   struct imageGrid *imageZ,
                                   struct jobParams *config);

/* Trace preparation on the device (the MATLAB writers' steps).
 *
 * Offset sort (scripts/ktmGetSortedOffTraces.m:17-22): perm[i] = index of
 * the trace with the i-th smallest half-offset magnitude sqrt(hx^2 + hy^2),
 * ties in trace order (stable); key[i] = that magnitude (may be NULL). */
psktm_status psktm_offset_sort(const float *off_x, const float *off_y, int n,
                               int *perm, float *key, void *stream);
/* Offset classes: class c holds traces with full offset 2|h| in
 * [c*width, (c+1)*width), c < n_classes (larger offsets are dropped).
 * perm lists the kept traces class by class (trace order within a class);
 * class_start[c] .. class_start[c+1] is class c's range (n_classes + 1
 * entries). Returns the number kept in *n_kept (host). */
psktm_status psktm_offset_classes(const float *off_x, const float *off_y, int n,
                                  float width, int n_classes, int *perm,
                                  int *class_start, int *n_kept, void *stream);
/* Constant-offset selection (scripts/ktmGetConstOffTraces.m:18-28): keeps
 * traces with h - tol < round(|h|) < h + tol, in trace order; *n_kept on
 * the host. */
psktm_status psktm_select_offset(const float *off_x, const float *off_y, int n,
                                 float h, float tol, int *selected, int *n_kept,
                                 void *stream);
/* Gather + zero padding (scripts/ktmWriteTraceBlockConstOffset.m:100-118):
 * out[r][s] = in[rows[r]][s] for r < n_rows, s < m; zero elsewhere, with
 * out of shape [ceil(n_rows/sze)*sze][ceil(m/sze)*sze] (rows NULL means
 * identity). */
psktm_status psktm_gather_pad(const float *in, int m, const int *rows, int n_rows,
                              int sze, float *out, void *stream);
/* Temporal upsampling by 2^N (scripts/ktmInterpTraceBlock.m:134-154): each
 * trace of m samples becomes (m-1)*2^N + 1 samples by linear interpolation
 * on the sample grid 0 : 2^-N : m-1. */
psktm_status psktm_upsample_traces(const float *in, int n, int m, int N, float *out,
                                   void *stream);

/* SEG-Y (rev 0) and SU ingest: the reference's MATLAB readers
 * scripts/ktmGetInputMeta.m, scripts/ktmGetInputTrace.m and
 * scripts/ktmGetTotalTraceNo.m, with the sample and header decode on the
 * device. A SEG-Y file is a 3200-byte EBCDIC header, a 400-byte binary reel
 * header and records of a 240-byte trace header plus n_samples samples; an
 * SU file (name ending in ".su") is records only. Trace ranges are 0-based
 * [first, first + count). */
#define PSKTM_SEGY_HEADER_COLUMNS 101

typedef struct {
    long long data_offset;  /* byte offset of the first record: 3600 or 0 (SU) */
    long long n_traces;     /* whole records in the file */
    long long record_bytes; /* 240 + n_samples * bytes_per_sample */
    int n_samples;          /* reel value 8 (SU: trace header bytes 115-116) */
    int sample_interval_us; /* reel value 6 (SU: trace header bytes 117-118) */
    int format;             /* 1 IBM float, 2 int32, 3 int16, 5 IEEE float */
    int bytes_per_sample;   /* 2 for format 3, else 4 */
    int big_endian;         /* byte order of trace headers and samples */
    int ragged;             /* 1 if the file size is not a whole number of records */
} psktm_segy_layout;

/* ktmGetInputMeta(file, en, 1) (scripts/ktmGetInputMeta.m:30-56): the EBCDIC
 * header through the script's table (unlisted codes -> '?'), 40 lines of 80
 * characters each followed by '\n', then a NUL: 3241 bytes. */
psktm_status psktm_segy_text_header(const char *path, char *out);
/* ktmGetInputMeta(file, en, 2) (scripts/ktmGetInputMeta.m:58-76): job, line
 * and reel numbers (4 bytes each) and 24 two-byte values, big-endian
 * unsigned as the script composes them. */
psktm_status psktm_segy_reel_header(const char *path, long long values[27]);
/* Record layout and trace count (scripts/ktmGetTotalTraceNo.m:18-32). For
 * SEG-Y the format and sample count come from the reel header; big_endian
 * gives the byte order of the records (SEG-Y: 1). */
psktm_status psktm_segy_open(const char *path, int big_endian, psktm_segy_layout *layout);
/* ktmGetInputMeta(file, en, 3, range) (scripts/ktmGetInputMeta.m:78-153):
 * count x 101 header columns (7 x int32, 4 x int16, 8 x int32, 2 x int16,
 * 4 x int32 = source X/Y, receiver X/Y, then 76 x int16), host memory.
 * saturate = 1 reproduces the script's int32()/int16() casts, which clamp
 * negative values to 2147483647 / 32767; 0 decodes two's complement. */
psktm_status psktm_segy_trace_headers(const psktm_segy_layout *layout, const char *path,
                                      long long first, long long count, int saturate,
                                      int *out);
/* Decode `count` raw records already in device memory (record_bytes apart)
 * into traces[count][n_samples] (FP32) and, if headers is not NULL,
 * headers[count][101] (two's complement). Format 1 uses the reference's
 * ktmIBM32toIEEE arithmetic when ibm_reference = 1 (bit-identical to it),
 * the exact IBM value otherwise; formats 2 and 3 convert the integers; 5 is
 * IEEE. Stream-ordered. */
psktm_status psktm_segy_decode(const psktm_segy_layout *layout, const void *records,
                               long long count, int ibm_reference, float *traces,
                               int *headers, void *stream);
/* File -> device: reads records [first, first + count) in chunks through
 * pinned host buffers, copies them to the device and decodes them there
 * (traces and optional headers as in psktm_segy_decode). Returns after the
 * last chunk has been decoded on `stream`. The chunk size is 64 MiB unless
 * the environment variable PSKTM_SEGY_CHUNK_BYTES sets another. */
psktm_status psktm_segy_load(const psktm_segy_layout *layout, const char *path,
                             long long first, long long count, int ibm_reference,
                             float *traces, int *headers, void *stream);

/* FP64 reductions for conjugate gradients, stream-ordered, no host sync:
 * result[0] = sum_i a[i]*b[i] accumulated in double. Workgroup sums meet in
 * an FP64 atomic add, so the last bits depend on their order. */
psktm_status psktm_dot_f64(const float *a, const float *b, size_t n,
                           double *result, void *stream);
/* The same sum in a fixed order: at most PSKTM_DOT_WORKSPACE workgroups write
 * one partial each into workspace[] (caller-owned device memory of
 * PSKTM_DOT_WORKSPACE doubles), then one workgroup adds them in index order,
 * so equal inputs give equal bits on every run. Calls sharing a workspace
 * must not overlap. */
#define PSKTM_DOT_WORKSPACE 1024
psktm_status psktm_dot_f64_ws(const float *a, const float *b, size_t n,
                              double *result, double *workspace, void *stream);
/* No counterpart in the reference: psktm_dot_f64_ws for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc) and pass no workspace. Same kernels, same
 * bits; the partials live in a static device array of PSKTM_DOT_WORKSPACE
 * doubles (8 KiB) that each GPU holds until the process ends. Any n fits (the
 * first pass uses at most PSKTM_DOT_WORKSPACE workgroups), so no size is
 * refused. Calls no hipMalloc/hipFree. All calls on one GPU share the array:
 * calls that overlap on one GPU race. */
psktm_status psktm_dot_f64_noalloc(const float *a, const float *b, size_t n,
                                   double *result, void *stream);
/* y[i] += (float)(num[0]/den[0]) * sign * x[i], scalars read on the device. */
psktm_status psktm_axpy_ratio(float *y, const float *x, size_t n,
                              const double *num, const double *den,
                              float sign, void *stream);
/* p[i] = r[i] + (float)(num[0]/den[0]) * p[i]. */
psktm_status psktm_xpby_ratio(float *p, const float *r, size_t n,
                              const double *num, const double *den,
                              void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PSKTM_H */
