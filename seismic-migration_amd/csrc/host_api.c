/*
 * host_api.c - the reference's C host API for the migration path, in
 * libpsktm.so: image axes (src/psktmFunctions.c), IBM decode and flat-binary
 * I/O (src/ioFunctions.c), plus a real reader for inp.txt.
 *
 * Compiled as C with contraction off: the axis values must be the reference's
 * binary32 start + (float)i * step with separate roundings.
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/psktm.h"

/* Contraction is disabled by -ffp-contract=off on the compile line. */

void psktm__set_error(const char *fmt, ...);
void psktm__clear_error(void);

static void fill_axis(const struct imageGrid *g, float *out)
{
    for (int i = 0; i < g->imagePnts; i++) {
        float fi = (float)i;
        float p = fi * g->imageStep;
        out[i] = g->imageStart + p;
    }
}

/* src/psktmFunctions.c:1-17 */
void initTwoAxes(struct imageGrid *imageY, struct imageGrid *imageZ,
                 float *imageYY, float *imageZZ)
{
    fill_axis(imageY, imageYY);
    fill_axis(imageZ, imageZZ);
}

/* src/psktmFunctions.c:19-39 */
void initThreeAxes(struct imageGrid *imageX, struct imageGrid *imageY,
                   struct imageGrid *imageZ, float *imageXX, float *imageYY,
                   float *imageZZ)
{
    fill_axis(imageX, imageXX);
    fill_axis(imageY, imageYY);
    fill_axis(imageZ, imageZZ);
}

/* src/ioFunctions.c:4-40. The reference evaluates
 *   sign * ((float)fraction / (float)FRACTION_MASK) * powf(16.0f, exp)
 * left to right in binary32; 16^e is an exact power of two (or 0 / inf at
 * the ends of the range), so ldexpf gives the same value powf does. */
void ktmIBM32toIEEE(unsigned int *input, float *output, int n)
{
    for (int i = 0; i < n; i++) {
        const uint32_t w = input[i];
        const float sign = (w & SIGN_MASK) ? -1.0f : 1.0f;
        const int e = (int)((w & EXPONENT_MASK) >> 24) - 64;
        const float q = (float)(w & FRACTION_MASK) / (float)FRACTION_MASK;
        const float p = ldexpf(1.0f, 4 * e);
        const float sq = sign * q;
        output[i] = sq * p;
    }
}

/* True IBM value: (-1)^s * fraction * 2^(4(e-64) - 24). (float)fraction is
 * exact (24 bits); ldexpf rounds once, only for subnormal results. */
void psktm_ibm32_to_ieee_exact(const unsigned int *input, float *output, int n)
{
    for (int i = 0; i < n; i++) {
        const uint32_t w = input[i];
        const int e = (int)((w & EXPONENT_MASK) >> 24) - 64;
        float v = ldexpf((float)(w & FRACTION_MASK), 4 * e - 24);
        output[i] = (w & SIGN_MASK) ? -v : v;
    }
}

/* src/ioFunctions.c:44-56: the reference counts 4-byte freads until EOF, with
 * z starting at -1 to drop the final failed read, so it returns the number of
 * whole words (a trailing partial word is not counted); then rewinds. */
int fileSizeFourBytes(FILE *fin)
{
    if (!fin) {
        psktm__set_error("fileSizeFourBytes: NULL file");
        return -1;
    }
    if (fseek(fin, 0, SEEK_END) != 0) {
        psktm__set_error("fileSizeFourBytes: seek failed");
        return -1;
    }
    long bytes = ftell(fin);
    fseek(fin, 0, SEEK_SET);
    if (bytes < 0) return -1;
    return (int)(bytes / 4);
}

/* src/ioFunctions.c:60-68 */
unsigned int *binReadUints(FILE *fin, int n)
{
    if (!fin || n < 0) return NULL;
    unsigned int *x = (unsigned int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(unsigned int));
    if (!x) return NULL;
    size_t got = fread(x, sizeof(unsigned int), (size_t)n, fin);
    if (got < (size_t)n) memset(x + got, 0, ((size_t)n - got) * sizeof(unsigned int));
    fseek(fin, 0, SEEK_SET);
    return x;
}

/* src/ioFunctions.c:70-78 */
float *binReadFloats(FILE *fin, int n)
{
    if (!fin || n < 0) return NULL;
    float *x = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    if (!x) return NULL;
    size_t got = fread(x, sizeof(float), (size_t)n, fin);
    if (got < (size_t)n) memset(x + got, 0, ((size_t)n - got) * sizeof(float));
    fseek(fin, 0, SEEK_SET);
    return x;
}

/* src/ioFunctions.c:80-84 */
void fileWrite(FILE *fout, float *ptr, int t)
{
    if (!fout || !ptr || t <= 0) return;
    fwrite(ptr, sizeof(float), (size_t)t, fout);
}

/* src/ioFunctions.c:88-91: the message becomes the last error (the library
 * does not print). */
void errorMessage(char *str)
{
    psktm__set_error("Can't open %s.", str ? str : "(null)");
}

static void open_one(const char *name, FILE **slot, struct fileInputPtrs *input)
{
    FILE *f = fopen(name, "r+b");
    if (!f) {
        errorMessage((char *)name);
        input->flag = 1;
    } else {
        *slot = f;
    }
}

/* src/ioFunctions.c:94-153: same relative paths, same flag semantics. */
void loadInputData(struct fileInputPtrs *input)
{
    if (!input) return;
    input->flag = 0;
    open_one("../data/inputData.bin", &input->inputData, input);
    open_one("../data/inputSrcX.bin", &input->inputSrcX, input);
    open_one("../data/inputSrcY.bin", &input->inputSrcY, input);
    open_one("../data/inputRecX.bin", &input->inputRecX, input);
    open_one("../data/inputRecY.bin", &input->inputRecY, input);
}

/* src/ioFunctions.c:155-175: empty in the reference. */
void loadRTM2dConfig(void) {}
void loadKTMConfig(void) {}
void loadKDMConfig(void) {}

psktm_status psktm_load_ktm_config(const char *path, struct imageGrid *imageX,
                                   struct imageGrid *imageY,
                                   struct imageGrid *imageZ,
                                   struct jobParams *config)
{
    psktm__clear_error();
    if (!path || !imageX || !imageY || !imageZ || !config) {
        psktm__set_error("psktm_load_ktm_config: NULL argument");
        return PSKTM_E_INVAL;
    }
    FILE *f = fopen(path, "r");
    if (!f) {
        psktm__set_error("Can't open %s.", path);
        return PSKTM_E_IO;
    }
    double a[12];
    int got = 0;
    while (got < 12 && fscanf(f, "%lf", &a[got]) == 1) got++;
    fclose(f);
    if (got < 12) {
        psktm__set_error("%s holds %d numbers; 12 expected", path, got);
        return PSKTM_E_IO;
    }
    /* a(1)=SAMPLES, a(2)=dt, a(3)=RECORDS (scripts/ktmWriteTraceBlockConstOffset.m:161-165);
     * a(4..12) = x, y, z start/step/points (scripts/ktmVisualizeOutputFile.m:29-47). */
    config->tracePts = (int)lround(a[0]);
    config->traceSampleRate = (float)a[1];
    config->traceNo = (int)lround(a[2]);
    imageX->imageStart = (float)a[3]; imageX->imageStep = (float)a[4]; imageX->imagePnts = (int)lround(a[5]);
    imageY->imageStart = (float)a[6]; imageY->imageStep = (float)a[7]; imageY->imagePnts = (int)lround(a[8]);
    imageZ->imageStart = (float)a[9]; imageZ->imageStep = (float)a[10]; imageZ->imagePnts = (int)lround(a[11]);
    if (config->tracePts <= 0 || config->traceNo < 0 || !(config->traceSampleRate > 0.0f) ||
        imageX->imagePnts <= 0 || imageY->imagePnts <= 0 || imageZ->imagePnts <= 0) {
        psktm__set_error("%s: invalid sizes", path);
        return PSKTM_E_IO;
    }
    return PSKTM_OK;
}
