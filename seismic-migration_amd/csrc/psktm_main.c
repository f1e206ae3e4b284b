/*
 * psktm_main - the file-driven PSTM job the reference implies but never
 * ships: read the on-disk job written by the MATLAB/Octave block writers,
 * migrate it on the GPU, write the image.
 *
 *   inp.txt            12 floats (scripts/ktmWriteTraceBlockConstOffset.m:154-167)
 *   inputData.bin      traces, trace-major, IBM words (or floats for impulse
 *                      data, --float) (:128, :141-147)
 *   input{Src,Rec}{X,Y}.bin   float per trace (:136-152)
 *   velModel.bin       float per output sample (scripts/ktmWriteVelModel.m)
 *   outputData.bin     float image, tau fastest, then y, then x
 *                      (scripts/ktmVisualizeOutputFile.m:16-28)
 *
 * Steps follow the reference's API: loadInputData (src/ioFunctions.c:94-153;
 * used as is when the data directory is the reference's ../data),
 * fileSizeFourBytes, binReadUints + ktmIBM32toIEEE or binReadFloats, midpoint
 * and half offset from source/receiver (scripts/ktmConvertXYtoOffMid.m:12),
 * ktmMigrationGPU (or ktmMigrationCPU with --cpu-entry: the same image),
 * fileWrite.
 *
 * usage: psktm_main [--data DIR] [--float] [--v0 V] [--out FILE] [--cpu-entry]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/psktm.h"

static int fail(const char *what)
{
    const char *e = psktm_last_error();
    fprintf(stderr, "psktm_main: %s%s%s\n", what, (e && e[0]) ? ": " : "", e ? e : "");
    return 1;
}

static FILE *open_in(const char *dir, const char *name, int *flag)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) {
        errorMessage(path);
        *flag = 1;
    }
    return f;
}

int main(int argc, char **argv)
{
    const char *dir = "../data", *out = NULL;
    int float_mode = 0, cpu_entry = 0;
    float v0 = 0.0f;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--data") && i + 1 < argc) dir = argv[++i];
        else if (!strcmp(argv[i], "--float")) float_mode = 1;
        else if (!strcmp(argv[i], "--v0") && i + 1 < argc) v0 = (float)atof(argv[++i]);
        else if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "--cpu-entry")) cpu_entry = 1;
        else {
            fprintf(stderr, "usage: %s [--data DIR] [--float] [--v0 V] [--out FILE] [--cpu-entry]\n",
                    argv[0]);
            return 2;
        }
    }

    char path[4096];
    struct imageGrid X, Y, Z;
    struct jobParams job;
    snprintf(path, sizeof(path), "%s/inp.txt", dir);
    if (psktm_load_ktm_config(path, &X, &Y, &Z, &job) != PSKTM_OK) return fail("config");

    struct fileInputPtrs in;
    memset(&in, 0, sizeof(in));
    if (!strcmp(dir, "../data")) {
        loadInputData(&in);
    } else {
        in.flag = 0;
        in.inputData = open_in(dir, "inputData.bin", &in.flag);
        in.inputSrcX = open_in(dir, "inputSrcX.bin", &in.flag);
        in.inputSrcY = open_in(dir, "inputSrcY.bin", &in.flag);
        in.inputRecX = open_in(dir, "inputRecX.bin", &in.flag);
        in.inputRecY = open_in(dir, "inputRecY.bin", &in.flag);
    }
    if (in.flag) return fail("input files");

    const int ntr = job.traceNo, ns = job.tracePts;
    const long long nsamp = (long long)ntr * ns;
    if (nsamp <= 0 || nsamp > 0x7fffffffLL) return fail("trace count x samples out of range");
    if (fileSizeFourBytes(in.inputData) < nsamp) return fail("inputData.bin holds fewer samples than inp.txt says");
    float *traces = NULL;
    if (float_mode) {
        traces = binReadFloats(in.inputData, (int)nsamp);
    } else {
        unsigned int *w = binReadUints(in.inputData, (int)nsamp);
        traces = (float *)malloc(sizeof(float) * (size_t)nsamp);
        if (w && traces) ktmIBM32toIEEE(w, traces, (int)nsamp);
        free(w);
    }
    float *sx = binReadFloats(in.inputSrcX, ntr), *sy = binReadFloats(in.inputSrcY, ntr);
    float *rx = binReadFloats(in.inputRecX, ntr), *ry = binReadFloats(in.inputRecY, ntr);
    float *mx = (float *)malloc(sizeof(float) * (size_t)ntr);
    float *my = (float *)malloc(sizeof(float) * (size_t)ntr);
    float *hx = (float *)malloc(sizeof(float) * (size_t)ntr);
    float *hy = (float *)malloc(sizeof(float) * (size_t)ntr);
    if (!traces || !sx || !sy || !rx || !ry || !mx || !my || !hx || !hy) return fail("out of host memory");
    for (int k = 0; k < ntr; k++) {        /* scripts/ktmConvertXYtoOffMid.m:12 */
        mx[k] = (rx[k] + sx[k]) / 2.0f;
        my[k] = (ry[k] + sy[k]) / 2.0f;
        hx[k] = (rx[k] - sx[k]) / 2.0f;
        hy[k] = (ry[k] - sy[k]) / 2.0f;
    }

    const int nt = Z.imagePnts;
    float *vel = NULL;
    snprintf(path, sizeof(path), "%s/velModel.bin", dir);
    FILE *fv = fopen(path, "rb");
    if (fv) {
        if (fileSizeFourBytes(fv) < nt) return fail("velModel.bin holds fewer values than output samples");
        vel = binReadFloats(fv, nt);
        fclose(fv);
    } else if (v0 > 0.0f) {
        vel = (float *)malloc(sizeof(float) * (size_t)nt);
        for (int l = 0; vel && l < nt; l++) vel[l] = v0;
    } else {
        return fail("no velModel.bin and no --v0");
    }

    /* the config loader leaves positive sizes; their product must fit in
     * size_t bytes (2^40 samples = 4 TiB, far past any host image) */
    if ((double)X.imagePnts * (double)Y.imagePnts * (double)Z.imagePnts > 1099511627776.0)
        return fail("image size out of range");
    const size_t nimg = (size_t)X.imagePnts * Y.imagePnts * Z.imagePnts;
    float *image = (float *)malloc(sizeof(float) * nimg);
    if (!image || !vel) return fail("out of host memory");
    if (cpu_entry)
        ktmMigrationCPU(&X, &Y, &Z, &job, mx, my, hx, hy, traces, vel, image);
    else
        ktmMigrationGPU(&X, &Y, &Z, &job, mx, my, hx, hy, traces, vel, image);
    if (psktm_last_error()[0]) return fail("migration");

    if (out) snprintf(path, sizeof(path), "%s", out);
    else snprintf(path, sizeof(path), "%s/outputData.bin", dir);
    FILE *fo = fopen(path, "wb");
    if (!fo) return fail("cannot create the output file");
    fileWrite(fo, image, (int)nimg);
    fclose(fo);
    fprintf(stderr, "psktm_main: %d traces x %d samples -> %d x %d x %d image written to %s\n",
            ntr, ns, X.imagePnts, Y.imagePnts, Z.imagePnts, path);
    fclose(in.inputData); fclose(in.inputSrcX); fclose(in.inputSrcY);
    fclose(in.inputRecX); fclose(in.inputRecY);
    free(traces); free(sx); free(sy); free(rx); free(ry);
    free(mx); free(my); free(hx); free(hy); free(vel); free(image);
    return 0;
}
