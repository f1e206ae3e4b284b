/*
 * psktm_oracle.c - CPU oracle for the PSTM migration path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, the smoke check
 * and the CPU-baseline leg of the benchmark load it; the product library
 * (libpsktm.so) never links, loads or calls it.
 *
 * Contents
 *   oracle_ktm_migration_cpu   restatement of ktmMigrationCPU
 *                              (src/psktmCPU.c:13-101) with the reference's
 *                              binary32 operation order; parallel over image
 *                              x so every image sample still receives its
 *                              trace contributions in k order: bit-identical
 *                              to the single-thread reference loop, free of
 *                              the reference's OpenMP race (src/psktmCPU.c:56,84).
 *   oracle_pstm_f32            the same arithmetic with the extensions this
 *                              build adds (aperture, v(x,y,tau), trace range).
 *   oracle_ktm_migration_racy  the reference's trace-parallel loop structure
 *                              as written (racy; timing only).
 *   oracle_pstm_f64            FP64 forward (TRUNC or LINEAR interpolation;
 *                              per-update weight or gained traces).
 *   oracle_pstm_adj_f64        FP64 adjoint (demigration) of the same operator.
 *   oracle_init_three_axes     restatement of initThreeAxes
 *                              (src/psktmFunctions.c:19-39).
 *   oracle_ibm32_to_ieee_ref   restatement of ktmIBM32toIEEE
 *                              (src/ioFunctions.c:4-40).
 *
 * Pins (what checks each function beyond its own formula) are listed in the
 * tests: golden-free closed forms (single-spike support and amplitude,
 * zero-offset t = 2 tau, weight identity), invariants (linearity, source-
 * receiver symmetry bit-exact, zero data, sample 0 never read), FP64 brute
 * force written from source/receiver positions, FP64 dot-product test for the
 * adjoint, and known IBM words.
 *
 * Build: gcc -O3 -fopenmp -ffp-contract=off (no fast-math). The reference's
 * arithmetic is binary32 with separate rounding of every + - * / and
 * correctly rounded sqrtf; contraction into FMA would change it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "../include/psktm.h"

/* Contraction is disabled by -ffp-contract=off on the compile line. */

/* initThreeAxes (src/psktmFunctions.c:28-34): start + (float)i * step. */
static void axis_f32(int n, float start, float step, float *out)
{
    for (int i = 0; i < n; i++) {
        float fi = (float)i;
        float prod = fi * step;
        out[i] = start + prod;
    }
}

void oracle_init_three_axes(const struct imageGrid *X, const struct imageGrid *Y,
                            const struct imageGrid *Z, float *xx, float *yy,
                            float *zz)
{
    axis_f32(X->imagePnts, X->imageStart, X->imageStep, xx);
    axis_f32(Y->imagePnts, Y->imageStart, Y->imageStep, yy);
    axis_f32(Z->imagePnts, Z->imageStart, Z->imageStep, zz);
}

/* Aperture rule shared with the HIP kernels: binary32, no contraction,
 * d2 = (mx-x)^2 + (my-y)^2 <= r2, r2 = R*R rounded to float. */
static inline int in_aperture(float mx, float my, float x, float y, float r2)
{
    float ex = mx - x;
    float ey = my - y;
    float ex2 = ex * ex;
    float ey2 = ey * ey;
    float d2 = ex2 + ey2;
    return d2 <= r2;
}

/*
 * Binary32 migration with the reference operation order
 * (src/psktmCPU.c:42-87):
 *   tauS[l] = zz[l]*zz[l]; sS = 1.0f/(v*v)
 *   tempC = ((mx-x)-hx)*((mx-x)-hx); tempD = ((mx-x)+hx)*((mx-x)+hx)
 *   tempA = tempC + ((my-y)-hy)*((my-y)-hy); tempB = tempD + (...)
 *   t = sqrtf(tauS + tempA*sS); t += sqrtf(tauS + tempB*sS)
 *   n = (int)(t/dt); if (n < ns && n > 0) img += (d[n]*t)*sqrtf(tauS/t)
 * vel3d = 0: vel[nt] (the reference); vel3d = 1: vel[(i*ny+j)*nt+l].
 * aperture <= 0: every trace updates every bin (the reference).
 * Traces k0..k1-1 are used, in increasing k.
 */
void oracle_pstm_f32(int nx, float x0, float dx, int ny, float y0, float dy,
                     int nt, float tau0, float dtau, int k0, int k1, int ns,
                     float dt, const float *midX, const float *midY,
                     const float *offX, const float *offY, const float *traces,
                     const float *vel, int vel3d, float aperture, float *image,
                     int nthreads)
{
    float *xx = (float *)malloc(sizeof(float) * (size_t)nx);
    float *yy = (float *)malloc(sizeof(float) * (size_t)ny);
    float *zz = (float *)malloc(sizeof(float) * (size_t)nt);
    float *tauS = (float *)malloc(sizeof(float) * (size_t)nt);
    axis_f32(nx, x0, dx, xx);
    axis_f32(ny, y0, dy, yy);
    axis_f32(nt, tau0, dtau, zz);
    for (int l = 0; l < nt; l++) tauS[l] = zz[l] * zz[l];

    size_t nsl = vel3d ? (size_t)nx * ny * nt : (size_t)nt;
    float *sS = (float *)malloc(sizeof(float) * nsl);
    for (size_t q = 0; q < nsl; q++) {
        float v2 = vel[q] * vel[q];
        sS[q] = 1.0f / v2;
    }
    const int use_ap = aperture > 0.0f;
    const float r2 = aperture * aperture;

    memset(image, 0, sizeof(float) * (size_t)nx * ny * nt);
    if (nthreads > 0) omp_set_num_threads(nthreads);

#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nx; i++) {
        for (int k = k0; k < k1; k++) {
            const float mx = midX[k], my = midY[k], hx = offX[k], hy = offY[k];
            const float *tr = traces + (size_t)k * ns;
            float ax = (mx - xx[i]) - hx;
            float bx = (mx - xx[i]) + hx;
            float tempC = ax * ax;
            float tempD = bx * bx;
            for (int j = 0; j < ny; j++) {
                if (use_ap && !in_aperture(mx, my, xx[i], yy[j], r2)) continue;
                float ay = (my - yy[j]) - hy;
                float by = (my - yy[j]) + hy;
                float ay2 = ay * ay, by2 = by * by;
                float tempA = tempC + ay2;
                float tempB = tempD + by2;
                float *img = image + ((size_t)i * ny + j) * nt;
                const float *s = vel3d ? sS + ((size_t)i * ny + j) * nt : sS;
                for (int l = 0; l < nt; l++) {
                    float pa = tempA * s[l];
                    float pb = tempB * s[l];
                    float t = sqrtf(tauS[l] + pa);
                    t += sqrtf(tauS[l] + pb);
                    float qn = t / dt;
                    /* (int)qn lies in (0, ns) exactly when 1 <= qn < ns; the
                     * guard keeps the conversion defined for inf and NaN. */
                    if (qn >= 1.0f && qn < (float)ns) {
                        int n = (int)qn;
                        float dtm = tr[n] * t;
                        float q = tauS[l] / t;
                        img[l] += dtm * sqrtf(q);
                    }
                }
            }
        }
    }
    free(xx); free(yy); free(zz); free(tauS); free(sS);
}

/* ktmMigrationCPU with the reference signature (src/psktmCPU.c:13-23). */
void oracle_ktm_migration_cpu(struct imageGrid *imageX, struct imageGrid *imageY,
                              struct imageGrid *imageZ, struct jobParams *config,
                              float *midX, float *midY, float *offX, float *offY,
                              float *traces, float *slowness, float *image,
                              int nthreads)
{
    oracle_pstm_f32(imageX->imagePnts, imageX->imageStart, imageX->imageStep,
                    imageY->imagePnts, imageY->imageStart, imageY->imageStep,
                    imageZ->imagePnts, imageZ->imageStart, imageZ->imageStep,
                    0, config->traceNo, config->tracePts,
                    config->traceSampleRate, midX, midY, offX, offY, traces,
                    slowness, 0, 0.0f, image, nthreads);
}

/* The reference's loop structure as written (src/psktmCPU.c:56-91): OpenMP
 * over traces with unsynchronised image updates. Timing only: with more than
 * one thread its result is not deterministic. */
void oracle_ktm_migration_racy(int nx, float x0, float dx, int ny, float y0,
                               float dy, int nt, float tau0, float dtau,
                               int ntr, int ns, float dt, const float *midX,
                               const float *midY, const float *offX,
                               const float *offY, const float *traces,
                               const float *vel, float *image, int nthreads)
{
    float *xx = (float *)malloc(sizeof(float) * (size_t)nx);
    float *yy = (float *)malloc(sizeof(float) * (size_t)ny);
    float *zz = (float *)malloc(sizeof(float) * (size_t)nt);
    float *tauS = (float *)malloc(sizeof(float) * (size_t)nt);
    float *sS = (float *)malloc(sizeof(float) * (size_t)nt);
    axis_f32(nx, x0, dx, xx);
    axis_f32(ny, y0, dy, yy);
    axis_f32(nt, tau0, dtau, zz);
    for (int l = 0; l < nt; l++) {
        tauS[l] = zz[l] * zz[l];
        float v2 = vel[l] * vel[l];
        sS[l] = 1.0f / v2;
    }
    memset(image, 0, sizeof(float) * (size_t)nx * ny * nt);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for
    for (int k = 0; k < ntr; k++) {
        const float *tr = traces + (size_t)k * ns;
        for (int i = 0; i < nx; i++) {
            float ax = (midX[k] - xx[i]) - offX[k];
            float bx = (midX[k] - xx[i]) + offX[k];
            float tempC = ax * ax, tempD = bx * bx;
            for (int j = 0; j < ny; j++) {
                float ay = (midY[k] - yy[j]) - offY[k];
                float by = (midY[k] - yy[j]) + offY[k];
                float ay2 = ay * ay, by2 = by * by;
                float tempA = tempC + ay2, tempB = tempD + by2;
                float *img = image + ((size_t)i * ny + j) * nt;
                for (int l = 0; l < nt; l++) {
                    float pa = tempA * sS[l], pb = tempB * sS[l];
                    float t = sqrtf(tauS[l] + pa);
                    t += sqrtf(tauS[l] + pb);
                    float qn = t / dt;
                    /* (int)qn lies in (0, ns) exactly when 1 <= qn < ns; the
                     * guard keeps the conversion defined for inf and NaN. */
                    if (qn >= 1.0f && qn < (float)ns) {
                        int n = (int)qn;
                        float dtm = tr[n] * t;
                        float q = tauS[l] / t;
                        img[l] += dtm * sqrtf(q);
                    }
                }
            }
        }
    }
    free(xx); free(yy); free(zz); free(tauS); free(sS);
}

/* ------------------------------------------------------------------------ */
/* FP64 operator: forward (migration) and adjoint (demigration).             */
/* ------------------------------------------------------------------------ */

/* One update in FP64: traveltime, sample position and weight. Returns 0 when
 * the sample is outside the trace for the interpolation rule.
 *   TRUNC : n = floor(t/dt), valid iff 0 < n < ns, weights (1, 0)
 *   LINEAR: n = floor(t/dt), f = t/dt - n, valid iff 0 <= n and n+1 < ns,
 *           weights (1-f, f) on samples n, n+1
 *   amplitude weight t*sqrt(tau^2/t) = tau*sqrt(t) (src/psktmCPU.c:84-85);
 *   gained != 0: tau only (the traces carry sqrt(n dt) per sample). */
static inline int update_f64(double tau, double s, double A, double B,
                             double dt, int ns, int interp, int gained, long *n_out,
                             double *w0, double *w1)
{
    double tau2 = tau * tau;
    double t = sqrt(tau2 + A * s) + sqrt(tau2 + B * s);
    double xi = t / dt;
    if (!(xi >= 0.0) || xi >= (double)ns) return 0;
    long n = (long)floor(xi);
    double amp = gained ? tau : ((t > 0.0) ? t * sqrt(tau2 / t) : 0.0);
    if (interp == PSKTM_INTERP_TRUNC) {
        if (!(n > 0 && n < ns)) return 0;
        *n_out = n; *w0 = amp; *w1 = 0.0;
        return 1;
    }
    if (!(n >= 0 && n + 1 < ns)) return 0;
    double f = xi - (double)n;
    *n_out = n; *w0 = amp * (1.0 - f); *w1 = amp * f;
    return 1;
}

static double *slowness_f64(int nx, int ny, int nt, const float *vel, int vel3d)
{
    size_t nsl = vel3d ? (size_t)nx * ny * nt : (size_t)nt;
    double *s = (double *)malloc(sizeof(double) * nsl);
    for (size_t q = 0; q < nsl; q++) {
        double v = (double)vel[q];
        s[q] = 1.0 / (v * v);
    }
    return s;
}

/*
 * FP64 migration of traces k in [0, ntr) into the bins [ix0,ix1) x [iy0,iy1)
 * (all nt samples). out[((i-ix0)*(iy1-iy0) + (j-iy0))*nt + l].
 * Bin coordinates and tau are the binary32 axis values (as the reference
 * computes them) promoted to double; everything after that is FP64. The
 * aperture decision uses the shared binary32 rule so that FP64 and GPU agree
 * on which (trace, bin) pairs exist. traces may be scaled per trace by
 * scale[k] (NULL = 1).
 */
/* One image column (bin i, j) of oracle_pstm_f64 into o[nt]. */
static void pstm_f64_column(int i, int j, int ny, int nt, const float *xx, const float *yy,
                            const float *zz, const double *s, int vel3d, int use_ap, float r2,
                            int ntr, int ns, float dt, const float *midX, const float *midY,
                            const float *offX, const float *offY, const float *traces,
                            const float *scale, int interp, int gained, double *o)
{
    const double x = xx[i], y = yy[j];
    const double *sb = vel3d ? s + ((size_t)i * ny + j) * nt : s;
    for (int k = 0; k < ntr; k++) {
        if (use_ap && !in_aperture(midX[k], midY[k], xx[i], yy[j], r2)) continue;
        const double ex = (double)midX[k] - x, ey = (double)midY[k] - y;
        const double hx = offX[k], hy = offY[k];
        const double A = (ex - hx) * (ex - hx) + (ey - hy) * (ey - hy);
        const double B = (ex + hx) * (ex + hx) + (ey + hy) * (ey + hy);
        const float *tr = traces + (size_t)k * ns;
        const double sc = scale ? (double)scale[k] : 1.0;
        for (int l = 0; l < nt; l++) {
            long n; double w0, w1;
            if (!update_f64(zz[l], sb[l], A, B, dt, ns, interp, gained, &n, &w0, &w1))
                continue;
            double d = w0 * (double)tr[n];
            if (w1 != 0.0) d += w1 * (double)tr[n + 1];
            o[l] += sc * d;
        }
    }
}

/* oracle_pstm_f64 on an explicit list of bins (bin_i[b], bin_j[b]),
 * out[b][nt]: scattered columns (edges, corners, interior) in one parallel
 * call. */
void oracle_pstm_f64_bins(int nx, float x0, float dx, int ny, float y0, float dy,
                          int nt, float tau0, float dtau, int ntr, int ns, float dt,
                          const float *midX, const float *midY, const float *offX,
                          const float *offY, const float *traces, const float *scale,
                          const float *vel, int vel3d, float aperture, int interp,
                          int gained, const int *bin_i, const int *bin_j, int nbins,
                          double *out, int nthreads)
{
    float *xx = (float *)malloc(sizeof(float) * (size_t)nx);
    float *yy = (float *)malloc(sizeof(float) * (size_t)ny);
    float *zz = (float *)malloc(sizeof(float) * (size_t)nt);
    axis_f32(nx, x0, dx, xx);
    axis_f32(ny, y0, dy, yy);
    axis_f32(nt, tau0, dtau, zz);
    double *s = slowness_f64(nx, ny, nt, vel, vel3d);
    const int use_ap = aperture > 0.0f;
    const float r2 = aperture * aperture;
    memset(out, 0, sizeof(double) * (size_t)nbins * nt);
    if (nthreads > 0) omp_set_num_threads(nthreads);

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbins; b++)
        pstm_f64_column(bin_i[b], bin_j[b], ny, nt, xx, yy, zz, s, vel3d, use_ap, r2, ntr, ns,
                        dt, midX, midY, offX, offY, traces, scale, interp, gained,
                        out + (size_t)b * nt);
    free(xx); free(yy); free(zz); free(s);
}

void oracle_pstm_f64(int nx, float x0, float dx, int ny, float y0, float dy,
                     int nt, float tau0, float dtau, int ntr, int ns, float dt,
                     const float *midX, const float *midY, const float *offX,
                     const float *offY, const float *traces, const float *scale,
                     const float *vel, int vel3d, float aperture, int interp,
                     int gained, int ix0, int ix1, int iy0, int iy1, double *out,
                     int nthreads)
{
    float *xx = (float *)malloc(sizeof(float) * (size_t)nx);
    float *yy = (float *)malloc(sizeof(float) * (size_t)ny);
    float *zz = (float *)malloc(sizeof(float) * (size_t)nt);
    axis_f32(nx, x0, dx, xx);
    axis_f32(ny, y0, dy, yy);
    axis_f32(nt, tau0, dtau, zz);
    double *s = slowness_f64(nx, ny, nt, vel, vel3d);
    const int use_ap = aperture > 0.0f;
    const float r2 = aperture * aperture;
    const int sy = iy1 - iy0;
    const int nbins = (ix1 - ix0) * sy;
    memset(out, 0, sizeof(double) * (size_t)nbins * nt);
    if (nthreads > 0) omp_set_num_threads(nthreads);

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbins; b++)
        pstm_f64_column(ix0 + b / sy, iy0 + b % sy, ny, nt, xx, yy, zz, s, vel3d, use_ap, r2,
                        ntr, ns, dt, midX, midY, offX, offY, traces, scale, interp, gained,
                        out + (size_t)b * nt);
    free(xx); free(yy); free(zz); free(s);
}

/*
 * FP64 adjoint of oracle_pstm_f64 over the full grid: traces_out[k][n] =
 * sum over in-aperture bins and tau of the same weights times image.
 */
void oracle_pstm_adj_f64(int nx, float x0, float dx, int ny, float y0, float dy,
                         int nt, float tau0, float dtau, int ntr, int ns,
                         float dt, const float *midX, const float *midY,
                         const float *offX, const float *offY,
                         const float *image, const float *vel, int vel3d,
                         float aperture, int interp, double *traces_out,
                         int nthreads)
{
    float *xx = (float *)malloc(sizeof(float) * (size_t)nx);
    float *yy = (float *)malloc(sizeof(float) * (size_t)ny);
    float *zz = (float *)malloc(sizeof(float) * (size_t)nt);
    axis_f32(nx, x0, dx, xx);
    axis_f32(ny, y0, dy, yy);
    axis_f32(nt, tau0, dtau, zz);
    double *s = slowness_f64(nx, ny, nt, vel, vel3d);
    const int use_ap = aperture > 0.0f;
    const float r2 = aperture * aperture;
    memset(traces_out, 0, sizeof(double) * (size_t)ntr * ns);
    if (nthreads > 0) omp_set_num_threads(nthreads);

#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < ntr; k++) {
        double *tr = traces_out + (size_t)k * ns;
        const double hx = offX[k], hy = offY[k];
        for (int i = 0; i < nx; i++) {
            for (int j = 0; j < ny; j++) {
                if (use_ap && !in_aperture(midX[k], midY[k], xx[i], yy[j], r2)) continue;
                const double ex = (double)midX[k] - (double)xx[i];
                const double ey = (double)midY[k] - (double)yy[j];
                const double A = (ex - hx) * (ex - hx) + (ey - hy) * (ey - hy);
                const double B = (ex + hx) * (ex + hx) + (ey + hy) * (ey + hy);
                const float *im = image + ((size_t)i * ny + j) * nt;
                const double *sb = vel3d ? s + ((size_t)i * ny + j) * nt : s;
                for (int l = 0; l < nt; l++) {
                    long n; double w0, w1;
                    if (!update_f64(zz[l], sb[l], A, B, dt, ns, interp, 0, &n, &w0, &w1))
                        continue;
                    tr[n] += w0 * (double)im[l];
                    if (w1 != 0.0) tr[n + 1] += w1 * (double)im[l];
                }
            }
        }
    }
    free(xx); free(yy); free(zz); free(s);
}

/* ktmIBM32toIEEE (src/ioFunctions.c:18-38): sign*(fraction/0xFFFFFF)*16^e. */
void oracle_ibm32_to_ieee_ref(const uint32_t *input, float *output, int n)
{
    for (int i = 0; i < n; i++) {
        float sign = (-2.0f * (float)((input[i] & SIGN_MASK) > 0)) + 1.0f;
        uint32_t exponent = input[i] & EXPONENT_MASK;
        uint32_t fraction = input[i] & FRACTION_MASK;
        float e = (float)((exponent >> 24) & 127u) - 64.0f;
        float q = (float)fraction / (float)FRACTION_MASK;
        output[i] = sign * q * powf(16.0f, e);
    }
}
