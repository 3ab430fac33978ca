/*
 * zola_beads.c -- FP64 CPU restatement of the calibration bead preparation.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle library, like zola_oracle.c):
 * the parity tests and the CPU-baseline leg of the benchmark call it; the
 * product never links or loads it.
 *
 * Restates (paths relative to src/main/java/org/pasteur/imagej/):
 *   zo_bead_convert        process/InitBackgroundAndPhotonNumber.java:50-77
 *   zo_bead_convert_scmos  process/InitBackgroundAndPhotonNumber.java:79-97
 *   zo_bead_prepare        process/InitBackgroundAndPhotonNumber.java:99-468
 *     refine_points        process/InitBackgroundAndPhotonNumber.java:563-618
 *     compute_background   process/InitBackgroundAndPhotonNumber.java:622-904
 *     poly2_fit            utils/PolynomialFit.java:39-60,92-153,315-353,557-606
 *     poly2_eval           utils/PolynomialFit.java:191-222
 *     gauss_lsolve         utils/GaussianElimination.java (lsolve(double[][], double[]))
 * Every sum runs in the reference's loop order, in FP64, with no contraction
 * (the Makefile builds ISO C99, which keeps a*b+c as two roundings).
 *
 * Layout: image [z][i][ii] with i the ImageJ x column (W) and ii the row (H),
 * as the reference's double[S][W][H].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "zola_oracle.h"

void zo_bead_convert(long n, double* image, double adu, double gain, double offset) {
    const double slope = adu / gain;
    const double intercept = -offset * adu / gain;
    for (long k = 0; k < n; k++) {
        double v = (image[k] * slope + intercept) + 1;
        image[k] = v < 0 ? 0 : v;
    }
}

void zo_bead_convert_scmos(int S, int W, int H, double* image, const double* offset, const double* gain) {
    const long wh = (long)W * H;
    for (int z = 0; z < S; z++)
        for (long k = 0; k < wh; k++) {
            double v = (image[z * wh + k] - offset[k]) / gain[k];
            image[z * wh + k] = v < 0 ? 0 : v;
        }
}

static int cmp_double(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Seed refinement: z-summed patch, 90th-percentile threshold, intensity
 * centroid over the inner half, moved only by less than P/4 per axis. */
static void refine_points(int S, int W, int H, const double* image, int P, int nb, float* xp, float* yp) {
    double* sub = (double*)malloc(sizeof(double) * P * P);
    double* vect = (double*)malloc(sizeof(double) * P * P);
    const long wh = (long)W * H;
    for (int p = 0; p < nb; p++) {
        const int x0 = (int)xp[p] - P / 2, y0 = (int)yp[p] - P / 2;
        memset(sub, 0, sizeof(double) * P * P);
        for (int i = 0; i < S; i++)
            for (int u = 0; u < P; u++)
                for (int uu = 0; uu < P; uu++)
                    if (x0 + u >= 0 && x0 + u < W && y0 + uu >= 0 && y0 + uu < H)
                        sub[u * P + uu] += image[i * wh + (long)(x0 + u) * H + (y0 + uu)];
        memcpy(vect, sub, sizeof(double) * P * P);
        qsort(vect, (size_t)P * P, sizeof(double), cmp_double);
        const double thr = vect[(int)((P * P - 1) * .9)];
        double mx = 0, my = 0, sum = 0;
        int count = 0;
        for (int u = P / 4; u < 3 * P / 4; u++)
            for (int uu = P / 4; uu < 3 * P / 4; uu++)
                if (sub[u * P + uu] > thr) {
                    mx += (double)(x0 + u) * sub[u * P + uu];
                    my += (double)(y0 + uu) * sub[u * P + uu];
                    sum += sub[u * P + uu];
                    count++;
                }
        if (count >= 1) {
            const float cx = (float)(mx / sum), cy = (float)(my / sum);
            const float dx = xp[p] - cx, dy = yp[p] - cy;
            const float dx2 = dx * dx, dy2 = dy * dy;
            if (sqrt((double)dx2) < (double)(P / 4)) xp[p] = cx;
            if (sqrt((double)dy2) < (double)(P / 4)) yp[p] = cy;
        }
    }
    free(sub);
    free(vect);
}

/* Monomial exponents of the order-2 fit in two variables, in the order
 * PolynomialFit.constructIndices enumerates them: u = e0 + 3*e1 for u = 0..8,
 * kept when e0 + e1 < 3. */
static const int kE0[6] = {0, 1, 2, 0, 1, 0};
static const int kE1[6] = {0, 0, 0, 1, 1, 2};

/* GaussianElimination.lsolve: partial pivoting (first row with the strictly
 * largest |a|), rows swapped with b, singular when |pivot| <= 1e-20. */
static int gauss_lsolve(int n, double* a, double* b, double* x) {
    for (int p = 0; p < n; p++) {
        int mx = p;
        for (int i = p + 1; i < n; i++)
            if (fabs(a[i * n + p]) > fabs(a[mx * n + p])) mx = i;
        for (int j = 0; j < n; j++) {
            const double t = a[p * n + j];
            a[p * n + j] = a[mx * n + j];
            a[mx * n + j] = t;
        }
        const double t = b[p];
        b[p] = b[mx];
        b[mx] = t;
        if (fabs(a[p * n + p]) <= 1e-20) return -1;
        for (int i = p + 1; i < n; i++) {
            const double alpha = a[i * n + p] / a[p * n + p];
            b[i] -= alpha * b[p];
            for (int j = p; j < n; j++) a[i * n + j] -= alpha * a[p * n + j];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0.0;
        for (int j = i + 1; j < n; j++) sum += a[i * n + j] * x[j];
        x[i] = (b[i] - sum) / a[i * n + i];
    }
    return 0;
}

/* PolynomialFit(2, {y, y}, {x0, x1}).run(): normal equations summed over the
 * data in order, then lsolve.  Both output dimensions carry the same data,
 * so one solve gives both coefficient vectors. */
static int poly2_fit(long nd, const double* x0, const double* x1, const double* y, double* coef) {
    double XX[36], YY[6];
    for (int i = 0; i < 6; i++) {
        for (int ai = 0; ai < 6; ai++) {
            double acc = 0;
            for (long k = 0; k < nd; k++) {
                double pl = 1;
                pl *= pow(x0[k], kE0[i]);
                pl *= pow(x1[k], kE1[i]);
                double pr = pl;
                pr *= pow(x0[k], kE0[ai]);
                pr *= pow(x1[k], kE1[ai]);
                acc += pr;
            }
            XX[i * 6 + ai] = acc;
        }
        double acc = 0;
        for (long k = 0; k < nd; k++) {
            double pr = 1;
            pr *= pow(x0[k], kE0[i]);
            pr *= pow(x1[k], kE1[i]);
            pr *= y[k];
            acc += pr;
        }
        YY[i] = acc;
    }
    return gauss_lsolve(6, XX, YY, coef);
}

static double poly2_eval(const double* coef, double v0, double v1) {
    double res = 0;
    for (int i = 0; i < 6; i++) {
        double prod = coef[i];
        prod *= pow(v0, kE0[i]);
        prod *= pow(v1, kE1[i]);
        res += prod;
    }
    return res;
}

/* computeBackground: stack moments, the high-intensity mask is2, the slice of
 * lowest filtered mean, the dilated mask is3 around it, per-pixel means over
 * is3 and the order-2 polynomial background.  info = {miniposition, ndata}. */
static int compute_background(int S, int W, int H, const double* image, int sF, double* poly, double* coef,
                              double* stats, int* info) {
    const long wh = (long)W * H, n = (long)S * wh;
    const int nint = S * W * H;  /* the reference divides by an int product */
    double mean0 = 0, std0 = 0;
    for (long k = 0; k < n; k++) mean0 += image[k];
    mean0 /= nint;
    for (long k = 0; k < n; k++) std0 += (image[k] - mean0) * (image[k] - mean0);
    std0 /= nint;
    std0 = sqrt(std0);
    double mean = 0, std = 0;
    int countMean0 = 0;
    for (long k = 0; k < n; k++)
        if (image[k] <= mean0 + 2 * std0) {
            mean += image[k];
            countMean0++;
        }
    if (countMean0 > 0) mean /= countMean0;
    for (long k = 0; k < n; k++)
        if (image[k] <= mean0 + 3 * std0) std += (image[k] - mean) * (image[k] - mean);
    if (countMean0 > 0) {
        std /= countMean0;
        std = sqrt(std);
    }
    if (stats) {
        stats[0] = mean0;
        stats[1] = std0;
        stats[2] = mean;
        stats[3] = std;
    }
    unsigned char* is2 = (unsigned char*)malloc((size_t)n);
    for (long k = 0; k < n; k++) is2[k] = image[k] > mean + 3 * std ? 0 : 1;
    double* meanStack = (double*)calloc((size_t)S, sizeof(double));
    double* filtMean = (double*)calloc((size_t)S, sizeof(double));
    for (int i = 0; i < S; i++) {
        int count = 0;
        for (long k = 0; k < wh; k++)
            if (is2[i * wh + k]) {
                meanStack[i] += image[i * wh + k];
                count++;
            }
        if (count != 0) meanStack[i] /= count;
    }
    int mp = 0;
    for (int i = 0; i < S; i++) {
        int count = 0;
        for (int u = -sF; u <= sF; u++)
            if (i + u >= 0 && i + u < S) {
                filtMean[i] += meanStack[i + u];
                count++;
            }
        if (count != 0) filtMean[i] /= count;
        if (filtMean[i] < filtMean[mp]) mp = i;
    }
    const int nN = sF;
    if (mp < nN) mp = nN;
    if (mp >= S - nN) mp = S - nN - 1;
    const int sD = 2;
    double* xa = (double*)malloc(sizeof(double) * wh);
    double* xb = (double*)malloc(sizeof(double) * wh);
    double* ym = (double*)malloc(sizeof(double) * wh);
    unsigned char* is3col = (unsigned char*)malloc((size_t)S);
    long nd = 0;
    for (int ii = 0; ii < W; ii++)
        for (int iii = 0; iii < H; iii++) {
            int count = 0;
            for (int i = 0; i < S; i++) {
                int ok = 1;
                if (i < mp - nN || i > mp + nN) ok = 0;
                else
                    for (int u = -sD; u <= sD && ok; u++)
                        for (int uu = -sD; uu <= sD && ok; uu++)
                            for (int uuu = -sD; uuu <= sD && ok; uuu++)
                                if (i + u >= 0 && i + u < S && ii + uu >= 0 && ii + uu < W && iii + uuu >= 0 &&
                                    iii + uuu < H && !is2[(i + u) * wh + (long)(ii + uu) * H + (iii + uuu)])
                                    ok = 0;
                is3col[i] = (unsigned char)ok;
                count += ok;
            }
            if (count >= 1) {
                double m = 0;
                for (int i = 0; i < S; i++)
                    if (is3col[i]) m += image[i * wh + (long)ii * H + iii];
                m /= count;
                xa[nd] = ii;
                xb[nd] = iii;
                ym[nd] = m;
                nd++;
            }
        }
    int rc = poly2_fit(nd, xa, xb, ym, coef);
    if (rc == 0)
        for (int ii = 0; ii < W; ii++)
            for (int iii = 0; iii < H; iii++) poly[(long)ii * H + iii] = poly2_eval(coef, ii, iii);
    if (info) {
        info[0] = mp;
        info[1] = (int)nd;
    }
    free(is2);
    free(meanStack);
    free(filtMean);
    free(xa);
    free(xb);
    free(ym);
    free(is3col);
    return rc;
}

int zo_bead_prepare(int S, int W, int H, const double* image, const double* vargain, double zstep, int P, int nb,
                    float* xp, float* yp, double* patches, double* B, double* A, double* scmos_patches,
                    double* poly, double* coef, double* stats, int* info) {
    if (S <= 0 || W <= 0 || H <= 0 || P <= 0 || nb < 0) return -2;
    const long wh = (long)W * H;
    const int sF = (int)(0.5 / zstep) / 2 > 1 ? (int)(0.5 / zstep) / 2 : 1;
    refine_points(S, W, H, image, P, nb, xp, yp);
    int bginfo[2] = {0, 0};
    const int rc = compute_background(S, W, H, image, sF, poly, coef, stats, bginfo);
    if (info) {
        info[0] = sF;
        info[1] = bginfo[0];
        info[2] = bginfo[1];
    }
    if (rc != 0) return -1;
    double* A0 = (double*)calloc((size_t)S * (nb > 0 ? nb : 1), sizeof(double));
    for (int p = 0; p < nb; p++) {
        const int x0 = (int)xp[p] - P / 2, y0 = (int)yp[p] - P / 2;
        double b = 0, cb = 0;
        for (int u = 0; u < P; u++)
            for (int uu = 0; uu < P; uu++)
                if (x0 + u >= 0 && x0 + u < W && y0 + uu >= 0 && y0 + uu < H) {
                    b += poly[(long)(x0 + u) * H + (y0 + uu)];
                    cb++;
                }
        if (cb >= 1) b /= cb;
        B[p] = b;
        for (int z = 0; z < S; z++) {
            double a = 0;
            for (int u = 0; u < P; u++)
                for (int uu = 0; uu < P; uu++)
                    if (x0 + u >= 0 && x0 + u < W && y0 + uu >= 0 && y0 + uu < H)
                        a += image[z * wh + (long)(x0 + u) * H + (y0 + uu)] - b;
            A0[z * nb + p] = a;
        }
    }
    for (int p = 0; p < nb; p++)
        for (int z = 0; z < S; z++) {
            double m = 0, c = 0;
            for (int f = -sF; f <= sF; f++)
                if (z + f >= 0 && z + f < S) {
                    m += A0[(z + f) * nb + p];
                    c++;
                }
            if (c >= 1) m /= c;
            A[(long)p * S + z] = m;
        }
    for (int p = 0; p < nb; p++) {
        const int x0 = (int)xp[p] - P / 2, y0 = (int)yp[p] - P / 2;
        for (int z = 0; z < S; z++)
            for (int u = 0; u < P; u++)
                for (int uu = 0; uu < P; uu++) {
                    const int in = x0 + u >= 0 && x0 + u < W && y0 + uu >= 0 && y0 + uu < H;
                    const long o = (((long)p * S + z) * P + u) * P + uu;
                    if (in) {
                        const long k = (long)(x0 + u) * H + (y0 + uu);
                        patches[o] = vargain ? image[z * wh + k] + vargain[k] : image[z * wh + k];
                    } else {
                        patches[o] = B[p];
                    }
                }
        /* the reference re-allocates its sCMOS patch array inside this bead
         * loop (433-447): only the last bead's patch survives, the others are
         * zero */
        if (scmos_patches)
            for (int u = 0; u < P; u++)
                for (int uu = 0; uu < P; uu++) {
                    const int in = x0 + u >= 0 && x0 + u < W && y0 + uu >= 0 && y0 + uu < H;
                    scmos_patches[((long)p * P + u) * P + uu] =
                        (in && vargain && p == nb - 1) ? vargain[(long)(x0 + u) * H + (y0 + uu)] : 0;
                }
    }
    free(A0);
    return 0;
}
