/*
 * zola_oracle.c -- FP64 CPU restatement of ZOLA-3D's localization and
 * calibration hot path.  TEST INFRASTRUCTURE ONLY (the checker): it is
 * called by the tests, the smoke check and the CPU-baseline leg of the
 * benchmark, never by the product library.
 *
 * Reference paths are relative to src/main/java/org/pasteur/imagej/.
 * The file follows the reference's arithmetic step by step (same grids,
 * index maps, float-rounded Gaussian, finite-difference steps and loop
 * order) but is written from scratch in C; the FFT is an own iterative
 * radix-2 Cooley-Tukey with exact twiddles instead of the reference's
 * recurrence (differences are at the 1e-15 level).
 */
#include "zola_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>

#define ZO_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* helpers                                                              */
/* ------------------------------------------------------------------ */

/* (int)Math.pow(2,(int)Math.ceil(Math.log(x)/Math.log(2))), as used by
 * PhaseParameters.java:296 and PSFPhaseMany.java:108. */
static int ref_pow2(int x) {
    return (int)pow(2.0, (int)ceil(log((double)x) / log(2.0)));
}

static void* xcalloc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s);
    if (!p) abort();
    return p;
}

/* ------------------------------------------------------------------ */
/* FFT: unitary, iterative radix-2 (utils/FastFourierTransform.java:680-785
 * fixes the convention: forward e^{-i}, inverse e^{+i}, 1/sqrt(n) per axis). */
/* ------------------------------------------------------------------ */

static void fft1d(int n, double* re, double* im, long stride, int sign) {
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i * stride]; re[i * stride] = re[j * stride]; re[j * stride] = t;
            t = im[i * stride]; im[i * stride] = im[j * stride]; im[j * stride] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            double ang = sign * 2.0 * ZO_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int s = 0; s < n; s += len) {
                long a = (long)(s + k) * stride, b = (long)(s + k + half) * stride;
                double tr = wr * re[b] - wi * im[b];
                double ti = wr * im[b] + wi * re[b];
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr;        im[a] += ti;
            }
        }
    }
    double sc = 1.0 / sqrt((double)n);
    for (int i = 0; i < n; i++) { re[i * stride] *= sc; im[i * stride] *= sc; }
}

void zo_fft2d(int n1, int n2, double* re, double* im, int sign) {
    for (int i = 0; i < n1; i++) fft1d(n2, re + (long)i * n2, im + (long)i * n2, 1, sign);
    for (int j = 0; j < n2; j++) fft1d(n1, re + j, im + j, n2, sign);
}

/* ------------------------------------------------------------------ */
/* Zernike basis: utils/Zernike.java:26-79 (ordering), 545-605 (values), */
/* 814-836 (factorials).                                                */
/* ------------------------------------------------------------------ */

static double ref_fact(int k) {
    /* Zernike.Factorial: exact long table up to 19!, then double products. */
    static long long f[20];
    static int init = 0;
    if (!init) { f[0] = 1; for (int i = 1; i < 20; i++) f[i] = f[i - 1] * i; init = 1; }
    if (k < 20) return (double)f[k];
    double r = (double)f[19];
    for (int i = 20; i <= k; i++) r *= i;
    return r;
}

static void zernike_nm(int J, int* n, int* m) {
    int idx = 0;
    for (int rang = 0; idx < J; rang++)
        for (int u = -rang; u <= rang && idx < J; u += 2) { n[idx] = rang; m[idx] = u; idx++; }
}

static double zernike_value(int n, int m, double theta, double rho) {
    int am = m < 0 ? -m : m;
    if (n < am) return -1.0;
    double som = 0.0;
    for (int k = 0; k <= (n - am) / 2; k++) {
        double truc = pow(-1.0, k) * ref_fact(n - k);
        double trac = ref_fact(k) * ref_fact((n + am) / 2 - k) * ref_fact((n - am) / 2 - k);
        som += (truc / trac) * pow(rho, n - 2 * k);
    }
    return m >= 0 ? som * cos(theta * m) : som * sin(theta * am);
}

/* ------------------------------------------------------------------ */
/* model                                                                */
/* ------------------------------------------------------------------ */

struct zo_model {
    zo_optics o;
    int N;          /* size_cpu */
    int R, s2, D, J;
    double center, ringsize, radius_px;
    int *di, *dj;   /* disk coordinates (row-major scan, PhaseParameters.java:312-362) */
    double *kx, *ky, *kz, *kzoil, *kzflag, *kzoilflag, *pupil;
    int *sx, *sy;   /* fftshift scatter (PSFPhaseMany.java:203-204) */
    double* Z;      /* J*D basis */
    double* coef;   /* J */
    double* phase;  /* D */
    double* gauss;  /* s2*s2, float-rounded (PSFPhaseMany.java:387-404) */
    int *shiftN, *nextS, *lastS;  /* 1-D index maps (separable) */
    /* work */
    double *re, *im, *re2, *im2, *bre, *bim;
    double *tre, *tim, *t2re, *t2im;
    /* last evaluation (for the psf cache) */
    double last[4];
    int have_last;
    double* last_psf;
};

static int psf_cache_enabled = 0;
void zo_set_psf_cache(int on) { psf_cache_enabled = on; }

/* Sensitivity probe (no reference counterpart): with eps > 0 every PSF gets a
 * deterministic perturbation of up to eps * max(psf) per pixel, a hash of the
 * pixel index and of the bits of (x, y, zoil, zwat), so the model stays a fixed
 * function of its parameters as an FP32 synthesis is.  Used to measure how far an
 * algorithm's end point moves under errors of the size of the GPU's FP32 PSF. */
static double psf_perturb_eps = 0.0;
void zo_set_psf_perturb(double eps) { psf_perturb_eps = eps; }
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static void psf_perturb(double* out, int n, double x, double y, double zoil, double zwat) {
    double v[4] = {x, y, zoil, zwat}, mx = 0;
    uint64_t h = 0x9e3779b97f4a7c15ULL;
    for (int k = 0; k < 4; k++) {
        uint64_t b;
        memcpy(&b, &v[k], 8);
        h = mix64(h ^ b);
    }
    for (int k = 0; k < n; k++) mx = fmax(mx, out[k]);
    for (int k = 0; k < n; k++) {
        double u = (double)(mix64(h + (uint64_t)k) >> 11) * (1.0 / 9007199254740992.0);
        out[k] += psf_perturb_eps * mx * (2 * u - 1);
    }
}

static void model_gauss(zo_model* m, double sigma) {
    int s2 = m->s2;
    double sigF = (1.0 / sigma) * s2 / (ZO_PI * 2.0);
    double sigpow = sigF * sigF;
    double c = (double)(s2 / 2);
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            double j = i, jj = ii;
            m->gauss[i * s2 + ii] =
                (double)(float)exp(-.5 * ((j - c) * (j - c) + (jj - c) * (jj - c)) / sigpow);
        }
}

static void model_phase(zo_model* m) {
    /* ZernikePhase.computeCombination (ZernikePhase.java:200-224) */
    for (int p = 0; p < m->D; p++) m->phase[p] = 0.0;
    for (int j = 0; j < m->J; j++)
        for (int p = 0; p < m->D; p++) m->phase[p] += m->Z[(long)j * m->D + p] * m->coef[j];
    m->have_last = 0;
}

zo_model* zo_model_create(const zo_optics* o, int R, const double* coeffs, int J) {
    zo_model* m = (zo_model*)xcalloc(1, sizeof(zo_model));
    m->o = *o;
    int size = o->size_fft;
    if (size % 2 != 0) size--;                       /* PhaseParameters.java:278-280 */
    m->N = ref_pow2(size);                            /* size_cpu, :296 */
    int N = m->N;
    m->center = (N / 2) - .5;                         /* :300 */
    m->ringsize = (o->na / o->wavelength) * (o->na / o->wavelength);  /* :303 */
    m->radius_px = sqrt(m->ringsize) * (N * o->xystep);
    m->R = R;
    /* PSFPhaseMany.java:102-108: s2 = pow2(min(2R, N)) */
    int s2 = R * 2;
    if (s2 > N) s2 = N;
    m->s2 = ref_pow2(s2);
    m->J = J;

    /* disk (PhaseParameters.java:312-362) */
    double c = m->center;
    int D = 0;
    for (int i = 0; i < N; i++)
        for (int ii = 0; ii < N; ii++) {
            double a = (i - c) / (N * o->xystep), b = (ii - c) / (N * o->xystep);
            if (a * a + b * b <= m->ringsize) D++;
        }
    m->D = D;
    m->di = (int*)xcalloc(D, sizeof(int));
    m->dj = (int*)xcalloc(D, sizeof(int));
    m->kx = (double*)xcalloc(D, sizeof(double));
    m->ky = (double*)xcalloc(D, sizeof(double));
    m->kz = (double*)xcalloc(D, sizeof(double));
    m->kzoil = (double*)xcalloc(D, sizeof(double));
    m->kzflag = (double*)xcalloc(D, sizeof(double));
    m->kzoilflag = (double*)xcalloc(D, sizeof(double));
    m->pupil = (double*)xcalloc(D, sizeof(double));
    m->sx = (int*)xcalloc(D, sizeof(int));
    m->sy = (int*)xcalloc(D, sizeof(int));
    m->phase = (double*)xcalloc(D, sizeof(double));

    /* pupil tables (PSFPhaseMany.java:145-210; resetKz 431-470) */
    double left = (o->nwat / o->wavelength) * (o->nwat / o->wavelength);
    double left_oil = (o->noil / o->wavelength) * (o->noil / o->wavelength);
    int id = 0;
    for (int i = 0; i < N; i++)
        for (int ii = 0; ii < N; ii++) {
            double a = (i - c) / (N * o->xystep), b = (ii - c) / (N * o->xystep);
            if (!(a * a + b * b <= m->ringsize)) continue;
            m->di[id] = i;
            m->dj[id] = ii;
            m->kx[id] = (i - c) * (2 * ZO_PI) / (o->xystep * N);
            m->ky[id] = (ii - c) * (2 * ZO_PI) / (o->xystep * N);
            double kxx = ((i - c) / (o->xystep * N)) * ((i - c) / (o->xystep * N));
            double kyy = ((ii - c) / (o->xystep * N)) * ((ii - c) / (o->xystep * N));
            double right = kxx + kyy;
            if (left > right) { m->kz[id] = 2 * ZO_PI * sqrt(left - right); m->kzflag[id] = 1; }
            else {
                m->kz[id] = o->reset_kz ? 2 * ZO_PI * sqrt(fabs(left - right)) : 0.0;
                m->kzflag[id] = 0;
            }
            if (left_oil > right) { m->kzoil[id] = 2 * ZO_PI * sqrt(left_oil - right); m->kzoilflag[id] = 1; }
            else { m->kzoil[id] = 0; m->kzoilflag[id] = 0; }
            if (o->with_apo) m->pupil[id] = left_oil > right ? 1.0 / pow(1 - (right / left_oil), .25) : 0.0;
            else m->pupil[id] = 1.0 / sqrt((double)D);
            m->sx[id] = (i + N / 2) % N;
            m->sy[id] = (ii + N / 2) % N;
            id++;
        }

    /* Zernike basis on the disk (ZernikePhase.java:52-141; Zernike.java:545-605) */
    m->Z = (double*)xcalloc((size_t)(J > 0 ? J : 1) * D, sizeof(double));
    m->coef = (double*)xcalloc(J > 0 ? J : 1, sizeof(double));
    if (J > 0) {
        int* zn = (int*)xcalloc(J, sizeof(int));
        int* zm = (int*)xcalloc(J, sizeof(int));
        zernike_nm(J, zn, zm);
        double cz = ((double)N) / 2. - .5;
        for (int j = 0; j < J; j++)
            for (int p = 0; p < D; p++) {
                double x = ((double)(m->di[p] - cz) / m->radius_px);
                double y = ((double)(m->dj[p] - cz) / m->radius_px);
                double theta = atan2(y, x), rho = sqrt(x * x + y * y);
                m->Z[(long)j * D + p] = rho < 1 ? zernike_value(zn[j], zm[j], theta, rho) : 0.0;
            }
        free(zn); free(zm);
        if (coeffs) memcpy(m->coef, coeffs, sizeof(double) * J);
    }
    model_phase(m);

    /* index maps (PSFPhaseMany.java:232-260), separable in the two axes */
    s2 = m->s2;
    m->shiftN = (int*)xcalloc(s2, sizeof(int));
    m->nextS = (int*)xcalloc(s2, sizeof(int));
    m->lastS = (int*)xcalloc(R, sizeof(int));
    for (int i = 0, j = (N / 2) - (s2 / 2); i < s2; i++, j++) {
        m->shiftN[i] = (j + N / 2) % N;
        m->nextS[i] = (i + s2 / 2) % s2;
    }
    for (int i = 0, j = (s2 / 2) - (R / 2); i < R; i++, j++) m->lastS[i] = (j + s2 / 2) % s2;

    m->gauss = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    model_gauss(m, o->sigma);

    m->re = (double*)xcalloc((size_t)N * N, sizeof(double));
    m->im = (double*)xcalloc((size_t)N * N, sizeof(double));
    m->tre = (double*)xcalloc((size_t)N * N, sizeof(double));
    m->tim = (double*)xcalloc((size_t)N * N, sizeof(double));
    m->re2 = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    m->im2 = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    m->t2re = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    m->t2im = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    m->bre = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    m->bim = (double*)xcalloc((size_t)s2 * s2, sizeof(double));
    m->last_psf = (double*)xcalloc((size_t)R * R, sizeof(double));
    return m;
}

void zo_model_free(zo_model* m) {
    if (!m) return;
    free(m->di); free(m->dj); free(m->kx); free(m->ky); free(m->kz); free(m->kzoil);
    free(m->kzflag); free(m->kzoilflag); free(m->pupil); free(m->sx); free(m->sy);
    free(m->Z); free(m->coef); free(m->phase); free(m->gauss); free(m->shiftN);
    free(m->nextS); free(m->lastS); free(m->re); free(m->im); free(m->tre); free(m->tim);
    free(m->re2); free(m->im2); free(m->t2re); free(m->t2im); free(m->bre); free(m->bim);
    free(m->last_psf);
    free(m);
}

int zo_model_set_coeffs(zo_model* m, const double* coeffs, int J) {
    if (J != m->J) return -1;
    memcpy(m->coef, coeffs, sizeof(double) * J);
    model_phase(m);
    return 0;
}

int zo_model_set_sigma(zo_model* m, double sigma) {
    m->o.sigma = sigma;
    model_gauss(m, sigma);
    m->have_last = 0;
    return 0;
}

int zo_model_set_basis_row(zo_model* m, int k, const double* row) {
    if (k < 0 || k >= m->J) return -1;
    memcpy(m->Z + (long)k * m->D, row, sizeof(double) * m->D);
    model_phase(m);
    return 0;
}

void zo_model_info(const zo_model* m, int* info) {
    info[0] = m->N; info[1] = m->D; info[2] = m->s2; info[3] = m->R; info[4] = m->J;
}

double zo_model_radius_px(const zo_model* m) { return m->radius_px; }

void zo_model_tables(const zo_model* m, int* di, int* dj, double* kx, double* ky, double* kz,
                     double* kzoil, double* kzflag, double* kzoilflag, double* pupil,
                     double* phase) {
    size_t D = (size_t)m->D;
    if (di) memcpy(di, m->di, D * sizeof(int));
    if (dj) memcpy(dj, m->dj, D * sizeof(int));
    if (kx) memcpy(kx, m->kx, D * sizeof(double));
    if (ky) memcpy(ky, m->ky, D * sizeof(double));
    if (kz) memcpy(kz, m->kz, D * sizeof(double));
    if (kzoil) memcpy(kzoil, m->kzoil, D * sizeof(double));
    if (kzflag) memcpy(kzflag, m->kzflag, D * sizeof(double));
    if (kzoilflag) memcpy(kzoilflag, m->kzoilflag, D * sizeof(double));
    if (pupil) memcpy(pupil, m->pupil, D * sizeof(double));
    if (phase) memcpy(phase, m->phase, D * sizeof(double));
}

void zo_model_basis(const zo_model* m, double* Z) {
    memcpy(Z, m->Z, sizeof(double) * (size_t)m->J * m->D);
}

void zo_model_gauss(const zo_model* m, double* g) {
    memcpy(g, m->gauss, sizeof(double) * (size_t)m->s2 * m->s2);
}

/* ------------------------------------------------------------------ */
/* PSF synthesis: PSFPhaseMany.computePSF (PSFPhaseMany.java:553-622)   */
/* ------------------------------------------------------------------ */

/* Pupil field -> E (N*N) -> |E|^2 on the s2 grid -> blurred B (complex,
 * s2*s2), keeps intermediates in the model's work arrays. */
static void psf_forward_core(zo_model* m, double x, double y, double zoil, double zwat) {
    int N = m->N, s2 = m->s2;
    memset(m->re, 0, sizeof(double) * N * N);
    memset(m->im, 0, sizeof(double) * N * N);
    for (int i = 0; i < m->D; i++) {
        double p = m->kx[i] * x + m->ky[i] * y + m->phase[i];
        double q = p + m->kz[i] * zwat - m->kzoil[i] * zoil;
        double r = p + m->kz[i] * zwat * m->kzflag[i] - m->kzoil[i] * zoil * m->kzoilflag[i];
        m->re[m->sx[i] * N + m->sy[i]] = m->pupil[i] * cos(q);
        m->im[m->sx[i] * N + m->sy[i]] = m->pupil[i] * sin(r);
    }
    zo_fft2d(N, N, m->re, m->im, +1);                /* ifft2D */
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            double p = m->re[m->shiftN[i] * N + m->shiftN[ii]];
            double q = m->im[m->shiftN[i] * N + m->shiftN[ii]];
            m->re2[m->nextS[i] * s2 + m->nextS[ii]] = p * p + q * q;
            m->im2[m->nextS[i] * s2 + m->nextS[ii]] = 0;
        }
    memcpy(m->bre, m->re2, sizeof(double) * s2 * s2);
    memcpy(m->bim, m->im2, sizeof(double) * s2 * s2);
    zo_fft2d(s2, s2, m->bre, m->bim, -1);            /* fft2D */
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            double g = m->gauss[m->nextS[i] * s2 + m->nextS[ii]];
            m->bre[i * s2 + ii] *= g;
            m->bim[i * s2 + ii] *= g;
        }
    zo_fft2d(s2, s2, m->bre, m->bim, +1);            /* ifft2D */
}

void zo_model_psf(zo_model* m, double x, double y, double zoil, double zwat, double* out) {
    int R = m->R, s2 = m->s2;
    if (psf_cache_enabled && m->have_last && m->last[0] == x && m->last[1] == y &&
        m->last[2] == zoil && m->last[3] == zwat) {
        memcpy(out, m->last_psf, sizeof(double) * R * R);
        return;
    }
    psf_forward_core(m, x, y, zoil, zwat);
    double r = 0;
    for (int i = 0; i < R; i++)
        for (int ii = 0; ii < R; ii++) {
            double p = m->bre[m->lastS[i] * s2 + m->lastS[ii]];
            double q = m->bim[m->lastS[i] * s2 + m->lastS[ii]];
            out[i * R + ii] = sqrt(p * p + q * q);
            r += out[i * R + ii];
        }
    for (int k = 0; k < R * R; k++) out[k] /= r;
    if (psf_perturb_eps > 0) psf_perturb(out, R * R, x, y, zoil, zwat);
    m->last[0] = x; m->last[1] = y; m->last[2] = zoil; m->last[3] = zwat;
    memcpy(m->last_psf, out, sizeof(double) * R * R);
    m->have_last = 1;
}

/* Analytic tangents of the same pipeline (forward mode). */
void zo_model_psf_jac(zo_model* m, double x, double y, double zoil, double zwat, double* out,
                      double* dout) {
    int N = m->N, s2 = m->s2, R = m->R, D = m->D;
    psf_forward_core(m, x, y, zoil, zwat);
    m->have_last = 0;
    double S = 0;
    double* mag = (double*)xcalloc((size_t)R * R, sizeof(double));
    for (int i = 0; i < R; i++)
        for (int ii = 0; ii < R; ii++) {
            double p = m->bre[m->lastS[i] * s2 + m->lastS[ii]];
            double q = m->bim[m->lastS[i] * s2 + m->lastS[ii]];
            mag[i * R + ii] = sqrt(p * p + q * q);
            S += mag[i * R + ii];
        }
    for (int k = 0; k < R * R; k++) out[k] = mag[k] / S;
    double* dmag = (double*)xcalloc((size_t)R * R, sizeof(double));
    for (int t = 0; t < 3; t++) {
        memset(m->tre, 0, sizeof(double) * N * N);
        memset(m->tim, 0, sizeof(double) * N * N);
        for (int i = 0; i < D; i++) {
            double p = m->kx[i] * x + m->ky[i] * y + m->phase[i];
            double q = p + m->kz[i] * zwat - m->kzoil[i] * zoil;
            double r = p + m->kz[i] * zwat * m->kzflag[i] - m->kzoil[i] * zoil * m->kzoilflag[i];
            double dq, dr;
            if (t == 0) { dq = m->kx[i]; dr = m->kx[i]; }
            else if (t == 1) { dq = m->ky[i]; dr = m->ky[i]; }
            else { dq = m->kz[i]; dr = m->kz[i] * m->kzflag[i]; }
            m->tre[m->sx[i] * N + m->sy[i]] = -m->pupil[i] * sin(q) * dq;
            m->tim[m->sx[i] * N + m->sy[i]] = m->pupil[i] * cos(r) * dr;
        }
        zo_fft2d(N, N, m->tre, m->tim, +1);
        for (int i = 0; i < s2; i++)
            for (int ii = 0; ii < s2; ii++) {
                long src = (long)m->shiftN[i] * N + m->shiftN[ii];
                long dst = (long)m->nextS[i] * s2 + m->nextS[ii];
                m->t2re[dst] = 2 * (m->re[src] * m->tre[src] + m->im[src] * m->tim[src]);
                m->t2im[dst] = 0;
            }
        zo_fft2d(s2, s2, m->t2re, m->t2im, -1);
        for (int i = 0; i < s2; i++)
            for (int ii = 0; ii < s2; ii++) {
                double g = m->gauss[m->nextS[i] * s2 + m->nextS[ii]];
                m->t2re[i * s2 + ii] *= g;
                m->t2im[i * s2 + ii] *= g;
            }
        zo_fft2d(s2, s2, m->t2re, m->t2im, +1);
        double dS = 0;
        for (int i = 0; i < R; i++)
            for (int ii = 0; ii < R; ii++) {
                long k = (long)m->lastS[i] * s2 + m->lastS[ii];
                double mg = mag[i * R + ii];
                double d = mg > 0 ? (m->bre[k] * m->t2re[k] + m->bim[k] * m->t2im[k]) / mg : 0.0;
                dmag[i * R + ii] = d;
                dS += d;
            }
        for (int k = 0; k < R * R; k++)
            dout[(long)t * R * R + k] = dmag[k] / S - mag[k] * dS / (S * S);
    }
    free(mag);
    free(dmag);
}

/* Reverse mode (adjoint) of the same pipeline. */
void zo_model_psf_vjp(zo_model* m, double x, double y, double zoil, double zwat, const double* w,
                      double* gpos, double* ga) {
    int N = m->N, s2 = m->s2, R = m->R, D = m->D;
    psf_forward_core(m, x, y, zoil, zwat);
    m->have_last = 0;
    double S = 0, wm = 0;
    double* mag = (double*)xcalloc((size_t)R * R, sizeof(double));
    for (int i = 0; i < R; i++)
        for (int ii = 0; ii < R; ii++) {
            long k = (long)m->lastS[i] * s2 + m->lastS[ii];
            mag[i * R + ii] = sqrt(m->bre[k] * m->bre[k] + m->bim[k] * m->bim[k]);
            S += mag[i * R + ii];
        }
    for (int k = 0; k < R * R; k++) wm += w[k] * mag[k];
    /* dL/dB on the s2 grid (zero outside the crop) */
    memset(m->t2re, 0, sizeof(double) * s2 * s2);
    memset(m->t2im, 0, sizeof(double) * s2 * s2);
    for (int i = 0; i < R; i++)
        for (int ii = 0; ii < R; ii++) {
            long k = (long)m->lastS[i] * s2 + m->lastS[ii];
            double gm = w[i * R + ii] / S - wm / (S * S);
            double mg = mag[i * R + ii];
            if (mg > 0) { m->t2re[k] += gm * m->bre[k] / mg; m->t2im[k] += gm * m->bim[k] / mg; }
        }
    /* adjoint of IFFT o G o FFT is IFFT o G o FFT (G real diagonal) */
    zo_fft2d(s2, s2, m->t2re, m->t2im, -1);
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            double g = m->gauss[m->nextS[i] * s2 + m->nextS[ii]];
            m->t2re[i * s2 + ii] *= g;
            m->t2im[i * s2 + ii] *= g;
        }
    zo_fft2d(s2, s2, m->t2re, m->t2im, +1);
    /* dL/dE = 2 * dL/dI * E on the cropped positions of the N grid */
    memset(m->tre, 0, sizeof(double) * N * N);
    memset(m->tim, 0, sizeof(double) * N * N);
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            long src = (long)m->shiftN[i] * N + m->shiftN[ii];
            long dst = (long)m->nextS[i] * s2 + m->nextS[ii];
            double gI = m->t2re[dst];
            m->tre[src] += 2 * gI * m->re[src];
            m->tim[src] += 2 * gI * m->im[src];
        }
    /* adjoint of the unitary inverse DFT is the unitary forward DFT */
    zo_fft2d(N, N, m->tre, m->tim, -1);
    double gx = 0, gy = 0, gzw = 0, gzo = 0;
    for (int j = 0; ga && j < m->J; j++) ga[j] = 0;
    for (int i = 0; i < D; i++) {
        double p = m->kx[i] * x + m->ky[i] * y + m->phase[i];
        double q = p + m->kz[i] * zwat - m->kzoil[i] * zoil;
        double r = p + m->kz[i] * zwat * m->kzflag[i] - m->kzoil[i] * zoil * m->kzoilflag[i];
        long k = (long)m->sx[i] * N + m->sy[i];
        double gq = -m->pupil[i] * sin(q) * m->tre[k];
        double gr = m->pupil[i] * cos(r) * m->tim[k];
        double gphi = gq + gr;
        gx += m->kx[i] * gphi;
        gy += m->ky[i] * gphi;
        gzw += m->kz[i] * (gq + m->kzflag[i] * gr);
        gzo += -m->kzoil[i] * (gq + m->kzoilflag[i] * gr);
        if (ga)
            for (int j = 0; j < m->J; j++) ga[j] += m->Z[(long)j * D + i] * gphi;
    }
    gpos[0] = gx; gpos[1] = gy; gpos[2] = gzw; gpos[3] = gzo;
    free(mag);
}

/* ------------------------------------------------------------------ */
/* Matrixe.inverse (utils/Matrixe.java:140-246): Laplace determinant and */
/* the transposed cofactor matrix.                                      */
/* ------------------------------------------------------------------ */

static double det_rec(int n, const double* a) {
    if (n == 1) return a[0];
    if (n == 2) return a[0] * a[3] - a[1] * a[2];
    double sum = 0.0;
    double sub[64];
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int rr = 1; rr < n; rr++, r++) {
            int c = 0;
            for (int cc = 0; cc < n; cc++) {
                if (cc == i) continue;
                sub[r * (n - 1) + c++] = a[rr * n + cc];
            }
        }
        sum += pow(-1, i + 1 + 1) * a[i] * det_rec(n - 1, sub);
    }
    return sum;
}

int zo_inverse_cofactor(int n, const double* a, double* ainv) {
    double det = 1.0 / det_rec(n, a);
    double sub[64];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int r = 0;
            for (int rr = 0; rr < n; rr++) {
                if (rr == i) continue;
                int c = 0;
                for (int cc = 0; cc < n; cc++) {
                    if (cc == j) continue;
                    sub[r * (n - 1) + c++] = a[rr * n + cc];
                }
                r++;
            }
            double cof = pow(-1, i + j + 2) * det_rec(n - 1, sub);
            ainv[j * n + i] = cof * det;   /* transpose */
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Likelihood and the coordinate-descent MLE (cpu/Localization.java)   */
/* ------------------------------------------------------------------ */

double zo_nll(int npx, const double* psf, const double* roi, double A, double B) {
    double lik = 0;
    for (int k = 0; k < npx; k++) {
        double mu = psf[k] * A + B;
        lik += mu - roi[k] * log(mu);
    }
    return lik;
}

typedef struct {
    zo_model* m;
    const double* roi;
    double* psf;
    double* model;
    double zfocus;
    double tA, tB, tX, tY, tZ;
    double h, eps, maxShift, minZ, maxZ, thr;
    int iterMax;
} loc_t;

/* sCMOS term var/gain^2 per ROI pixel (R x R, utils/SCMOScamera.java:95) added
 * to the model of both fit forms, or NULL (Localization.java:1131-1132). */
static const double* scmos_map = NULL;
void zo_set_scmos(const double* varg) { scmos_map = varg; }

/* Localization.computeLikelihood (1125-1143) */
static double L(loc_t* s, double x, double y, double z, double a, double b) {
    int R = s->m->R;
    zo_model_psf(s->m, x, y, s->zfocus, z, s->psf);
    double lik = 0;
    for (int k = 0; k < R * R; k++) {
        s->model[k] = s->psf[k] * a + b + (scmos_map ? scmos_map[k] : 0.0);
        lik += s->model[k] - s->roi[k] * log(s->model[k]);
    }
    return lik;
}

static double newton_step(double lik1, double lik2, double lik3, double h) {
    double grad;
    if (fabs((lik3 + lik1 - 2 * lik2) / (h * h)) == 0) grad = ((lik3 - lik1) * (2 * h));
    else grad = ((lik3 - lik1) / (2 * h)) / fabs((lik3 + lik1 - 2 * lik2) / (h * h));
    return grad;
}

/* minimizeZ / minimizeX / minimizeY (672-906): one coordinate, clamp to
 * maxShift, gamma in {1, 0.1} with break on the first improvement. */
static double minimize_pos(loc_t* s, int which, double prevLik) {
    double* t = which == 0 ? &s->tZ : which == 1 ? &s->tX : &s->tY;
    double save = *t, h = s->h;
    double lik1, lik3;
    if (which == 0) {
        lik1 = L(s, s->tX, s->tY, s->tZ - h, s->tA, s->tB);
        lik3 = L(s, s->tX, s->tY, s->tZ + h, s->tA, s->tB);
    } else if (which == 1) {
        lik1 = L(s, s->tX - h, s->tY, s->tZ, s->tA, s->tB);
        lik3 = L(s, s->tX + h, s->tY, s->tZ, s->tA, s->tB);
    } else {
        lik1 = L(s, s->tX, s->tY - h, s->tZ, s->tA, s->tB);
        lik3 = L(s, s->tX, s->tY + h, s->tZ, s->tA, s->tB);
    }
    double lik2 = prevLik;
    double grad = newton_step(lik1, lik2, lik3, h);
    if (grad > s->maxShift) grad = s->maxShift;
    else if (grad < -s->maxShift) grad = -s->maxShift;
    for (double gamma = 1; gamma > .02; gamma /= 10) {
        *t -= gamma * grad;
        double lik = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);
        if (lik < lik2) { prevLik = lik; break; }
        *t = save;
    }
    return prevLik;
}

/* minimizeA (929-977) h*1000 and minimizeB (997-1058) h*10: no break in
 * the gamma loop, accepted values carry over. */
static double minimize_amp(loc_t* s, int isB) {
    double h = isB ? s->h * 10 : s->h * 1000;
    double* t = isB ? &s->tB : &s->tA;
    double save = *t;
    double lik1, lik2, lik3;
    if (!isB) {
        lik1 = L(s, s->tX, s->tY, s->tZ, s->tA - h, s->tB);
        lik2 = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);
        lik3 = L(s, s->tX, s->tY, s->tZ, s->tA + h, s->tB);
    } else {
        lik1 = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB - h);
        lik2 = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);
        lik3 = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB + h);
        *t = save;
    }
    double grad = newton_step(lik1, lik2, lik3, h);
    for (double gamma = 1; gamma > .02; gamma /= 10) {
        *t -= gamma * grad;
        double lik = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);
        if (lik < lik2) { lik2 = lik; save = *t; }
        else *t = save;
    }
    return L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);
}

static void test_photon(loc_t* s) {
    s->tA = fmax(1, s->tA);
    s->tB = fmax(.01, s->tB);
}

/* computeFisher (317-403) with the reference's finite differences, including
 * its A-derivative built on the model f1 (f0[3] = f1*(a-h)+b). */
static void fisher_fd(loc_t* s, double* F) {
    zo_model* m = s->m;
    int n = m->R * m->R;
    double x = s->tX, y = s->tY, z = s->tZ, a = s->tA, b = s->tB, hd = s->h;
    double* f1 = (double*)xcalloc(n, sizeof(double));
    double* f0 = (double*)xcalloc((size_t)5 * n, sizeof(double));
    double* f2 = (double*)xcalloc((size_t)5 * n, sizeof(double));
    double* ps = (double*)xcalloc(n, sizeof(double));
    zo_model_psf(m, x, y, s->zfocus, z, ps);
    for (int k = 0; k < n; k++) f1[k] = ps[k] * a + b;
    const double pos[6][3] = {{x + hd, y, z}, {x - hd, y, z}, {x, y + hd, z},
                              {x, y - hd, z}, {x, y, z + hd}, {x, y, z - hd}};
    for (int q = 0; q < 6; q++) {
        zo_model_psf(m, pos[q][0], pos[q][1], s->zfocus, pos[q][2], ps);
        double* dst = (q % 2 == 0 ? f2 : f0) + (long)(q / 2) * n;
        for (int k = 0; k < n; k++) dst[k] = ps[k] * a + b;
    }
    for (int k = 0; k < n; k++) {
        f0[3 * n + k] = f1[k] * (a - hd) + b;
        f2[3 * n + k] = f1[k] * (a + hd) + b;
        f0[4 * n + k] = f1[k] * a + (b - hd);
        f2[4 * n + k] = f1[k] * a + (b + hd);
    }
    for (int p = 0; p < 5; p++)
        for (int pp = 0; pp < 5; pp++) {
            double acc = 0;
            for (int k = 0; k < n; k++) {
                double d1 = (f2[p * n + k] - f0[p * n + k]) / (2 * fabs(hd));
                double d2 = (f2[pp * n + k] - f0[pp * n + k]) / (2 * fabs(hd));
                acc += (1 / f1[k]) * d1 * d2;
            }
            F[p * 5 + pp] = acc;
        }
    free(f1); free(f0); free(f2); free(ps);
}

/* computeCRLB (408-451) */
static void crlb_from_fisher(const double* F, double* crlb) {
    double inv[25];
    zo_inverse_cofactor(5, F, inv);
    for (int i = 0; i < 3; i++) crlb[i] = sqrt(inv[i * 5 + i]);
}

int zo_localize(zo_model* m, const double* roi, double* theta, double zfocus, double minZ,
                double maxZ, int iterMax, double photonThreshold, double* out_crlb,
                double* out_lik, int* out_iter, double* model, int* nan_flag) {
    int n = m->R * m->R;
    loc_t s;
    s.m = m; s.roi = roi; s.zfocus = zfocus;
    s.psf = (double*)xcalloc(n, sizeof(double));
    s.model = (double*)xcalloc(n, sizeof(double));
    s.tA = theta[0]; s.tB = theta[1]; s.tX = theta[2]; s.tY = theta[3]; s.tZ = theta[4];
    s.h = .0005; s.eps = .0001; s.maxShift = m->o.xystep / 2.;  /* 29-31, 115 */
    s.minZ = minZ; s.maxZ = maxZ; s.iterMax = iterMax; s.thr = photonThreshold;
    if (nan_flag) *nan_flag = 0;

    /* localize() 491-653 */
    test_photon(&s);
    double likelihood = L(&s, s.tX, s.tY, s.tZ, s.tA, s.tB);
    double oldX = DBL_MAX, oldY = DBL_MAX, oldZ = DBL_MAX;
    int nbIter = iterMax, brk = 0, ok = 1;
    likelihood = minimize_amp(&s, 0);
    test_photon(&s);
    likelihood = minimize_amp(&s, 1);
    test_photon(&s);
    int i;
    for (i = 0; i < iterMax; i++) {
        likelihood = minimize_pos(&s, 0, likelihood);
        if (s.tZ < s.minZ) s.tZ = s.minZ;
        if (s.tZ > s.maxZ) s.tZ = s.maxZ;
        likelihood = minimize_pos(&s, 1, likelihood);
        likelihood = minimize_pos(&s, 2, likelihood);
        likelihood = minimize_amp(&s, 0);
        test_photon(&s);
        likelihood = minimize_amp(&s, 1);
        test_photon(&s);
        double distance = sqrt((s.tX - oldX) * (s.tX - oldX) + (s.tY - oldY) * (s.tY - oldY) +
                               (s.tZ - oldZ) * (s.tZ - oldZ));
        int breaking = 0;
        if (i > 4) breaking = !(s.tA > s.thr);
        if (breaking) brk = 1;
        if (distance < s.eps || breaking) { nbIter = i + 1; break; }
        if (isnan(likelihood)) {
            ok = 0;
            if (nan_flag) *nan_flag = 1;
            break;
        }
        oldX = s.tX; oldY = s.tY; oldZ = s.tZ;
    }
    if (i == iterMax) nbIter = iterMax;
    double crlb[3] = {NAN, NAN, NAN};
    if (ok) {
        /* finish() 295-311 */
        double F[25];
        fisher_fd(&s, F);
        likelihood = L(&s, s.tX, s.tY, s.tZ, s.tA, s.tB);
        crlb_from_fisher(F, crlb);
        ok = !brk;
    }
    theta[0] = s.tA; theta[1] = s.tB; theta[2] = s.tX; theta[3] = s.tY; theta[4] = s.tZ;
    if (out_crlb) memcpy(out_crlb, crlb, sizeof(crlb));
    if (out_lik) *out_lik = likelihood;
    if (out_iter) *out_iter = nbIter;
    if (model) memcpy(model, s.model, sizeof(double) * n);
    free(s.psf);
    free(s.model);
    return ok;
}

typedef struct {
    const zo_optics* o;
    int R, J, b0, b1;
    const double* coeffs;
    const double* rois;
    double* theta;
    double zfocus, minZ, maxZ, thr;
    int iterMax;
    double *crlb, *lik;
    int *iters, *status;
} batch_job;

static void* batch_worker(void* arg) {
    batch_job* j = (batch_job*)arg;
    zo_model* m = zo_model_create(j->o, j->R, j->coeffs, j->J);
    int n = j->R * j->R;
    for (int b = j->b0; b < j->b1; b++) {
        int nanf = 0;
        int st = zo_localize(m, j->rois + (long)b * n, j->theta + (long)b * 5, j->zfocus,
                             j->minZ, j->maxZ, j->iterMax, j->thr,
                             j->crlb ? j->crlb + (long)b * 3 : NULL, j->lik ? j->lik + b : NULL,
                             j->iters ? j->iters + b : NULL, NULL, &nanf);
        if (j->status) j->status[b] = st;
    }
    zo_model_free(m);
    return NULL;
}

int zo_localize_batch(const zo_optics* o, int R, const double* coeffs, int J, int B,
                      const double* rois, double* theta, double zfocus, double minZ,
                      double maxZ, int iterMax, double photonThreshold, double* crlb,
                      double* lik, int* iters, int* status, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > B) nthreads = B > 0 ? B : 1;
    pthread_t* th = (pthread_t*)xcalloc(nthreads, sizeof(pthread_t));
    batch_job* jobs = (batch_job*)xcalloc(nthreads, sizeof(batch_job));
    for (int t = 0; t < nthreads; t++) {
        batch_job* j = &jobs[t];
        j->o = o; j->R = R; j->J = J; j->coeffs = coeffs; j->rois = rois; j->theta = theta;
        j->zfocus = zfocus; j->minZ = minZ; j->maxZ = maxZ; j->thr = photonThreshold;
        j->iterMax = iterMax; j->crlb = crlb; j->lik = lik; j->iters = iters; j->status = status;
        j->b0 = (int)((long)B * t / nthreads);
        j->b1 = (int)((long)B * (t + 1) / nthreads);
        pthread_create(&th[t], NULL, batch_worker, j);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 Levenberg-Marquardt with analytic tangents (no reference         */
/* counterpart: the converged Poisson MLE used as the accuracy anchor).  */
/* ------------------------------------------------------------------ */

/* theta order everywhere in this section: [A, B, x, y, z]. */

/* Likelihood resolution relative to A: a rejected trial whose predicted
 * decrease is at most lm_noise_floor * A ends the fit (converged at the noise
 * level).  0 for the FP64 optimum; 1e-8 mirrors the FP32 GPU fit, whose
 * likelihood differs from FP64 by about 1e-9 A. */
static double lm_noise_floor = 0.0;
void zo_set_lm_noise_floor(double v) { lm_noise_floor = v; }
static double lm_eval(zo_model* m, const double* roi, const double* th, double zfocus,
                      double* psf, double* dpsf, double* g, double* H) {
    int n = m->R * m->R;
    zo_model_psf_jac(m, th[2], th[3], zfocus, th[4], psf, dpsf);
    double nll = 0;
    if (g) for (int i = 0; i < 5; i++) g[i] = 0;
    if (H) for (int i = 0; i < 25; i++) H[i] = 0;
    for (int k = 0; k < n; k++) {
        double mu = th[0] * psf[k] + th[1] + (scmos_map ? scmos_map[k] : 0.0);
        nll += mu - roi[k] * log(mu);
        double J[5] = {psf[k], 1.0, th[0] * dpsf[k], th[0] * dpsf[n + k], th[0] * dpsf[2 * n + k]};
        double wg = 1.0 - roi[k] / mu, wh = 1.0 / mu;
        for (int p = 0; p < 5; p++) {
            if (g) g[p] += wg * J[p];
            if (H) for (int q = 0; q < 5; q++) H[p * 5 + q] += wh * J[p] * J[q];
        }
    }
    return nll;
}

static double lm_nll(zo_model* m, const double* roi, const double* th, double zfocus, double* psf) {
    int n = m->R * m->R;
    zo_model_psf(m, th[2], th[3], zfocus, th[4], psf);
    double lik = 0;
    for (int k = 0; k < n; k++) {
        double mu = psf[k] * th[0] + th[1] + (scmos_map ? scmos_map[k] : 0.0);
        lik += mu - roi[k] * log(mu);
    }
    return lik;
}

void zo_fisher_analytic(zo_model* m, const double* theta, double zfocus, double* fisher) {
    int n = m->R * m->R;
    double* psf = (double*)xcalloc(n, sizeof(double));
    double* dpsf = (double*)xcalloc((size_t)3 * n, sizeof(double));
    double* roi = (double*)xcalloc(n, sizeof(double));
    double H[25];
    lm_eval(m, roi, theta, zfocus, psf, dpsf, NULL, H);
    memcpy(fisher, H, sizeof(H));
    free(psf); free(dpsf); free(roi);
}

/* Cholesky solve of the symmetric n x n system A x = b (A full, row-major;
 * only the lower triangle is read).  Returns -1 unless A is positive definite. */
static int chol_solve(int n, const double* A, double* b) {
    double L[25];
    for (int j = 0; j < n; j++) {
        double s = A[j * n + j];
        for (int k = 0; k < j; k++) s -= L[j * n + k] * L[j * n + k];
        if (!(s > 0)) return -1;
        L[j * n + j] = sqrt(s);
        for (int i = j + 1; i < n; i++) {
            double t = A[i * n + j];
            for (int k = 0; k < j; k++) t -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = t / L[j * n + j];
        }
    }
    for (int i = 0; i < n; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= L[i * n + k] * b[k];
        b[i] = t / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double t = b[i];
        for (int k = i + 1; k < n; k++) t -= L[k * n + i] * b[k];
        b[i] = t / L[i * n + i];
    }
    return 0;
}

/* Structured secant correction of the Fisher matrix (NL2SOL's idea, Dennis,
 * Gay & Welsch 1981, for a Poisson likelihood): the Hessian of the NLL is the
 * Fisher matrix F plus S = sum (1 - n/mu) d2mu, which Fisher scoring drops.
 * After an accepted step s, S is sized by min(1, |s'y|/|s'Ss|) and updated
 * with Powell's symmetric Broyden formula so that S s = y = g+ - g - F+ s,
 * all in the variables scaled by sqrt(diag F+). */
static void lm_secant_update(double* S, const double* F, const double* s, const double* gnew,
                             const double* gold) {
    double D[5], sh[5], yh[5], Sh[25];
    for (int p = 0; p < 5; p++) {
        if (!(F[p * 5 + p] > 0)) return;
        D[p] = sqrt(F[p * 5 + p]);
    }
    for (int p = 0; p < 5; p++) {
        double fs = 0.0;
        for (int q = 0; q < 5; q++) fs += F[p * 5 + q] * s[q];
        sh[p] = D[p] * s[p];
        yh[p] = (gnew[p] - gold[p] - fs) / D[p];
    }
    for (int p = 0; p < 5; p++)
        for (int q = 0; q < 5; q++) Sh[p * 5 + q] = S[p * 5 + q] / (D[p] * D[q]);
    double sSs = 0.0, sy = 0.0, ss = 0.0;
    for (int p = 0; p < 5; p++) {
        double t = 0.0;
        for (int q = 0; q < 5; q++) t += Sh[p * 5 + q] * sh[q];
        sSs += sh[p] * t;
        sy += sh[p] * yh[p];
        ss += sh[p] * sh[p];
    }
    if (!(ss > 0)) return;
    const double tau = sSs != 0.0 ? fmin(1.0, fabs(sy) / fabs(sSs)) : 1.0;
    double r[5], rs = 0.0;
    for (int p = 0; p < 5; p++) {
        double t = 0.0;
        for (int q = 0; q < 5; q++) t += Sh[p * 5 + q] * sh[q];
        r[p] = yh[p] - tau * t;
        rs += r[p] * sh[p];
    }
    for (int p = 0; p < 5; p++)
        for (int q = 0; q < 5; q++) {
            const double v = tau * Sh[p * 5 + q] + (r[p] * sh[q] + sh[p] * r[q]) / ss -
                             rs * sh[p] * sh[q] / (ss * ss);
            S[p * 5 + q] = v * D[p] * D[q];
        }
}

/* Trial point of an LM step d from th, clamped as every trial is: x, y by at
 * most one pixel dx, z by at most 0.2 um, A >= 1, B >= 0.01, zmin <= z <= zmax. */
static void lm_clamp_trial(const double* th, const double* d, double dx, double zmin, double zmax,
                           double* tn) {
    for (int p = 0; p < 5; p++) tn[p] = th[p] + d[p];
    for (int p = 2; p < 4; p++) {
        if (d[p] > dx) tn[p] = th[p] + dx;
        if (d[p] < -dx) tn[p] = th[p] - dx;
    }
    if (d[4] > 0.2) tn[4] = th[4] + 0.2;
    if (d[4] < -0.2) tn[4] = th[4] - 0.2;
    tn[0] = fmax(tn[0], 1.0);
    tn[1] = fmax(tn[1], 0.01);
    tn[4] = fmin(fmax(tn[4], zmin), zmax);
}

/* The trial tc from th is below tol in x, y, z and below 1% of the statistical
 * precision 1/sqrt(F_ii) in A and B (F row-major 5 x 5). */
static int lm_small_step(const double* th, const double* tc, const double* F, double tol) {
    return fmax(fabs(tc[2] - th[2]), fmax(fabs(tc[3] - th[3]), fabs(tc[4] - th[4]))) < tol &&
           fabs(tc[0] - th[0]) * sqrt(F[0]) < 0.01 && fabs(tc[1] - th[1]) * sqrt(F[6]) < 0.01;
}

int zo_lm_fit(zo_model* m, const double* roi, double* theta, double zfocus, double zmin,
              double zmax, int maxit, double tol, double* out_nll, int* out_iter,
              double* out_crlb) {
    /* Levenberg-Marquardt on the Fisher matrix plus a structured secant term S
     * (lm_secant_update), damped with lambda * diag(F), Nielsen's update:
     * success: lambda *= max(1/3, 1-(2 rho-1)^3), nu = 2 (rho = actual/predicted
     * decrease); failure or a system that is not positive definite: lambda *= nu,
     * nu *= 2.  A trial is accepted when the likelihood does not rise; a
     * rejected trial predicted to gain at most lm_noise_floor * A ends the fit. */
    int n = m->R * m->R;
    double* psf = (double*)xcalloc(n, sizeof(double));
    double* dpsf = (double*)xcalloc((size_t)3 * n, sizeof(double));
    double th[5], g[5], H[25], S[25] = {0};
    memcpy(th, theta, sizeof(th));
    th[0] = fmax(th[0], 1.0);
    th[1] = fmax(th[1], 0.01);
    double lambda = 1e-3, nu = 2.0;
    double dx = m->o.xystep;
    double nll = lm_eval(m, roi, th, zfocus, psf, dpsf, g, H);
    int it, conv = 0;
    for (it = 0; it < maxit && !conv; it++) {
        int accepted = 0;
        for (int tries = 0; tries < 12 && !accepted; tries++) {
            double A[25], d[5], tn[5];
            for (int p = 0; p < 25; p++) A[p] = H[p] + S[p];
            for (int p = 0; p < 5; p++) { A[p * 5 + p] += lambda * H[p * 5 + p]; d[p] = -g[p]; }
            if (chol_solve(5, A, d) != 0) { lambda *= nu; nu *= 2; continue; }
            if (tries == 0) {
                /* converged at the point just accepted when the step from it,
                 * almost undamped (lambda at most 1e-4, raised tenfold while the
                 * system is not positive definite) and clamped as trials are, is
                 * below tol in x, y, z and below 1% of the statistical precision
                 * in A and B; a damped step understates the distance along weakly
                 * determined directions.  The almost undamped step is solved only
                 * when the damped one already passes (it is the longer one). */
                double tc[5];
                lm_clamp_trial(th, d, dx, zmin, zmax, tc);
                int small = lm_small_step(th, tc, H, tol);
                if (small && lambda > 1e-4) {
                    double dc[5], lc = 1e-4;
                    int have = 0;
                    for (int k = 0; k < 8 && !have && lc < lambda; k++) {
                        double Ac[25];
                        for (int p = 0; p < 25; p++) Ac[p] = H[p] + S[p];
                        for (int p = 0; p < 5; p++) { Ac[p * 5 + p] += lc * H[p * 5 + p]; dc[p] = -g[p]; }
                        if (chol_solve(5, Ac, dc) == 0) have = 1;
                        else lc *= 10.0;
                    }
                    if (have) {
                        lm_clamp_trial(th, dc, dx, zmin, zmax, tc);
                        small = lm_small_step(th, tc, H, tol);
                    }
                }
                if (small) {
                    conv = 1;
                    accepted = 1;
                    break;
                }
            }
            /* predicted decrease -(g'd + d'(F + S)d / 2); with
             * (F + S + lambda diag F) d = -g this is -g'd / 2 + lambda sum F_pp d_p^2 / 2 */
            double gd = 0.0, dd = 0.0;
            for (int p = 0; p < 5; p++) {
                gd += g[p] * d[p];
                dd += H[p * 5 + p] * d[p] * d[p];
            }
            double pred = -0.5 * gd + 0.5 * lambda * dd;
            lm_clamp_trial(th, d, dx, zmin, zmax, tn);
            double nn = lm_nll(m, roi, tn, zfocus, psf);
            if (nn <= nll) {
                double step = fmax(fabs(tn[2] - th[2]), fmax(fabs(tn[3] - th[3]), fabs(tn[4] - th[4])));
                double rho = pred > 0 ? (nll - nn) / pred : 0.0;
                double t3 = 2.0 * rho - 1.0;
                lambda *= fmax(1.0 / 3.0, 1.0 - t3 * t3 * t3);
                nu = 2.0;
                double sv[5], gold[5];
                for (int p = 0; p < 5; p++) sv[p] = tn[p] - th[p];
                memcpy(gold, g, sizeof(gold));
                memcpy(th, tn, sizeof(th));
                accepted = 1;
                /* at a bound the clamped steps shrink to zero while the unclamped
                 * step does not: a step below tol ends the fit there */
                if (step < tol && (th[4] <= zmin || th[4] >= zmax || th[0] <= 1.0 || th[1] <= 0.01)) conv = 1;
                nll = lm_eval(m, roi, th, zfocus, psf, dpsf, g, H);
                lm_secant_update(S, H, sv, g, gold);
            } else if (pred <= lm_noise_floor * th[0]) {
                break;             /* the gain left is below the likelihood noise */
            } else {
                lambda *= nu;
                nu *= 2;
            }
        }
        if (!accepted) conv = 1;   /* no decrease possible: numerical optimum */
    }
    memcpy(theta, th, sizeof(th));
    if (out_nll) *out_nll = nll;
    if (out_iter) *out_iter = it;
    if (out_crlb) {
        /* CRLB of x, y, z from the analytic Fisher matrix at the optimum */
        double F[25];
        zo_fisher_analytic(m, th, zfocus, F);
        for (int c = 2; c < 5; c++) {
            double A[25], e[5] = {0, 0, 0, 0, 0};
            memcpy(A, F, sizeof(A));
            e[c] = 1;
            if (chol_solve(5, A, e) == 0) out_crlb[c - 2] = sqrt(e[c]);
            else out_crlb[c - 2] = NAN;
        }
    }
    free(psf);
    free(dpsf);
    return conv;
}

/* ------------------------------------------------------------------ */
/* getAandB (LocalizationPipeline.java:1119-1162), score (1099-1115)     */
/* ------------------------------------------------------------------ */

void zo_get_a_and_b(int npx, const double* patch, const double* psf, double thr, double* res) {
    double maxPatch = -INFINITY, maxi = -INFINITY;
    for (int k = 0; k < npx; k++) {
        if (psf[k] > maxi) maxi = psf[k];
        if (patch[k] > maxPatch) maxPatch = patch[k];
    }
    double threshold = maxi * thr;
    double countB = 0, B = 0, A = 0, countA = 0;
    for (int k = 0; k < npx; k++) {
        if (psf[k] < threshold) { B += patch[k]; countB++; }
        else { A += patch[k]; countA++; }
    }
    if (countB != 0) B /= countB; else B = 0;
    A -= countA * B;
    res[0] = A; res[1] = B; res[2] = maxPatch;
}

double zo_score(int npx, const double* patch, const double* model) {
    double sum = 0, cnt = 0;
    for (int k = 0; k < npx; k++) {
        sum += (patch[k] - model[k]) * (patch[k] - model[k]) / model[k];
        cnt++;
    }
    return sum / cnt;
}

/* ------------------------------------------------------------------ */
/* Detector (cpu/Predetection.java:63-593)                              */
/* ------------------------------------------------------------------ */

struct zo_detector {
    int W, H, FX, FY, R, nZ;
    double thr, zstep, xystep;
    double* range;
    double* varPSF;
    double* psf;          /* nZ*R*R (normalised templates) */
    double *dR, *dI;      /* nZ*FX*FY template spectra */
    double *norm, *full, *sum1, *sq1, *sum2, *sq2, *mean, *var;
    double *fr, *fi, *cr, *ci;
    double* corr;         /* nZ*W*H */
};

zo_detector* zo_detector_create(zo_model* m, int W, int H, double zmin, double zmax,
                                double zstep, double thr, double zfocus) {
    zo_detector* d = (zo_detector*)xcalloc(1, sizeof(zo_detector));
    d->W = W; d->H = H; d->R = m->R; d->thr = thr; d->zstep = zstep; d->xystep = m->o.xystep;
    d->FX = ref_pow2(2 * W);
    d->FY = ref_pow2(2 * H);
    int R = d->R, FX = d->FX, FY = d->FY;
    size_t WH = (size_t)W * H;
    d->norm = (double*)xcalloc(WH, sizeof(double));
    d->full = (double*)xcalloc(WH, sizeof(double));
    d->sum1 = (double*)xcalloc(WH, sizeof(double));
    d->sq1 = (double*)xcalloc(WH, sizeof(double));
    d->sum2 = (double*)xcalloc(WH, sizeof(double));
    d->sq2 = (double*)xcalloc(WH, sizeof(double));
    d->mean = (double*)xcalloc(WH, sizeof(double));
    d->var = (double*)xcalloc(WH, sizeof(double));
    /* normImage (104-133) including its mixed use of the X/Y sizes */
    for (int i = 0; i < W; i++)
        for (int ii = 0; ii < H; ii++) {
            int vx1 = R / 2, vx2 = R / 2, vy1 = R / 2, vy2 = R / 2;
            if (i < R / 2 - 1) vx1 = i + 1;
            if (i > W - R / 2 - 2) vx2 = H - i - 1;
            if (ii < R / 2 - 1) vy1 = ii + 1;
            if (ii > W - R / 2 - 2) vy2 = H - ii - 2;
            d->norm[i * H + ii] = (double)((vx1 + vx2) * (vy1 + vy2));
        }
    /* range (118-125): accumulated z */
    int number = (int)ceil((zmax - zmin) / zstep);
    d->nZ = number;
    d->range = (double*)xcalloc(number, sizeof(double));
    double z = zmin;
    for (int i = 0; i < number; i++) { d->range[i] = z; z += zstep; }
    /* dictionary (154-260) */
    d->varPSF = (double*)xcalloc(number, sizeof(double));
    d->psf = (double*)xcalloc((size_t)number * R * R, sizeof(double));
    d->dR = (double*)xcalloc((size_t)number * FX * FY, sizeof(double));
    d->dI = (double*)xcalloc((size_t)number * FX * FY, sizeof(double));
    double* pd = (double*)xcalloc((size_t)R * R, sizeof(double));
    for (int r = 0; r < number; r++) {
        zo_model_psf(m, 0, 0, zfocus, d->range[r], pd);
        double mean = 0;
        for (int k = 0; k < R * R; k++) { d->psf[(long)r * R * R + k] = pd[k]; mean += pd[k]; }
        mean /= (double)R * R;
        double v = 0;
        for (int k = 0; k < R * R; k++) { pd[k] -= mean; v += pd[k] * pd[k]; }
        d->varPSF[r] = v / (double)(R * R);
        double* re = d->dR + (long)r * FX * FY;
        double* im = d->dI + (long)r * FX * FY;
        for (int u = 0; u < R; u++)
            for (int uu = 0; uu < R; uu++) {
                int x = (u - R / 2 + FX / 2 + FX / 2) % FX;
                int y = (uu - R / 2 + FY / 2 + FY / 2) % FY;
                re[x * FY + y] = pd[u * R + uu];
                im[x * FY + y] = 0;
            }
        zo_fft2d(FX, FY, re, im, -1);
    }
    free(pd);
    d->fr = (double*)xcalloc((size_t)FX * FY, sizeof(double));
    d->fi = (double*)xcalloc((size_t)FX * FY, sizeof(double));
    d->cr = (double*)xcalloc((size_t)FX * FY, sizeof(double));
    d->ci = (double*)xcalloc((size_t)FX * FY, sizeof(double));
    d->corr = (double*)xcalloc((size_t)number * WH, sizeof(double));
    return d;
}

void zo_detector_free(zo_detector* d) {
    if (!d) return;
    free(d->range); free(d->varPSF); free(d->psf); free(d->dR); free(d->dI); free(d->norm);
    free(d->full); free(d->sum1); free(d->sq1); free(d->sum2); free(d->sq2); free(d->mean);
    free(d->var); free(d->fr); free(d->fi); free(d->cr); free(d->ci); free(d->corr);
    free(d);
}

int zo_detector_nz(const zo_detector* d) { return d->nZ; }
void zo_detector_range(const zo_detector* d, double* range) {
    memcpy(range, d->range, sizeof(double) * d->nZ);
}
void zo_detector_templates(const zo_detector* d, double* psf) {
    memcpy(psf, d->psf, sizeof(double) * (size_t)d->nZ * d->R * d->R);
}

/* setImage (292-410) */
static void det_set_image(zo_detector* d, const double* image) {
    int W = d->W, H = d->H, FX = d->FX, FY = d->FY, R = d->R;
    memset(d->fr, 0, sizeof(double) * FX * FY);
    memset(d->fi, 0, sizeof(double) * FX * FY);
    for (int i = 0; i < W; i++)
        for (int ii = 0; ii < H; ii++) {
            d->full[i * H + ii] = (double)(float)image[i * H + ii];
            int x = (i - W / 2 + FX / 2 + FX / 2) % FX;
            int y = (ii - H / 2 + FY / 2 + FY / 2) % FY;
            d->fr[x * FY + y] = image[i * H + ii];
        }
    zo_fft2d(FX, FY, d->fr, d->fi, -1);
    double* f = d->full;
    int h2 = R / 2;
    for (int i = 0; i < W; i++) {
        double* s = d->sum1 + (long)i * H;
        double* q = d->sq1 + (long)i * H;
        s[0] = 0; q[0] = 0;
        for (int t = 0; t < h2 + 1; t++)
            if (t < H) { s[0] += f[i * H + t]; q[0] += f[i * H + t] * f[i * H + t]; }
        for (int ii = 1; ii < H; ii++) {
            s[ii] = s[ii - 1];
            if (ii - h2 >= 0) s[ii] -= f[i * H + ii - h2];
            if (ii + h2 < H) s[ii] += f[i * H + ii + h2];
            q[ii] = q[ii - 1];
            if (ii - h2 >= 0) q[ii] -= f[i * H + ii - h2] * f[i * H + ii - h2];
            if (ii + h2 < H) q[ii] += f[i * H + ii + h2] * f[i * H + ii + h2];
        }
    }
    for (int ii = 0; ii < H; ii++) {
        d->sum2[ii] = 0; d->sq2[ii] = 0;
        for (int t = 0; t < h2 + 1; t++)
            if (t < W) { d->sum2[ii] += d->sum1[t * H + ii]; d->sq2[ii] += d->sq1[t * H + ii]; }
        for (int i = 1; i < W; i++) {
            d->sum2[i * H + ii] = d->sum2[(i - 1) * H + ii];
            if (i - h2 >= 0) d->sum2[i * H + ii] -= d->sum1[(i - h2) * H + ii];
            if (i + h2 < W) d->sum2[i * H + ii] += d->sum1[(i + h2) * H + ii];
            d->sq2[i * H + ii] = d->sq2[(i - 1) * H + ii];
            if (i - h2 >= 0) d->sq2[i * H + ii] -= d->sq1[(i - h2) * H + ii];
            if (i + h2 < W) d->sq2[i * H + ii] += d->sq1[(i + h2) * H + ii];
        }
        for (int i = 0; i < W; i++) d->mean[i * H + ii] = d->sum2[i * H + ii] / d->norm[i * H + ii];
    }
    for (int k = 0; k < W * H; k++) d->var[k] = (d->sq2[k] / d->norm[k]) - d->mean[k] * d->mean[k];
}

static double top_parabola(double xa, double ya, double xb, double yb, double xc, double yc) {
    double a = (yc - ya) / ((xc - xa) * (xc - xb)) - (yb - ya) / ((xb - xa) * (xc - xb));
    double b = ((yb - ya) / (xb - xa)) - a * (xb + xa);
    return (-b / (2 * a));
}

typedef struct { double v[7]; long order; } cand_t;

static int cand_cmp(const void* a, const void* b) {
    const cand_t* x = (const cand_t*)a;
    const cand_t* y = (const cand_t*)b;
    if (x->v[0] > y->v[0]) return -1;
    if (x->v[0] < y->v[0]) return 1;
    return x->order < y->order ? -1 : (x->order > y->order);
}

int zo_detector_run(zo_detector* d, const double* image, double* out, int cap, double* corr_out) {
    int W = d->W, H = d->H, FX = d->FX, FY = d->FY, R = d->R, nZ = d->nZ;
    det_set_image(d, image);
    /* convolveNormalizedInFourierDomain (424-496) */
    for (int r = 0; r < nZ; r++) {
        const double* dr = d->dR + (long)r * FX * FY;
        const double* di = d->dI + (long)r * FX * FY;
        for (int k = 0; k < FX * FY; k++) {
            d->cr[k] = d->fi[k] * di[k] + d->fr[k] * dr[k];
            d->ci[k] = d->fi[k] * dr[k] - d->fr[k] * di[k];
        }
        zo_fft2d(FX, FY, d->cr, d->ci, +1);
        double* res = d->corr + (long)r * W * H;
        for (int ii = 0; ii < W; ii++)
            for (int iii = 0; iii < H; iii++) {
                int x = (ii - W / 2 + FX / 2 + FX / 2) % FX;
                int y = (iii - H / 2 + FY / 2 + FY / 2) % FY;
                double v = d->varPSF[r] * d->var[ii * H + iii];
                if (sqrt(v) > 0) res[ii * H + iii] = d->cr[x * FY + y] / sqrt(v);
                else if (d->varPSF[r] <= 0) res[ii * H + iii] = -1;
                else if (d->var[ii * H + iii] <= 0) res[ii * H + iii] = -2;
            }
    }
    if (corr_out) memcpy(corr_out, d->corr, sizeof(double) * (size_t)nZ * W * H);
    /* getVectMaxPosition (500-586) */
    int dec = 3, lo = dec > R / 2 ? dec : R / 2;
    double* tmp = (double*)xcalloc((size_t)W * H, sizeof(double));
    size_t ncap = 1024, nc = 0;
    cand_t* cands = (cand_t*)xcalloc(ncap, sizeof(cand_t));
    for (int r = 0; r < nZ; r++) {
        const double* c = d->corr + (long)r * W * H;
        for (int i = lo; i < W - lo; i++)
            for (int j = lo; j < H - lo; j++)
                if (c[i * H + j] > 0) {
                    double t = -INFINITY;
                    for (int a = -dec; a <= dec; a++) t = fmax(c[i * H + j + a], t);
                    tmp[i * H + j] = t;
                }
        for (int j = lo; j < H - lo; j++)
            for (int i = lo; i < W - lo; i++) {
                double v = c[i * H + j];
                if (!(v > d->thr)) continue;
                if (!(v == tmp[i * H + j])) continue;
                double mx = -INFINITY;
                for (int a = -dec; a <= dec; a++) mx = fmax(tmp[(i + a) * H + j], mx);
                if (!(mx == v)) continue;
                double topZ = 0, topX = 0, topY = 0;
                if (r > 1 && r < nZ - 1)
                    topZ = top_parabola(-1. * d->zstep, d->corr[(long)(r - 1) * W * H + i * H + j], 0, v,
                                        1. * d->zstep, d->corr[(long)(r + 1) * W * H + i * H + j]);
                if (i > 1 && i < W - 1)
                    topX = top_parabola(-1. * d->xystep, c[(i - 1) * H + j], 0, v, 1. * d->xystep,
                                        c[(i + 1) * H + j]);
                if (j > 1 && j < H - 1)
                    topY = top_parabola(-1. * d->xystep, c[i * H + j - 1], 0, v, 1. * d->xystep,
                                        c[i * H + j + 1]);
                if (nc == ncap) {
                    ncap *= 2;
                    cands = (cand_t*)realloc(cands, ncap * sizeof(cand_t));
                    if (!cands) abort();
                }
                cand_t* cc = &cands[nc];
                cc->v[0] = v; cc->v[1] = r; cc->v[2] = i; cc->v[3] = j;
                cc->v[4] = topZ; cc->v[5] = topX; cc->v[6] = topY;
                cc->order = (long)nc;
                nc++;
            }
    }
    qsort(cands, nc, sizeof(cand_t), cand_cmp);
    for (size_t k = 0; k < nc && (long)k < cap; k++) memcpy(out + 7 * k, cands[k].v, 7 * sizeof(double));
    free(cands);
    free(tmp);
    return (int)nc;
}

/* ------------------------------------------------------------------ */
/* Focus centre (cpu/SearchPSFcenter.java:37-81, 86-214, 219-351)       */
/* ------------------------------------------------------------------ */

static double sc_crlb(zo_model* m, double zfocus, double z) {
    /* computeFisher(0, 0, z, .005): photons 3000, background 50 */
    const double h = .005, ph = 3000, bg = 50;
    int n = m->R * m->R;
    double* f1 = (double*)xcalloc(n, sizeof(double));
    double* a = (double*)xcalloc(n, sizeof(double));
    double* b = (double*)xcalloc(n, sizeof(double));
    double* d = (double*)xcalloc((size_t)5 * n, sizeof(double));
    zo_model_psf(m, 0, 0, zfocus, z, f1);
    for (int k = 0; k < n; k++) f1[k] = f1[k] * ph + bg;
    for (int p = 0; p < 3; p++) {
        double dx = p == 0 ? h : 0, dy = p == 1 ? h : 0, dz = p == 2 ? h : 0;
        zo_model_psf(m, dx, dy, zfocus, z + dz, a);
        zo_model_psf(m, -dx, -dy, zfocus, z - dz, b);
        for (int k = 0; k < n; k++) d[p * n + k] = ((a[k] * ph + bg) - (b[k] * ph + bg)) / (2 * fabs(h));
    }
    for (int k = 0; k < n; k++) {
        /* the photon derivative on the model f1 with step 100 h (SearchPSFcenter.java:271-276) */
        d[3 * n + k] = (f1[k] * (ph + h * 100) + bg - (f1[k] * (ph - h * 100) + bg)) / (2 * fabs(h));
        d[4 * n + k] = (f1[k] * ph + (bg + h) - (f1[k] * ph + (bg - h))) / (2 * fabs(h));
    }
    double I[25], inv[25];
    for (int p = 0; p < 5; p++)
        for (int q = 0; q < 5; q++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += (1 / f1[k]) * d[p * n + k] * d[q * n + k];
            I[p * 5 + q] = s;
        }
    zo_inverse_cofactor(5, I, inv);
    free(f1); free(a); free(b); free(d);
    double sx = sqrt(inv[0]), sy = sqrt(inv[6]);
    return sqrt(sx * sx + sy * sy);
}

static double sc_min_position(zo_model* m, double zfocus, double minZ, double maxZ, double stepZ) {
    double mv = INFINITY, mp = 0;
    for (double u = minZ; u < maxZ; u += stepZ) {
        double c = sc_crlb(m, zfocus, u);
        if (mv > c) { mv = c; mp = u; }
    }
    return mp;
}

static double sc_min_range(zo_model* m, double zfocus, double minZ, double maxZ, double stepZ, double axial) {
    axial /= 2;
    int cnt = 0;
    for (double u = minZ - axial / 2; u <= maxZ + axial / 2; u += stepZ) cnt++;
    double* crlb = (double*)xcalloc(cnt ? cnt : 1, sizeof(double));
    cnt = 0;
    for (double u = minZ - axial / 2; u <= maxZ + axial / 2; u += stepZ) crlb[cnt++] = sc_crlb(m, zfocus, u);
    double mv = INFINITY, mp = 0;
    int nb = (int)(axial / (2 * stepZ));
    for (double u = minZ; u < maxZ; u += stepZ) {
        double C = 0;
        int nb2 = -(int)(axial / (2 * stepZ));
        for (double i = -axial / 2; i <= axial / 2; i += stepZ) {
            int k = nb + nb2;
            if (k >= 0 && k < cnt) C += crlb[k];
            nb2++;
        }
        if (mv > C) { mv = C; mp = u; }
        nb++;
    }
    free(crlb);
    return mp;
}

double zo_search_psf_center(zo_model* m, double zfocus, double axialRange) {
    double stepZ = m->o.xystep * 5;
    double minZ = zfocus - 2 * axialRange - fabs(zfocus);
    double maxZ = zfocus + fabs(zfocus) + 2 * axialRange;
    double position = sc_min_position(m, zfocus, minZ, maxZ, stepZ);
    double old = INFINITY;
    for (int i = 0; i < 10; i++) {
        minZ = position - stepZ * 5;
        maxZ = position + stepZ * 5;
        stepZ = stepZ / 2;
        position = sc_min_position(m, zfocus, minZ, maxZ, stepZ);
        if (fabs(old - position) < 0.00001) break;
        old = position;
    }
    return sc_min_range(m, zfocus, position - axialRange * 2, position + axialRange * 2, m->o.xystep,
                        axialRange);
}

/* ------------------------------------------------------------------ */
/* Frame walk (cpu/LocalizationPipeline.java:721-1073, 1099-1250)        */
/* ------------------------------------------------------------------ */

/* getAandB of the sCMOS branch (1168-1211): the sums and the maximum are of
 * patch - scmos. */
static void get_a_and_b_scmos(int npx, const double* patch, const double* psf, const double* sc, double thr,
                              double* res) {
    double maxPatch = -INFINITY, maxi = -INFINITY;
    for (int k = 0; k < npx; k++) {
        if (psf[k] > maxi) maxi = psf[k];
        if (patch[k] - sc[k] > maxPatch) maxPatch = patch[k] - sc[k];
    }
    double threshold = maxi * thr;
    double countB = 0, B = 0, A = 0, countA = 0;
    for (int k = 0; k < npx; k++) {
        if (psf[k] < threshold) { B += patch[k] - sc[k]; countB++; }
        else { A += patch[k] - sc[k]; countA++; }
    }
    if (countB != 0) B /= countB; else B = 0;
    A -= countA * B;
    res[0] = A; res[1] = B; res[2] = maxPatch;
}

static int frame_walk(zo_model* m, zo_detector* d, const double* det_image, const double* image,
                      const double* varg, double zfocus, double minZ, double maxZ, int iterMax,
                      double photonThr, double* out, int cap, int* iters);

int zo_frame_localize(zo_model* m, zo_detector* d, const double* image, double zfocus,
                      double minZ, double maxZ, int iterMax, double photonThr,
                      double* out, int cap) {
    return frame_walk(m, d, image, image, NULL, zfocus, minZ, maxZ, iterMax, photonThr, out, cap, NULL);
}

int zo_frame_localize_iters(zo_model* m, zo_detector* d, const double* image, double zfocus,
                            double minZ, double maxZ, int iterMax, double photonThr,
                            double* out, int cap, int* iters) {
    return frame_walk(m, d, image, image, NULL, zfocus, minZ, maxZ, iterMax, photonThr, out, cap, iters);
}

int zo_frame_localize_scmos(zo_model* m, zo_detector* d, const double* det_image, const double* image,
                            const double* varg, double zfocus, double minZ, double maxZ, int iterMax,
                            double photonThr, double* out, int cap) {
    return frame_walk(m, d, det_image, image, varg, zfocus, minZ, maxZ, iterMax, photonThr, out, cap, NULL);
}

/* The walk on `image`, candidates from `det_image`; with an sCMOS map varg
 * (W x H, var/gain^2), the patch's map enters getAandB and the fit's model
 * (LocalizationPipeline.java:797-850). */
static int frame_walk(zo_model* m, zo_detector* d, const double* det_image, const double* image,
                      const double* varg, double zfocus, double minZ, double maxZ, int iterMax,
                      double photonThr, double* out, int cap, int* iters) {
    int W = d->W, H = d->H, R = m->R, n = R * R, h = R / 2;
    double dx = m->o.xystep;
    int ccap = 1 << 16;
    double* cand = (double*)xcalloc((size_t)ccap * 7, sizeof(double));
    int nc = zo_detector_run(d, det_image, cand, ccap, NULL);
    if (nc > ccap) {
        free(cand);
        ccap = nc;
        cand = (double*)xcalloc((size_t)ccap * 7, sizeof(double));
        nc = zo_detector_run(d, det_image, cand, ccap, NULL);
    }
    unsigned char* mask = (unsigned char*)xcalloc((size_t)W * H, 1);
    double* patch = (double*)xcalloc(n, sizeof(double));
    double* pvar = (double*)xcalloc(n, sizeof(double));
    double* model = (double*)xcalloc(n, sizeof(double));
    int ptr = 0, nloc = 0;
    for (;;) {
        /* myGetMaxPositions (1218-1250) */
        int k = -1;
        for (int i = ptr; i < nc; i++) {
            int ci = (int)floor(cand[i * 7 + 2] + 0.5), cj = (int)floor(cand[i * 7 + 3] + 0.5);
            if (!mask[ci * H + cj] && ci > h + 1 && cj > h + 1 && ci < W - h - 1 && cj < H - h - 1) {
                k = i;
                ptr = i + 1;
                break;
            }
        }
        if (k < 0) break;
        int r = (int)floor(cand[k * 7 + 1] + 0.5);
        int ci = (int)floor(cand[k * 7 + 2] + 0.5), cj = (int)floor(cand[k * 7 + 3] + 0.5);
        /* patch, stopping at the first masked pixel (757-770) */
        int already = 0;
        for (int a = 0; a < R && !already; a++)
            for (int b = 0; b < R; b++) {
                if (mask[(ci + a - h) * H + cj + b - h]) { already = 1; break; }
                patch[a * R + b] = image[(ci + a - h) * H + cj + b - h];
                if (varg) pvar[a * R + b] = varg[(ci + a - h) * H + cj + b - h];
            }
        if (already) continue;
        double ab[3];
        if (varg) get_a_and_b_scmos(n, patch, d->psf + (long)r * n, pvar, .1, ab);
        else zo_get_a_and_b(n, patch, d->psf + (long)r * n, .1, ab);
        int treated = 0;
        if ((ab[2] / ab[1] > 1) && (ab[1] > 0)) {
            double theta[5] = {ab[0], ab[1], -cand[k * 7 + 5], -cand[k * 7 + 6],
                               d->range[r] + cand[k * 7 + 4]};
            double crlb[3], lik;
            int it, nanf;
            if (varg) zo_set_scmos(pvar);
            int ok = zo_localize(m, patch, theta, zfocus, minZ - .3, maxZ + .3, iterMax, 0.0, crlb,
                                 &lik, &it, model, &nanf);
            if (varg) zo_set_scmos(NULL);
            if (ok) {
                double B = theta[1], thr = 0;
                for (int p = 0; p < n; p++)
                    if (thr < (model[p] - B)) thr = model[p] - B;
                thr *= .2;
                if (theta[0] > photonThr && fabs(theta[2]) < 10 * dx && fabs(theta[3]) < 10 * dx) {
                    int px = (int)(ci - (theta[2] / dx)), py = (int)(cj - (theta[3] / dx));
                    if (px >= 0 && px < W && py >= 0 && py < H && !mask[px * H + py] &&
                        theta[4] > minZ && theta[4] < maxZ) {
                        if (nloc < cap) {
                            double* o = out + (long)nloc * 12;
                            o[0] = 0;
                            o[1] = 1000 * (ci * dx - theta[2]);
                            o[2] = 1000 * (cj * dx - theta[3]);
                            o[3] = 1000 * theta[4];
                            o[4] = theta[0];
                            o[5] = theta[1];
                            o[6] = zo_score(n, patch, model);
                            o[7] = 1000 * crlb[0];
                            o[8] = 1000 * crlb[1];
                            o[9] = 1000 * crlb[2];
                            o[10] = lik;
                            o[11] = k;
                            if (iters) iters[nloc] = it;
                        }
                        nloc++;
                        for (int a = 0; a < R; a++)
                            for (int b = 0; b < R; b++)
                                if ((model[a * R + b] - B) > thr) mask[(ci + a - h) * H + cj + b - h] = 1;
                        treated = 1;
                    }
                }
            }
        }
        if (!treated)
            for (int a = 0; a < R; a++)
                for (int b = 0; b < R; b++)
                    if (patch[a * R + b] > ab[2] * .5) mask[(ci + a - h) * H + cj + b - h] = 1;
    }
    free(cand); free(mask); free(patch); free(pvar); free(model);
    return nloc;
}

/* ------------------------------------------------------------------ */
/* Calibration model (ZernikePhaseRetrieval.Model, 1842-1879)           */
/* ------------------------------------------------------------------ */

double zo_calib_nll(zo_model* m, int S, const double* stack, const double* dz, const double* xs,
                    const double* ys, const double* zs, const double* Aslice, const double* bg,
                    double zwat) {
    int n = m->R * m->R;
    double* psf = (double*)xcalloc(n, sizeof(double));
    double lik = 0;
    for (int s = 0; s < S; s++) {
        zo_model_psf(m, xs[s], ys[s], dz[s] + zs[s], zwat, psf);
        for (int k = 0; k < n; k++) {
            double mod = psf[k] * Aslice[s] + bg[k];
            if (mod <= 0) mod = 0.00000001;
            lik += mod - stack[(long)s * n + k] * log(mod);
        }
    }
    free(psf);
    return lik;
}
