/*
 * zola_calib_ref.c -- FP64 restatement of the reference CPU calibration loop.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline of the calibration benchmark (one
 * timed iteration of the reference algorithm) and the tests.  Nothing in the
 * product links it.
 *
 * Restates process/cpu/ZernikePhaseRetrieval.java:
 *   constructor 78-185: basis slots 0..2 replaced by kx, ky, kz_oil (93-95), slice
 *     defocus deltaZ[s] = (s - S/2) zstep (131-140), per-stack registration,
 *     fit_a, fit_a_each and an order-2 polynomial background per stack whose
 *     constant term starts at paramImage.B (166-177);
 *   one iteration of phase_retrieve_zernike's loop (257-300):
 *     updateRegistrationStacks 1207-1290, updateSigmaGaussianKernel 985-1075,
 *     updatePhotonBpoly 861-983, updatePhotonA 780-859, updatePhotonAeach
 *     673-778, updatePhaseZernike 1081-1205, then the stack likelihood;
 *   Model.computePSF / computeLikelihood 1830-1879, computePSFLik /
 *     computeLikOnly / computeLikOnlyStack / computePSFOnly 1417-1445,
 *     switchToStack 1725-1732, computeBackground2D 1518-1530 and
 *     PolynomialBackground 1548-1630 (monomials 1, i, i^2, ii, i ii, ii^2).
 * The reference computes the slices of a stack on one thread each; here the
 * slices are spread over nthreads threads, each with its own model.  The step
 * rules keep the reference's quirks: the registration line search moves by the
 * whole step at every gamma (1262), updatePhotonA moves twice the accepted step
 * (829-836), and the Zernike line search starts at gamma = 100 (1163).
 */
#include "zola_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

struct zo_calib_ref {
    int nb, S, R, J, D, T;
    zo_model** m;          /* T models (work buffers per thread)           */
    double* coef;          /* J coefficients (slots 0..2 set per stack)    */
    double sigma;
    const double* img;     /* nb x S x R x R                               */
    double* dz;            /* S                                            */
    const double* A;       /* nb x S                                       */
    const double* scmos;   /* nb x R x R or NULL                           */
    double* reg;           /* 3 x nb                                       */
    double* poly;          /* nb x 6                                       */
    double* bg;            /* nb x R x R                                   */
    double* fa;            /* nb                                           */
    double* fae;           /* nb x S                                       */
    double* psf;           /* S x R x R: PSFs of the current stack          */
    double* lik;           /* S: per-slice likelihoods                      */
    int cur;               /* stack the PSFs belong to                      */
};

static void* xc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s);
    if (!p) abort();
    return p;
}

static void background2d(zo_calib_ref* c, int z) {
    int R = c->R;
    const double* a = c->poly + 6 * z;
    for (int i = 0; i < R; i++)
        for (int ii = 0; ii < R; ii++)
            c->bg[((size_t)z * R + i) * R + ii] =
                a[0] + a[1] * i + a[2] * (double)i * i + a[3] * ii + a[4] * (double)i * ii + a[5] * (double)ii * ii;
}

static void set_model_coeffs(zo_calib_ref* c) {
    for (int t = 0; t < c->T; t++) zo_model_set_coeffs(c->m[t], c->coef, c->J);
}

static void set_model_sigma(zo_calib_ref* c, double s) {
    for (int t = 0; t < c->T; t++) zo_model_set_sigma(c->m[t], s);
}

/* switchToStack: registration of stack z into slots 0..2, phase recomputed */
static void switch_to_stack(zo_calib_ref* c, int z) {
    for (int p = 0; p < 3; p++) c->coef[p] = c->reg[p * c->nb + z];
    set_model_coeffs(c);
}

typedef struct {
    zo_calib_ref* c;
    int t, z, with_psf;
} job_t;

static void slice_lik(zo_calib_ref* c, int z, int s) {
    int n = c->R * c->R;
    const double* p = c->psf + (size_t)s * n;
    const double* im = c->img + ((size_t)z * c->S + s) * n;
    const double* bg = c->bg + (size_t)z * n;
    const double* sc = c->scmos ? c->scmos + (size_t)z * n : NULL;
    double amp = c->A[(size_t)z * c->S + s] + c->fa[z] + c->fae[(size_t)z * c->S + s];
    double l = 0;
    for (int k = 0; k < n; k++) {
        double mod = p[k] * amp + bg[k] + (sc ? sc[k] : 0.0);
        if (mod <= 0) mod = 0.00000001;
        l += mod - im[k] * log(mod);
    }
    c->lik[s] = l;
}

static void* worker(void* arg) {
    job_t* j = (job_t*)arg;
    zo_calib_ref* c = j->c;
    for (int s = j->t; s < c->S; s += c->T) {
        if (j->with_psf)
            zo_model_psf(c->m[j->t], 0.0, 0.0, c->dz[s], 0.0, c->psf + (size_t)s * c->R * c->R);
        slice_lik(c, j->z, s);
    }
    return NULL;
}

/* All slices of stack z: PSFs (computePSF) when with_psf, then the per-slice
 * likelihoods with the current amplitudes and background. */
static void run_stack(zo_calib_ref* c, int z, int with_psf) {
    pthread_t th[64];
    job_t jobs[64];
    for (int t = 0; t < c->T; t++) {
        jobs[t].c = c;
        jobs[t].t = t;
        jobs[t].z = z;
        jobs[t].with_psf = with_psf;
        pthread_create(&th[t], NULL, worker, &jobs[t]);
    }
    for (int t = 0; t < c->T; t++) pthread_join(th[t], NULL);
    if (with_psf) c->cur = z;
}

static double sum_lik(zo_calib_ref* c) {
    double l = 0;
    for (int s = 0; s < c->S; s++) l += c->lik[s];
    return l;
}

static double psf_lik(zo_calib_ref* c, int z) {     /* computePSFLik */
    switch_to_stack(c, z);
    run_stack(c, z, 1);
    return sum_lik(c);
}

static double lik_only(zo_calib_ref* c, int z) {    /* computeLikOnly */
    switch_to_stack(c, z);
    run_stack(c, z, 0);
    return sum_lik(c);
}

static double newton(double l1, double l2, double l3, double h) {
    double curv = fabs((l3 + l1 - 2 * l2) / (h * h));
    return curv == 0 ? (l3 - l1) / (2 * h) : ((l3 - l1) / (2 * h)) / curv;
}

zo_calib_ref* zo_calib_ref_create(const zo_optics* o, int R, const double* coeffs, int J, int nb, int S,
                                  const double* stacks, double zstep, const double* A, const double* scmos,
                                  const double* B0, double sigma, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    if (nthreads > S) nthreads = S;
    zo_calib_ref* c = (zo_calib_ref*)xc(1, sizeof(zo_calib_ref));
    c->nb = nb; c->S = S; c->R = R; c->J = J; c->T = nthreads;
    c->coef = (double*)xc(J, sizeof(double));
    memcpy(c->coef, coeffs, sizeof(double) * J);
    c->m = (zo_model**)xc(nthreads, sizeof(zo_model*));
    zo_optics oo = *o;
    oo.sigma = sigma;
    for (int t = 0; t < nthreads; t++) c->m[t] = zo_model_create(&oo, R, c->coef, J);
    int info[5];
    zo_model_info(c->m[0], info);
    c->D = info[1];
    /* basis slots 0..2 <- kx, ky, kz_oil (ZernikePhaseRetrieval.java:93-95) */
    double* kx = (double*)xc(c->D, sizeof(double));
    double* ky = (double*)xc(c->D, sizeof(double));
    double* kzo = (double*)xc(c->D, sizeof(double));
    zo_model_tables(c->m[0], NULL, NULL, kx, ky, NULL, kzo, NULL, NULL, NULL, NULL);
    for (int t = 0; t < nthreads; t++) {
        zo_model_set_basis_row(c->m[t], 0, kx);
        zo_model_set_basis_row(c->m[t], 1, ky);
        zo_model_set_basis_row(c->m[t], 2, kzo);
    }
    free(kx); free(ky); free(kzo);
    c->sigma = sigma;
    c->img = stacks;
    c->A = A;
    c->scmos = scmos;
    c->dz = (double*)xc(S, sizeof(double));
    int center = S / 2;
    for (int s = 0; s < S; s++) c->dz[s] = (((double)s) - center) * zstep;
    c->reg = (double*)xc(3 * (size_t)nb, sizeof(double));
    c->poly = (double*)xc(6 * (size_t)nb, sizeof(double));
    c->bg = (double*)xc((size_t)nb * R * R, sizeof(double));
    c->fa = (double*)xc(nb, sizeof(double));
    c->fae = (double*)xc((size_t)nb * S, sizeof(double));
    c->psf = (double*)xc((size_t)S * R * R, sizeof(double));
    c->lik = (double*)xc(S, sizeof(double));
    for (int z = 0; z < nb; z++) {
        c->poly[6 * z] = B0[z];
        background2d(c, z);
    }
    c->cur = -1;
    return c;
}

void zo_calib_ref_free(zo_calib_ref* c) {
    if (!c) return;
    for (int t = 0; t < c->T; t++) zo_model_free(c->m[t]);
    free(c->m); free(c->coef); free(c->dz); free(c->reg); free(c->poly); free(c->bg); free(c->fa);
    free(c->fae); free(c->psf); free(c->lik);
    free(c);
}

double zo_calib_ref_likelihood(zo_calib_ref* c) {
    double l = 0;
    for (int z = 0; z < c->nb; z++) l += psf_lik(c, z);
    return l;
}

static void update_registration(zo_calib_ref* c) {
    double h = .01;
    for (int p = 0; p < 3; p++)
        for (int z = 0; z < c->nb; z++) {
            double* r = &c->reg[p * c->nb + z];
            double save = *r;
            *r = save - h;
            double l1 = psf_lik(c, z);
            *r = save;
            double l2 = psf_lik(c, z);
            *r = save + h;
            double l3 = psf_lik(c, z);
            double grad = newton(l1, l2, l3, h);
            int found = 0;
            for (double gamma = 1; gamma > .02; gamma /= 10) {
                *r = save - grad;
                double l = psf_lik(c, z);
                if (l < l2) { found = 1; break; }
            }
            if (!found) { *r = save; psf_lik(c, z); }
        }
}

static void update_sigma(zo_calib_ref* c) {
    double h = .001;
    double l1 = 0, l2 = 0, l3 = 0;
    set_model_sigma(c, c->sigma - h);
    for (int z = 0; z < c->nb; z++) l1 += psf_lik(c, z);
    set_model_sigma(c, c->sigma);
    for (int z = 0; z < c->nb; z++) l2 += psf_lik(c, z);
    set_model_sigma(c, c->sigma + h);
    for (int z = 0; z < c->nb; z++) l3 += psf_lik(c, z);
    double save = c->sigma;
    double grad = newton(l1, l2, l3, h);
    int found = 0;
    for (double gamma = 1; gamma > .02; gamma /= 10) {
        c->sigma = save - gamma * grad;
        set_model_sigma(c, c->sigma);
        double l = 0;
        for (int z = 0; z < c->nb; z++) l += psf_lik(c, z);
        if (l < l2) { found = 1; break; }
        c->sigma = save;
    }
    if (!found) set_model_sigma(c, c->sigma);
}

static void update_bpoly(zo_calib_ref* c) {
    double h = .0001;
    static const int order_idx[3][3] = {{0, -1, -1}, {1, 3, -1}, {2, 4, 5}};
    for (int z = 0; z < c->nb; z++) {
        psf_lik(c, z);                                   /* computePSFOnly */
        for (int o = 0; o <= 2; o++)
            for (int q = 0; q < 3 && order_idx[o][q] >= 0; q++) {
                double* a = &c->poly[6 * z + order_idx[o][q]];
                double save = *a;
                *a = save - h; background2d(c, z);
                double l1 = lik_only(c, z);
                *a = save; background2d(c, z);
                double l2 = lik_only(c, z);
                *a = save + h; background2d(c, z);
                double l3 = lik_only(c, z);
                double grad = newton(l1, l2, l3, h);
                int found = 0;
                for (double gamma = 1; gamma > .02; gamma /= 10) {
                    *a = save - gamma * grad;
                    background2d(c, z);
                    double l = lik_only(c, z);
                    if (l < l2) { found = 1; break; }
                }
                if (!found) { *a = save; background2d(c, z); }
            }
    }
}

static void update_photon_a(zo_calib_ref* c) {
    double h = 1;
    for (int z = 0; z < c->nb; z++) {
        psf_lik(c, z);
        double save = c->fa[z];
        c->fa[z] = save - h;
        double l1 = lik_only(c, z);
        c->fa[z] = save;
        double l2 = lik_only(c, z);
        c->fa[z] = save + h;
        double l3 = lik_only(c, z);
        double grad = newton(l1, l2, l3, h);
        int found = 0;
        for (double gamma = 1; gamma > .002; gamma /= 10) {
            c->fa[z] = save - gamma * grad;
            double l = lik_only(c, z);
            if (l < l2) {
                c->fa[z] = c->fa[z] - gamma * grad;
                found = 1;
                break;
            }
        }
        if (!found) { c->fa[z] = save; lik_only(c, z); }
    }
}

static void update_photon_a_each(zo_calib_ref* c) {
    double h = 1;
    int S = c->S;
    double* save = (double*)xc(S, sizeof(double));
    double* l1 = (double*)xc(S, sizeof(double));
    double* l2 = (double*)xc(S, sizeof(double));
    double* l3 = (double*)xc(S, sizeof(double));
    double* grad = (double*)xc(S, sizeof(double));
    int* found = (int*)xc(S, sizeof(int));
    for (int z = 0; z < c->nb; z++) {
        double* fae = c->fae + (size_t)z * S;
        psf_lik(c, z);
        for (int s = 0; s < S; s++) save[s] = fae[s];
        for (int s = 0; s < S; s++) fae[s] = save[s] - h;
        lik_only(c, z);
        memcpy(l1, c->lik, sizeof(double) * S);
        for (int s = 0; s < S; s++) fae[s] = save[s];
        lik_only(c, z);
        memcpy(l2, c->lik, sizeof(double) * S);
        for (int s = 0; s < S; s++) fae[s] = save[s] + h;
        lik_only(c, z);
        memcpy(l3, c->lik, sizeof(double) * S);
        for (int s = 0; s < S; s++) {
            grad[s] = newton(l1[s], l2[s], l3[s], h);
            found[s] = 0;
        }
        for (double gamma = 1; gamma > .002; gamma /= 10) {
            for (int s = 0; s < S; s++)
                if (!found[s]) fae[s] = save[s] - gamma * grad[s];
            lik_only(c, z);
            int nf = 0;
            for (int s = 0; s < S; s++)
                if (!found[s]) {
                    if (c->lik[s] < l2[s]) found[s] = 1;
                    else { fae[s] = save[s]; nf++; }
                }
            if (nf == 0) break;
        }
        lik_only(c, z);
    }
    free(save); free(l1); free(l2); free(l3); free(grad); free(found);
}

static void update_zernike(zo_calib_ref* c) {
    double h = .01;
    for (int p = 3; p < c->J; p++) {
        double save = c->coef[p];
        double l1 = 0, l2 = 0, l3 = 0;
        c->coef[p] = save - h;
        for (int z = 0; z < c->nb; z++) l1 += psf_lik(c, z);
        c->coef[p] = save;
        for (int z = 0; z < c->nb; z++) l2 += psf_lik(c, z);
        c->coef[p] = save + h;
        for (int z = 0; z < c->nb; z++) l3 += psf_lik(c, z);
        double grad = newton(l1, l2, l3, h);
        int found = 0;
        for (double gamma = 100; gamma > .02; gamma /= 10) {
            c->coef[p] = save - gamma * grad;
            double l = 0;
            for (int z = 0; z < c->nb; z++) l += psf_lik(c, z);
            if (l < l2) { found = 1; break; }
            c->coef[p] = save;
        }
        if (!found) set_model_coeffs(c);
    }
}

double zo_calib_ref_iteration(zo_calib_ref* c) {
    update_registration(c);
    update_sigma(c);
    update_bpoly(c);
    update_photon_a(c);
    update_photon_a_each(c);
    update_zernike(c);
    return zo_calib_ref_likelihood(c);
}

void zo_calib_ref_state(const zo_calib_ref* c, double* coeffs, double* reg, double* poly, double* fa, double* fae,
                        double* sigma) {
    if (coeffs) memcpy(coeffs, c->coef, sizeof(double) * c->J);
    if (reg) memcpy(reg, c->reg, sizeof(double) * 3 * c->nb);
    if (poly) memcpy(poly, c->poly, sizeof(double) * 6 * c->nb);
    if (fa) memcpy(fa, c->fa, sizeof(double) * c->nb);
    if (fae) memcpy(fae, c->fae, sizeof(double) * c->nb * c->S);
    if (sigma) *sigma = c->sigma;
}
