/*
 * zola_oracle.h -- FP64 CPU restatement of ZOLA-3D's localization and
 * calibration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it.  The
 * product (libzola_hip.so) never links, loads or calls it.
 *
 * Every function restates a reference file (paths relative to
 * src/main/java/org/pasteur/imagej/ of the reference tree):
 *   zo_model_create      process/PhaseParameters.java:265-366,
 *                        utils/Zernike.java:26-79,545-605,
 *                        process/cpu/ZernikePhase.java:52-141,200-224,
 *                        process/cpu/PSFPhaseMany.java:100-271,387-404,431-470
 *   zo_model_psf         process/cpu/PSFPhaseMany.java:553-622
 *   zo_fft2d             utils/FastFourierTransform.java:680-785 (unitary, +i inverse)
 *   zo_localize          process/cpu/Localization.java:29-31,115,317-653,672-1143
 *   zo_inverse_cofactor  utils/Matrixe.java:140-246
 *   zo_get_a_and_b       process/cpu/LocalizationPipeline.java:1119-1162
 *   zo_score             process/cpu/LocalizationPipeline.java:1099-1115
 *   zo_detector_*        process/cpu/Predetection.java:63-593
 *   zo_frame_localize    process/cpu/LocalizationPipeline.java:721-1073,1099-1250
 *   zo_search_psf_center process/cpu/SearchPSFcenter.java:37-351
 *   zo_calib_nll         process/cpu/ZernikePhaseRetrieval.java:131-140,1729-1736,1842-1879
 *   zo_bead_*            process/InitBackgroundAndPhotonNumber.java:50-468,563-904,
 *                        utils/PolynomialFit.java, utils/GaussianElimination.java
 *                        (zola_beads.c)
 * Functions with no reference counterpart (analytic tangents, the adjoint,
 * the FP64 Levenberg-Marquardt optimum) are derived from the same model and
 * are pinned by central finite differences of zo_model_psf.
 */
#ifndef ZOLA_ORACLE_H
#define ZOLA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zo_optics {
    int    size_fft;     /* sizeFFT (pupil grid before the pow2 rounding)   */
    double xystep;       /* camera pixel size, um                           */
    double zstep;        /* calibration z step, um                          */
    double wavelength;   /* um                                              */
    double noil;         /* immersion index                                 */
    double nwat;         /* sample index                                    */
    double na;
    double sigma;        /* Gaussian blur sigma, camera pixels              */
    int    with_apo;     /* apodisation factor instead of 1/sqrt(D)         */
    int    reset_kz;     /* 1: kz as after PSFPhase.resetKz (templates)     */
} zo_optics;

typedef struct zo_model zo_model;

/* Model construction.  J Zernike coefficients in the reference's OSA order
 * (n = 0.., m = -n..n step 2), R = sizeoutput (ROI side). */
zo_model* zo_model_create(const zo_optics* o, int R, const double* coeffs, int J);
void      zo_model_free(zo_model* m);
int       zo_model_set_coeffs(zo_model* m, const double* coeffs, int J);
int       zo_model_set_sigma(zo_model* m, double sigma);
/* Overwrite basis slot k (disk-sized vector) as calibration does with kx/ky/kz_oil. */
int       zo_model_set_basis_row(zo_model* m, int k, const double* row);
/* info[0]=size_cpu info[1]=D info[2]=s2 info[3]=R info[4]=J */
void      zo_model_info(const zo_model* m, int* info);
/* Disk-sized tables (length D each); any pointer may be NULL. */
void      zo_model_tables(const zo_model* m, int* di, int* dj, double* kx, double* ky,
                          double* kz, double* kzoil, double* kzflag, double* kzoilflag,
                          double* pupil, double* phase);
void      zo_model_basis(const zo_model* m, double* Z /* J*D */);
void      zo_model_gauss(const zo_model* m, double* g /* s2*s2 */);
double    zo_model_radius_px(const zo_model* m);

/* PSF (R*R, row-major [i][ii] = first index = ImageJ x column), sum 1. */
void zo_model_psf(zo_model* m, double x, double y, double zoil, double zwat, double* out);
/* PSF plus analytic tangents d/dx, d/dy, d/dzwat (3*R*R). */
void zo_model_psf_jac(zo_model* m, double x, double y, double zoil, double zwat,
                      double* out, double* dout);
/* Adjoint: given w = dL/dpsf (R*R), returns dL/d(x,y,zwat,zoil) in gpos[4]
 * and dL/da_j in ga[J] (ga may be NULL). */
void zo_model_psf_vjp(zo_model* m, double x, double y, double zoil, double zwat,
                      const double* w, double* gpos, double* ga);

/* Unitary 2-D DFT of an n1*n2 complex array (n1, n2 powers of two).
 * sign = -1: forward (e^{-i}); sign = +1: inverse (e^{+i}). */
void zo_fft2d(int n1, int n2, double* re, double* im, int sign);

/* 5x5 (n x n) inverse by cofactors, as Matrixe.inverse. Returns 0. */
int  zo_inverse_cofactor(int n, const double* a, double* ainv);

/* Poisson NLL without ln n!: sum(mu - n ln mu), mu = A*psf + B. */
double zo_nll(int npx, const double* psf, const double* roi, double A, double B);

/* Reference coordinate-descent MLE (cpu/Localization.localize).
 * theta in/out: [A, B, x, y, z] (um).  zfocus enters as z_oil.
 * out_crlb[3] (um), out_lik (likelihood after the loop), out_iter
 * (iterations performed), model (R*R, may be NULL) = final A*psf+B.
 * Returns 1 when localize() returns true, 0 when it returns false
 * (photon-count break or NaN likelihood), and sets *nan_flag. */
int zo_localize(zo_model* m, const double* roi, double* theta, double zfocus,
                double minZ, double maxZ, int iterMax, double photonThreshold,
                double* out_crlb, double* out_lik, int* out_iter, double* model,
                int* nan_flag);

void zo_set_psf_cache(int on);  /* reuse the PSF when (x,y,zoil,zwat) repeat */
/* Sensitivity probe: deterministic perturbation of up to eps * max(psf) per PSF
 * pixel (0 = off), a fixed function of the parameters like an FP32 synthesis. */
void zo_set_psf_perturb(double eps);

/* Same over a batch with nthreads POSIX threads (one model copy per thread).
 * rois: B*R*R, theta: B*5 in/out. */
int zo_localize_batch(const zo_optics* o, int R, const double* coeffs, int J, int B,
                      const double* rois, double* theta, double zfocus, double minZ,
                      double maxZ, int iterMax, double photonThreshold, double* crlb,
                      double* lik, int* iters, int* status, int nthreads);

/* FP64 Poisson-MLE Levenberg-Marquardt with analytic tangents (the converged
 * optimum the batched GPU fit is compared to).  theta [A,B,x,y,z] in/out. */
int zo_lm_fit(zo_model* m, const double* roi, double* theta, double zfocus, double zmin,
              double zmax, int maxit, double tol, double* out_nll, int* out_iter,
              double* out_crlb);

void zo_set_lm_noise_floor(double v);
/* sCMOS map (R x R, var/gain^2 per pixel) added to mu by zo_localize and
 * zo_lm_fit; NULL to clear.  The pointer is kept, not copied. */
void zo_set_scmos(const double* varg);

/* Fisher matrix sum(J J^T / mu) by analytic tangents, 5x5 in [A,B,x,y,z] order. */
void zo_fisher_analytic(zo_model* m, const double* theta, double zfocus, double* fisher);

/* getAandB: res = [A, B, maxPatch]. */
void zo_get_a_and_b(int npx, const double* patch, const double* psf, double thr, double* res);
double zo_score(int npx, const double* patch, const double* model);

/* Detector (cpu/Predetection). */
typedef struct zo_detector zo_detector;
zo_detector* zo_detector_create(zo_model* m, int W, int H, double zmin, double zmax,
                                double zstep, double thr, double zfocus);
void zo_detector_free(zo_detector* d);
int  zo_detector_nz(const zo_detector* d);
void zo_detector_range(const zo_detector* d, double* range);
void zo_detector_templates(const zo_detector* d, double* psf /* nZ*R*R */);
/* Runs one frame (W*H, [i][ii], i < W). Returns the number of candidates;
 * writes up to cap records of 7 doubles [score, r, i, j, dz, dx, dy],
 * sorted descending by score (stable).  corr (nZ*W*H) may be NULL. */
int  zo_detector_run(zo_detector* d, const double* image, double* out, int cap, double* corr);

/* One frame through LocalizationThread.run: detector candidates in score order,
 * myGetMaxPositions, mask test over the patch, getAandB with the candidate's
 * template, zo_localize (z clamp [minZ-0.3, maxZ+0.3], photon break threshold 0
 * as the CPU pipeline leaves it), acceptance on photons (> photonThr), shift
 * (< 10 xystep), centre pixel and z in (minZ, maxZ), masking of the model core
 * (model - B > 0.2 max) or of the bright patch pixels (> 0.5 maxPatch).
 * Writes up to cap records of 12 doubles [frame(0), x_nm, y_nm, z_nm, A, B,
 * score, crlbx_nm, crlby_nm, crlbz_nm, lik, candidate index]; returns the
 * number of localizations. */
int  zo_frame_localize(zo_model* m, zo_detector* d, const double* image, double zfocus,
                       double minZ, double maxZ, int iterMax, double photonThr,
                       double* out, int cap);

/* The walk with an sCMOS camera: candidates from det_image (converted with the
 * constant var/gain^2 of pixel 0 and clamped at 0), patches from image
 * ((adu - offset)/gain + var/gain^2 per pixel), varg the W x H map var/gain^2
 * used by getAandB and the fit (LocalizationPipeline.java:597-633, 797-850,
 * 1168-1211). */
/* As zo_frame_localize, also writing each record's fit iteration count (cap ints). */
int  zo_frame_localize_iters(zo_model* m, zo_detector* d, const double* image, double zfocus,
                             double minZ, double maxZ, int iterMax, double photonThr,
                             double* out, int cap, int* iters);
int  zo_frame_localize_scmos(zo_model* m, zo_detector* d, const double* det_image, const double* image,
                             const double* varg, double zfocus, double minZ, double maxZ, int iterMax,
                             double photonThr, double* out, int cap);

/* Depth of smallest lateral CRLB, refined, then the best window of width
 * axialRange/2 (SearchPSFcenter.getPosition).  m must have R = N (the
 * reference crops the PSF to the FFT size here). */
double zo_search_psf_center(zo_model* m, double zfocus, double axialRange);

/* Calibration negative log-likelihood of one bead stack (ZernikePhaseRetrieval.Model). */
double zo_calib_nll(zo_model* m, int S, const double* stack /* S*R*R */, const double* dz,
                    const double* xs, const double* ys, const double* zs,
                    const double* Aslice, const double* bg /* R*R */, double zwat);

/* Calibration bead preparation (zola_beads.c).  image S x W x H [z][i][ii].
 * zo_bead_convert: (v*adu/gain - offset*adu/gain) + 1, clamped at 0 (CCD/EMCCD
 * constructor); zo_bead_convert_scmos: (v - offset)/gain, clamped at 0.
 * zo_bead_prepare refines the seeds xp/yp (nb, in place), fits the order-2
 * polynomial background and writes patches nb x S x P x P (+ vargain where
 * given, B outside the frame), B nb, A nb x S, scmos_patches nb x P x P (may
 * be NULL), poly W x H, coef 6, stats {mean0, std0, mean, std}, info
 * {sizeFilt, miniposition, ndata}.  Returns 0, -1 (singular normal
 * equations) or -2 (bad sizes). */
void zo_bead_convert(long n, double* image, double adu, double gain, double offset);
void zo_bead_convert_scmos(int S, int W, int H, double* image, const double* offset, const double* gain);
int  zo_bead_prepare(int S, int W, int H, const double* image, const double* vargain, double zstep, int P,
                     int nb, float* xp, float* yp, double* patches, double* B, double* A,
                     double* scmos_patches, double* poly, double* coef, double* stats, int* info);

#ifdef __cplusplus
}
#endif
/* FP64 restatement of the reference CPU calibration loop (zola_calib_ref.c):
 * stacks nb x S x R x R photons, A nb x S photon curves, scmos nb x R x R or
 * NULL, B0 nb background levels; slices on nthreads threads. */
typedef struct zo_calib_ref zo_calib_ref;
zo_calib_ref* zo_calib_ref_create(const zo_optics* o, int R, const double* coeffs, int J, int nb, int S,
                                  const double* stacks, double zstep, const double* A, const double* scmos,
                                  const double* B0, double sigma, int nthreads);
void   zo_calib_ref_free(zo_calib_ref* c);
double zo_calib_ref_likelihood(zo_calib_ref* c);
double zo_calib_ref_iteration(zo_calib_ref* c);
void   zo_calib_ref_state(const zo_calib_ref* c, double* coeffs, double* reg, double* poly, double* fa,
                          double* fae, double* sigma);

#endif
