/*
 * zola_hip.h -- C ABI of the MI355X (gfx950) localization and calibration
 * hot path of ZOLA-3D.
 *
 * This is the drop-in boundary that replaces the reference's jcuda layer
 * (src/main/java/org/pasteur/imagej/cuda/MyCudaStream.java:36-132,
 * cuda/MyDefaultVecKernels.java:120-141,483-545, cuda/MyVecDouble.java and the
 * direct JCufft/JCublas2/JCusparse calls inside process/gpu/PSFmany_float_.java,
 * process/gpu/Modelmany_double_.java, process/gpu/Predetection_cross_.java and
 * process/gpu/ZernikePhaseRetrieval_.java).  One call carries a whole batch,
 * so the Java side no longer needs the per-evaluation thread barrier of
 * Modelmany_double_.getLikelihood (Modelmany_double_.java:865-940).
 *
 * Conventions
 *  - Plain C types only.  Every entry point returns ZOLA_OK (0) or a negative
 *    ZOLA_E* code; zola_last_error() gives a message for the context.
 *  - Entry points named *_dev take device pointers (memory already resident
 *    in HBM) and run asynchronously on the context's stream; the others take
 *    host pointers, copy, run and synchronise.
 *  - Units: um for positions, photons for intensities.  Array layouts are
 *    row-major [i][ii] with the first index the ImageJ x column, as in the
 *    reference (process/cpu/LocalizationPipeline.java:592-596).
 *  - Threads: every entry point locks its context (a detector locks the
 *    context it was created on), so host threads may share a handle and their
 *    calls run one after another, each on the context's stream.
 *    zola_last_error returns the calling thread's copy of the context's last
 *    message.  zola_ctx_destroy and zola_detector_destroy must not overlap
 *    other calls on the same handle.
 */
#ifndef ZOLA_HIP_H
#define ZOLA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ZOLA_OK 0
#define ZOLA_EINVAL (-1)      /* bad argument                             */
#define ZOLA_EUNSUPPORTED (-2)/* optics/ROI outside what the kernels cover */
#define ZOLA_EHIP (-3)        /* HIP runtime error                        */
#define ZOLA_ENOMEM (-4)

/* Optics of the calibration file (process/cpu/DataPhase.java:587-760 keys
 * sizeFFT, sizePix, NA, index, wavelength, sigmaGaussianKernel, withApoFactor;
 * nwat as set by DataPhase.setNwat). */
typedef struct zola_optics {
    int    size_fft;     /* sizeFFT; rounded down to even, then up to a power of two */
    double xystep;       /* camera pixel, um                                  */
    double zstep;        /* calibration z step, um                            */
    double wavelength;   /* um                                                */
    double noil;         /* immersion index                                   */
    double nwat;         /* sample index                                      */
    double na;
    double sigma;        /* Gaussian blur sigma, camera pixels                */
    int    with_apo;     /* 1: apodisation factor (PSFPhaseMany.java:189-200) */
    int    reserved;     /* must be 0                                         */
} zola_optics;

typedef struct zola_ctx zola_ctx;

/* Creates a context for ROI side R (sizeoutput) with J Zernike coefficients
 * in the reference's order (utils/Zernike.java:26-79).  Pupil grid N =
 * pow2(size_fft rounded down to even) (PhaseParameters.java:278-296); blur grid
 * s2 = pow2(min(2R, N)).  Covered: N = 64 with R in {17, 21, 31, 41}; N = 128
 * with R in {16, 20, 32, 41} (s2 = 32, 64, 64, 128).  Other shapes return
 * ZOLA_EUNSUPPORTED.  Replaces
 * DataPhase_(sizeFFT, json) + setSizeoutput + setMany
 * (process/gpu/DataPhase_.java:275-297) and MyCudaStream.init. */
int  zola_ctx_create(const zola_optics* optics, int R, const double* coeffs, int J,
                     zola_ctx** out);
int  zola_ctx_destroy(zola_ctx* ctx);
const char* zola_last_error(const zola_ctx* ctx);
/* Geometry: info[0]=N (pupil grid) info[1]=D (disk pixels) info[2]=s2 (blur
 * grid) info[3]=R info[4]=J. */
int  zola_ctx_info(const zola_ctx* ctx, int* info);
/* Replaces ZernikePhase_.setA + computeCombination + PSFmany_float_.updatePhase
 * (process/gpu/ZernikePhase_.java:602-611, PSFmany_float_.java:1338). */
int  zola_set_coefficients(zola_ctx* ctx, const double* coeffs, int J);
/* Replaces PSFmany_float_.updateSigmaGaussianKernel. */
int  zola_set_sigma(zola_ctx* ctx, double sigma);
/* Runs subsequent work on this HIP stream (hipStream_t passed as void*).  Waits
 * for the context's work on the previous stream first. */
int  zola_set_stream(zola_ctx* ctx, void* hip_stream);

/* Batched PSF synthesis.  params: B x 4 floats [x, y, z, zoil] (um; z is the
 * emitter depth z_wat, zoil the focus term), out: B x R x R floats, each PSF
 * normalised to sum 1 over the ROI.  Replaces PSFmany_float_.computePSF
 * (process/gpu/PSFmany_float_.java:1626-1712) and the kernels it launches
 * (vec_computePSF_phaseNManywithOil_f, vec_computePSF_signalN2Many_f,
 * vec_mulMany_f, vec_computePSF_signalNsqrtMany_fcrop, vec_divScalarMany_f,
 * three cuFFT executions and a cublasSgemv). */
int  zola_psf_forward(zola_ctx* ctx, int B, const float* params, float* out);
int  zola_psf_forward_dev(zola_ctx* ctx, int B, const float* params, float* out);

/* Adjoint (vector-Jacobian product) of zola_psf_forward.  Given w = dL/dpsf
 * (B x R x R), writes dparams (B x 4: dL/dx, dL/dy, dL/dz, dL/dzoil) and, when
 * da is not NULL, accumulates sum over the batch of dL/da_j into da (J
 * doubles, device memory for the _dev form; zeroed by the host form).  No
 * reference counterpart: it replaces finite-difference PSF re-evaluation in
 * process/gpu/ZernikePhaseRetrieval_.java:1108-1187 and Localization_.java. */
int  zola_psf_backward(zola_ctx* ctx, int B, const float* params, const float* w,
                       float* dparams, double* da);
int  zola_psf_backward_dev(zola_ctx* ctx, int B, const float* params, const float* w,
                           float* dparams, double* da);

/* Allocation-free form of zola_psf_backward_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or calibration loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as zola_psf_backward_dev; the per-PSF
 * scratch (dL/dphase on the pupil window, 1024 floats per PSF, and its 1024
 * FP64 sums over the PSFs) lives in a fixed device array of
 * ZOLA_NOALLOC_MAX_PSF x 1024 floats + 8 KiB (8 MiB + 8 KiB) that every GPU the
 * library runs on holds until the process ends, instead of the context's
 * grow-only buffer.  B above ZOLA_NOALLOC_MAX_PSF returns ZOLA_EINVAL.  All calls
 * on one GPU share that array: calls that overlap on one GPU (two streams or two
 * contexts at once) race. */
#define ZOLA_NOALLOC_MAX_PSF 2048
int  zola_psf_backward_noalloc_dev(zola_ctx* ctx, int B, const float* params, const float* w,
                                   float* dparams, double* da);

/* Fit modes */
#define ZOLA_FIT_REFERENCE 0  /* reference coordinate descent (cpu/Localization.java:491-653) */
#define ZOLA_FIT_LM        1  /* batched Poisson-MLE Levenberg-Marquardt with analytic tangents */

/* Status bits (out_status) */
#define ZOLA_ST_CONVERGED 1
#define ZOLA_ST_MAXITER   2
#define ZOLA_ST_NAN       4
#define ZOLA_ST_LOWPHOTON 8   /* reference photon-count break (Localization.java:597-611) */
/* Bits 8..31 of out_status hold the number of pupil fields (N x N inverse
 * transforms) the fit computed for that ROI: status >> 8. */
#define ZOLA_ST_NFIELD_SHIFT 8

/* Work layout of the 64-grid fit kernels (no counterpart in the reference,
 * whose GPU batch shape is fixed by the 100-ROI slots of
 * process/gpu/Modelmany_double_.java:271-387):
 * ZOLA_LAYOUT_WAVE one wave per ROI; ZOLA_LAYOUT_QUAD four waves (one block)
 * per ROI, ceil(R/4) ROI rows per lane; ZOLA_LAYOUT_PAIR two waves (one block)
 * per ROI, LM form at R <= 21 only; ZOLA_LAYOUT_AUTO (the default) two waves
 * for the LM form at R <= 21, one wave otherwise (the fastest layouts measured
 * on MI355X).  The 128-grid fits run four waves per ROI (ZOLA_LAYOUT_QUAD);
 * their LM form also has ZOLA_LAYOUT_OCT, eight waves (one block of 512
 * threads) per ROI, its default (ZOLA_LAYOUT_AUTO). */
#define ZOLA_LAYOUT_AUTO 0
#define ZOLA_LAYOUT_WAVE 1
#define ZOLA_LAYOUT_QUAD 2
#define ZOLA_LAYOUT_PAIR 3
#define ZOLA_LAYOUT_OCT 4

typedef struct zola_fit_config {
    int    mode;             /* ZOLA_FIT_*                                    */
    int    max_iter;         /* iterMax (reference default 20)                */
    double zmin, zmax;       /* clamp of z (reference: minZ-0.3, maxZ+0.3)    */
    double zfocus;           /* z_oil of every ROI (param.Zfocus)             */
    double photon_threshold; /* reference photonThreshold for the break       */
    double tol;              /* LM: stop when |dx|,|dy|,|dz| < tol (um)        */
} zola_fit_config;

/* Fits B ROIs of side R.  pixels: B x R x R floats (photons).  init: B x 5
 * doubles [A, B, x, y, z].  Outputs: theta B x 5 doubles, nll B doubles,
 * crlb B x 3 doubles (x, y, z um), iters B ints, status B ints.  Replaces
 * Localization_.localize / localizeWithoutFinishing + finish1
 * (process/gpu/Localization_.java:633-1011) driven through
 * Modelmany_double_ (process/gpu/Modelmany_double_.java:545-791). */
int  zola_fit_batch(zola_ctx* ctx, const zola_fit_config* cfg, int B, const float* pixels,
                    const double* init, double* theta, double* nll, double* crlb,
                    int* iters, int* status);
int  zola_fit_batch_dev(zola_ctx* ctx, const zola_fit_config* cfg, int B,
                        const float* pixels, const double* init, double* theta,
                        double* nll, double* crlb, int* iters, int* status);

/* As zola_fit_batch_dev with a focus term per ROI: zoil B floats (um) in place
 * of cfg->zfocus, e.g. emitters from fields at different focal depths under a
 * refractive-index mismatch.  The reference fixes one Zfocus per data set
 * (process/cpu/Localization.java through PSFPhaseMany.computePSF's z_oil). */
int  zola_fit_batch_zoil_dev(zola_ctx* ctx, const zola_fit_config* cfg, int B,
                             const float* pixels, const double* init, const float* zoil,
                             double* theta, double* nll, double* crlb, int* iters, int* status);

/* Sets the work layout (ZOLA_LAYOUT_*) of the fits of ROI side R in fit mode
 * `mode` for this context (64 grid: every side; 128 grid: the context's side).
 * Results of the layouts agree to FP32 rounding of the blur sums (a different
 * summation order).  A layout that does not exist for (R, mode) returns
 * ZOLA_EUNSUPPORTED. */
int  zola_set_fit_layout(zola_ctx* ctx, int R, int mode, int layout);

/* Autotuner of the fit layout: fits a synthetic batch of B ROIs of side R
 * (noise-free model images, started off the truth) with each layout on the
 * context's stream, keeps the fastest for (R, mode) and returns it in
 * *layout, with the kernel times in ms[0] (one wave per ROI), ms[1] (four
 * waves per ROI) and ms[2] (two waves per ROI; 0 where that layout does not
 * exist): ms holds three doubles.  Synchronises the stream; not for use under
 * graph capture. */
int  zola_autotune_fit(zola_ctx* ctx, int R, int mode, int B, int* layout, double* ms);

/* As zola_fit_batch_dev for an sCMOS camera: varg holds B x R x R floats, the
 * per-pixel var/gain^2 of each ROI (utils/SCMOScamera.java:95), added to the
 * model mu = A psf + B as in Localization.computeLikelihood
 * (process/cpu/Localization.java:1131-1132); pixels are the converted
 * (adu - offset)/gain + var/gain^2 (process/cpu/LocalizationPipeline.java:633).
 * The reference form keeps the reference's Fisher weights 1/(A psf + B)
 * (computeFisher, 317-403, has no sCMOS term); the LM form uses the full mu. */
int  zola_fit_batch_scmos_dev(zola_ctx* ctx, const zola_fit_config* cfg, int B,
                              const float* pixels, const float* varg, const double* init,
                              double* theta, double* nll, double* crlb, int* iters, int* status);

/* General form of the three above: zoil and varg may each be NULL. */
int  zola_fit_batch_ext_dev(zola_ctx* ctx, const zola_fit_config* cfg, int B,
                            const float* pixels, const double* init, const float* zoil,
                            const float* varg, double* theta, double* nll, double* crlb,
                            int* iters, int* status);

/* Fits G groups of ROIs of different sides R[g] (device pointers, one set per
 * group; 1 <= G <= 8) in one graph launch on the context's stream: one kernel
 * node per group, the nodes independent so the groups run concurrently.  On the
 * 64 grid every instantiated side (17, 21, 31, 41) shares the context's model
 * (the blur grid is 64 for all of them); on the 128 grid all groups must have
 * the context's R.  Results are those of zola_fit_batch_dev per group.  No
 * reference counterpart: the plugin runs one fixed patch size per run
 * (LocalizationPipeline sizePatch); this replaces several zola_fit_batch_dev
 * launches when patches of several sizes are fitted together. */
int  zola_fit_mixed_dev(zola_ctx* ctx, const zola_fit_config* cfg, int G, const int* R, const int* B,
                        const float* const* pixels, const double* const* init, double* const* theta,
                        double* const* nll, double* const* crlb, int* const* iters,
                        int* const* status);

/* Host-array fit of ROIs of several sides in one call (ragged batch): ROI b is
 * the roi_sizes[b] x roi_sizes[b] block of pixels starting at
 * pixels[roi_offsets[b]] (row-major, n_pixels floats in all); init, theta, nll,
 * crlb, iters and status are B-row arrays in ROI order.  The ROIs are grouped
 * by side (at most 8 sides), copied in once, fitted by zola_fit_mixed_dev and
 * copied back in ROI order; the results equal those of zola_fit_batch on each
 * side's ROIs.  The reference fixes one patch side per model object and
 * pipeline (sizeImage in process/gpu/Modelmany_double_.java:271-300, sizePatch
 * in process/gpu/LocalizationPipeline_.java:61,109); this form lets one host
 * call carry patches of several sides. */
int  zola_fit_ragged(zola_ctx* ctx, const zola_fit_config* cfg, int B, const long long* roi_offsets,
                     const int* roi_sizes, const float* pixels, long long n_pixels, const double* init,
                     double* theta, double* nll, double* crlb, int* iters, int* status);

/* Per-ROI initial values as the reference pipeline forms them for an isolated
 * ROI: z from the best zero-shift normalised correlation against nZ templates
 * PSF(0,0,z_r) (z_r = zmin_t + r*zstep_t), then getAandB with that template
 * (process/cpu/LocalizationPipeline.java:1119-1162, threshold 0.1).  Writes
 * init B x 5 doubles [A, B, 0, 0, z_r].  The template stack is built on the
 * first call and kept while the model (coefficients, sigma) and (nZ, zmin_t,
 * zstep_t, zfocus) stay the same; a call that rebuilds it waits for the
 * stream, a call that reuses it only enqueues the kernel. */
int  zola_fit_init_dev(zola_ctx* ctx, int B, const float* pixels, int nZ, double zmin_t,
                       double zstep_t, double zfocus, double* init);

/* The context's HIP stream (hipStream_t as void*). */
int  zola_get_stream(const zola_ctx* ctx, void** stream);

/* One calibration evaluation over B bead slices (device pointers): the model
 * of process/cpu/ZernikePhaseRetrieval.java:1842-1879, mu = amp*psf + bg
 * (clamped at 1e-8), psf at params B x 4 [x, y, z, zoil] (calibration slices use
 * zoil = slice defocus + bead z0, z = 0).  Writes nll B, dparams B x 4 (dNLL/dx,
 * dNLL/dy, dNLL/dz, dNLL/dzoil), damp B, dbg B and, when da is not NULL,
 * accumulates dNLL/da_j into da (J doubles).  One forward + adjoint pass replaces
 * the per-coefficient finite differences of
 * process/gpu/ZernikePhaseRetrieval_.java:1108-1187 and getLikelihood 2838-2862. */
int  zola_calib_eval_dev(zola_ctx* ctx, int B, const float* params, const float* pixels,
                         const double* amp, const double* bg, double* nll, float* dparams,
                         double* damp, double* dbg, double* da);

/* PSF calibration from nb bead z-stacks of S slices (host arrays), replacing
 * ZernikePhaseRetrieval_.run (process/gpu/ZernikePhaseRetrieval_.java:362-474).
 * pixels nb*S*R*R photons, dz S slice defocus (um), bead nb*5 in/out
 * [x, y, z0, A, bg], coeffs J in/out (context coefficients are updated),
 * fit_mask J (1 = fitted; NULL = all).  FP64 L-BFGS on exact adjoint gradients;
 * stops at max_iter, when the largest scaled gradient component is below gtol, or
 * when the relative NLL decrease falls below 1e-13. */
int  zola_calib_fit(zola_ctx* ctx, int nb, int S, const float* pixels, const double* dz,
                    double* bead, double* coeffs, const int* fit_mask, int max_iter,
                    double gtol, double* out_nll, int* out_iters, int* out_evals);

/* ---------------------------------------------------------------------------
 * The reference's full calibration model (process/cpu/ZernikePhaseRetrieval.java:
 * 1842-1879; GPU twin process/gpu/ZernikePhaseRetrieval_.java:362-474):
 *   mu(b, s, i, ii) = psf_bs(i, ii) * amp(b, s) + background_b(i, ii) [+ scmos_b(i, ii)],
 * replaced by 1e-8 where <= 0, NLL = sum mu - n ln mu, with
 *   psf_bs = PSF at (x_b, y_b), z_oil = dz_s + z0_b, z_wat = 0 (computePSF 1830-1835),
 *   amp(b, s) = A[b][s] + fit_a[b] + fit_a_each[b][s] (the bead preparation's photon
 *     curve plus the fitted offsets, 1859-1867),
 *   background_b = c0 + c1 i + c2 i^2 + c3 ii + c4 i ii + c5 ii^2 in patch
 *     coordinates (PolynomialBackground of order 2, 1518-1630, same monomial order),
 *   and the Gaussian blur width sigma (px) a free parameter (updateSigmaGaussianKernel).
 * PSF b*S + s is slice s of bead b. */

/* One forward + adjoint evaluation of the full model (device pointers, context
 * stream).  params nb*S x 4 [x, y, z, zoil] (slices: z = 0, zoil = dz_s + z0_b),
 * pixels nb*S x R x R, amp nb*S, poly nb x 6, offset nb x R x R or NULL.
 * Writes nll nb*S, dparams nb*S x 4 (dNLL/dx, dy, dz, dzoil), damp nb*S, dpoly
 * nb*S x 6 (may be NULL), dsigma nb*S (dNLL/dsigma, may be NULL) and adds
 * dNLL/da_j to da (J, may be NULL).  Replaces the per-parameter finite
 * differences of updatePhaseZernike, updateRegistrationStacks, updatePhotonA,
 * updatePhotonAeach, updatePhotonBpoly and updateSigmaGaussianKernel
 * (ZernikePhaseRetrieval.java:673-1205). */
int  zola_calib_eval_full_dev(zola_ctx* ctx, int nb, int S, const float* params, const float* pixels,
                              const double* amp, const double* poly, const double* offset, double* nll,
                              float* dparams, double* damp, double* dpoly, double* dsigma, double* da);

/* Parameters fitted by zola_calib_fit_full besides the Zernike coefficients */
#define ZOLA_CALIB_FIT_REGISTRATION 1   /* bead x, y, z0 (updateRegistrationStacks)        */
#define ZOLA_CALIB_FIT_SIGMA        2   /* blur width (updateSigmaGaussianKernel)           */
#define ZOLA_CALIB_FIT_BACKGROUND   4   /* background polynomial (updatePhotonBpoly)        */
#define ZOLA_CALIB_FIT_PHOTONS      8   /* fit_a, fit_a_each (updatePhotonA, ...Aeach)      */
#define ZOLA_CALIB_FIT_ALL          15

/* Bead stacks and their per-bead parameters (host arrays) as the reference holds
 * them after InitBackgroundAndPhotonNumber (paramImage) and in its fit state. */
typedef struct zola_calib_beads {
    int nb, S;
    const float* pixels;   /* nb x S x R x R photons (paramImage.image)                  */
    const double* dz;      /* S slice defocus, um (deltaZ, ZernikePhaseRetrieval.java:131-140) */
    const double* A;       /* nb x S photon curves (paramImage.A)                         */
    const double* offset;  /* nb x R x R sCMOS var/gain^2 (paramImage.scmos), or NULL      */
    double* xyz;           /* nb x 3 in/out: registration x, y, z0 (um)                    */
    double* poly;          /* nb x 6 in/out: background (start: c0 = paramImage.B, rest 0) */
    double* fit_a;         /* nb in/out                                                    */
    double* fit_a_each;    /* nb x S in/out                                                */
} zola_calib_beads;

/* Phase retrieval on the full model, replacing ZernikePhaseRetrieval.run /
 * phase_retrieve_zernike (process/cpu/ZernikePhaseRetrieval.java:227-363; GPU
 * process/gpu/ZernikePhaseRetrieval_.java:362-474): the coefficients with
 * fit_mask[j] = 1 (NULL: all), sigma (in/out, when ZOLA_CALIB_FIT_SIGMA) and the
 * bead parameters selected by flags, by FP64 L-BFGS on the exact gradients of
 * zola_calib_eval_full_dev.  Stops at max_iter iterations, when the largest
 * scaled gradient component is below gtol, or when the relative NLL decrease
 * falls below 1e-13.  The context's coefficients and sigma are left at the
 * result. */
int  zola_calib_fit_full(zola_ctx* ctx, zola_calib_beads* beads, double* sigma, double* coeffs,
                         const int* fit_mask, int flags, int max_iter, double gtol, double* out_nll,
                         int* out_iters, int* out_evals);

/* ---------------------------------------------------------------------------
 * Frame detector: normalised FFT cross-correlation of whole frames against a
 * z-stack of PSF templates, local maxima and parabola refinement, in FP64.
 * Replaces process/cpu/Predetection.java:63-593 (the CPU path of the plugin;
 * process/gpu/Predetection_cross_.java:104-777 is its FP32 GPU twin).
 * Frames are W x H floats in photons, laid out [i][ii] with i the ImageJ x
 * column, after the plugin's rescale and clamp at 0
 * (process/cpu/LocalizationPipeline.java:611-630). */
typedef struct zola_detector zola_detector;

/* Candidate record: 7 doubles [score, r, i, j, dz, dx, dy] as built by
 * Predetection.getVectMaxPosition (Predetection.java:553-563): correlation
 * score, template index r, pixel (i, j), parabola offsets (um).  Lists are
 * ordered by score, descending, ties in the reference's scan order. */
#define ZOLA_CAND_FIELDS 7

/* Builds the template dictionary for W x H frames: nZ = ceil((zmax-zmin)/zstep)
 * templates PSF(0, 0, zoil = zfocus, z = zmin + r*zstep) of side R (the
 * context's), padded FFT size pow2(2W) x pow2(2H) (at most 2048 each).
 * threshold is thresholdCrossCorrelation (LocalizationPipeline.java:59, 0.1).
 * Replaces new Predetection(width, height, dp, minZ, maxZ, stepZ, threshold)
 * (LocalizationPipeline.java:455). */
int  zola_detector_create(zola_ctx* ctx, int W, int H, double zmin, double zmax, double zstep,
                          double threshold, double zfocus, zola_detector** out);
int  zola_detector_destroy(zola_detector* det);
/* info[0]=nZ info[1]=FX info[2]=FY info[3]=W info[4]=H */
int  zola_detector_info(const zola_detector* det, int* info);
/* Template z values (Predetection.getRange), nZ doubles. */
int  zola_detector_range(const zola_detector* det, double* range);
/* Templates nZ x R x R, each normalised to sum 1 (Predetection.getPSFNonNormalized). */
int  zola_detector_templates(const zola_detector* det, double* psf);
/* Candidates of F frames (host arrays): frames F x W x H, cand F x cap x 7
 * (frame f's list at cand + f*cap*7, at most cap records), counts F (number
 * found, may exceed cap).  Replaces Predetection.setImage +
 * convolveNormalizedInFourierDomain (Predetection.java:292-496). */
int  zola_detect_frames(zola_detector* det, int F, const float* frames, int cap, double* cand,
                        int* counts);
/* Centre of the template z range: the depth whose window of width
 * axial_range/2 has the smallest summed lateral CRLB for a 3000-photon
 * emitter on background 50, found by the reference's coarse scan, halving
 * refinement and window scan (process/cpu/SearchPSFcenter.java:37-351; the PSF
 * is cropped to the FFT size, the Fisher matrix by the reference's finite
 * differences).  FP64 on the host cores; called once per run by the pipeline
 * (LocalizationPipeline.java:159-172: minZ/maxZ = centre -/+ axial_range/2). */
int  zola_search_psf_center(zola_ctx* ctx, double zfocus, double axial_range, double* center);

/* The normalised correlation maps of one frame, nZ x W x H doubles
 * (Predetection.resultCorrelation). */
int  zola_detector_correlation(zola_detector* det, const float* frame, double* corr);

/* ---------------------------------------------------------------------------
 * Frame localization: detection, then the per-frame greedy walk of
 * process/cpu/LocalizationPipeline.java:721-1073 (LocalizationThread.run) over
 * the ordered candidates: skip candidates whose patch touches the mask, fit
 * with getAandB + parabola initial values, accept on photons, shift, centre
 * pixel and z range, mask the model core (accepted) or the bright patch pixels
 * (not accepted).  Fits run batched on the GPU; decisions are replayed in the
 * reference's order on the host. */
typedef struct zola_accept_config {
    double photon_threshold; /* photonThreshold (ZOLA.java:248, default 500)          */
    double min_z, max_z;     /* accepted z: ZfocusCenter -/+ axialRange/2              */
    double snr_min;          /* maxSNR: fit only if maxPatch/B > snr_min (1)           */
    double model_mask_frac;  /* accepted: mask model-B > frac*max(model-B) (0.2)       */
    double patch_mask_frac;  /* not accepted: mask patch > frac*maxPatch (0.5)          */
    double max_shift_px;     /* accept |x|,|y| < max_shift_px * xystep (10)             */
} zola_accept_config;

/* Localization record: 12 doubles [frame, x_nm, y_nm, z_nm, A, B, score,
 * crlb_x_nm, crlb_y_nm, crlb_z_nm, nll, candidate index] -- the fields of
 * PLocalization as LocalizationPipeline.java:890-905 fills them (x_nm =
 * 1000 (i*xystep - x), score = scoreCompute), plus the fit's NLL and the
 * position of the candidate in the frame's ordered list. */
#define ZOLA_LOC_FIELDS 12

/* Localizes F frames (host arrays).  fit: mode and iteration settings of the
 * per-candidate fit (the reference CPU pipeline: mode ZOLA_FIT_REFERENCE,
 * max_iter 20, z clamp [min_z-0.3, max_z+0.3], photon_threshold 0 -- the
 * pipeline never sets Localization.setMinimumPhotonNumber).  locs: cap x 12,
 * frames in order, each frame's records in acceptance order; n_locs = total
 * found (may exceed cap); per_frame (F ints, may be NULL); rounds (may be
 * NULL) = number of batched fit launches. */
int  zola_localize_frames(zola_detector* det, const zola_fit_config* fit,
                          const zola_accept_config* acc, int F, const float* frames, int cap,
                          double* locs, int* n_locs, int* per_frame, int* rounds);

/* As zola_localize_frames for an sCMOS camera (the isSCMOS branch of
 * LocalizationThread.run, LocalizationPipeline.java:597-634, 797-850,
 * 1168-1211): adu the raw frames (F x W x H), offset, gain and variance the
 * camera maps (W x H, utils/SCMOScamera.java:78-96).  Detection runs on
 * (adu - offset)/gain + variance[0]/gain[0]^2 clamped at 0; the walk on
 * (adu - offset)/gain + variance/gain^2 per pixel, with the map patch in
 * getAandB (patch - map), in the fit's model and in the accepted model's
 * mask and score. */
int  zola_localize_frames_scmos(zola_detector* det, const zola_fit_config* fit,
                                const zola_accept_config* acc, int F, const float* adu,
                                const float* offset, const float* gain, const float* variance, int cap,
                                double* locs, int* n_locs, int* per_frame, int* rounds);

/* ---------------------------------------------------------------------------
 * Calibration bead preparation, replacing InitBackgroundAndPhotonNumber
 * (process/InitBackgroundAndPhotonNumber.java: constructors 50-97,
 * initBackgroundAndPhotonNumberMany 99-468, refinePoints 563-618,
 * computeBackground 622-904, with utils/PolynomialFit.java at order 2 and
 * utils/GaussianElimination.java), which ZOLA.java:2671-2675 builds from the
 * bead z-stack before the phase retrieval.  Host arrays; the stack is
 * processed in FP64 on the context's device and stream.
 *
 * adu: S x W x H raw stack [z][i][ii] (i the ImageJ x column, ii fastest),
 * S*W*H < 2^31.  xp, yp: nb seeds (float pixels), refined in place.  P: patch
 * side (sizePatchPhaseRet).  Outputs: patches nb x S x P x P (converted stack
 * + var/gain^2 for sCMOS inside the frame, B outside), B nb (mean of the
 * polynomial background over the bead square), A nb x S (photon sum of each
 * slice minus B, averaged over +-max((int)(0.5/zstep)/2, 1) slices),
 * background W x H (the fitted polynomial; may be NULL), diag (may be NULL).
 * Returns ZOLA_EINVAL, with the message of zola_last_error, when the
 * background normal equations are singular (the reference fails there too). */
typedef struct zola_bead_diag {
    double mean0, std0;   /* mean and std of the whole stack                   */
    double mean, std;     /* mean over v <= mean0 + 2 std0; std over v <= mean0 + 3 std0 */
    double coef[6];       /* background a0 + a1 x + a2 x^2 + a3 y + a4 xy + a5 y^2, x = i */
    int size_filt;        /* slice filter half-width                           */
    int miniposition;     /* centre slice of the background window             */
    int ndata;            /* pixels in the background fit                      */
} zola_bead_diag;

/* CCD/EMCCD constructor (50-77): photons = (adu*adu_factor/gain -
 * offset*adu_factor/gain) + 1, clamped at 0. */
int  zola_bead_prepare(zola_ctx* ctx, int S, int W, int H, const float* adu, double adu_factor,
                       double gain, double offset, double zstep, int P, int nb, float* xp, float* yp,
                       double* patches, double* B, double* A, double* background,
                       zola_bead_diag* diag);

/* sCMOS constructor (79-97): photons = (adu - offset)/gain, clamped at 0;
 * offset, gain, variance W x H maps (utils/SCMOScamera.java:90-96, var/gain^2
 * formed in FP64).  scmos_patches nb x P x P (may be NULL) as the reference
 * leaves them: var/gain^2 for the last bead only, zero for the others (the
 * array is re-allocated inside its bead loop, 433-447). */
int  zola_bead_prepare_scmos(zola_ctx* ctx, int S, int W, int H, const float* adu, const float* offset,
                             const float* gain, const float* variance, double zstep, int P, int nb,
                             float* xp, float* yp, double* patches, double* B, double* A,
                             double* scmos_patches, double* background, zola_bead_diag* diag);

#ifdef __cplusplus
}
#endif
#endif
