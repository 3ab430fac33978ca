// zola_ctx.hpp -- the context behind the opaque zola_ctx handle, shared by
// the translation units of the library (C ABI, detector, calibration).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_psf64.hpp"
#include "zola_psf128.hpp"

struct zola_ctx {
    zola_optics o{};
    int N = 0, R = 0, s2 = 0, D = 0, J = 0;
    double center = 0, ringsize = 0, radius_px = 0, kappa = 0, kz0 = 0, kzo0 = 0;
    std::vector<int> di, dj, wu, wv;
    std::vector<double> kz, kzo, amp, Z, coef;
    // phase on the disk points (FP64); after a coefficient update that went
    // straight to the device tables it is recomputed on first host use
    mutable std::vector<double> phase;
    mutable bool phase_stale = false;
    zola::Model64 model{};
    float* d_tab = nullptr;  // 4 x 1024 floats: phi, amp, kz', kzo'
    zola::Model128 model128{};
    float4* d_tab128 = nullptr;  // 64 x 64 window {phi, amp, kz', kzo'} (N = 128)
    float* d_taps128 = nullptr;  // h1[128], hd[128], kwin[64][128] (N = 128)
    std::vector<float> taps128;  // host copy of d_taps128
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    // grow-only staging for the host-pointer entry points
    void* d_stage = nullptr;
    size_t stage_bytes = 0;
    float* d_tmpl = nullptr;      // template stack for zola_fit_init_dev, then its per-template sums
    size_t tmpl_bytes = 0;
    // the template stack is kept while the model and its z grid are unchanged
    unsigned model_ver = 0;       // bumped whenever the pupil tables or blur taps change
    unsigned tmpl_ver = ~0u;
    int tmpl_nZ = 0;
    double tmpl_key[3] = {};      // zmin_t, zstep_t, zfocus
    double* d_Zw = nullptr;       // Zernike basis on the 32x32 window, J x 1024 (FP64)
    float* d_qbar = nullptr;      // per-PSF dL/dphase for the adjoint, grow-only
    size_t qbar_bytes = 0;
    int* d_queue = nullptr;       // ROI counters of the persistent fits (one per concurrent group)
    float* scratch = nullptr;     // device address of the fixed allocation-free scratch (zola_noalloc.cpp)
    // N = 256 (zola_grid256.hip): FP64 pupil window, Gaussian spectra, twiddles, work buffers
    double4* d_tab256 = nullptr;
    double* d_gk256 = nullptr;
    double2* d_tw256 = nullptr;
    char* d_work256 = nullptr;
    size_t work256_bytes = 0;
    // graph of the last mixed-size fit (one kernel node per ROI-size group)
    hipGraphExec_t mixed_exec = nullptr;
    hipStream_t side[8] = {};
    hipStream_t cap = nullptr;    // private capture stream (the context stream may be the legacy default)
    hipEvent_t fork = nullptr, join[8] = {};
    // work layout of the 64-grid fits per ROI side (17, 21, 31, 41) and mode
    // (reference form, LM): ZOLA_LAYOUT_*
    int layout[4][2] = {};
    // LM-form layout of the 128-grid fits (ZOLA_LAYOUT_AUTO, QUAD or OCT)
    int layout128 = 0;
    // set by the calibration loops around their evaluations: the full-model
    // evaluation writes dL/da instead of adding to it (no clearing pass)
    bool da_assign = false;
    // Entry points on one context run one at a time (CtxLock): host threads may
    // share a handle, their calls are serialised.
    mutable std::recursive_mutex mu;
};

// Holds a context's lock for the duration of an entry point (no-op for a
// null context, which the entry point then rejects).
class CtxLock {
  public:
    explicit CtxLock(const zola_ctx* c) : m_(c ? &c->mu : nullptr) {
        if (m_) m_->lock();
    }
    ~CtxLock() {
        if (m_) m_->unlock();
    }
    CtxLock(const CtxLock&) = delete;
    CtxLock& operator=(const CtxLock&) = delete;

  private:
    std::recursive_mutex* m_;
};

// Looks up the device address of the fixed scratch array of the allocation-free
// entry points for the current device (once per device) into c->scratch.
int zola_scratch_setup(zola_ctx* c);

// Launches without argument checks, on stream st (which may be capturing): the
// PSF synthesis of zola_psf_forward_dev and the fit of zola_fit_batch_ext_dev.
int zola_psf_launch(zola_ctx* c, int B, const float* params, float* out, hipStream_t st);
int zola_fit_launch(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels, const double* init,
                    const float* zoil, const float* varg, double* theta, double* nll, double* crlb, int* iters,
                    int* status, hipStream_t st);

// FD Fisher matrices (nz x 25) of the focus-centre search at depths zs, FP64 on
// the device (zola_center_dev.hip).
int zola_center_fisher_dev(zola_ctx* c, int nz, const double* zs, double zfocus, double photons, double bg,
                           double h, double* fisher);

// Kernel layout of a 64-grid fit: 0 one wave per ROI, 1 four waves per ROI,
// 2 two waves per ROI (LM form, R <= 21).
int zola_fit_layout(const zola_ctx* c, int R, int mode);
// Kernel layout of a 128-grid fit: 1 four waves per ROI, 3 eight waves (LM form).
int zola_fit_layout128(const zola_ctx* c, int mode);

constexpr int kGrid256 = 256;  // pupil grid of the calibration-only contexts

// phase(p) = sum_j Z_j(p) a_j on the disk points, summed in j order
// (ZernikePhase.java:200-224); the device update of the tables forms the same
// products and sums in the same order.
inline void compute_phase(const zola_ctx* c) {
    c->phase.assign(c->D, 0.0);
    for (int j = 0; j < c->J; j++)
        for (int p = 0; p < c->D; p++) c->phase[p] += c->Z[(size_t)j * c->D + p] * c->coef[j];
    c->phase_stale = false;
}
inline void ensure_phase(const zola_ctx* c) {
    if (c->phase_stale) compute_phase(c);
}

// The 64-grid adjoint writes B x 1024 floats of dL/dphase; the 1024 FP64 sums
// over the PSFs that the Zernike projection forms follow them in the same
// buffer (B x 4096 bytes from its start, so 8-byte aligned).
constexpr size_t kQBytes = 1024 * sizeof(double);
inline double* qsum_slot(float* qbar, int B) { return (double*)(qbar + (size_t)B * 1024); }

inline int set_err(zola_ctx* c, int code, const char* fmt, ...);

// Entry points of the localization path (fits, initial values, detection, the
// focus-centre search) refuse N = 256 contexts, which serve PSF synthesis, the
// adjoint and calibration.
#define ZOLA_REFUSE_GRID256(c)                                                                   \
    do {                                                                                         \
        if ((c)->N == kGrid256)                                                                  \
            return set_err((c), ZOLA_EUNSUPPORTED,                                               \
                           "N=256 contexts serve PSF synthesis, the adjoint and calibration");   \
    } while (0)

inline int set_err(zola_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIP_TRY(ctx, expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err(ctx, ZOLA_EHIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)


inline int stage(zola_ctx* c, size_t bytes) {
    if (bytes <= c->stage_bytes) return ZOLA_OK;
    if (c->d_stage) HIP_TRY(c, hipFree(c->d_stage));
    c->d_stage = nullptr;
    c->stage_bytes = 0;
    HIP_TRY(c, hipMalloc(&c->d_stage, bytes));
    c->stage_bytes = bytes;
    return ZOLA_OK;
}


// FP64 PSF of the reference pipeline on the host (process/cpu/PSFPhaseMany.java:553-622):
// N x N pupil field, unitary inverse FFT, |E|^2 on the s2 grid, Gaussian blur by
// FFT with the float-rounded kernel of PSFPhaseMany.java:387-404, modulus, R x R
// crop, normalised to sum 1.  Used where a result must agree with an FP64
// computation to rounding (detector templates).
void zola_psf_f64_host(const zola_ctx* c, double x, double y, double zoil, double zwat, double* out);
// Same for an R x R crop of any side R <= N (blur grid pow2(min(2R, N))).
void zola_psf_f64_host_r(const zola_ctx* c, int R, double x, double y, double zoil, double zwat, double* out);
