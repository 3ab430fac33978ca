// zola_frames.cpp -- frame detector and frame localization pipeline (host side).
//
// zola_detector: process/cpu/Predetection.java:63-593.  Templates are the
// reference PSF in FP64 (zola_psf_f64_host), the correlation maps and the
// candidate search run on the GPU in FP64 (zola_detect.hip).
//
// zola_localize_frames: process/cpu/LocalizationPipeline.java:721-1073
// (LocalizationThread.run), myGetMaxPositions (1218-1250), getAandB
// (1119-1162) and scoreCompute (1099-1115).  The reference walks each frame's
// candidates in score order and fits one at a time; whether a candidate is
// fitted depends on the mask left by earlier ones, but the fit itself depends
// only on the frame and the candidate.  So the walk is replayed exactly on the
// host while the fits run in batches on the GPU: each round replays every
// frame up to its first candidate without a fit result, queues that fit plus
// the later candidates whose own decision cannot depend on any pending one,
// and fits the whole queue in one launch.  The replay, and therefore every
// decision, is the reference's; the look-ahead only chooses what to fit early.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_ctx.hpp"
#include "zola_detect.h"

namespace {

constexpr double PI = 3.14159265358979323846;
using cd = std::complex<double>;

int ref_pow2(int x) { return (int)std::pow(2.0, (int)std::ceil(std::log((double)x) / std::log(2.0))); }

// Unitary radix-2 FFT of one strided line (forward sign -1, inverse +1).
void fft_line(int n, cd* a, long stride, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i * stride], a[j * stride]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int k = 0; k < half; k++) {
            const double ang = sign * 2.0 * PI * (double)k / (double)len;
            const cd w(std::cos(ang), std::sin(ang));
            for (int s = 0; s < n; s += len) {
                cd& u = a[(s + k) * stride];
                cd& v = a[(s + k + half) * stride];
                const cd t = w * v;
                v = u - t;
                u = u + t;
            }
        }
    }
    const double sc = 1.0 / std::sqrt((double)n);
    for (int i = 0; i < n; i++) a[i * stride] *= sc;
}

void fft2(int n1, int n2, cd* a, int sign) {
    for (int i = 0; i < n1; i++) fft_line(n2, a + (long)i * n2, 1, sign);
    for (int j = 0; j < n2; j++) fft_line(n1, a + j, n2, sign);
}

}  // namespace

void zola_psf_f64_host(const zola_ctx* c, double x, double y, double zoil, double zwat, double* out) {
    zola_psf_f64_host_r(c, c->R, x, y, zoil, zwat, out);
}

void zola_psf_f64_host_r(const zola_ctx* c, int R, double x, double y, double zoil, double zwat, double* out) {
    const int N = c->N, s2 = ref_pow2(std::min(2 * R, N));
    const double cc = c->center, xs = c->o.xystep;
    std::vector<cd> E((size_t)N * N, cd(0, 0));
    ensure_phase(c);
    for (int p = 0; p < c->D; p++) {
        const int i = c->di[p], ii = c->dj[p];
        const double kx = (i - cc) * (2 * PI) / (xs * N), ky = (ii - cc) * (2 * PI) / (xs * N);
        const double pp = kx * x + ky * y + c->phase[p];
        const double q = pp + c->kz[p] * zwat - c->kzo[p] * zoil;
        E[(size_t)((i + N / 2) % N) * N + (ii + N / 2) % N] = cd(c->amp[p] * std::cos(q), c->amp[p] * std::sin(q));
    }
    fft2(N, N, E.data(), +1);
    std::vector<int> shiftN(s2), nextS(s2), lastS(R);
    for (int i = 0, j = N / 2 - s2 / 2; i < s2; i++, j++) {
        shiftN[i] = (j + N / 2) % N;
        nextS[i] = (i + s2 / 2) % s2;
    }
    for (int i = 0, j = s2 / 2 - R / 2; i < R; i++, j++) lastS[i] = (j + s2 / 2) % s2;
    std::vector<cd> B((size_t)s2 * s2);
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            const cd e = E[(size_t)shiftN[i] * N + shiftN[ii]];
            B[(size_t)nextS[i] * s2 + nextS[ii]] = cd(e.real() * e.real() + e.imag() * e.imag(), 0.0);
        }
    fft2(s2, s2, B.data(), -1);
    const double sigF = (1.0 / c->o.sigma) * s2 / (PI * 2.0), sigpow = sigF * sigF, c2 = (double)(s2 / 2);
    for (int i = 0; i < s2; i++)
        for (int ii = 0; ii < s2; ii++) {
            const double u = nextS[i], v = nextS[ii];
            const double g = (double)(float)std::exp(-.5 * ((u - c2) * (u - c2) + (v - c2) * (v - c2)) / sigpow);
            B[(size_t)i * s2 + ii] *= g;
        }
    fft2(s2, s2, B.data(), +1);
    double sum = 0;
    for (int i = 0; i < R; i++)
        for (int ii = 0; ii < R; ii++) {
            const cd b = B[(size_t)lastS[i] * s2 + lastS[ii]];
            out[i * R + ii] = std::sqrt(b.real() * b.real() + b.imag() * b.imag());
            sum += out[i * R + ii];
        }
    for (int k = 0; k < R * R; k++) out[k] /= sum;
}

struct zola_detector {
    zola_ctx* ctx = nullptr;
    zola::DetGeom g{};
    std::vector<double> range, psf, var_psf, norm;
    double2* d_tw = nullptr;
    double2* d_tspec = nullptr;
    double* d_varpsf = nullptr;
    double* d_norm = nullptr;
    // per-batch buffers (grow-only)
    int fb = 0, nr = 0;  // nr: template pairs per correlation pass
    float* d_frames = nullptr;
    double *d_sum1 = nullptr, *d_sq1 = nullptr, *d_var = nullptr, *d_corr = nullptr;
    double2 *d_spec = nullptr, *d_work = nullptr;
    int cap = 0, cap_fb = 0;
    zola::DetCand* d_cand = nullptr;
    int* d_count = nullptr;
    double* d_sorted = nullptr;
    // frame-pipeline buffers (grow-only)
    int fit_cap = 0;
    int* d_roi = nullptr;
    double *d_init = nullptr, *d_th = nullptr, *d_nll = nullptr, *d_cr = nullptr;
    int *d_it = nullptr, *d_st = nullptr;
    float *d_px = nullptr, *d_par = nullptr, *d_psf = nullptr;
    double* d_mstat = nullptr;    // per fit: max(mu - B), chi2 score
    uint32_t* d_mbits = nullptr;  // per fit: model mask bits
    // sCMOS walk: frames of the walk (the detection frames differ), map, map patches
    float* d_locf = nullptr;
    int locf_fb = 0;
    float* d_vmap = nullptr;
    float* d_vpx = nullptr;
    // one fit round as a graph (captured on cap_st, replayed on the context stream)
    // and its pinned host staging: roi/init in, theta/nll/crlb/status/model
    // statistics/mask bits out
    hipStream_t cap_st = nullptr;
    hipGraphExec_t round_exec = nullptr;
    int* h_roi = nullptr;
    double* h_init = nullptr;
    double* h_res = nullptr;      // per fit: theta 5, nll, crlb 3
    int* h_st = nullptr;
    double* h_mstat = nullptr;
    uint32_t* h_bits = nullptr;
};

namespace {

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void free_batch(zola_detector* d) {
    dfree(d->d_frames); dfree(d->d_sum1); dfree(d->d_sq1); dfree(d->d_var); dfree(d->d_corr);
    dfree(d->d_spec); dfree(d->d_work); dfree(d->d_count);
    d->fb = 0;
}

void free_cand(zola_detector* d) {
    dfree(d->d_cand); dfree(d->d_sorted);
    d->cap = 0;
    d->cap_fb = 0;
}

template <class T>
void hfree(T*& p) {
    if (p) (void)hipHostFree((void*)p);
    p = nullptr;
}

void free_fit(zola_detector* d) {
    if (d->round_exec) {
        (void)hipStreamSynchronize(d->ctx->stream);
        (void)hipGraphExecDestroy(d->round_exec);
        d->round_exec = nullptr;
    }
    dfree(d->d_roi); dfree(d->d_init); dfree(d->d_th); dfree(d->d_nll); dfree(d->d_cr);
    dfree(d->d_it); dfree(d->d_st); dfree(d->d_px); dfree(d->d_par); dfree(d->d_psf);
    dfree(d->d_mstat); dfree(d->d_mbits); dfree(d->d_vpx);
    hfree(d->h_roi); hfree(d->h_init); hfree(d->h_res); hfree(d->h_st); hfree(d->h_mstat); hfree(d->h_bits);
    d->fit_cap = 0;
}

// Frames per detection batch: correlation maps dominate (nZ x W x H doubles per frame).
int batch_frames(const zola_detector* d, int F) {
    const zola::DetGeom& g = d->g;
    const double per = (double)g.nZ * g.W * g.H * 8 + (double)g.W * g.H * (4 + 24) + (double)g.FX * g.FY * 16;
    // up to 2048 frames in flight (at most 64 GiB of correlation maps and
    // spectra, a fraction of the 288 GB of HBM): more frames per batch give each
    // fit round more ROIs and the walk more threads
    int fb = (int)std::min(2048.0, std::max(1.0, 64.0 * (1u << 30) / per));
    return std::max(1, std::min(fb, F));
}

int ensure_batch(zola_detector* d, int fb) {
    zola_ctx* c = d->ctx;
    if (fb <= d->fb) return ZOLA_OK;
    free_batch(d);
    const zola::DetGeom& g = d->g;
    const size_t WH = (size_t)g.W * g.H;
    int nr = (int)std::max<size_t>(1, (size_t)(512u << 20) / ((size_t)fb * g.W * g.FY * 16));
    d->nr = std::min(nr, (g.nZ + 1) / 2);  // template pairs per correlation pass
    HIP_TRY(c, hipMalloc((void**)&d->d_frames, fb * WH * sizeof(float)));
    HIP_TRY(c, hipMalloc((void**)&d->d_sum1, fb * WH * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_sq1, fb * WH * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_var, fb * WH * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_corr, (size_t)fb * g.nZ * WH * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_spec, (size_t)fb * g.FX * g.FY * sizeof(double2)));
    HIP_TRY(c, hipMalloc((void**)&d->d_work, (size_t)fb * d->nr * g.W * g.FY * sizeof(double2)));
    HIP_TRY(c, hipMalloc((void**)&d->d_count, fb * sizeof(int)));
    d->fb = fb;
    return ZOLA_OK;
}

int ensure_cand(zola_detector* d, int fb, int cap) {
    zola_ctx* c = d->ctx;
    if (cap <= d->cap && fb <= d->cap_fb) return ZOLA_OK;
    cap = std::max(cap, d->cap);
    fb = std::max(fb, d->cap_fb);
    free_cand(d);
    HIP_TRY(c, hipMalloc((void**)&d->d_cand, (size_t)fb * cap * sizeof(zola::DetCand)));
    HIP_TRY(c, hipMalloc((void**)&d->d_sorted, (size_t)fb * cap * 7 * sizeof(double)));
    d->cap = cap;
    d->cap_fb = fb;
    return ZOLA_OK;
}

// Correlation maps of the fb frames already in d_frames.
int run_correlation(zola_detector* d, int fb) {
    zola_ctx* c = d->ctx;
    const zola::DetGeom& g = d->g;
    hipStream_t st = c->stream;
    HIP_TRY(c, zola_launch_window_stats(g, fb, d->d_frames, d->d_sum1, d->d_sq1, d->d_norm, d->d_var, st));
    HIP_TRY(c, zola_launch_frame_spectrum(g, fb, d->d_frames, d->d_spec, d->d_tw, st));
    for (int r0 = 0; r0 < g.nZ; r0 += 2 * d->nr) {
        const int np = std::min(d->nr, (g.nZ - r0 + 1) / 2);
        HIP_TRY(c, zola_launch_correlation(g, fb, r0, np, d->d_spec, d->d_tspec, d->d_var, d->d_varpsf,
                                           d->d_work, d->d_corr, d->d_tw, st));
    }
    return ZOLA_OK;
}

// Candidates of the fb frames in d_frames, ordered into d_sorted (fb x cap x 7);
// counts[f] = number found.
int run_detect(zola_detector* d, int fb, std::vector<int>& counts) {
    zola_ctx* c = d->ctx;
    hipStream_t st = c->stream;
    int rc = run_correlation(d, fb);
    if (rc) return rc;
    rc = ensure_cand(d, fb, std::max(d->cap, 4096));
    if (rc) return rc;
    counts.assign(fb, 0);
    for (int pass = 0; pass < 2; pass++) {
        HIP_TRY(c, hipMemsetAsync(d->d_count, 0, fb * sizeof(int), st));
        HIP_TRY(c, zola_launch_local_max(d->g, fb, d->d_corr, d->cap, d->d_cand, d->d_count, st));
        HIP_TRY(c, hipMemcpyAsync(counts.data(), d->d_count, fb * sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
        const int mx = *std::max_element(counts.begin(), counts.end());
        if (mx <= d->cap) break;
        rc = ensure_cand(d, fb, mx + mx / 4);
        if (rc) return rc;
    }
    const int mx = *std::max_element(counts.begin(), counts.end());
    if (mx > d->cap) return set_err(c, ZOLA_EHIP, "candidate buffer overflow");
    HIP_TRY(c, zola_launch_rank_sort(fb, d->cap, mx, d->d_cand, d->d_count, d->d_sorted, st));
    return ZOLA_OK;
}

}  // namespace

extern "C" {

int zola_detector_create(zola_ctx* c, int W, int H, double zmin, double zmax, double zstep,
                         double threshold, double zfocus, zola_detector** out) {
    CtxLock lk_(c);
    if (!c || !out) return ZOLA_EINVAL;
    *out = nullptr;
    ZOLA_REFUSE_GRID256(c);
    if (W <= 0 || H <= 0 || !(zstep > 0) || !(zmax > zmin))
        return set_err(c, ZOLA_EINVAL, "bad detector geometry or z range");
    const int R = c->R;
    if (W <= 2 * std::max(3, R / 2) || H <= 2 * std::max(3, R / 2))
        return set_err(c, ZOLA_EINVAL, "frame %dx%d too small for ROI side %d", W, H, R);
    const int FX = ref_pow2(2 * W), FY = ref_pow2(2 * H);
    if (FX > zola::kFftElems || FY > zola::kFftElems || FX < 16 || FY < 16)
        return set_err(c, ZOLA_EUNSUPPORTED, "padded frame %dx%d outside 16..%d", FX, FY, zola::kFftElems);
    const int number = (int)std::ceil((zmax - zmin) / zstep);
    if (number <= 0) return set_err(c, ZOLA_EINVAL, "empty template range");
    zola_detector* d = new zola_detector();
    d->ctx = c;
    d->g = zola::DetGeom{W, H, FX, FY, R, number, threshold, zstep, c->o.xystep};
    // range (Predetection.java:118-125): accumulated z
    d->range.resize(number);
    double z = zmin;
    for (int i = 0; i < number; i++) {
        d->range[i] = z;
        z += zstep;
    }
    // dictionary (Predetection.java:154-260)
    const int n = R * R;
    d->psf.resize((size_t)number * n);
    d->var_psf.resize(number);
    std::vector<double> centred((size_t)number * n);
    for (int r = 0; r < number; r++) {
        double* p = d->psf.data() + (size_t)r * n;
        zola_psf_f64_host(c, 0, 0, zfocus, d->range[r], p);
        double mean = 0;
        for (int k = 0; k < n; k++) mean += p[k];
        mean /= (double)R * R;
        double v = 0;
        for (int k = 0; k < n; k++) {
            const double t = p[k] - mean;
            centred[(size_t)r * n + k] = t;
            v += t * t;
        }
        d->var_psf[r] = v / (double)(R * R);
    }
    // normImage (Predetection.java:104-133), including its mixed use of the two sizes
    d->norm.resize((size_t)W * H);
    for (int i = 0; i < W; i++)
        for (int ii = 0; ii < H; ii++) {
            int vx1 = R / 2, vx2 = R / 2, vy1 = R / 2, vy2 = R / 2;
            if (i < R / 2 - 1) vx1 = i + 1;
            if (i > W - R / 2 - 2) vx2 = H - i - 1;
            if (ii < R / 2 - 1) vy1 = ii + 1;
            if (ii > W - R / 2 - 2) vy2 = H - ii - 2;
            d->norm[(size_t)i * H + ii] = (double)((vx1 + vx2) * (vy1 + vy2));
        }
    *out = d;
    std::vector<double2> tw(zola::kTwN);
    zola_twiddle_table(tw.data());
    HIP_TRY(c, hipMalloc((void**)&d->d_tw, tw.size() * sizeof(double2)));
    HIP_TRY(c, hipMemcpy(d->d_tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMalloc((void**)&d->d_varpsf, number * sizeof(double)));
    HIP_TRY(c, hipMemcpy(d->d_varpsf, d->var_psf.data(), number * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMalloc((void**)&d->d_norm, d->norm.size() * sizeof(double)));
    HIP_TRY(c, hipMemcpy(d->d_norm, d->norm.data(), d->norm.size() * sizeof(double), hipMemcpyHostToDevice));
    double* d_tmpl = nullptr;
    HIP_TRY(c, hipMalloc((void**)&d_tmpl, centred.size() * sizeof(double)));
    hipError_t e = hipMemcpy(d_tmpl, centred.data(), centred.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&d->d_tspec, (size_t)number * FX * FY * sizeof(double2));
    if (e == hipSuccess) e = zola_launch_template_spectrum(d->g, d_tmpl, d->d_tspec, d->d_tw, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_tmpl);
    if (e != hipSuccess) return set_err(c, ZOLA_EHIP, "template spectra: %s", hipGetErrorString(e));
    return ZOLA_OK;
}

int zola_detector_destroy(zola_detector* d) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d) return ZOLA_OK;
    free_batch(d);
    free_cand(d);
    free_fit(d);
    dfree(d->d_tw);
    dfree(d->d_tspec);
    dfree(d->d_varpsf);
    dfree(d->d_norm);
    dfree(d->d_locf);
    dfree(d->d_vmap);
    if (d->cap_st) (void)hipStreamDestroy(d->cap_st);
    delete d;
    return ZOLA_OK;
}

int zola_detector_info(const zola_detector* d, int* info) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d || !info) return ZOLA_EINVAL;
    info[0] = d->g.nZ; info[1] = d->g.FX; info[2] = d->g.FY; info[3] = d->g.W; info[4] = d->g.H;
    return ZOLA_OK;
}

int zola_detector_range(const zola_detector* d, double* range) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d || !range) return ZOLA_EINVAL;
    std::memcpy(range, d->range.data(), d->range.size() * sizeof(double));
    return ZOLA_OK;
}

int zola_detector_templates(const zola_detector* d, double* psf) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d || !psf) return ZOLA_EINVAL;
    std::memcpy(psf, d->psf.data(), d->psf.size() * sizeof(double));
    return ZOLA_OK;
}

int zola_detector_correlation(zola_detector* d, const float* frame, double* corr) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d) return ZOLA_EINVAL;
    zola_ctx* c = d->ctx;
    if (!frame || !corr) return set_err(c, ZOLA_EINVAL, "null buffer");
    int rc = ensure_batch(d, 1);
    if (rc) return rc;
    const zola::DetGeom& g = d->g;
    const size_t WH = (size_t)g.W * g.H;
    HIP_TRY(c, hipMemcpyAsync(d->d_frames, frame, WH * sizeof(float), hipMemcpyHostToDevice, c->stream));
    rc = run_correlation(d, 1);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(corr, d->d_corr, (size_t)g.nZ * WH * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

int zola_detect_frames(zola_detector* d, int F, const float* frames, int cap, double* cand, int* counts) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d) return ZOLA_EINVAL;
    zola_ctx* c = d->ctx;
    if (F < 0 || cap < 0 || (F > 0 && (!frames || !counts || (cap > 0 && !cand))))
        return set_err(c, ZOLA_EINVAL, "bad detection arguments");
    if (F == 0) return ZOLA_OK;
    const zola::DetGeom& g = d->g;
    const size_t WH = (size_t)g.W * g.H;
    const int fbmax = batch_frames(d, F);
    int rc = ensure_batch(d, fbmax);
    if (rc) return rc;
    std::vector<int> cnt;
    for (int f0 = 0; f0 < F; f0 += fbmax) {
        const int fb = std::min(fbmax, F - f0);
        HIP_TRY(c, hipMemcpyAsync(d->d_frames, frames + f0 * WH, fb * WH * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream));
        rc = run_detect(d, fb, cnt);
        if (rc) return rc;
        for (int f = 0; f < fb; f++) {
            counts[f0 + f] = cnt[f];
            const int m = std::min(cnt[f], cap);
            if (m > 0)
                HIP_TRY(c, hipMemcpyAsync(cand + (size_t)(f0 + f) * cap * 7, d->d_sorted + (size_t)f * d->cap * 7,
                                          (size_t)m * 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return ZOLA_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ frame pipeline
namespace {

// A fit and the statistics of its model that the walk uses (k_model_stats):
// max(mu - B), the chi2 score and the model mask bits (P.bits, nw words).
struct FitRes {
    double th[5], crlb[3], nll;
    int status;
    double mmax, score;
};

struct FrameRun {
    int f = 0;
    const float* img = nullptr;
    int n = 0;
    std::vector<double> cand;          // n x 7, ordered
    std::vector<uint8_t> mask;         // W x H
    int ptr = 0;
    bool done = false;
    std::vector<int> res_idx;          // n: index into results, -1 none
    std::vector<uint8_t> queued;       // n
    std::vector<double> ab;            // n x 3 (A, B, max patch)
    std::vector<uint8_t> ab_ok;        // n
    std::vector<double> locs;          // records
    std::vector<int> q;                // candidates queued for the next fit round
};

struct Pipe {
    const zola_detector* d;
    const zola_fit_config* fit;
    const zola_accept_config* acc;
    int W, H, R, half, foot;
    std::vector<FitRes> results;
    std::vector<uint32_t> bits;        // nw words per result
    int nw = 0;
    const float* vmap = nullptr;       // sCMOS: W x H var/gain^2 (host), or null
    std::vector<int> q_frame, q_k;     // fit queue (run index, candidate)
};

inline int cand_i(const FrameRun& fr, int k) { return (int)std::floor(fr.cand[(size_t)k * 7 + 2] + 0.5); }
inline int cand_j(const FrameRun& fr, int k) { return (int)std::floor(fr.cand[(size_t)k * 7 + 3] + 0.5); }
inline int cand_r(const FrameRun& fr, int k) { return (int)std::floor(fr.cand[(size_t)k * 7 + 1] + 0.5); }

// myGetMaxPositions bounds (LocalizationPipeline.java:1226)
inline bool in_bounds(const Pipe& P, int ci, int cj) {
    return ci > P.R / 2 + 1 && cj > P.R / 2 + 1 && ci < P.W - P.R / 2 - 1 && cj < P.H - P.R / 2 - 1;
}

bool patch_touches(const Pipe& P, const std::vector<uint8_t>& m, int ci, int cj) {
    for (int a = -P.half; a < P.R - P.half; a++) {
        const uint8_t* row = m.data() + (size_t)(ci + a) * P.H + cj - P.half;
        uint8_t any = 0;  // OR over the row without an early exit: vectorises
        for (int b = 0; b < P.R; b++) any |= row[b];
        if (any) return true;
    }
    return false;
}

const double* get_ab(const Pipe& P, FrameRun& fr, int k) {
    if (!fr.ab_ok[k]) {
        const int ci = cand_i(fr, k), cj = cand_j(fr, k), r = cand_r(fr, k);
        const int n = P.R * P.R;
        const double* psf = P.d->psf.data() + (size_t)r * n;
        double maxPatch = -INFINITY, maxi = -INFINITY;
        // with an sCMOS map the sums and the maximum are of patch - map (1168-1211)
        auto val = [&](int a, int b) {
            const size_t q = (size_t)(ci + a - P.half) * P.H + cj + b - P.half;
            return P.vmap ? (double)fr.img[q] - (double)P.vmap[q] : (double)fr.img[q];
        };
        for (int a = 0; a < P.R; a++)
            for (int b = 0; b < P.R; b++) {
                const double v = val(a, b);
                if (psf[a * P.R + b] > maxi) maxi = psf[a * P.R + b];
                if (v > maxPatch) maxPatch = v;
            }
        const double thr = maxi * .1;
        double cB = 0, Bs = 0, As = 0, cA = 0;
        for (int a = 0; a < P.R; a++)
            for (int b = 0; b < P.R; b++) {
                const double v = val(a, b);
                if (psf[a * P.R + b] < thr) {
                    Bs += v;
                    cB++;
                } else {
                    As += v;
                    cA++;
                }
            }
        if (cB != 0) Bs /= cB;
        else Bs = 0;
        As -= cA * Bs;
        fr.ab[(size_t)k * 3 + 0] = As;
        fr.ab[(size_t)k * 3 + 1] = Bs;
        fr.ab[(size_t)k * 3 + 2] = maxPatch;
        fr.ab_ok[k] = 1;
    }
    return fr.ab.data() + (size_t)k * 3;
}

inline bool snr_gate(const Pipe& P, const double* ab) { return (ab[2] / ab[1] > P.acc->snr_min) && (ab[1] > 0); }

void queue_fit(Pipe&, FrameRun& fr, int, int k) {
    if (fr.queued[k] || fr.res_idx[k] >= 0) return;
    fr.queued[k] = 1;
    fr.q.push_back(k);
}

// The accept branch of LocalizationThread.run (LocalizationPipeline.java:857-948)
// on a fitted candidate; updates the mask and, when emit, appends the record.
bool apply_fit(const Pipe& P, FrameRun& fr, int k, const FitRes& res, std::vector<uint8_t>& mask, bool emit) {
    const bool lowph = P.fit->mode == ZOLA_FIT_REFERENCE && (res.status & ZOLA_ST_LOWPHOTON);
    if ((res.status & ZOLA_ST_NAN) || lowph) return false;
    const int ci = cand_i(fr, k), cj = cand_j(fr, k);
    const double A = res.th[0], Bg = res.th[1], x = res.th[2], y = res.th[3], z = res.th[4];
    const double dx = P.d->g.xystep;
    if (!(A > P.acc->photon_threshold)) return false;
    if (!(std::fabs(x) < P.acc->max_shift_px * dx && std::fabs(y) < P.acc->max_shift_px * dx)) return false;
    const int px = (int)(ci - (x / dx)), py = (int)(cj - (y / dx));
    if (!(px >= 0 && px < P.W && py >= 0 && py < P.H && !mask[(size_t)px * P.H + py])) return false;
    if (!(z > P.acc->min_z && z < P.acc->max_z)) return false;
    if (emit) {
        const double rec[ZOLA_LOC_FIELDS] = {(double)fr.f,
                                             1000 * (ci * dx - x),
                                             1000 * (cj * dx - y),
                                             1000 * z,
                                             A,
                                             Bg,
                                             res.score,
                                             1000 * res.crlb[0],
                                             1000 * res.crlb[1],
                                             1000 * res.crlb[2],
                                             res.nll,
                                             (double)k};
        fr.locs.insert(fr.locs.end(), rec, rec + ZOLA_LOC_FIELDS);
    }
    // pixels of the model above 0.2 of its peak (mu - B > 0.2 max(mu - B))
    const uint32_t* bits = P.bits.data() + (size_t)fr.res_idx[k] * P.nw;
    for (int a = 0; a < P.R; a++)
        for (int b = 0; b < P.R; b++) {
            const int p = a * P.R + b;
            if ((bits[p >> 5] >> (p & 31)) & 1u) mask[(size_t)(ci + a - P.half) * P.H + cj + b - P.half] = 1;
        }
    return true;
}

void mask_patch(const Pipe& P, const FrameRun& fr, int k, double maxPatch, std::vector<uint8_t>& mask) {
    const int ci = cand_i(fr, k), cj = cand_j(fr, k);
    for (int a = 0; a < P.R; a++)
        for (int b = 0; b < P.R; b++) {
            const size_t idx = (size_t)(ci + a - P.half) * P.H + cj + b - P.half;
            if (fr.img[idx] > maxPatch * P.acc->patch_mask_frac) mask[idx] = 1;
        }
}

// Coarse map of pixels whose mask state is not yet known (8 x 8 blocks).
struct Unc {
    int W, H, BW, BH;
    std::vector<uint8_t> blk, px;
    void init(int w, int h) {
        W = w; H = h; BW = (w + 7) / 8; BH = (h + 7) / 8;
        blk.assign((size_t)BW * BH, 0);
        px.assign((size_t)w * h, 0);
    }
    void mark(int i0, int i1, int j0, int j1) {  // inclusive pixel box
        i0 = std::max(i0, 0); j0 = std::max(j0, 0); i1 = std::min(i1, W - 1); j1 = std::min(j1, H - 1);
        for (int i = i0; i <= i1; i++) std::memset(px.data() + (size_t)i * H + j0, 1, j1 - j0 + 1);
        for (int bi = i0 / 8; bi <= i1 / 8; bi++)
            for (int bj = j0 / 8; bj <= j1 / 8; bj++) blk[(size_t)bi * BH + bj] = 1;
    }
    bool any(int i0, int i1, int j0, int j1) const {
        i0 = std::max(i0, 0); j0 = std::max(j0, 0); i1 = std::min(i1, W - 1); j1 = std::min(j1, H - 1);
        for (int bi = i0 / 8; bi <= i1 / 8; bi++)
            for (int bj = j0 / 8; bj <= j1 / 8; bj++)
                if (blk[(size_t)bi * BH + bj]) return true;
        return false;
    }
};

// Host threads for the walk replay: the cores this process may use, at most 16.
int host_threads() {
    const unsigned n = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(n, 16u));
}

// Look-ahead after the walk stopped at k0: queue the fits that the coming
// rounds will need, as far as the known mask can tell.
void lookahead(Pipe& P, FrameRun& fr, int run, int k0) {
    std::vector<uint8_t> pm = fr.mask;
    Unc unc;
    unc.init(P.W, P.H);
    {
        const int ci = cand_i(fr, k0), cj = cand_j(fr, k0);
        unc.mark(ci - P.half, ci - P.half + P.R - 1, cj - P.half, cj - P.half + P.R - 1);
    }
    for (int k = k0 + 1; k < fr.n; k++) {
        const int ci = cand_i(fr, k), cj = cand_j(fr, k);
        if (!in_bounds(P, ci, cj)) continue;
        if (pm[(size_t)ci * P.H + cj]) continue;  // the mask only grows: skipped for certain
        const bool uncertain = unc.any(ci - P.foot, ci + P.foot, cj - P.foot, cj + P.foot);
        if (uncertain) {
            if (!unc.px[(size_t)ci * P.H + cj]) {
                const double* ab = get_ab(P, fr, k);
                if (snr_gate(P, ab)) queue_fit(P, fr, run, k);
            }
            unc.mark(ci - P.half, ci - P.half + P.R - 1, cj - P.half, cj - P.half + P.R - 1);
            continue;
        }
        if (patch_touches(P, pm, ci, cj)) continue;
        const double* ab = get_ab(P, fr, k);
        if (!snr_gate(P, ab)) {
            mask_patch(P, fr, k, ab[2], pm);
            continue;
        }
        if (fr.res_idx[k] >= 0) {
            if (!apply_fit(P, fr, k, P.results[fr.res_idx[k]], pm, false)) mask_patch(P, fr, k, ab[2], pm);
            continue;
        }
        queue_fit(P, fr, run, k);
        unc.mark(ci - P.half, ci - P.half + P.R - 1, cj - P.half, cj - P.half + P.R - 1);
    }
}

// Replays LocalizationThread.run from the frame's saved state until the walk
// ends or needs a fit that is not available yet.
void replay(Pipe& P, FrameRun& fr, int run) {
    while (!fr.done) {
        int k = -1;
        for (int i = fr.ptr; i < fr.n; i++) {
            const int ci = cand_i(fr, i), cj = cand_j(fr, i);
            if (!fr.mask[(size_t)ci * P.H + cj] && in_bounds(P, ci, cj)) {
                k = i;
                break;
            }
        }
        if (k < 0) {
            fr.done = true;
            fr.ptr = fr.n;
            return;
        }
        const int ci = cand_i(fr, k), cj = cand_j(fr, k);
        if (patch_touches(P, fr.mask, ci, cj)) {
            fr.ptr = k + 1;
            continue;
        }
        const double* ab = get_ab(P, fr, k);
        bool treated = false;
        if (snr_gate(P, ab)) {
            if (fr.res_idx[k] < 0) {
                fr.ptr = k;
                queue_fit(P, fr, run, k);
                lookahead(P, fr, run, k);
                return;
            }
            treated = apply_fit(P, fr, k, P.results[fr.res_idx[k]], fr.mask, true);
        }
        if (!treated) mask_patch(P, fr, k, ab[2], fr.mask);
        fr.ptr = k + 1;
    }
}

int ensure_fit(zola_detector* d, int B) {
    zola_ctx* c = d->ctx;
    if (B <= d->fit_cap) return ZOLA_OK;
    free_fit(d);
    B = std::max(B, 1024);
    const size_t R2 = (size_t)c->R * c->R;
    HIP_TRY(c, hipMalloc((void**)&d->d_roi, (size_t)B * 3 * sizeof(int)));
    HIP_TRY(c, hipMalloc((void**)&d->d_init, (size_t)B * 5 * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_th, (size_t)B * 5 * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_nll, (size_t)B * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_cr, (size_t)B * 3 * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_it, (size_t)B * sizeof(int)));
    HIP_TRY(c, hipMalloc((void**)&d->d_st, (size_t)B * sizeof(int)));
    HIP_TRY(c, hipMalloc((void**)&d->d_px, (size_t)B * R2 * sizeof(float)));
    HIP_TRY(c, hipMalloc((void**)&d->d_par, (size_t)B * 4 * sizeof(float)));
    HIP_TRY(c, hipMalloc((void**)&d->d_psf, (size_t)B * R2 * sizeof(float)));
    HIP_TRY(c, hipMalloc((void**)&d->d_mstat, (size_t)B * 2 * sizeof(double)));
    HIP_TRY(c, hipMalloc((void**)&d->d_mbits, (size_t)B * ((R2 + 31) / 32) * sizeof(uint32_t)));
    HIP_TRY(c, hipMalloc((void**)&d->d_vpx, (size_t)B * R2 * sizeof(float)));
    HIP_TRY(c, hipHostMalloc((void**)&d->h_roi, (size_t)B * 3 * sizeof(int), hipHostMallocDefault));
    HIP_TRY(c, hipHostMalloc((void**)&d->h_init, (size_t)B * 5 * sizeof(double), hipHostMallocDefault));
    HIP_TRY(c, hipHostMalloc((void**)&d->h_res, (size_t)B * 9 * sizeof(double), hipHostMallocDefault));
    HIP_TRY(c, hipHostMalloc((void**)&d->h_st, (size_t)B * sizeof(int), hipHostMallocDefault));
    HIP_TRY(c, hipHostMalloc((void**)&d->h_mstat, (size_t)B * 2 * sizeof(double), hipHostMallocDefault));
    HIP_TRY(c, hipHostMalloc((void**)&d->h_bits, (size_t)B * ((R2 + 31) / 32) * sizeof(uint32_t),
                             hipHostMallocDefault));
    if (!d->cap_st) HIP_TRY(c, hipStreamCreateWithFlags(&d->cap_st, hipStreamNonBlocking));
    d->fit_cap = B;
    return ZOLA_OK;
}

// Fits the queued candidates of this round on the GPU.
int fit_queue(zola_detector* d, Pipe& P, std::vector<FrameRun>& runs) {
    zola_ctx* c = d->ctx;
    hipStream_t st = c->stream;
    const int B = (int)P.q_k.size();
    int rc = ensure_fit(d, B);
    if (rc) return rc;
    const size_t R2 = (size_t)P.R * P.R;
    std::vector<int> roi((size_t)B * 3);
    std::vector<double> init((size_t)B * 5);
    for (int b = 0; b < B; b++) {
        FrameRun& fr = runs[P.q_frame[b]];
        const int k = P.q_k[b];
        const double* ab = get_ab(P, fr, k);
        const double* cv = fr.cand.data() + (size_t)k * 7;
        roi[b * 3 + 0] = P.q_frame[b];
        roi[b * 3 + 1] = cand_i(fr, k);
        roi[b * 3 + 2] = cand_j(fr, k);
        // loc.init(A, B, -dx, -dy, range[r] + dz) (LocalizationPipeline.java:849)
        init[b * 5 + 0] = ab[0];
        init[b * 5 + 1] = ab[1];
        init[b * 5 + 2] = -cv[5];
        init[b * 5 + 3] = -cv[6];
        init[b * 5 + 4] = d->range[cand_r(fr, k)] + cv[4];
    }
    std::memcpy(d->h_roi, roi.data(), roi.size() * sizeof(int));
    std::memcpy(d->h_init, init.data(), init.size() * sizeof(double));
    // The round's device work as one graph, captured on a private stream and
    // replayed on the context stream: ROI centres and starts in, patch gathers,
    // the fit, the model at the fitted parameters (Localization.getPSF: the last
    // likelihood's model) and the statistics the walk needs (max, score, mask
    // bits; not the PSFs), results out.  One synchronisation per round.
    const int nw = (int)((R2 + 31) / 32);
    hipStream_t cs = d->cap_st;
    HIP_TRY(c, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    auto record = [&]() -> int {
        HIP_TRY(c, hipMemcpyAsync(d->d_roi, d->h_roi, (size_t)B * 3 * sizeof(int), hipMemcpyHostToDevice, cs));
        HIP_TRY(c, hipMemcpyAsync(d->d_init, d->h_init, (size_t)B * 5 * sizeof(double), hipMemcpyHostToDevice, cs));
        HIP_TRY(c, zola_launch_gather_rois(P.W, P.H, P.R, B, d->d_roi, P.vmap ? d->d_locf : d->d_frames, d->d_px, cs));
        if (P.vmap) HIP_TRY(c, zola_launch_gather_map(P.W, P.H, P.R, B, d->d_roi, d->d_vmap, d->d_vpx, cs));
        int r = zola_fit_launch(c, P.fit, B, d->d_px, d->d_init, nullptr, P.vmap ? d->d_vpx : nullptr, d->d_th,
                                d->d_nll, d->d_cr, d->d_it, d->d_st, cs);
        if (r) return r;
        HIP_TRY(c, zola_launch_theta_params(B, d->d_th, P.fit->zfocus, d->d_par, cs));
        r = zola_psf_launch(c, B, d->d_par, d->d_psf, cs);
        if (r) return r;
        HIP_TRY(c, zola_launch_model_stats(B, P.R, d->d_psf, d->d_px, P.vmap ? d->d_vpx : nullptr, d->d_th,
                                           P.acc->model_mask_frac, d->d_mstat, d->d_mstat + B, d->d_mbits, cs));
        HIP_TRY(c, hipMemcpyAsync(d->h_res, d->d_th, (size_t)B * 5 * sizeof(double), hipMemcpyDeviceToHost, cs));
        HIP_TRY(c, hipMemcpyAsync(d->h_res + (size_t)B * 5, d->d_nll, (size_t)B * sizeof(double),
                                  hipMemcpyDeviceToHost, cs));
        HIP_TRY(c, hipMemcpyAsync(d->h_res + (size_t)B * 6, d->d_cr, (size_t)B * 3 * sizeof(double),
                                  hipMemcpyDeviceToHost, cs));
        HIP_TRY(c, hipMemcpyAsync(d->h_st, d->d_st, (size_t)B * sizeof(int), hipMemcpyDeviceToHost, cs));
        HIP_TRY(c, hipMemcpyAsync(d->h_mstat, d->d_mstat, (size_t)B * 2 * sizeof(double), hipMemcpyDeviceToHost, cs));
        HIP_TRY(c, hipMemcpyAsync(d->h_bits, d->d_mbits, (size_t)B * nw * sizeof(uint32_t), hipMemcpyDeviceToHost, cs));
        return ZOLA_OK;
    };
    rc = record();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : set_err(c, ZOLA_EHIP, "fit-round capture: %s", hipGetErrorString(ec));
    }
    // the previous round has been synchronised: reuse its executable when only
    // the parameters changed, else instantiate
    bool updated = false;
    if (d->round_exec) {
        hipGraphExecUpdateResult res;
        hipGraphNode_t bad = nullptr;
        updated = hipGraphExecUpdate(d->round_exec, graph, &bad, &res) == hipSuccess &&
                  res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGraphExecDestroy(d->round_exec);
            d->round_exec = nullptr;
        }
    }
    hipError_t eg = hipSuccess;
    if (!updated) eg = hipGraphInstantiate(&d->round_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (eg != hipSuccess) {
        d->round_exec = nullptr;
        return set_err(c, ZOLA_EHIP, "fit-round graph: %s", hipGetErrorString(eg));
    }
    HIP_TRY(c, hipGraphLaunch(d->round_exec, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    const double* th = d->h_res;
    const double* nll = d->h_res + (size_t)B * 5;
    const double* cr = d->h_res + (size_t)B * 6;
    const int* stt = d->h_st;
    const double* mstat = d->h_mstat;
    const size_t base = P.bits.size();
    P.nw = nw;
    P.bits.resize(base + (size_t)B * nw);
    std::memcpy(P.bits.data() + base, d->h_bits, (size_t)B * nw * sizeof(uint32_t));
    for (int b = 0; b < B; b++) {
        FitRes r;
        std::memcpy(r.th, &th[b * 5], 5 * sizeof(double));
        std::memcpy(r.crlb, &cr[b * 3], 3 * sizeof(double));
        r.nll = nll[b];
        r.status = stt[b];
        r.mmax = mstat[b];
        r.score = mstat[B + b];
        FrameRun& fr = runs[P.q_frame[b]];
        fr.res_idx[P.q_k[b]] = (int)P.results.size();
        fr.queued[P.q_k[b]] = 0;
        P.results.push_back(r);
    }
    return ZOLA_OK;
}

}  // namespace

namespace {

// The pipeline on detection frames det (uploaded for the detector) and walk
// frames img (patches, fits; the same array without an sCMOS map), vmap the
// sCMOS var/gain^2 map or null.
int localize_impl(zola_detector* d, const zola_fit_config* fit, const zola_accept_config* acc, int F,
                  const float* det, const float* img, const float* vmap, int cap, double* locs, int* n_locs,
                  int* per_frame, int* rounds) {
    zola_ctx* c = d->ctx;
    *n_locs = 0;
    if (rounds) *rounds = 0;
    if (F == 0) return ZOLA_OK;
    const zola::DetGeom& g = d->g;
    const size_t WH = (size_t)g.W * g.H;
    const int fbmax = batch_frames(d, F);
    int rc = ensure_batch(d, fbmax);
    if (rc) return rc;
    if (vmap) {
        if (d->locf_fb < fbmax) {
            dfree(d->d_locf);
            HIP_TRY(c, hipMalloc((void**)&d->d_locf, (size_t)fbmax * WH * sizeof(float)));
            d->locf_fb = fbmax;
        }
        if (!d->d_vmap) HIP_TRY(c, hipMalloc((void**)&d->d_vmap, WH * sizeof(float)));
        HIP_TRY(c, hipMemcpyAsync(d->d_vmap, vmap, WH * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    Pipe P;
    P.vmap = vmap;
    P.d = d;
    P.fit = fit;
    P.acc = acc;
    P.W = g.W;
    P.H = g.H;
    P.R = g.R;
    P.half = g.R / 2;
    P.foot = std::max(g.R / 2, (int)std::ceil(acc->max_shift_px)) + 1;
    long total = 0;
    std::vector<int> cnt;
    for (int f0 = 0; f0 < F; f0 += fbmax) {
        const int fb = std::min(fbmax, F - f0);
        HIP_TRY(c, hipMemcpyAsync(d->d_frames, det + f0 * WH, fb * WH * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream));
        if (vmap)
            HIP_TRY(c, hipMemcpyAsync(d->d_locf, img + f0 * WH, fb * WH * sizeof(float), hipMemcpyHostToDevice,
                                      c->stream));
        rc = run_detect(d, fb, cnt);
        if (rc) return rc;
        std::vector<FrameRun> runs(fb);
        for (int f = 0; f < fb; f++) {
            FrameRun& fr = runs[f];
            fr.f = f0 + f;
            fr.img = img + (size_t)(f0 + f) * WH;
            fr.n = cnt[f];
            fr.cand.resize((size_t)fr.n * 7);
            if (fr.n > 0)
                HIP_TRY(c, hipMemcpyAsync(fr.cand.data(), d->d_sorted + (size_t)f * d->cap * 7,
                                          (size_t)fr.n * 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            fr.mask.assign(WH, 0);
            fr.res_idx.assign(fr.n, -1);
            fr.queued.assign(fr.n, 0);
            fr.ab.assign((size_t)fr.n * 3, 0.0);
            fr.ab_ok.assign(fr.n, 0);
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        P.results.clear();
        P.bits.clear();
        for (;;) {
            // Frames are independent: their walks replay on host threads; the
            // queued fits are then gathered in frame order.
            std::atomic<int> next{0};
            auto work = [&]() {
                for (int f; (f = next.fetch_add(1)) < fb;) replay(P, runs[f], f);
            };
            const int nt = std::max(1, std::min(fb, host_threads()));
            std::vector<std::thread> pool;
            for (int t = 1; t < nt; t++) pool.emplace_back(work);
            work();
            for (auto& t : pool) t.join();
            P.q_frame.clear();
            P.q_k.clear();
            for (int f = 0; f < fb; f++) {
                for (int k : runs[f].q) {
                    P.q_frame.push_back(f);
                    P.q_k.push_back(k);
                }
                runs[f].q.clear();
            }
            if (P.q_k.empty()) break;
            rc = fit_queue(d, P, runs);
            if (rc) return rc;
            if (rounds) (*rounds)++;
        }
        for (int f = 0; f < fb; f++) {
            const FrameRun& fr = runs[f];
            const int m = (int)(fr.locs.size() / ZOLA_LOC_FIELDS);
            if (per_frame) per_frame[f0 + f] = m;
            for (int q = 0; q < m; q++, total++)
                if (total < cap)
                    std::memcpy(locs + (size_t)total * ZOLA_LOC_FIELDS, fr.locs.data() + (size_t)q * ZOLA_LOC_FIELDS,
                                ZOLA_LOC_FIELDS * sizeof(double));
        }
    }
    *n_locs = (int)total;
    return ZOLA_OK;
}

}  // namespace

extern "C" int zola_localize_frames(zola_detector* d, const zola_fit_config* fit, const zola_accept_config* acc,
                                    int F, const float* frames, int cap, double* locs, int* n_locs,
                                    int* per_frame, int* rounds) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d) return ZOLA_EINVAL;
    if (!fit || !acc || F < 0 || cap < 0 || !n_locs || (F > 0 && !frames) || (cap > 0 && !locs))
        return set_err(d->ctx, ZOLA_EINVAL, "bad frame-pipeline arguments");
    return localize_impl(d, fit, acc, F, frames, frames, nullptr, cap, locs, n_locs, per_frame, rounds);
}

extern "C" int zola_localize_frames_scmos(zola_detector* d, const zola_fit_config* fit,
                                          const zola_accept_config* acc, int F, const float* adu,
                                          const float* offset, const float* gain, const float* variance, int cap,
                                          double* locs, int* n_locs, int* per_frame, int* rounds) {
    CtxLock lk_(d ? d->ctx : nullptr);
    if (!d) return ZOLA_EINVAL;
    zola_ctx* c = d->ctx;
    if (!fit || !acc || F < 0 || cap < 0 || !n_locs || (F > 0 && (!adu || !offset || !gain || !variance)) ||
        (cap > 0 && !locs))
        return set_err(c, ZOLA_EINVAL, "bad sCMOS frame-pipeline arguments");
    const size_t WH = (size_t)d->g.W * d->g.H;
    // vargain = variance / gain^2 (SCMOScamera.java:95); detection frames use the
    // constant vargain of pixel 0 and are clamped at 0, the walk's frames use the
    // pixel's own and are not (LocalizationPipeline.java:597-603, 628-634)
    std::vector<float> vmap(WH);
    for (size_t q = 0; q < WH; q++) vmap[q] = (float)((double)variance[q] / ((double)gain[q] * (double)gain[q]));
    const double v0 = (double)variance[0] / ((double)gain[0] * (double)gain[0]);
    std::vector<float> det((size_t)F * WH), img((size_t)F * WH);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int f; (f = next.fetch_add(1)) < F;)
            for (size_t q = 0; q < WH; q++) {
                const size_t k = (size_t)f * WH + q;
                const double e = ((double)adu[k] - (double)offset[q]) / (double)gain[q];
                const double dv = e + v0;
                det[k] = (float)(dv < 0 ? 0.0 : dv);
                img[k] = (float)(e + (double)vmap[q]);
            }
    };
    const int nt = std::max(1, std::min(F, host_threads()));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return localize_impl(d, fit, acc, F, det.data(), img.data(), vmap.data(), cap, locs, n_locs, per_frame, rounds);
}
