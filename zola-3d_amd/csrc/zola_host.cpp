// zola_host.cpp -- context, pupil tables and the C ABI (include/zola_hip.h).
//
// The optics tables are built on the host in FP64 exactly as the reference
// builds them (process/PhaseParameters.java:265-366 for the grid and disk,
// utils/Zernike.java:26-79,545-605 for the basis, process/cpu/
// PSFPhaseMany.java:145-210 for kx, ky, kz, kz_oil and the amplitude,
// PSFPhaseMany.java:387-404 for the Gaussian), then reduced to the compact
// device form used by the kernels: a 32x32 window of the fftshifted pupil
// and the separable spatial blur kernel h1.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_kernels.h"
#include "zola_adjoint64.hpp"
#include "zola_ctx.hpp"
#include "zola_grid256.h"

namespace {

constexpr double PI = 3.14159265358979323846;

int ref_pow2(int x) { return (int)std::pow(2.0, (int)std::ceil(std::log((double)x) / std::log(2.0))); }

double ref_fact(int k) {
    double r = 1.0;
    long long f = 1;
    for (int i = 1; i <= k && i < 20; i++) f *= i;
    r = (double)f;
    for (int i = 20; i <= k; i++) r *= i;
    return r;
}

double zernike_value(int n, int m, double theta, double rho) {
    int am = m < 0 ? -m : m;
    double som = 0.0;
    for (int k = 0; k <= (n - am) / 2; k++) {
        double truc = std::pow(-1.0, k) * ref_fact(n - k);
        double trac = ref_fact(k) * ref_fact((n + am) / 2 - k) * ref_fact((n - am) / 2 - k);
        som += (truc / trac) * std::pow(rho, n - 2 * k);
    }
    return m >= 0 ? som * std::cos(theta * m) : som * std::sin(theta * am);
}

}  // namespace

static int compute_h1(zola_ctx* c) {
    // the 256 grid blurs by FFT with the 2-D Gaussian spectrum (zola256_upload_gauss)
    if (c->N == kGrid256) {
        c->model_ver++;
        return ZOLA_OK;
    }
    // G1(k) = exp(-(k' - s2/2)^2 / (2 sigF^2)), k' = (k + s2/2) % s2, sigF = s2/(2 pi sigma);
    // h1(d) = (1/s2) sum_k G1(k) cos(2 pi k d / s2)  (G1 is even, so h1 is real).
    const int s2 = c->s2;
    const double sigF = (1.0 / c->o.sigma) * s2 / (PI * 2);
    const double sigpow = sigF * sigF;
    std::vector<double> G(s2);
    for (int k = 0; k < s2; k++) {
        double kk = (double)((k + s2 / 2) % s2) - (double)(s2 / 2);
        G[k] = std::exp(-.5 * kk * kk / sigpow);
    }
    // cos and sin of 2 pi m / s2, m = k d mod s2 (the calibration sets sigma on every evaluation)
    std::vector<double> ct(s2), st(s2);
    for (int m = 0; m < s2; m++) {
        ct[m] = std::cos(2 * PI * (double)m / s2);
        st[m] = std::sin(2 * PI * (double)m / s2);
    }
    // dG1/dsigma = -G1 kk^2 / (sigF^2 sigma)  (sigF = s2/(2 pi sigma)), formed
    // once per k (the same operations in the same order as inside the sums);
    // s2 is a power of two, so (k d) mod s2 is a mask
    std::vector<double> Gs(s2);
    for (int k = 0; k < s2; k++) {
        const double kk = (double)((k + s2 / 2) % s2) - (double)(s2 / 2);
        Gs[k] = G[k] * kk * kk / (sigpow * c->o.sigma);
    }
    std::vector<double> h1(s2), hd(s2), hs(s2);
    for (int d = 0; d < s2; d++) {
        double acc = 0, accs = 0;
        for (int k = 0; k < s2; k++) {
            const double cs = ct[(k * d) & (s2 - 1)];
            acc += G[k] * cs;
            accs -= Gs[k] * cs;
        }
        h1[d] = acc / s2;
        hs[d] = accs / s2;
    }
    // Spectral derivative of the blurred image along one axis, per um of
    // emitter shift.  Shifting the emitter by x moves the (band-limited,
    // |k| <= 31) intensity by x/dx samples, so d(blur)/dx = (1/dx) hd (*) I with
    // hd = IDFT(i 2 pi k_s / s2 . G1); the Nyquist bin carries no intensity.
    std::vector<double> Gd(s2);
    for (int k = 0; k < s2; k++) {
        int ks = k < s2 / 2 ? k : k - s2;
        if (k == s2 / 2) ks = 0;
        Gd[k] = G[k] * (2 * PI * ks / s2);
    }
    for (int d = 0; d < s2; d++) {
        double acc = 0;
        for (int k = 0; k < s2; k++) acc -= Gd[k] * st[(k * d) & (s2 - 1)];
        hd[d] = acc / s2 / c->o.xystep;
    }
    c->model_ver++;
    if (c->N == zola::kN) {
        for (int d = 0; d < s2; d++) {
            c->model.h1[d] = (float)h1[d];
            c->model.hd[d] = (float)hd[d];
            c->model.hs[d] = (float)hs[d];
        }
        return ZOLA_OK;
    }
    // 128 grid: the s2 taps of both kernels at offsets 0 and 128 (s2 = 32, 64 or
    // 128; the unused tail stays 0), then for s2 < 128 the crop-derivative kernel
    // kwin[c][g] = sum_{o in [-s2/2, s2/2)} h1(c - R/2 - o) dN(o - g) (c < R, rows
    // padded to 64 with zeros), dN the 128-periodic spectral derivative with its
    // Nyquist bin dropped, per um; all uploaded by upload_taps128
    const int N = zola::kN2;
    c->taps128.assign(2 * N + 64 * N, 0.f);
    for (int d = 0; d < s2; d++) {
        c->taps128[d] = (float)h1[d];
        c->taps128[N + d] = (float)hd[d];
    }
    if (s2 < N) {
        std::vector<double> dN(N);
        for (int d = 0; d < N; d++) {
            double acc = 0;
            for (int k = 0; k < N; k++) {
                int ks = k < N / 2 ? k : k - N;
                if (k == N / 2) ks = 0;
                acc -= (2 * PI * ks / N) * std::sin(2 * PI * (double)((k * d) % N) / N);
            }
            dN[d] = acc / N / c->o.xystep;
        }
        const int R = c->R;
        for (int cc = 0; cc < R; cc++)
            for (int g = 0; g < N; g++) {
                const int goff = g < N / 2 ? g : g - N;
                double acc = 0;
                for (int o = -s2 / 2; o < s2 / 2; o++)
                    acc += h1[(((cc - R / 2 - o) % s2) + s2) % s2] * dN[(((o - goff) % N) + N) % N];
                c->taps128[2 * N + cc * N + g] = (float)acc;
            }
    }
    return ZOLA_OK;
}

static int upload_taps128(zola_ctx* c) {
    HIP_TRY(c, hipMemcpyAsync(c->d_taps128, c->taps128.data(), c->taps128.size() * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

static int upload_tables(zola_ctx* c) {
    c->model_ver++;
    ensure_phase(c);
    if (c->N == kGrid256) return zola256_upload_tables(c);
    if (c->N == zola::kN2) {
        std::vector<float4> tab(zola::kW2 * zola::kW2, make_float4(0.f, 0.f, 0.f, 0.f));
        for (int p = 0; p < c->D; p++)
            tab[c->wu[p] * zola::kW2 + c->wv[p]] =
                make_float4((float)c->phase[p], (float)c->amp[p], (float)(c->kz[p] - c->kz0),
                            (float)(c->kzo[p] - c->kzo0));
        HIP_TRY(c, hipMemcpyAsync(c->d_tab128, tab.data(), tab.size() * sizeof(float4), hipMemcpyHostToDevice,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return ZOLA_OK;
    }
    std::vector<float> tab(8 * 1024, 0.f);
    for (int p = 0; p < c->D; p++) {
        int w = c->wu[p] * 32 + c->wv[p];
        tab[w] = (float)c->phase[p];
        tab[1024 + w] = (float)c->amp[p];
        tab[2048 + w] = (float)(c->kz[p] - c->kz0);
        tab[3072 + w] = (float)(c->kzo[p] - c->kzo0);
        for (int k = 0; k < 4; k++) tab[4096 + 4 * w + k] = tab[1024 * k + w];
    }
    HIP_TRY(c, hipMemcpyAsync(c->d_tab, tab.data(), tab.size() * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

namespace {
int roi_slot(int R) { return R == 17 ? 0 : R == 21 ? 1 : R == 31 ? 2 : R == 41 ? 3 : -1; }
}  // namespace

int zola_fit_layout(const zola_ctx* c, int R, int mode) {
    const int i = roi_slot(R);
    const int l = i >= 0 ? c->layout[i][mode == ZOLA_FIT_LM ? 1 : 0] : ZOLA_LAYOUT_AUTO;
    if (l == ZOLA_LAYOUT_WAVE) return 0;
    if (l == ZOLA_LAYOUT_QUAD) return 1;
    if (l == ZOLA_LAYOUT_PAIR) return 2;
    // the fastest layouts on 32,768 (LM) and 8,192 (reference form) ROIs:
    // two waves for the LM form up to 21, one wave otherwise (four waves
    // took 1.4x (LM) and 2.4x (reference form) the time at 31 and 1.04x and
    // 1.7x at 41)
    return (mode == ZOLA_FIT_LM && R <= 21) ? 2 : 0;
}

// Kernel layout of a 128-grid fit (zola_launch_fit128): 1 four waves per ROI,
// 3 eight waves per ROI (LM form).
int zola_fit_layout128(const zola_ctx* c, int mode) {
    if (mode != ZOLA_FIT_LM) return 1;
    if (c->layout128 == ZOLA_LAYOUT_QUAD) return 1;
    // default: eight waves (65,536 41^2 ROIs: 66.3 ms against 77.1 ms for four)
    return 3;
}

// Whether a layout (ZOLA_LAYOUT_WAVE/QUAD/PAIR) exists for (R, mode).
static bool layout_available(int R, int mode, int layout) {
    if (layout == ZOLA_LAYOUT_PAIR) return mode == ZOLA_FIT_LM && R <= 21;
    if (layout == ZOLA_LAYOUT_OCT) return false;
    return layout == ZOLA_LAYOUT_WAVE || layout == ZOLA_LAYOUT_QUAD;
}

int zola_psf_launch(zola_ctx* c, int B, const float* params, float* out, hipStream_t st) {
    if (c->N == kGrid256) {
        Eval256 e;
        e.B = B;
        e.params = params;
        e.out = out;
        return zola256_run(c, e, st);
    }
    if (c->N == zola::kN2) HIP_TRY(c, zola_launch_psf_forward128(c->model128, c->R, B, params, out, st));
    else HIP_TRY(c, zola_launch_psf_forward64(c->model, c->R, B, params, out, st));
    return ZOLA_OK;
}

int zola_fit_launch(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels, const double* init,
                    const float* zoil, const float* varg, double* theta, double* nll, double* crlb, int* iters,
                    int* status, hipStream_t st) {
    zola::FitArgs A{};
    A.B = B;
    A.max_iter = cfg->max_iter;
    A.xystep = (float)c->o.xystep;
    A.zmin = cfg->zmin;
    A.zmax = cfg->zmax;
    A.zfocus = cfg->zfocus;
    A.thr = cfg->photon_threshold;
    A.tol = cfg->tol;
    A.noise_rel = 1e-8;
    A.pixels = pixels;
    A.init = init;
    A.zoil = zoil;
    A.varg = varg;
    A.queue = c->d_queue;
    A.theta = theta;
    A.nll = nll;
    A.crlb = crlb;
    A.iters = iters;
    A.status = status;
    if (c->N == zola::kN2)
        HIP_TRY(c, zola_launch_fit128(c->model128, c->R, cfg->mode, zola_fit_layout128(c, cfg->mode), A, st));
    else HIP_TRY(c, zola_launch_fit64(c->model, c->R, cfg->mode, zola_fit_layout(c, c->R, cfg->mode), A, st));
    return ZOLA_OK;
}

extern "C" {

int zola_ctx_create(const zola_optics* o, int R, const double* coeffs, int J, zola_ctx** out) {
    if (!o || !out || R <= 0 || J < 0 || (J > 0 && !coeffs)) return ZOLA_EINVAL;
    *out = nullptr;
    zola_ctx* c = new zola_ctx();
    c->o = *o;
    int size = o->size_fft;
    if (size % 2 != 0) size--;
    if (size <= 0 || o->xystep <= 0 || o->wavelength <= 0 || o->sigma <= 0) {
        *out = c;
        return set_err(c, ZOLA_EINVAL, "invalid optics");
    }
    c->N = ref_pow2(size);
    c->R = R;
    c->J = J;
    int s2 = 2 * R;
    if (s2 > c->N) s2 = c->N;
    c->s2 = ref_pow2(s2);
    const int N = c->N;
    c->center = (N / 2) - .5;
    c->ringsize = (o->na / o->wavelength) * (o->na / o->wavelength);
    c->radius_px = std::sqrt(c->ringsize) * (N * o->xystep);
    c->kappa = 2 * PI / (o->xystep * N);
    *out = c;
    if (N != zola::kN && N != zola::kN2 && N != kGrid256)
        return set_err(c, ZOLA_EUNSUPPORTED, "pupil grid N=%d: this build covers N=64, 128 and 256", N);
    // the 64-grid kernels blur on the whole grid; the 128-grid ones on s2 = 32, 64 or 128;
    // the 256 grid (calibration) on any s2 of 16..256
    if (N == zola::kN && c->s2 != N)
        return set_err(c, ZOLA_EUNSUPPORTED, "blur grid s2=%d must equal N=%d (R >= N/4)", c->s2, N);
    if (N == kGrid256 ? (R < 8 || R > kGrid256)
                      : (N == zola::kN ? !zola_kernels_support_roi(R) : !zola_kernels_support_roi128(R)))
        return set_err(c, ZOLA_EUNSUPPORTED, "ROI side %d not instantiated for N=%d", R, N);
    const int halfw = N == zola::kN ? zola::kW / 2 : N == zola::kN2 ? zola::kW2 / 2 : N / 4;

    const double left = (o->nwat / o->wavelength) * (o->nwat / o->wavelength);
    const double left_oil = (o->noil / o->wavelength) * (o->noil / o->wavelength);
    const double cc = c->center;
    int D = 0;
    for (int i = 0; i < N; i++)
        for (int ii = 0; ii < N; ii++) {
            double a = (i - cc) / (N * o->xystep), b = (ii - cc) / (N * o->xystep);
            if (a * a + b * b <= c->ringsize) D++;
        }
    c->D = D;
    c->di.resize(D); c->dj.resize(D); c->wu.resize(D); c->wv.resize(D);
    c->kz.resize(D); c->kzo.resize(D); c->amp.resize(D);
    int id = 0;
    double kzmin = 1e300, kzmax = -1e300, kzomin = 1e300, kzomax = -1e300;
    for (int i = 0; i < N; i++)
        for (int ii = 0; ii < N; ii++) {
            double a = (i - cc) / (N * o->xystep), b = (ii - cc) / (N * o->xystep);
            if (!(a * a + b * b <= c->ringsize)) continue;
            c->di[id] = i;
            c->dj[id] = ii;
            double kxx = ((i - cc) / (o->xystep * N)) * ((i - cc) / (o->xystep * N));
            double kyy = ((ii - cc) / (o->xystep * N)) * ((ii - cc) / (o->xystep * N));
            double right = kxx + kyy;
            c->kz[id] = left > right ? 2 * PI * std::sqrt(left - right) : 0.0;
            c->kzo[id] = left_oil > right ? 2 * PI * std::sqrt(left_oil - right) : 0.0;
            if (o->with_apo) c->amp[id] = left_oil > right ? 1.0 / std::pow(1 - (right / left_oil), .25) : 0.0;
            else c->amp[id] = 1.0 / std::sqrt((double)D);
            int sx = (i + N / 2) % N, sy = (ii + N / 2) % N;
            int nu = sx < N / 2 ? sx : sx - N, nv = sy < N / 2 ? sy : sy - N;
            c->wu[id] = nu + halfw;
            c->wv[id] = nv + halfw;
            if (c->wu[id] < 0 || c->wu[id] >= 2 * halfw || c->wv[id] < 0 || c->wv[id] >= 2 * halfw)
                return set_err(c, ZOLA_EUNSUPPORTED,
                               "pupil radius %.3f px exceeds N/4 = %d px", c->radius_px, N / 4);
            kzmin = std::fmin(kzmin, c->kz[id]); kzmax = std::fmax(kzmax, c->kz[id]);
            kzomin = std::fmin(kzomin, c->kzo[id]); kzomax = std::fmax(kzomax, c->kzo[id]);
            id++;
        }
    c->kz0 = 0.5 * (kzmin + kzmax);
    c->kzo0 = 0.5 * (kzomin + kzomax);

    // Zernike basis on the disk (order of utils/Zernike.java:26-79)
    c->Z.assign((size_t)(J > 0 ? J : 1) * D, 0.0);
    c->coef.assign(J > 0 ? J : 1, 0.0);
    {
        std::vector<int> zn(J), zm(J);
        int idx = 0;
        for (int rang = 0; idx < J; rang++)
            for (int u = -rang; u <= rang && idx < J; u += 2) { zn[idx] = rang; zm[idx] = u; idx++; }
        double cz = ((double)N) / 2. - .5;
        for (int j = 0; j < J; j++)
            for (int p = 0; p < D; p++) {
                double x = ((double)(c->di[p] - cz) / c->radius_px);
                double y = ((double)(c->dj[p] - cz) / c->radius_px);
                double theta = std::atan2(y, x), rho = std::sqrt(x * x + y * y);
                c->Z[(size_t)j * D + p] = rho < 1 ? zernike_value(zn[j], zm[j], theta, rho) : 0.0;
            }
        for (int j = 0; j < J; j++) c->coef[j] = coeffs[j];
    }
    compute_phase(c);
    int rc0 = compute_h1(c);
    if (rc0) return rc0;

    HIP_TRY(c, hipStreamCreate(&c->stream));
    c->own_stream = true;
    {
        int rs = zola_scratch_setup(c);
        if (rs) return rs;
    }
    HIP_TRY(c, hipMalloc((void**)&c->d_queue, 8 * sizeof(int)));
    if (N == kGrid256) {
        int rc = zola256_setup(c);
        if (rc) return rc;
        c->err.clear();
        return ZOLA_OK;
    }
    if (N == zola::kN2) {
        HIP_TRY(c, hipMalloc((void**)&c->d_tab128, zola::kW2 * zola::kW2 * sizeof(float4)));
        HIP_TRY(c, hipMalloc((void**)&c->d_taps128, (2 + 64) * zola::kN2 * sizeof(float)));
        c->model128.tab = c->d_tab128;
        c->model128.h1 = c->d_taps128;
        c->model128.hd = c->d_taps128 + zola::kN2;
        c->model128.kwin = c->d_taps128 + 2 * zola::kN2;
        c->model128.kappa = (float)c->kappa;
        int rc = upload_tables(c);
        if (rc) return rc;
        rc = upload_taps128(c);
        if (rc) return rc;
        c->err.clear();
        return ZOLA_OK;
    }
    // four tables of 1024 floats, then the same interleaved as 1024 float4
    HIP_TRY(c, hipMalloc((void**)&c->d_tab, 8 * 1024 * sizeof(float)));
    c->model.phi = c->d_tab;
    c->model.amp = c->d_tab + 1024;
    c->model.kz = c->d_tab + 2048;
    c->model.kzo = c->d_tab + 3072;
    c->model.tab4 = reinterpret_cast<const float4*>(c->d_tab + 4096);
    c->model.kappa = (float)c->kappa;
    int rc = upload_tables(c);
    if (rc) return rc;
    if (J > 0) {
        std::vector<double> zw((size_t)J * 1024, 0.0);
        for (int jz = 0; jz < J; jz++)
            for (int p = 0; p < D; p++)
                zw[(size_t)jz * 1024 + c->wu[p] * 32 + c->wv[p]] = c->Z[(size_t)jz * D + p];
        HIP_TRY(c, hipMalloc((void**)&c->d_Zw, zw.size() * sizeof(double)));
        HIP_TRY(c, hipMemcpy(c->d_Zw, zw.data(), zw.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c->err.clear();
    return ZOLA_OK;
}

int zola_ctx_destroy(zola_ctx* c) {
    if (!c) return ZOLA_OK;
    if (c->d_tab) (void)hipFree(c->d_tab);
    if (c->d_tab128) (void)hipFree(c->d_tab128);
    if (c->d_taps128) (void)hipFree(c->d_taps128);
    if (c->d_stage) (void)hipFree(c->d_stage);
    if (c->d_tmpl) (void)hipFree(c->d_tmpl);
    if (c->d_Zw) (void)hipFree(c->d_Zw);
    if (c->d_qbar) (void)hipFree(c->d_qbar);
    if (c->d_queue) (void)hipFree(c->d_queue);
    zola256_free(c);
    if (c->mixed_exec) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipGraphExecDestroy(c->mixed_exec);
    }
    for (int g = 0; g < 8; g++) {
        if (c->side[g]) (void)hipStreamDestroy(c->side[g]);
        if (c->join[g]) (void)hipEventDestroy(c->join[g]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->cap) (void)hipStreamDestroy(c->cap);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ZOLA_OK;
}

// The message is copied for the calling thread under the context's lock, so the
// pointer stays valid until this thread's next zola_last_error call.
const char* zola_last_error(const zola_ctx* c) {
    if (!c) return "null context";
    static thread_local std::string msg;
    CtxLock lk_(c);
    msg = c->err;
    return msg.c_str();
}

int zola_ctx_info(const zola_ctx* c, int* info) {
    CtxLock lk_(c);
    if (!c || !info) return ZOLA_EINVAL;
    info[0] = c->N; info[1] = c->D; info[2] = c->s2; info[3] = c->R; info[4] = c->J;
    return ZOLA_OK;
}

int zola_set_coefficients(zola_ctx* c, const double* coeffs, int J) {
    CtxLock lk_(c);
    if (!c || !coeffs || J != c->J) return set_err(c, ZOLA_EINVAL, "coefficient count must be %d", c ? c->J : 0);
    for (int j = 0; j < J; j++) c->coef[j] = coeffs[j];
    // calibration loops change the coefficients every evaluation: the 64 and
    // 256 grids form the phase tables on the device from the Zernike basis
    // already there (no host pass over the disk, no copy, no synchronisation)
    if (J > 0 && J <= kMaxCoefDev && c->d_Zw && (c->N == zola::kN || c->N == kGrid256)) {
        c->model_ver++;
        c->phase_stale = true;
        if (c->N == kGrid256) return zola256_set_phase(c);
        HIP_TRY(c, zola_launch_phase_tab64(J, c->coef.data(), c->d_Zw, c->d_tab, c->stream));
        return ZOLA_OK;
    }
    compute_phase(c);
    return upload_tables(c);
}

int zola_set_sigma(zola_ctx* c, double sigma) {
    CtxLock lk_(c);
    if (!c || !(sigma > 0)) return set_err(c, ZOLA_EINVAL, "sigma must be > 0");
    c->o.sigma = sigma;
    int rc = compute_h1(c);
    if (rc == ZOLA_OK && c->N == zola::kN2) rc = upload_taps128(c);
    if (rc == ZOLA_OK && c->N == kGrid256) rc = zola256_set_gauss(c);
    return rc;
}

int zola_set_stream(zola_ctx* c, void* s) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    // the context's buffers (templates, staging) may still be in use by work on
    // the previous stream: finish it before work on the new stream can reuse them
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->own_stream && c->stream) HIP_TRY(c, hipStreamDestroy(c->stream));
    c->own_stream = false;
    c->stream = (hipStream_t)s;
    return ZOLA_OK;
}

int zola_psf_forward_dev(zola_ctx* c, int B, const float* params, float* out) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (B < 0 || (B > 0 && (!params || !out))) return set_err(c, ZOLA_EINVAL, "bad batch");
    if (B == 0) return ZOLA_OK;
    return zola_psf_launch(c, B, params, out, c->stream);
}

int zola_psf_forward(zola_ctx* c, int B, const float* params, float* out) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (B < 0 || (B > 0 && (!params || !out))) return set_err(c, ZOLA_EINVAL, "bad batch");
    if (B == 0) return ZOLA_OK;
    size_t pb = (size_t)B * 4 * sizeof(float), ob = (size_t)B * c->R * c->R * sizeof(float);
    int rc = stage(c, pb + ob);
    if (rc) return rc;
    float* dp = (float*)c->d_stage;
    float* dout = (float*)((char*)c->d_stage + pb);
    HIP_TRY(c, hipMemcpyAsync(dp, params, pb, hipMemcpyHostToDevice, c->stream));
    rc = zola_psf_forward_dev(c, B, dp, dout);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

int zola_psf_backward_dev(zola_ctx* c, int B, const float* params, const float* w,
                          float* dparams, double* da) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (c->N != zola::kN && c->N != kGrid256)
        return set_err(c, ZOLA_EUNSUPPORTED, "the adjoint covers the 64 and 256 pupil grids");
    if (B < 0 || (B > 0 && (!params || !w || !dparams))) return set_err(c, ZOLA_EINVAL, "bad batch");
    if (B == 0) return ZOLA_OK;
    if (c->N == kGrid256) {
        Eval256 e;
        e.B = B;
        e.params = params;
        e.w = w;
        e.dparams = dparams;
        e.da = da;
        return zola256_run(c, e, c->stream);
    }
    float* qbar = nullptr;
    if (da && c->J > 0) {
        const size_t need = (size_t)B * 1024 * sizeof(float) + kQBytes;
        if (need > c->qbar_bytes) {
            if (c->d_qbar) HIP_TRY(c, hipFree(c->d_qbar));
            c->d_qbar = nullptr;
            c->qbar_bytes = 0;
            HIP_TRY(c, hipMalloc((void**)&c->d_qbar, need));
            c->qbar_bytes = need;
        }
        qbar = c->d_qbar;
    }
    zola::AdjArgs A{};
    A.B = B;
    A.params = params;
    A.w = w;
    A.dparams = dparams;
    A.qbar = qbar;
    HIP_TRY(c, zola_launch_adjoint64(c->model, c->R, A, c->stream));
    if (qbar) HIP_TRY(c, zola_launch_project_zernike(B, c->J, c->d_Zw, qbar, qsum_slot(qbar, B), da, c->stream));
    return ZOLA_OK;
}

int zola_psf_backward(zola_ctx* c, int B, const float* params, const float* w, float* dparams,
                      double* da) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (c->N != zola::kN && c->N != kGrid256)
        return set_err(c, ZOLA_EUNSUPPORTED, "the adjoint covers the 64 and 256 pupil grids");
    if (B < 0 || (B > 0 && (!params || !w || !dparams))) return set_err(c, ZOLA_EINVAL, "bad batch");
    if (B == 0) return ZOLA_OK;
    const size_t R2 = (size_t)c->R * c->R;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bp = (size_t)B * 4 * sizeof(float), bw = B * R2 * sizeof(float);
    const size_t bd = (size_t)B * 4 * sizeof(float), ba = (size_t)(c->J > 0 ? c->J : 1) * sizeof(double);
    int rc = stage(c, al(bp) + al(bw) + al(bd) + al(ba));
    if (rc) return rc;
    char* base = (char*)c->d_stage;
    float* dpar = (float*)base; base += al(bp);
    float* dw = (float*)base; base += al(bw);
    float* ddp = (float*)base; base += al(bd);
    double* dda = (double*)base;
    HIP_TRY(c, hipMemcpyAsync(dpar, params, bp, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dw, w, bw, hipMemcpyHostToDevice, c->stream));
    if (da) HIP_TRY(c, hipMemsetAsync(dda, 0, ba, c->stream));
    rc = zola_psf_backward_dev(c, B, dpar, dw, ddp, da ? dda : nullptr);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(dparams, ddp, bd, hipMemcpyDeviceToHost, c->stream));
    if (da && c->J > 0)
        HIP_TRY(c, hipMemcpyAsync(da, dda, (size_t)c->J * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

int zola_get_stream(const zola_ctx* c, void** stream) {
    CtxLock lk_(c);
    if (!c || !stream) return ZOLA_EINVAL;
    *stream = (void*)c->stream;
    return ZOLA_OK;
}

int zola_calib_eval_dev(zola_ctx* c, int B, const float* params, const float* pixels,
                        const double* amp, const double* bg, double* nll, float* dparams,
                        double* damp, double* dbg, double* da) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (c->N != zola::kN && c->N != kGrid256)
        return set_err(c, ZOLA_EUNSUPPORTED, "the adjoint covers the 64 and 256 pupil grids");
    if (B < 0) return set_err(c, ZOLA_EINVAL, "bad batch");
    if (B == 0) return ZOLA_OK;
    if (!params || !pixels || !amp || !bg || !nll || !dparams || !damp || !dbg)
        return set_err(c, ZOLA_EINVAL, "null buffer");
    if (c->N == kGrid256) {
        Eval256 e;
        e.B = B;
        e.params = params;
        e.pixels = pixels;
        e.amp = amp;
        e.bg = bg;
        e.nll = nll;
        e.dparams = dparams;
        e.damp = damp;
        e.dbg = dbg;
        e.da = da;
        e.da_assign = c->da_assign;
        return zola256_run(c, e, c->stream);
    }
    float* qbar = nullptr;
    if (da && c->J > 0) {
        const size_t need = (size_t)B * 1024 * sizeof(float) + kQBytes;
        if (need > c->qbar_bytes) {
            if (c->d_qbar) HIP_TRY(c, hipFree(c->d_qbar));
            c->d_qbar = nullptr;
            c->qbar_bytes = 0;
            HIP_TRY(c, hipMalloc((void**)&c->d_qbar, need));
            c->qbar_bytes = need;
        }
        qbar = c->d_qbar;
    }
    zola::AdjArgs A{};
    A.B = B;
    A.params = params;
    A.dparams = dparams;
    A.qbar = qbar;
    A.pixels = pixels;
    A.amp = amp;
    A.bg = bg;
    A.nll = nll;
    A.damp = damp;
    A.dbg = dbg;
    A.S = 1;
    HIP_TRY(c, zola_launch_adjoint64(c->model, c->R, A, c->stream));
    if (qbar) HIP_TRY(c, zola_launch_project_zernike(B, c->J, c->d_Zw, qbar, qsum_slot(qbar, B), da, c->stream, c->da_assign));
    return ZOLA_OK;
}

int zola_calib_eval_full_dev(zola_ctx* c, int nb, int S, const float* params, const float* pixels,
                             const double* amp, const double* poly, const double* offset, double* nll,
                             float* dparams, double* damp, double* dpoly, double* dsigma, double* da) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (c->N != zola::kN && c->N != kGrid256)
        return set_err(c, ZOLA_EUNSUPPORTED, "the adjoint covers the 64 and 256 pupil grids");
    if (nb < 0 || S <= 0 || (long)nb * S > (1L << 30)) return set_err(c, ZOLA_EINVAL, "bad batch");
    const int B = nb * S;
    if (B == 0) return ZOLA_OK;
    if (!params || !pixels || !amp || !poly || !nll || !dparams || !damp)
        return set_err(c, ZOLA_EINVAL, "null buffer");
    if (c->N == kGrid256) {
        Eval256 e;
        e.B = B;
        e.params = params;
        e.pixels = pixels;
        e.amp = amp;
        e.S = S;
        e.poly = poly;
        e.offset = offset;
        e.nll = nll;
        e.dparams = dparams;
        e.damp = damp;
        e.dpoly = dpoly;
        e.dsig = dsigma;
        e.da = da;
        e.da_assign = c->da_assign;
        return zola256_run(c, e, c->stream);
    }
    float* qbar = nullptr;
    if (da && c->J > 0) {
        const size_t need = (size_t)B * 1024 * sizeof(float) + kQBytes;
        if (need > c->qbar_bytes) {
            if (c->d_qbar) HIP_TRY(c, hipFree(c->d_qbar));
            c->d_qbar = nullptr;
            c->qbar_bytes = 0;
            HIP_TRY(c, hipMalloc((void**)&c->d_qbar, need));
            c->qbar_bytes = need;
        }
        qbar = c->d_qbar;
    }
    zola::AdjArgs A{};
    A.B = B;
    A.params = params;
    A.dparams = dparams;
    A.qbar = qbar;
    A.pixels = pixels;
    A.amp = amp;
    A.nll = nll;
    A.damp = damp;
    A.S = S;
    A.poly = poly;
    A.offset = offset;
    A.dpoly = dpoly;
    A.dsig = dsigma;
    HIP_TRY(c, zola_launch_adjoint64(c->model, c->R, A, c->stream));
    if (qbar) HIP_TRY(c, zola_launch_project_zernike(B, c->J, c->d_Zw, qbar, qsum_slot(qbar, B), da, c->stream, c->da_assign));
    return ZOLA_OK;
}

int zola_fit_batch_dev(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels,
                       const double* init, double* theta, double* nll, double* crlb, int* iters,
                       int* status) {
    CtxLock lk_(c);
    return zola_fit_batch_zoil_dev(c, cfg, B, pixels, init, nullptr, theta, nll, crlb, iters, status);
}

int zola_fit_batch_zoil_dev(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels,
                            const double* init, const float* zoil, double* theta, double* nll, double* crlb,
                            int* iters, int* status) {
    CtxLock lk_(c);
    return zola_fit_batch_ext_dev(c, cfg, B, pixels, init, zoil, nullptr, theta, nll, crlb, iters, status);
}

int zola_fit_batch_scmos_dev(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels,
                             const float* varg, const double* init, double* theta, double* nll, double* crlb,
                             int* iters, int* status) {
    CtxLock lk_(c);
    if (c && !varg) return set_err(c, ZOLA_EINVAL, "null sCMOS map");
    return zola_fit_batch_ext_dev(c, cfg, B, pixels, init, nullptr, varg, theta, nll, crlb, iters, status);
}

int zola_fit_batch_ext_dev(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels, const double* init,
                           const float* zoil, const float* varg, double* theta, double* nll, double* crlb,
                           int* iters, int* status) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    ZOLA_REFUSE_GRID256(c);
    if (!cfg || B < 0) return set_err(c, ZOLA_EINVAL, "bad fit arguments");
    if (B == 0) return ZOLA_OK;
    if (!pixels || !init || !theta || !nll || !crlb || !iters || !status)
        return set_err(c, ZOLA_EINVAL, "null buffer");
    if (cfg->mode != ZOLA_FIT_REFERENCE && cfg->mode != ZOLA_FIT_LM)
        return set_err(c, ZOLA_EINVAL, "unknown fit mode %d", cfg->mode);
    if (cfg->max_iter < 0 || !(cfg->zmax >= cfg->zmin))
        return set_err(c, ZOLA_EINVAL, "bad iteration count or z range");
    return zola_fit_launch(c, cfg, B, pixels, init, zoil, varg, theta, nll, crlb, iters, status, c->stream);
}

int zola_set_fit_layout(zola_ctx* c, int R, int mode, int layout) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (mode != ZOLA_FIT_REFERENCE && mode != ZOLA_FIT_LM) return set_err(c, ZOLA_EINVAL, "unknown fit mode %d", mode);
    if (layout < ZOLA_LAYOUT_AUTO || layout > ZOLA_LAYOUT_OCT) return set_err(c, ZOLA_EINVAL, "unknown layout %d", layout);
    if (c->N == zola::kN2) {
        // 128 grid: four waves per ROI for both forms, eight for the LM form
        if (R != c->R) return set_err(c, ZOLA_EUNSUPPORTED, "ROI side %d not available", R);
        const bool ok = layout == ZOLA_LAYOUT_AUTO || layout == ZOLA_LAYOUT_QUAD ||
                        (layout == ZOLA_LAYOUT_OCT && mode == ZOLA_FIT_LM);
        if (!ok) return set_err(c, ZOLA_EUNSUPPORTED, "layout %d not available on the 128 grid, mode %d", layout, mode);
        if (mode == ZOLA_FIT_LM) c->layout128 = layout;
        return ZOLA_OK;
    }
    if (c->N != zola::kN) return set_err(c, ZOLA_EUNSUPPORTED, "fit layouts apply to the 64 and 128 grids");
    const int i = roi_slot(R);
    if (i < 0) return set_err(c, ZOLA_EUNSUPPORTED, "ROI side %d not available", R);
    if (layout == ZOLA_LAYOUT_OCT) return set_err(c, ZOLA_EUNSUPPORTED, "eight waves per ROI serve the 128 grid");
    if (layout != ZOLA_LAYOUT_AUTO && !layout_available(R, mode, layout))
        return set_err(c, ZOLA_EUNSUPPORTED, "layout %d not available for side %d, mode %d", layout, R, mode);
    c->layout[i][mode == ZOLA_FIT_LM ? 1 : 0] = layout;
    return ZOLA_OK;
}

int zola_autotune_fit(zola_ctx* c, int R, int mode, int B, int* layout, double* ms) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (c->N != zola::kN) return set_err(c, ZOLA_EUNSUPPORTED, "fit layouts apply to the 64 grid");
    if (roi_slot(R) < 0) return set_err(c, ZOLA_EUNSUPPORTED, "ROI side %d not available", R);
    if (mode != ZOLA_FIT_REFERENCE && mode != ZOLA_FIT_LM) return set_err(c, ZOLA_EINVAL, "unknown fit mode %d", mode);
    if (B < 64 || B > (1 << 20)) return set_err(c, ZOLA_EINVAL, "autotune batch %d outside [64, 2^20]", B);
    // Synthetic batch: noise-free model images A psf + bg at spread (x, y, z),
    // started off the truth, so that both layouts run typical iteration counts.
    const size_t R2 = (size_t)R * R;
    std::vector<float> par((size_t)B * 4);
    std::vector<double> init((size_t)B * 5);
    uint32_t st = 0x9E3779B9u;
    auto uni = [&]() {
        st = st * 1664525u + 1013904223u;
        return (double)(st >> 8) / (double)(1u << 24);
    };
    for (int b = 0; b < B; b++) {
        const double x = (2 * uni() - 1) * c->o.xystep, y = (2 * uni() - 1) * c->o.xystep, z = 2 * uni() - 1;
        par[b * 4 + 0] = (float)x;
        par[b * 4 + 1] = (float)y;
        par[b * 4 + 2] = (float)z;
        par[b * 4 + 3] = 0.f;
        init[b * 5 + 0] = 1500.0;
        init[b * 5 + 1] = 15.0;
        init[b * 5 + 2] = x + 0.3 * c->o.xystep;
        init[b * 5 + 3] = y - 0.3 * c->o.xystep;
        init[b * 5 + 4] = z + 0.05;
    }
    float *d_par = nullptr, *d_px = nullptr;
    double *d_init = nullptr, *d_th = nullptr, *d_nll = nullptr, *d_cr = nullptr;
    int *d_it = nullptr, *d_st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ZOLA_OK;
    double t[3] = {0, 0, 0};
    auto run = [&]() -> int {
        HIP_TRY(c, hipMalloc((void**)&d_par, par.size() * sizeof(float)));
        HIP_TRY(c, hipMalloc((void**)&d_px, (size_t)B * R2 * sizeof(float)));
        HIP_TRY(c, hipMalloc((void**)&d_init, init.size() * sizeof(double)));
        HIP_TRY(c, hipMalloc((void**)&d_th, (size_t)B * 5 * sizeof(double)));
        HIP_TRY(c, hipMalloc((void**)&d_nll, (size_t)B * sizeof(double)));
        HIP_TRY(c, hipMalloc((void**)&d_cr, (size_t)B * 3 * sizeof(double)));
        HIP_TRY(c, hipMalloc((void**)&d_it, (size_t)B * sizeof(int)));
        HIP_TRY(c, hipMalloc((void**)&d_st, (size_t)B * sizeof(int)));
        HIP_TRY(c, hipEventCreate(&e0));
        HIP_TRY(c, hipEventCreate(&e1));
        HIP_TRY(c, hipMemcpyAsync(d_par, par.data(), par.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(d_init, init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, zola_launch_psf_forward64(c->model, R, B, d_par, d_px, c->stream));
        std::vector<float> px((size_t)B * R2);
        HIP_TRY(c, hipMemcpyAsync(px.data(), d_px, px.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < px.size(); k++) px[k] = 1000.f * px[k] + 10.f;
        HIP_TRY(c, hipMemcpyAsync(d_px, px.data(), px.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        zola::FitArgs A{};
        A.B = B;
        A.max_iter = 30;
        A.xystep = (float)c->o.xystep;
        A.zmin = -1.3;
        A.zmax = 1.3;
        A.tol = 1e-4;
        A.noise_rel = 1e-8;
        A.pixels = d_px;
        A.init = d_init;
        A.queue = c->d_queue;
        A.theta = d_th;
        A.nll = d_nll;
        A.crlb = d_cr;
        A.iters = d_it;
        A.status = d_st;
        for (int q = 0; q < 3; q++) {
            // kernel layout q: 0 wave, 1 quad, 2 pair (ZOLA_LAYOUT_* is q + 1)
            if (!layout_available(R, mode, q + 1)) continue;
            HIP_TRY(c, zola_launch_fit64(c->model, R, mode, q, A, c->stream));  // warm-up
            float best = 1e30f;
            for (int rep = 0; rep < 2; rep++) {
                HIP_TRY(c, hipEventRecord(e0, c->stream));
                HIP_TRY(c, zola_launch_fit64(c->model, R, mode, q, A, c->stream));
                HIP_TRY(c, hipEventRecord(e1, c->stream));
                HIP_TRY(c, hipEventSynchronize(e1));
                float m = 0.f;
                HIP_TRY(c, hipEventElapsedTime(&m, e0, e1));
                best = std::min(best, m);
            }
            t[q] = best;
        }
        return ZOLA_OK;
    };
    rc = run();
    for (void* p : {(void*)d_par, (void*)d_px, (void*)d_init, (void*)d_th, (void*)d_nll, (void*)d_cr, (void*)d_it,
                    (void*)d_st})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    int best = 0;
    for (int q = 1; q < 3; q++)
        if (t[q] > 0 && t[q] < t[best]) best = q;
    const int choice = best + 1;  // ZOLA_LAYOUT_WAVE, _QUAD, _PAIR
    c->layout[roi_slot(R)][mode == ZOLA_FIT_LM ? 1 : 0] = choice;
    if (layout) *layout = choice;
    if (ms)
        for (int q = 0; q < 3; q++) ms[q] = t[q];
    return ZOLA_OK;
}

int zola_fit_mixed_dev(zola_ctx* c, const zola_fit_config* cfg, int G, const int* R, const int* B,
                       const float* const* pixels, const double* const* init, double* const* theta,
                       double* const* nll, double* const* crlb, int* const* iters, int* const* status) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    ZOLA_REFUSE_GRID256(c);
    if (!cfg || G <= 0 || G > 8 || !R || !B || !pixels || !init || !theta || !nll || !crlb || !iters || !status)
        return set_err(c, ZOLA_EINVAL, "bad mixed-fit arguments (1 <= G <= 8)");
    if (cfg->mode != ZOLA_FIT_REFERENCE && cfg->mode != ZOLA_FIT_LM)
        return set_err(c, ZOLA_EINVAL, "unknown fit mode %d", cfg->mode);
    if (cfg->max_iter < 0 || !(cfg->zmax >= cfg->zmin))
        return set_err(c, ZOLA_EINVAL, "bad iteration count or z range");
    for (int g = 0; g < G; g++) {
        if (B[g] < 0) return set_err(c, ZOLA_EINVAL, "negative batch in group %d", g);
        const bool ok = c->N == zola::kN ? zola_kernels_support_roi(R[g]) : (R[g] == c->R && zola_kernels_support_roi128(R[g]));
        if (!ok) return set_err(c, ZOLA_EUNSUPPORTED, "ROI side %d not available on the %d grid", R[g], c->N);
        if (B[g] > 0 && (!pixels[g] || !init[g] || !theta[g] || !nll[g] || !crlb[g] || !iters[g] || !status[g]))
            return set_err(c, ZOLA_EINVAL, "null buffer in group %d", g);
    }
    if (!c->fork) HIP_TRY(c, hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
    if (!c->cap) HIP_TRY(c, hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking));
    for (int g = 0; g < G; g++) {
        if (!c->side[g]) HIP_TRY(c, hipStreamCreateWithFlags(&c->side[g], hipStreamNonBlocking));
        if (!c->join[g]) HIP_TRY(c, hipEventCreateWithFlags(&c->join[g], hipEventDisableTiming));
    }
    // capture on a private stream (nothing runs during capture; the graph is
    // launched on the context's stream, after the work already queued there):
    // fork to one side stream per group, one fit launch each, join
    HIP_TRY(c, hipStreamBeginCapture(c->cap, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipEventRecord(c->fork, c->cap);
    for (int g = 0; g < G && err == hipSuccess; g++) {
        err = hipStreamWaitEvent(c->side[g], c->fork, 0);
        if (err != hipSuccess) break;
        if (B[g] > 0) {
            zola::FitArgs A{};
            A.B = B[g];
            A.max_iter = cfg->max_iter;
            A.xystep = (float)c->o.xystep;
            A.zmin = cfg->zmin;
            A.zmax = cfg->zmax;
            A.zfocus = cfg->zfocus;
            A.thr = cfg->photon_threshold;
            A.tol = cfg->tol;
            A.noise_rel = 1e-8;
            A.pixels = pixels[g];
            A.init = init[g];
            A.queue = c->d_queue + g;
            A.theta = theta[g];
            A.nll = nll[g];
            A.crlb = crlb[g];
            A.iters = iters[g];
            A.status = status[g];
            err = c->N == zola::kN2 ? zola_launch_fit128(c->model128, R[g], cfg->mode, zola_fit_layout128(c, cfg->mode), A, c->side[g])
                                    : zola_launch_fit64(c->model, R[g], cfg->mode, zola_fit_layout(c, R[g], cfg->mode),
                                                        A, c->side[g]);
            if (err != hipSuccess) break;
        }
        err = hipEventRecord(c->join[g], c->side[g]);
        if (err == hipSuccess) err = hipStreamWaitEvent(c->cap, c->join[g], 0);
    }
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->cap, &graph);
    if (err != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_err(c, ZOLA_EHIP, "mixed-fit capture: %s", hipGetErrorString(err != hipSuccess ? err : e2));
    }
    // reuse the previous executable graph when only the parameters changed
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    bool updated = false;
    if (c->mixed_exec) {
        hipGraphExecUpdateResult res;
        hipGraphNode_t bad = nullptr;
        updated = hipGraphExecUpdate(c->mixed_exec, graph, &bad, &res) == hipSuccess &&
                  res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGraphExecDestroy(c->mixed_exec);
            c->mixed_exec = nullptr;
        }
    }
    if (!updated) err = hipGraphInstantiate(&c->mixed_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) {
        c->mixed_exec = nullptr;
        return set_err(c, ZOLA_EHIP, "mixed-fit instantiate: %s", hipGetErrorString(err));
    }
    HIP_TRY(c, hipGraphLaunch(c->mixed_exec, c->stream));
    return ZOLA_OK;
}

int zola_fit_batch(zola_ctx* c, const zola_fit_config* cfg, int B, const float* pixels,
                   const double* init, double* theta, double* nll, double* crlb, int* iters,
                   int* status) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (!cfg || B < 0) return set_err(c, ZOLA_EINVAL, "bad fit arguments");
    if (B == 0) return ZOLA_OK;
    const size_t R2 = (size_t)c->R * c->R;
    const size_t bp = B * R2 * sizeof(float), bi = (size_t)B * 5 * sizeof(double);
    const size_t bt = bi, bn = (size_t)B * sizeof(double), bc = (size_t)B * 3 * sizeof(double);
    const size_t bk = (size_t)B * sizeof(int);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = al(bp) + al(bi) + al(bt) + al(bn) + al(bc) + 2 * al(bk);
    int rc = stage(c, total);
    if (rc) return rc;
    char* base = (char*)c->d_stage;
    float* dpx = (float*)base; base += al(bp);
    double* dini = (double*)base; base += al(bi);
    double* dth = (double*)base; base += al(bt);
    double* dnll = (double*)base; base += al(bn);
    double* dcr = (double*)base; base += al(bc);
    int* dit = (int*)base; base += al(bk);
    int* dst = (int*)base;
    HIP_TRY(c, hipMemcpyAsync(dpx, pixels, bp, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dini, init, bi, hipMemcpyHostToDevice, c->stream));
    rc = zola_fit_batch_dev(c, cfg, B, dpx, dini, dth, dnll, dcr, dit, dst);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(theta, dth, bt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(nll, dnll, bn, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(crlb, dcr, bc, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(iters, dit, bk, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(status, dst, bk, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

int zola_fit_ragged(zola_ctx* c, const zola_fit_config* cfg, int B, const long long* roi_offsets,
                    const int* roi_sizes, const float* pixels, long long n_pixels, const double* init,
                    double* theta, double* nll, double* crlb, int* iters, int* status) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    ZOLA_REFUSE_GRID256(c);
    if (!cfg || B < 0 || n_pixels < 0) return set_err(c, ZOLA_EINVAL, "bad ragged-fit arguments");
    if (B == 0) return ZOLA_OK;
    if (!roi_offsets || !roi_sizes || !pixels || !init || !theta || !nll || !crlb || !iters || !status)
        return set_err(c, ZOLA_EINVAL, "null buffer");
    // group the ROIs by side, in ROI order within a group
    std::vector<int> sides;
    std::vector<int> grp(B);
    for (int b = 0; b < B; b++) {
        const int R = roi_sizes[b];
        const long long off = roi_offsets[b];
        if (R <= 0 || off < 0 || off + (long long)R * R > n_pixels)
            return set_err(c, ZOLA_EINVAL, "ROI %d (side %d at offset %lld) outside the %lld pixels", b, R, off,
                           n_pixels);
        int g = 0;
        while (g < (int)sides.size() && sides[g] != R) g++;
        if (g == (int)sides.size()) {
            if (g == 8) return set_err(c, ZOLA_EINVAL, "more than 8 ROI sides in one ragged batch");
            const bool ok = c->N == zola::kN ? zola_kernels_support_roi(R) : (R == c->R && zola_kernels_support_roi128(R));
            if (!ok) return set_err(c, ZOLA_EUNSUPPORTED, "ROI side %d not available on the %d grid", R, c->N);
            sides.push_back(R);
        }
        grp[b] = g;
    }
    const int G = (int)sides.size();
    std::vector<int> cnt(G, 0);
    for (int b = 0; b < B; b++) cnt[grp[b]]++;
    // device layout: per group [pixels][init][theta][nll][crlb][iters][status], 256 B aligned
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<size_t> gpx(G), gin(G);
    size_t px_total = 0, in_total = 0, out_total = 0;
    for (int g = 0; g < G; g++) {
        const size_t n = (size_t)cnt[g];
        gpx[g] = px_total;
        px_total += al(n * sides[g] * sides[g] * sizeof(float));
        gin[g] = in_total;
        in_total += al(n * 5 * sizeof(double));
        out_total += al(n * 5 * sizeof(double)) + al(n * sizeof(double)) + al(n * 3 * sizeof(double)) +
                     2 * al(n * sizeof(int));
    }
    int rc = stage(c, px_total + in_total + out_total);
    if (rc) return rc;
    // host packing: one contiguous image of the device pixel and init sections
    std::vector<char> hin(px_total + in_total);
    std::vector<int> pos(B);
    {
        std::vector<int> k(G, 0);
        for (int b = 0; b < B; b++) {
            const int g = grp[b], R = sides[g];
            pos[b] = k[g]++;
            std::memcpy(hin.data() + gpx[g] + (size_t)pos[b] * R * R * sizeof(float), pixels + roi_offsets[b],
                        (size_t)R * R * sizeof(float));
            std::memcpy(hin.data() + px_total + gin[g] + (size_t)pos[b] * 5 * sizeof(double), init + (size_t)b * 5,
                        5 * sizeof(double));
        }
    }
    char* base = (char*)c->d_stage;
    HIP_TRY(c, hipMemcpyAsync(base, hin.data(), hin.size(), hipMemcpyHostToDevice, c->stream));
    const float* gp[8];
    const double* gi[8];
    double *gt[8], *gn[8], *gc[8];
    int *gk[8], *gs[8];
    char* o = base + px_total + in_total;
    for (int g = 0; g < G; g++) {
        const size_t n = (size_t)cnt[g];
        gp[g] = (const float*)(base + gpx[g]);
        gi[g] = (const double*)(base + px_total + gin[g]);
        gt[g] = (double*)o; o += al(n * 5 * sizeof(double));
        gn[g] = (double*)o; o += al(n * sizeof(double));
        gc[g] = (double*)o; o += al(n * 3 * sizeof(double));
        gk[g] = (int*)o; o += al(n * sizeof(int));
        gs[g] = (int*)o; o += al(n * sizeof(int));
    }
    rc = zola_fit_mixed_dev(c, cfg, G, sides.data(), cnt.data(), gp, gi, gt, gn, gc, gk, gs);
    if (rc) return rc;
    std::vector<char> hout(out_total);
    HIP_TRY(c, hipMemcpyAsync(hout.data(), base + px_total + in_total, out_total, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; b++) {
        const int g = grp[b], p = pos[b];
        auto host = [&](const void* dev) { return hout.data() + ((const char*)dev - (base + px_total + in_total)); };
        std::memcpy(theta + (size_t)b * 5, (const double*)host(gt[g]) + (size_t)p * 5, 5 * sizeof(double));
        nll[b] = ((const double*)host(gn[g]))[p];
        std::memcpy(crlb + (size_t)b * 3, (const double*)host(gc[g]) + (size_t)p * 3, 3 * sizeof(double));
        iters[b] = ((const int*)host(gk[g]))[p];
        status[b] = ((const int*)host(gs[g]))[p];
    }
    return ZOLA_OK;
}

int zola_fit_init_dev(zola_ctx* c, int B, const float* pixels, int nZ, double zmin_t,
                      double zstep_t, double zfocus, double* init) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    ZOLA_REFUSE_GRID256(c);
    if (B < 0 || nZ <= 0 || !(zstep_t > 0)) return set_err(c, ZOLA_EINVAL, "bad init arguments");
    if (B == 0) return ZOLA_OK;
    if (!pixels || !init) return set_err(c, ZOLA_EINVAL, "null buffer");
    const size_t R2 = (size_t)c->R * c->R;
    // [nZ R^2 template floats][nZ x 4 parameter floats][pad to 16 B][nZ x 2 doubles of sums]
    const size_t stat_off = ((size_t)nZ * (R2 + 4) * sizeof(float) + 15) & ~(size_t)15;
    const size_t need = stat_off + (size_t)nZ * 2 * sizeof(double);
    const bool cached = c->d_tmpl && c->tmpl_ver == c->model_ver && c->tmpl_nZ == nZ &&
                        c->tmpl_key[0] == zmin_t && c->tmpl_key[1] == zstep_t && c->tmpl_key[2] == zfocus;
    double* tstat = nullptr;
    if (!cached) {
        c->tmpl_ver = ~0u;
        if (need > c->tmpl_bytes) {
            if (c->d_tmpl) HIP_TRY(c, hipFree(c->d_tmpl));
            c->d_tmpl = nullptr;
            c->tmpl_bytes = 0;
            HIP_TRY(c, hipMalloc((void**)&c->d_tmpl, need));
            c->tmpl_bytes = need;
        }
        tstat = reinterpret_cast<double*>(reinterpret_cast<char*>(c->d_tmpl) + stat_off);
        // templates PSF(0, 0, z_r) with z_r = zmin_t + r * zstep_t (Predetection.java:118-125,190)
        std::vector<float> par((size_t)nZ * 4, 0.f);
        for (int r = 0; r < nZ; r++) {
            par[(size_t)r * 4 + 2] = (float)(zmin_t + r * zstep_t);
            par[(size_t)r * 4 + 3] = (float)zfocus;
        }
        float* dpar = c->d_tmpl + (size_t)nZ * R2;
        HIP_TRY(c, hipMemcpyAsync(dpar, par.data(), par.size() * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream));
        int rc = zola_psf_forward_dev(c, nZ, dpar, c->d_tmpl);
        if (rc) return rc;
        HIP_TRY(c, zola_launch_template_stats64(c->R, nZ, c->d_tmpl, tstat, c->stream));
        // the parameter upload must not be overwritten before the kernels read it
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->tmpl_ver = c->model_ver;
        c->tmpl_nZ = nZ;
        c->tmpl_key[0] = zmin_t;
        c->tmpl_key[1] = zstep_t;
        c->tmpl_key[2] = zfocus;
    } else {
        tstat = reinterpret_cast<double*>(reinterpret_cast<char*>(c->d_tmpl) + stat_off);
    }
    HIP_TRY(c, zola_launch_fit_init64(c->R, B, pixels, nZ, c->d_tmpl, tstat, zmin_t, zstep_t, init,
                                      c->stream));
    return ZOLA_OK;
}

}  // extern "C"
