// zola_calib.cpp -- PSF calibration from bead z-stacks on the adjoint.
//
// Replaces the reference's calibration loop (process/gpu/ZernikePhaseRetrieval_.java:362-474,
// 1108-1187, 2838-2862; CPU twin process/cpu/ZernikePhaseRetrieval.java:78-363, 1742-1879).
// The reference probes every Zernike coefficient by finite differences (+-0.01 rad,
// then gamma = 100, 10, 1, 0.1), i.e. several full-stack likelihood evaluations per
// coefficient per iteration.  Here one kernel launch returns the stack likelihood and
// its exact gradient with respect to every coefficient and every bead's registration,
// photon count and background; an FP64 L-BFGS drives the fit.
//
// Model (ZernikePhaseRetrieval.Model.computeLikelihood, 1859-1867): slice s of bead b is
// psf(x_b, y_b, z_oil = dz_s + z0_b, z_wat = 0) * A_b + bg_b, clamped at 1e-8.  The
// reference registers beads by overwriting basis slots 0..2 with kx, ky, kz_oil
// (1729-1736); the same phase terms enter here through the PSF position, so those
// slots are not fitted.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "../../include/zola_hip.h"

namespace {

struct CalibBuffers {
    float* params = nullptr;
    float* pixels = nullptr;
    double* amp = nullptr;
    double* bg = nullptr;
    double* nll = nullptr;
    float* dparams = nullptr;
    double* damp = nullptr;
    double* dbg = nullptr;
    double* da = nullptr;
    void* base = nullptr;
};

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

// Full calibration of nb bead stacks of S slices (host arrays).
//   pixels  nb*S*R*R photons; dz S slice defocus (um)
//   bead    nb*5 in/out [x, y, z0, A, bg]
//   coeffs  J in/out; fit_mask J (1 = fitted)
int zola_calib_fit(zola_ctx* ctx, int nb, int S, const float* pixels, const double* dz,
                   double* bead, double* coeffs, const int* fit_mask, int max_iter, double gtol,
                   double* out_nll, int* out_iters, int* out_evals) {
    if (!ctx || nb <= 0 || S <= 0 || !pixels || !dz || !bead || !coeffs || max_iter < 0)
        return ZOLA_EINVAL;
    int info[5];
    if (zola_ctx_info(ctx, info) != ZOLA_OK) return ZOLA_EINVAL;
    const int R = info[3], J = info[4];
    const int B = nb * S;
    const size_t R2 = (size_t)R * R;
    void* sp = nullptr;
    zola_get_stream(ctx, &sp);
    hipStream_t st = (hipStream_t)sp;

    // One device block: [inputs: params, amp, bg | outputs: nll, dparams, damp,
    // dbg, da | pixels].  The input and output spans are mirrored in pinned host
    // memory with the same offsets, so each evaluation moves one copy each way.
    CalibBuffers d;
    const size_t jz1 = (size_t)(J > 0 ? J : 1);
    const size_t o_par = 0, o_amp = o_par + al((size_t)B * 4 * sizeof(float)),
                 o_bg = o_amp + al((size_t)B * sizeof(double)), in_bytes = o_bg + al((size_t)B * sizeof(double));
    const size_t o_nll = 0, o_dp = o_nll + al((size_t)B * sizeof(double)),
                 o_dA = o_dp + al((size_t)B * 4 * sizeof(float)), o_dB = o_dA + al((size_t)B * sizeof(double)),
                 o_da = o_dB + al((size_t)B * sizeof(double)), out_bytes = o_da + al(jz1 * sizeof(double));
    const size_t total = in_bytes + out_bytes + al(B * R2 * sizeof(float));
    if (hipMalloc(&d.base, total) != hipSuccess) return ZOLA_ENOMEM;
    char* hin = nullptr;
    char* hout = nullptr;
    if (hipHostMalloc((void**)&hin, in_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&hout, out_bytes, hipHostMallocDefault) != hipSuccess) {
        if (hin) (void)hipHostFree(hin);
        (void)hipFree(d.base);
        return ZOLA_ENOMEM;
    }
    char* din = (char*)d.base;
    char* dout = din + in_bytes;
    d.params = (float*)(din + o_par);
    d.amp = (double*)(din + o_amp);
    d.bg = (double*)(din + o_bg);
    d.nll = (double*)(dout + o_nll);
    d.dparams = (float*)(dout + o_dp);
    d.damp = (double*)(dout + o_dA);
    d.dbg = (double*)(dout + o_dB);
    d.da = (double*)(dout + o_da);
    d.pixels = (float*)(dout + out_bytes);
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        (void)hipHostFree(hin);
        (void)hipHostFree(hout);
        (void)hipFree(d.base);
    };
    int rc = ZOLA_OK;
    if (hipMemcpy(d.pixels, pixels, B * R2 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        release();
        return ZOLA_EHIP;
    }

    // variables: fitted coefficients, then per bead x, y, z0, A, bg (scaled)
    std::vector<int> fit;
    for (int jz = 0; jz < J; jz++)
        if (!fit_mask || fit_mask[jz]) fit.push_back(jz);
    const int nc = (int)fit.size();
    const int nv = nc + 5 * nb;
    std::vector<double> scale(nv, 1.0);
    for (int b = 0; b < nb; b++) {
        scale[nc + 5 * b + 0] = 0.1;
        scale[nc + 5 * b + 1] = 0.1;
        scale[nc + 5 * b + 2] = 0.1;
        scale[nc + 5 * b + 3] = std::fmax(std::fabs(bead[5 * b + 3]), 1.0);
        scale[nc + 5 * b + 4] = std::fmax(std::fabs(bead[5 * b + 4]), 1.0);
    }
    std::vector<double> u(nv), coef(coeffs, coeffs + J), bd(bead, bead + 5 * (size_t)nb);
    for (int k = 0; k < nc; k++) u[k] = coef[fit[k]] / scale[k];
    for (int k = 0; k < 5 * nb; k++) u[nc + k] = bd[k] / scale[nc + k];

    float* hpar = (float*)(hin + o_par);
    double* hamp = (double*)(hin + o_amp);
    double* hbg = (double*)(hin + o_bg);
    const double* hnll = (const double*)(hout + o_nll);
    const float* hdp = (const float*)(hout + o_dp);
    const double* hdA = (const double*)(hout + o_dA);
    const double* hdB = (const double*)(hout + o_dB);
    const double* hda = (const double*)(hout + o_da);
    int evals = 0;
    // f(u) and its gradient
    auto evaluate = [&](const std::vector<double>& uu, std::vector<double>& g, double& f) -> int {
        for (int k = 0; k < nc; k++) coef[fit[k]] = uu[k] * scale[k];
        int r = zola_set_coefficients(ctx, coef.data(), J);
        if (r) return r;
        for (int b = 0; b < nb; b++)
            for (int s = 0; s < S; s++) {
                const size_t i = (size_t)b * S + s;
                const double* bb = &uu[nc + 5 * b];
                hpar[i * 4 + 0] = (float)(bb[0] * scale[nc + 5 * b + 0]);
                hpar[i * 4 + 1] = (float)(bb[1] * scale[nc + 5 * b + 1]);
                hpar[i * 4 + 2] = 0.f;
                hpar[i * 4 + 3] = (float)(dz[s] + bb[2] * scale[nc + 5 * b + 2]);
                hamp[i] = bb[3] * scale[nc + 5 * b + 3];
                hbg[i] = bb[4] * scale[nc + 5 * b + 4];
            }
        // the previous evaluation ended with a stream synchronisation, so the
        // pinned input span is free to be rewritten above
        if (hipMemcpyAsync(din, hin, in_bytes, hipMemcpyHostToDevice, st) ||
            hipMemsetAsync(d.da, 0, jz1 * sizeof(double), st))
            return ZOLA_EHIP;
        r = zola_calib_eval_dev(ctx, B, d.params, d.pixels, d.amp, d.bg, d.nll, d.dparams, d.damp,
                                d.dbg, d.da);
        if (r) return r;
        if (hipMemcpyAsync(hout, dout, out_bytes, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st))
            return ZOLA_EHIP;
        evals++;
        f = 0.0;
        g.assign(nv, 0.0);
        for (int i = 0; i < B; i++) f += hnll[i];
        for (int k = 0; k < nc; k++) g[k] = hda[fit[k]] * scale[k];
        for (int b = 0; b < nb; b++)
            for (int s = 0; s < S; s++) {
                const size_t i = (size_t)b * S + s;
                double* gb = &g[nc + 5 * b];
                gb[0] += (double)hdp[i * 4 + 0];
                gb[1] += (double)hdp[i * 4 + 1];
                gb[2] += (double)hdp[i * 4 + 3];
                gb[3] += hdA[i];
                gb[4] += hdB[i];
            }
        for (int b = 0; b < nb; b++)
            for (int q = 0; q < 5; q++) g[nc + 5 * b + q] *= scale[nc + 5 * b + q];
        return ZOLA_OK;
    };

    std::vector<double> g, gn, un;
    double f = 0, fn = 0;
    rc = evaluate(u, g, f);
    std::deque<std::vector<double>> Sv, Yv;
    int it = 0;
    for (it = 0; it < max_iter && rc == ZOLA_OK; it++) {
        double gmax = 0;
        for (double v : g) gmax = std::fmax(gmax, std::fabs(v));
        if (gmax < gtol) break;
        // L-BFGS two-loop recursion
        std::vector<double> q = g;
        std::vector<double> alpha(Sv.size());
        for (int k = (int)Sv.size() - 1; k >= 0; k--) {
            double sy = 0, sq = 0;
            for (int v = 0; v < nv; v++) { sy += Sv[k][v] * Yv[k][v]; sq += Sv[k][v] * q[v]; }
            alpha[k] = sq / sy;
            for (int v = 0; v < nv; v++) q[v] -= alpha[k] * Yv[k][v];
        }
        double gamma = 1.0;
        if (!Sv.empty()) {
            double sy = 0, yy = 0;
            for (int v = 0; v < nv; v++) { sy += Sv.back()[v] * Yv.back()[v]; yy += Yv.back()[v] * Yv.back()[v]; }
            gamma = sy / yy;
        } else {
            gamma = 1e-3 / std::fmax(gmax, 1e-30);
        }
        for (int v = 0; v < nv; v++) q[v] *= gamma;
        for (size_t k = 0; k < Sv.size(); k++) {
            double sy = 0, yq = 0;
            for (int v = 0; v < nv; v++) { sy += Sv[k][v] * Yv[k][v]; yq += Yv[k][v] * q[v]; }
            const double beta = yq / sy;
            for (int v = 0; v < nv; v++) q[v] += Sv[k][v] * (alpha[k] - beta);
        }
        double gd = 0;
        for (int v = 0; v < nv; v++) gd -= g[v] * q[v];
        if (!(gd < 0)) {  // not a descent direction: restart from the gradient
            Sv.clear();
            Yv.clear();
            for (int v = 0; v < nv; v++) q[v] = g[v] * 1e-3 / std::fmax(gmax, 1e-30);
            gd = 0;
            for (int v = 0; v < nv; v++) gd -= g[v] * q[v];
        }
        // backtracking line search (Armijo)
        double step = 1.0;
        bool ok = false;
        for (int ls = 0; ls < 20; ls++) {
            un = u;
            for (int v = 0; v < nv; v++) un[v] -= step * q[v];
            for (int b = 0; b < nb; b++) {
                un[nc + 5 * b + 3] = std::fmax(un[nc + 5 * b + 3], 1.0 / scale[nc + 5 * b + 3]);
                un[nc + 5 * b + 4] = std::fmax(un[nc + 5 * b + 4], 1e-3 / scale[nc + 5 * b + 4]);
            }
            rc = evaluate(un, gn, fn);
            if (rc) break;
            if (fn <= f + 1e-4 * step * gd) { ok = true; break; }
            step *= 0.5;
        }
        if (rc || !ok) break;
        std::vector<double> s(nv), y(nv);
        double sy = 0;
        for (int v = 0; v < nv; v++) { s[v] = un[v] - u[v]; y[v] = gn[v] - g[v]; sy += s[v] * y[v]; }
        if (sy > 1e-12) {
            Sv.push_back(s);
            Yv.push_back(y);
            if (Sv.size() > 10) { Sv.pop_front(); Yv.pop_front(); }
        }
        const double rel = (f - fn) / std::fmax(std::fabs(f), 1.0);
        u.swap(un);
        g.swap(gn);
        f = fn;
        if (rel >= 0 && rel < 1e-13) { it++; break; }
    }
    if (rc == ZOLA_OK) {
        for (int k = 0; k < nc; k++) coeffs[fit[k]] = u[k] * scale[k];
        for (int k = 0; k < 5 * nb; k++) bead[k] = u[nc + k] * scale[nc + k];
        rc = zola_set_coefficients(ctx, coeffs, J);
    }
    if (out_nll) *out_nll = f;
    if (out_iters) *out_iters = it;
    if (out_evals) *out_evals = evals;
    release();
    return rc;
}

}  // extern "C"
