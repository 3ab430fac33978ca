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
// Two models.  zola_calib_fit: slice s of bead b is psf(x_b, y_b, z_oil = dz_s + z0_b,
// z_wat = 0) * A_b + bg_b, clamped at 1e-8, with scalar photons and background per bead.
// zola_calib_fit_full: the reference's own model (ZernikePhaseRetrieval.Model.
// computeLikelihood, 1859-1867): psf * (A[b][s] + fit_a[b] + fit_a_each[b][s]) + an
// order-2 polynomial background per bead (+ the bead's sCMOS map), with the blur width
// sigma fitted too.  The reference registers beads by overwriting basis slots 0..2 with
// kx, ky, kz_oil (1729-1736); the same phase terms enter here through the PSF position,
// so those slots are not fitted.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_ctx.hpp"

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

// FP64 L-BFGS (memory 10) with an Armijo backtracking line search on scaled
// variables u.  evaluate(u, g, f) returns f(u) and its gradient; project(u)
// applies the variables' bounds to a trial point.  Stops after max_iter
// iterations, when the largest gradient component is below gtol, when the line
// search fails, or when the relative decrease of f falls below 1e-13.
template <class Eval, class Proj>
int lbfgs_minimize(std::vector<double>& u, Eval evaluate, Proj project, int max_iter, double gtol, double& f,
                   int& iters) {
    const int nv = (int)u.size();
    std::vector<double> g, gn, un;
    double fn = 0;
    int rc = evaluate(u, g, f);
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
            project(un);
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
    iters = it;
    return rc;
}

}  // namespace

extern "C" {

// Full calibration of nb bead stacks of S slices (host arrays).
//   pixels  nb*S*R*R photons; dz S slice defocus (um)
//   bead    nb*5 in/out [x, y, z0, A, bg]
//   coeffs  J in/out; fit_mask J (1 = fitted)
int zola_calib_fit(zola_ctx* ctx, int nb, int S, const float* pixels, const double* dz,
                   double* bead, double* coeffs, const int* fit_mask, int max_iter, double gtol,
                   double* out_nll, int* out_iters, int* out_evals) {
    CtxLock lk_(ctx);
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
        if (hipMemcpyAsync(din, hin, in_bytes, hipMemcpyHostToDevice, st)) return ZOLA_EHIP;
        // dL/da is written, not accumulated: no clearing pass before the evaluation
        ctx->da_assign = true;
        r = zola_calib_eval_dev(ctx, B, d.params, d.pixels, d.amp, d.bg, d.nll, d.dparams, d.damp,
                                d.dbg, d.da);
        ctx->da_assign = false;
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

    double f = 0;
    int it = 0;
    rc = lbfgs_minimize(
        u, evaluate,
        [&](std::vector<double>& un) {
            for (int b = 0; b < nb; b++) {
                un[nc + 5 * b + 3] = std::fmax(un[nc + 5 * b + 3], 1.0 / scale[nc + 5 * b + 3]);
                un[nc + 5 * b + 4] = std::fmax(un[nc + 5 * b + 4], 1e-3 / scale[nc + 5 * b + 4]);
            }
        },
        max_iter, gtol, f, it);
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

// The reference's full calibration model (declared in include/zola_hip.h).
// Variables, each scaled to its natural size: the fitted coefficients (rad);
// per bead x, y, z0 (0.1 um); sigma (0.1 px); per bead the six background
// coefficients (c0 at the bead's background level, the monomials of degree d
// divided by R^d); per bead fit_a (0.1 of the mean photon curve) and per slice
// fit_a_each (0.1 of that slice's A).  sigma is kept >= 0.05 px.
int zola_calib_fit_full(zola_ctx* ctx, zola_calib_beads* bd, double* sigma, double* coeffs, const int* fit_mask,
                        int flags, int max_iter, double gtol, double* out_nll, int* out_iters, int* out_evals) {
    CtxLock lk_(ctx);
    if (!ctx || !bd || !coeffs || !sigma || max_iter < 0 || flags < 0 || flags > ZOLA_CALIB_FIT_ALL)
        return ZOLA_EINVAL;
    const int nb = bd->nb, S = bd->S;
    if (nb <= 0 || S <= 0 || !bd->pixels || !bd->dz || !bd->A || !bd->xyz || !bd->poly || !bd->fit_a ||
        !bd->fit_a_each || !(*sigma > 0))
        return ZOLA_EINVAL;
    int info[5];
    if (zola_ctx_info(ctx, info) != ZOLA_OK) return ZOLA_EINVAL;
    const int R = info[3], J = info[4];
    const int B = nb * S;
    const size_t R2 = (size_t)R * R;
    void* sp = nullptr;
    zola_get_stream(ctx, &sp);
    hipStream_t st = (hipStream_t)sp;
    const bool f_reg = flags & ZOLA_CALIB_FIT_REGISTRATION, f_sig = flags & ZOLA_CALIB_FIT_SIGMA,
               f_bg = flags & ZOLA_CALIB_FIT_BACKGROUND, f_ph = flags & ZOLA_CALIB_FIT_PHOTONS;

    // device block [inputs: params, amp, poly | outputs: nll, dparams, damp, dpoly,
    // dsig, da | pixels, offset]; inputs and outputs mirrored in pinned host memory
    const size_t jz1 = (size_t)(J > 0 ? J : 1);
    const size_t o_par = 0, o_amp = o_par + al((size_t)B * 4 * sizeof(float)),
                 o_poly = o_amp + al((size_t)B * sizeof(double)), in_bytes = o_poly + al((size_t)nb * 6 * sizeof(double));
    const size_t o_nll = 0, o_dp = o_nll + al((size_t)B * sizeof(double)),
                 o_dA = o_dp + al((size_t)B * 4 * sizeof(float)), o_dpoly = o_dA + al((size_t)B * sizeof(double)),
                 o_dsig = o_dpoly + al((size_t)B * 6 * sizeof(double)), o_da = o_dsig + al((size_t)B * sizeof(double)),
                 out_bytes = o_da + al(jz1 * sizeof(double));
    const size_t px_bytes = al((size_t)B * R2 * sizeof(float)), off_bytes = bd->offset ? al((size_t)nb * R2 * sizeof(double)) : 0;
    void* base = nullptr;
    if (hipMalloc(&base, in_bytes + out_bytes + px_bytes + off_bytes) != hipSuccess) return ZOLA_ENOMEM;
    char* hin = nullptr;
    char* hout = nullptr;
    if (hipHostMalloc((void**)&hin, in_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&hout, out_bytes, hipHostMallocDefault) != hipSuccess) {
        if (hin) (void)hipHostFree(hin);
        (void)hipFree(base);
        return ZOLA_ENOMEM;
    }
    char* din = (char*)base;
    char* dout = din + in_bytes;
    float* d_px = (float*)(dout + out_bytes);
    double* d_off = bd->offset ? (double*)((char*)d_px + px_bytes) : nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        (void)hipHostFree(hin);
        (void)hipHostFree(hout);
        (void)hipFree(base);
    };
    if (hipMemcpy(d_px, bd->pixels, (size_t)B * R2 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        (d_off && hipMemcpy(d_off, bd->offset, (size_t)nb * R2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)) {
        release();
        return ZOLA_EHIP;
    }

    // variable layout and scales
    std::vector<int> fit;
    for (int jz = 0; jz < J; jz++)
        if (!fit_mask || fit_mask[jz]) fit.push_back(jz);
    const int nc = (int)fit.size();
    const int o_reg = nc, o_sig = o_reg + (f_reg ? 3 * nb : 0), o_bg = o_sig + (f_sig ? 1 : 0),
              o_ph = o_bg + (f_bg ? 6 * nb : 0), nv = o_ph + (f_ph ? nb * (S + 1) : 0);
    std::vector<double> scale(nv, 1.0), u(nv);
    std::vector<double> coef(coeffs, coeffs + J), xyz(bd->xyz, bd->xyz + 3 * (size_t)nb),
        poly(bd->poly, bd->poly + 6 * (size_t)nb), fa(bd->fit_a, bd->fit_a + nb),
        fae(bd->fit_a_each, bd->fit_a_each + (size_t)nb * S);
    double sig = *sigma;
    for (int k = 0; k < nc; k++) u[k] = coef[fit[k]];
    if (f_reg)
        for (int k = 0; k < 3 * nb; k++) { scale[o_reg + k] = 0.1; u[o_reg + k] = xyz[k] / 0.1; }
    if (f_sig) { scale[o_sig] = 0.1; u[o_sig] = sig / 0.1; }
    if (f_bg)
        for (int b = 0; b < nb; b++) {
            const double s0 = std::fmax(std::fabs(poly[6 * b]), 1.0);
            const double deg[6] = {0, 1, 2, 1, 2, 2};
            for (int c = 0; c < 6; c++) {
                scale[o_bg + 6 * b + c] = s0 / std::pow((double)R, deg[c]);
                u[o_bg + 6 * b + c] = poly[6 * b + c] / scale[o_bg + 6 * b + c];
            }
        }
    if (f_ph)
        for (int b = 0; b < nb; b++) {
            double m = 0;
            for (int s = 0; s < S; s++) m += bd->A[(size_t)b * S + s];
            m /= S;
            const int ob = o_ph + b * (S + 1);
            scale[ob] = 0.1 * std::fmax(std::fabs(m), 1.0);
            u[ob] = fa[b] / scale[ob];
            for (int s = 0; s < S; s++) {
                scale[ob + 1 + s] = 0.1 * std::fmax(std::fabs(bd->A[(size_t)b * S + s]), 1.0);
                u[ob + 1 + s] = fae[(size_t)b * S + s] / scale[ob + 1 + s];
            }
        }

    float* hpar = (float*)(hin + o_par);
    double* hamp = (double*)(hin + o_amp);
    double* hpoly = (double*)(hin + o_poly);
    const double* hnll = (const double*)(hout + o_nll);
    const float* hdp = (const float*)(hout + o_dp);
    const double* hdA = (const double*)(hout + o_dA);
    const double* hdpoly = (const double*)(hout + o_dpoly);
    const double* hdsig = (const double*)(hout + o_dsig);
    const double* hda = (const double*)(hout + o_da);
    int evals = 0;
    auto unpack = [&](const std::vector<double>& uu) {
        for (int k = 0; k < nc; k++) coef[fit[k]] = uu[k];
        if (f_reg)
            for (int k = 0; k < 3 * nb; k++) xyz[k] = uu[o_reg + k] * scale[o_reg + k];
        if (f_sig) sig = uu[o_sig] * scale[o_sig];
        if (f_bg)
            for (int k = 0; k < 6 * nb; k++) poly[k] = uu[o_bg + k] * scale[o_bg + k];
        if (f_ph)
            for (int b = 0; b < nb; b++) {
                const int ob = o_ph + b * (S + 1);
                fa[b] = uu[ob] * scale[ob];
                for (int s = 0; s < S; s++) fae[(size_t)b * S + s] = uu[ob + 1 + s] * scale[ob + 1 + s];
            }
    };
    auto evaluate = [&](const std::vector<double>& uu, std::vector<double>& g, double& f) -> int {
        unpack(uu);
        int r = nc > 0 ? zola_set_coefficients(ctx, coef.data(), J) : ZOLA_OK;
        if (r == ZOLA_OK && f_sig) r = zola_set_sigma(ctx, sig);
        if (r) return r;
        for (int b = 0; b < nb; b++)
            for (int s = 0; s < S; s++) {
                const size_t i = (size_t)b * S + s;
                hpar[i * 4 + 0] = (float)xyz[3 * b + 0];
                hpar[i * 4 + 1] = (float)xyz[3 * b + 1];
                hpar[i * 4 + 2] = 0.f;
                hpar[i * 4 + 3] = (float)(bd->dz[s] + xyz[3 * b + 2]);
                hamp[i] = bd->A[i] + fa[b] + fae[i];
            }
        std::memcpy(hpoly, poly.data(), 6 * (size_t)nb * sizeof(double));
        if (hipMemcpyAsync(din, hin, in_bytes, hipMemcpyHostToDevice, st)) return ZOLA_EHIP;
        // dL/da is written, not accumulated: no clearing pass before the evaluation
        ctx->da_assign = true;
        r = zola_calib_eval_full_dev(ctx, nb, S, (const float*)(din + o_par), d_px, (const double*)(din + o_amp),
                                     (const double*)(din + o_poly), d_off, (double*)(dout + o_nll),
                                     (float*)(dout + o_dp), (double*)(dout + o_dA),
                                     f_bg ? (double*)(dout + o_dpoly) : nullptr,
                                     f_sig ? (double*)(dout + o_dsig) : nullptr,
                                     nc > 0 ? (double*)(dout + o_da) : nullptr);
        ctx->da_assign = false;
        if (r) return r;
        if (hipMemcpyAsync(hout, dout, out_bytes, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st))
            return ZOLA_EHIP;
        evals++;
        f = 0.0;
        g.assign(nv, 0.0);
        for (int i = 0; i < B; i++) f += hnll[i];
        for (int k = 0; k < nc; k++) g[k] = hda[fit[k]];
        for (int b = 0; b < nb; b++)
            for (int s = 0; s < S; s++) {
                const size_t i = (size_t)b * S + s;
                if (f_reg) {
                    g[o_reg + 3 * b + 0] += (double)hdp[i * 4 + 0];
                    g[o_reg + 3 * b + 1] += (double)hdp[i * 4 + 1];
                    g[o_reg + 3 * b + 2] += (double)hdp[i * 4 + 3];
                }
                if (f_sig) g[o_sig] += hdsig[i];
                if (f_bg)
                    for (int c = 0; c < 6; c++) g[o_bg + 6 * b + c] += hdpoly[i * 6 + c];
                if (f_ph) {
                    const int ob = o_ph + b * (S + 1);
                    g[ob] += hdA[i];
                    g[ob + 1 + s] += hdA[i];
                }
            }
        for (int v = nc; v < nv; v++) g[v] *= scale[v];
        return ZOLA_OK;
    };

    double f = 0;
    int it = 0;
    int rc = lbfgs_minimize(
        u, evaluate,
        [&](std::vector<double>& un) {
            if (f_sig) un[o_sig] = std::fmax(un[o_sig], 0.05 / scale[o_sig]);
        },
        max_iter, gtol, f, it);
    if (rc == ZOLA_OK) {
        unpack(u);
        std::memcpy(coeffs, coef.data(), J * sizeof(double));
        std::memcpy(bd->xyz, xyz.data(), 3 * (size_t)nb * sizeof(double));
        std::memcpy(bd->poly, poly.data(), 6 * (size_t)nb * sizeof(double));
        std::memcpy(bd->fit_a, fa.data(), nb * sizeof(double));
        std::memcpy(bd->fit_a_each, fae.data(), (size_t)nb * S * sizeof(double));
        *sigma = sig;
        rc = zola_set_coefficients(ctx, coeffs, J);
        if (rc == ZOLA_OK) rc = zola_set_sigma(ctx, sig);
    }
    if (out_nll) *out_nll = f;
    if (out_iters) *out_iters = it;
    if (out_evals) *out_evals = evals;
    release();
    return rc;
}

}  // extern "C"
