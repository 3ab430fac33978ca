// zola_grid256.h -- the 256 x 256 pupil grid (zola_grid256.hip): PSF synthesis,
// adjoint and calibration likelihood for contexts with N = 256.
#pragma once
#include <hip/hip_runtime.h>

struct zola_ctx;

// One batched evaluation on the 256 grid (device pointers, B PSFs at params
// B x 4 [x, y, z, zoil]).  Modes, by the pointers given:
//  * out:    PSF synthesis (B x R x R floats, sum 1);
//  * w:      adjoint of the PSF for a given dL/dpsf (B x R x R floats);
//  * pixels: calibration likelihood mu = amp psf + background (poly per bead of
//            S slices, or bg per PSF when poly is null, plus offset), clamped
//            at 1e-8, and its adjoint.
// dparams (B x 4 floats), da (J doubles, accumulated), dsig (B), nll, damp,
// dbg, dpoly (B x 6) are written when given.
struct Eval256 {
    int B = 0;
    const float* params = nullptr;
    float* out = nullptr;
    const float* w = nullptr;
    const float* pixels = nullptr;
    const double* amp = nullptr;
    const double* bg = nullptr;
    int S = 1;
    const double* poly = nullptr;
    const double* offset = nullptr;
    double* nll = nullptr;
    float* dparams = nullptr;
    double* damp = nullptr;
    double* dbg = nullptr;
    double* dpoly = nullptr;
    double* dsig = nullptr;
    double* da = nullptr;
    bool da_assign = false;   // write dL/da instead of adding to it
};

int zola256_setup(zola_ctx* c);           // device tables of a new N = 256 context
int zola256_upload_tables(zola_ctx* c);   // pupil window {phase, amp, kz, kzo} after a coefficient change
int zola256_upload_gauss(zola_ctx* c);    // Gaussian spectrum and its sigma derivative after a sigma change
// The same two updates formed on the device (stream-ordered, no host pass and
// no synchronisation): the phase from the coefficients and the basis already
// there, the Gaussian from sigma.
int zola256_set_phase(zola_ctx* c);
int zola256_set_gauss(zola_ctx* c);
void zola256_free(zola_ctx* c);
int zola256_run(zola_ctx* c, const Eval256& e, hipStream_t st);
