// zola_noalloc.cpp -- allocation-free form of the PSF adjoint.
//
// zola_psf_backward_dev writes each PSF's dL/dphase on the 32x32 pupil window
// (1024 floats) to a grow-only context buffer, then projects it on the Zernike
// basis.  Growing that buffer calls hipMalloc, which callers under HIP graph
// capture, in tight serving or calibration loops, or in harnesses that forbid
// allocation cannot do.  zola_psf_backward_noalloc_dev uses instead the fixed
// device array `scratch` below: ZOLA_NOALLOC_MAX_PSF x 1024 floats (8 MiB) for
// dL/dphase plus 1024 doubles (8 KiB) for its sums over the PSFs, held by every
// GPU the library runs on until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory, so its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) when a context is created there, and only that address is
// passed to kernels.  Every call on one GPU shares the array; the adjoint
// kernel overwrites all 1024 entries of each PSF it handles, and the sum kernel
// all 1024 sums, before the projection reads them, so nothing needs clearing
// between calls, but calls that overlap on one GPU (two streams, two contexts)
// race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/zola_hip.h"
#include "zola_adjoint64.hpp"
#include "zola_ctx.hpp"
#include "zola_kernels.h"

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kScratchFloats = (size_t)ZOLA_NOALLOC_MAX_PSF * 1024 + kQBytes / sizeof(float);

static __device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};   // device address of `scratch` per device

}  // namespace

// Called by zola_ctx_create with the context's device current.
int zola_scratch_setup(zola_ctx* c) {
    int dev = -1;
    HIP_TRY(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_err(c, ZOLA_EUNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(c, ZOLA_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[dev] = (float*)p;
    }
    c->scratch = g_scratch[dev];
    return ZOLA_OK;
}

extern "C" int zola_psf_backward_noalloc_dev(zola_ctx* c, int B, const float* params, const float* w,
                                             float* dparams, double* da) {
    CtxLock lk_(c);
    if (!c) return ZOLA_EINVAL;
    if (c->N != zola::kN) return set_err(c, ZOLA_EUNSUPPORTED, "the adjoint covers the 64 pupil grid");
    if (B < 0 || (B > 0 && (!params || !w || !dparams))) return set_err(c, ZOLA_EINVAL, "bad batch");
    if (B > ZOLA_NOALLOC_MAX_PSF)
        return set_err(c, ZOLA_EINVAL, "batch of %d PSFs exceeds the allocation-free limit of %d", B,
                       ZOLA_NOALLOC_MAX_PSF);
    if (B == 0) return ZOLA_OK;
    float* qbar = nullptr;
    if (da && c->J > 0) {
        if (!c->scratch) return set_err(c, ZOLA_EHIP, "scratch address not set up for this context");
        qbar = c->scratch;
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
