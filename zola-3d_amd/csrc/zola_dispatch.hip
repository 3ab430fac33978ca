// zola_dispatch.hip -- runtime dispatch of the per-ROI-size kernel instantiations.
#include <hip/hip_runtime.h>

#include "zola_kernels.h"
#include "zola_fit64.hpp"

namespace zola {
template <int R>
hipError_t launch_psf_forward_r(const Model64& M, int B, const float* params, float* out,
                                hipStream_t st);
template <int R>
hipError_t launch_fit_r(const Model64& M, int mode, const FitArgs& A, hipStream_t st);
template <int R>
hipError_t launch_init_r(int B, const float* pixels, int nZ, const float* tmpl, double zmin_t,
                         double zstep_t, double* init, hipStream_t st);
}  // namespace zola

using namespace zola;

#define ZOLA_R_DISPATCH(R, CALL) \
    switch (R) {                 \
        case 17: return CALL(17); \
        case 21: return CALL(21); \
        case 31: return CALL(31); \
        case 41: return CALL(41); \
        default: return hipErrorInvalidValue; \
    }

bool zola_kernels_support_roi(int R) {
    return R == 17 || R == 21 || R == 31 || R == 41;
}

hipError_t zola_launch_psf_forward64(const Model64& M, int R, int B, const float* params,
                                     float* out, hipStream_t st) {
#define CALL(RR) launch_psf_forward_r<RR>(M, B, params, out, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}


hipError_t zola_launch_fit64(const Model64& M, int R, int mode, const FitArgs& A, hipStream_t st) {
#define CALL(RR) launch_fit_r<RR>(M, mode, A, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}

hipError_t zola_launch_fit_init64(int R, int B, const float* pixels, int nZ, const float* tmpl,
                                  double zmin_t, double zstep_t, double* init, hipStream_t st) {
#define CALL(RR) launch_init_r<RR>(B, pixels, nZ, tmpl, zmin_t, zstep_t, init, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}
