// zola_dispatch.hip -- runtime dispatch of the per-ROI-size kernel instantiations.
#include <hip/hip_runtime.h>

#include "zola_kernels.h"
#include "zola_adjoint64.hpp"
#include "zola_project.hpp"
#include "zola_psf128.hpp"

namespace zola {
template <int R>
hipError_t launch_psf_forward_r(const Model64& M, int B, const float* params, float* out,
                                hipStream_t st);
template <int R>
hipError_t launch_fit_r(const Model64& M, int mode, const FitArgs& A, hipStream_t st);
template <int R>
hipError_t launch_adjoint_r(const Model64& M, const AdjArgs& A, hipStream_t st);
template <int R>
hipError_t launch_fit64q_r(const Model64& M, int mode, const FitArgs& A, hipStream_t st);
template <int R>
hipError_t launch_fit64p_r(const Model64& M, const FitArgs& A, hipStream_t st);
template <int R>
hipError_t launch_psf_forward128_r(const Model128& M, int B, const float* params, float* out, hipStream_t st);
template <int R>
hipError_t launch_fit128_r(const Model128& M, int mode, const FitArgs& A, hipStream_t st);
template <int R>
hipError_t launch_fit128o_r(const Model128& M, const FitArgs& A, hipStream_t st);
template <int R>
hipError_t launch_init_r(int B, const float* pixels, int nZ, const float* tmpl, const double* tstat,
                         double zmin_t, double zstep_t, double* init, hipStream_t st);
template <int R>
hipError_t launch_tstats_r(int nZ, const float* tmpl, double* tstat, hipStream_t st);

// phase(w) = sum_j Z_j(w) a_j, products and sums in j order, not fused:
// the host's FP64 phase bit for bit, rounded to float as the host upload does
// (zero outside the disk, where the basis is zero).
__global__ __launch_bounds__(256) void k_phase_tab64(int J, CoefVals cv, const double* __restrict__ Zw,
                                                     float* __restrict__ tab) {
#pragma clang fp contract(off)
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= 1024) return;
    double ph = 0.0;
    // 32 basis loads in flight, then the products and sums in j order
    for (int j0 = 0; j0 < J; j0 += 32) {
        double zv[32];
#pragma unroll
        for (int k = 0; k < 32; k++) zv[k] = j0 + k < J ? Zw[(size_t)(j0 + k) * 1024 + w] : 0.0;
#pragma unroll
        for (int k = 0; k < 32; k++)
            if (j0 + k < J) ph = ph + zv[k] * cv.a[j0 + k];
    }
    const float f = (float)ph;
    tab[w] = f;
    tab[4096 + 4 * w] = f;
}

}  // namespace zola

using namespace zola;

hipError_t zola_launch_phase_tab64(int J, const double* a, const double* Zw, float* tab, hipStream_t st) {
    if (J <= 0 || J > kMaxCoefDev) return hipErrorInvalidValue;
    CoefVals cv{};
    for (int j = 0; j < J; j++) cv.a[j] = a[j];
    hipLaunchKernelGGL(k_phase_tab64, dim3(4), dim3(256), 0, st, J, cv, Zw, tab);
    return hipGetLastError();
}

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


hipError_t zola_launch_fit64(const Model64& M, int R, int mode, int layout, const FitArgs& A, hipStream_t st) {
    if (layout == 2) {
        if (mode != 1) return hipErrorInvalidValue;
        if (R == 17) return launch_fit64p_r<17>(M, A, st);
        if (R == 21) return launch_fit64p_r<21>(M, A, st);
        return hipErrorInvalidValue;
    }
    if (layout == 1) {
#define CALL(RR) launch_fit64q_r<RR>(M, mode, A, st)
        ZOLA_R_DISPATCH(R, CALL)
#undef CALL
    }
#define CALL(RR) launch_fit_r<RR>(M, mode, A, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}

hipError_t zola_launch_fit_init64(int R, int B, const float* pixels, int nZ, const float* tmpl,
                                  const double* tstat, double zmin_t, double zstep_t, double* init,
                                  hipStream_t st) {
#define CALL(RR) launch_init_r<RR>(B, pixels, nZ, tmpl, tstat, zmin_t, zstep_t, init, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}

hipError_t zola_launch_template_stats64(int R, int nZ, const float* tmpl, double* tstat, hipStream_t st) {
#define CALL(RR) launch_tstats_r<RR>(nZ, tmpl, tstat, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}

hipError_t zola_launch_adjoint64(const Model64& M, int R, const AdjArgs& A, hipStream_t st) {
#define CALL(RR) launch_adjoint_r<RR>(M, A, st)
    ZOLA_R_DISPATCH(R, CALL)
#undef CALL
}

hipError_t zola_launch_project_zernike(int B, int J, const double* Zw, const float* qbar, double* Q, double* da,
                                       hipStream_t st, bool assign) {
    return launch_window_projection<float, 256>(B, J, kW * kW, Zw, qbar, Q, da, st, assign);
}

// 128 grid: the plugin's default patch (32, blur grid 64), smaller even patches
// 16 and 20 (blur grids 32 and 64) and 41 (blur grid 128)
#define ZOLA_R128_DISPATCH(R, CALL) \
    switch (R) {                    \
        case 16: return CALL(16);   \
        case 20: return CALL(20);   \
        case 32: return CALL(32);   \
        case 41: return CALL(41);   \
        default: return hipErrorInvalidValue; \
    }

bool zola_kernels_support_roi128(int R) { return R == 16 || R == 20 || R == 32 || R == 41; }

hipError_t zola_launch_psf_forward128(const Model128& M, int R, int B, const float* params, float* out,
                                      hipStream_t st) {
#define CALL(RR) launch_psf_forward128_r<RR>(M, B, params, out, st)
    ZOLA_R128_DISPATCH(R, CALL)
#undef CALL
}

hipError_t zola_launch_fit128(const Model128& M, int R, int mode, int layout, const FitArgs& A, hipStream_t st) {
    if (layout == 3) {
        if (mode != 1) return hipErrorInvalidValue;
#define CALL(RR) launch_fit128o_r<RR>(M, A, st)
        ZOLA_R128_DISPATCH(R, CALL)
#undef CALL
    }
#define CALL(RR) launch_fit128_r<RR>(M, mode, A, st)
    ZOLA_R128_DISPATCH(R, CALL)
#undef CALL
}

namespace {
__global__ void k_zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
}  // namespace

hipError_t zola_launch_zero_ints(int* p, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_ints, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

int zola_persistent_blocks(const void* kernel, int threads) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess) return 0;
    return (per > 0 ? per : 1) * cus;
}
