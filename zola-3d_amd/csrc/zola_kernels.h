// zola_kernels.h -- host-side declarations of the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

namespace zola {
struct Model64;
struct Model128;
struct FitArgs;
struct AdjArgs;
}

bool zola_kernels_support_roi(int R);

hipError_t zola_launch_psf_forward64(const zola::Model64& M, int R, int B, const float* params,
                                     float* out, hipStream_t st);

// layout: 0 one wave per ROI (zola_psf64.hpp), 1 four waves per ROI (zola_psf64q.hpp),
// 2 two waves per ROI (zola_psf64p.hpp; LM form, R = 17 and 21)
hipError_t zola_launch_fit64(const zola::Model64& M, int R, int mode, int layout, const zola::FitArgs& A,
                             hipStream_t st);

// tstat: per-template {sum t, sum t^2 - (sum t)^2 / R^2} from zola_launch_template_stats64
hipError_t zola_launch_fit_init64(int R, int B, const float* pixels, int nZ, const float* tmpl,
                                  const double* tstat, double zmin_t, double zstep_t, double* init,
                                  hipStream_t st);

hipError_t zola_launch_template_stats64(int R, int nZ, const float* tmpl, double* tstat, hipStream_t st);

hipError_t zola_launch_adjoint64(const zola::Model64& M, int R, const zola::AdjArgs& A,
                                 hipStream_t st);

// Zernike coefficients by value for the device update of the phase tables.
constexpr int kMaxCoefDev = 128;
struct CoefVals {
    double a[kMaxCoefDev];
};

// Phase tables of the 64 grid from the coefficients a[0..J) and the basis Zw
// (J x 1024 FP64 on the 32x32 window): tab[w] and the interleaved copy
// tab[4096 + 4 w] (the table layout of Model64).
hipError_t zola_launch_phase_tab64(int J, const double* a, const double* Zw, float* tab, hipStream_t st);

// da[0..J) += sum_p Z_j(p) Q(p) over the 32x32 window, Q(p) = sum_b qbar[b][p]
// formed into Q (1024 doubles of device memory).
hipError_t zola_launch_project_zernike(int B, int J, const double* Zw, const float* qbar, double* Q,
                                       double* da, hipStream_t st, bool assign = false);

// 128x128 pupil grid
bool zola_kernels_support_roi128(int R);
hipError_t zola_launch_psf_forward128(const zola::Model128& M, int R, int B, const float* params,
                                      float* out, hipStream_t st);
// layout: 1 four waves per ROI (both forms), 3 eight waves per ROI (LM form only).
hipError_t zola_launch_fit128(const zola::Model128& M, int R, int mode, int layout, const zola::FitArgs& A,
                              hipStream_t st);

// Zeroes n ints (a kernel, so that it can sit in a captured graph).
hipError_t zola_launch_zero_ints(int* p, int n, hipStream_t st);
// Blocks of a persistent launch: resident blocks per CU (occupancy) times CUs.
int zola_persistent_blocks(const void* kernel, int threads);
