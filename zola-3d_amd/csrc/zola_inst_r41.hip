// Kernel instantiations for ROI side 41.
#include "zola_kernels_impl.hpp"

namespace zola {
template hipError_t launch_psf_forward_r<41>(const Model64&, int, const float*, float*, hipStream_t);
template hipError_t launch_fit_r<41>(const Model64&, int, const FitArgs&, hipStream_t);
template hipError_t launch_init_r<41>(int, const float*, int, const float*, const double*, double, double,
                                      double*, hipStream_t);
template hipError_t launch_tstats_r<41>(int, const float*, double*, hipStream_t);
template hipError_t launch_adjoint_r<41>(const Model64&, const AdjArgs&, hipStream_t);
}  // namespace zola
