// Kernel instantiations for the 128x128 pupil grid (ROI side 16).
#include "zola_kernels_impl.hpp"

namespace zola {
template hipError_t launch_psf_forward128_r<16>(const Model128&, int, const float*, float*, hipStream_t);
template hipError_t launch_fit128_r<16>(const Model128&, int, const FitArgs&, hipStream_t);
}  // namespace zola
