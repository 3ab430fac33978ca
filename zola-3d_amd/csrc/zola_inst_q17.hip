// Kernel instantiations for ROI side 17, 64 grid, four waves per ROI.
#include "zola_kernels_impl.hpp"

namespace zola {
template hipError_t launch_fit64q_r<17>(const Model64&, int, const FitArgs&, hipStream_t);
}  // namespace zola
