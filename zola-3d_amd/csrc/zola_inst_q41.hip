// Kernel instantiations for ROI side 41, 64 grid, four waves per ROI.
#include "zola_kernels_impl.hpp"

namespace zola {
template hipError_t launch_fit64q_r<41>(const Model64&, int, const FitArgs&, hipStream_t);
}  // namespace zola
