// Kernel instantiations for ROI side 21, 64 grid, two waves per ROI (LM form).
#include "zola_kernels_impl.hpp"

namespace zola {
template hipError_t launch_fit64p_r<21>(const Model64&, const FitArgs&, hipStream_t);
}  // namespace zola
