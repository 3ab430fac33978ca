// LM fit of the 128x128 pupil grid by eight waves per ROI (ROI side 32), in a
// file of its own so that it compiles in parallel with the other kernels.
#include "zola_kernels_impl.hpp"

namespace zola {
template hipError_t launch_fit128o_r<32>(const Model128&, const FitArgs&, hipStream_t);
}  // namespace zola
