// G, P kernels with float storage and 2 components.
#include "step_kernels.hpp"

namespace mepgl {

template hipError_t launch_step_nc<float, 2>(const StepArgs& a, int which, hipStream_t st);

}  // namespace mepgl
