// G, P and update kernels with float storage and 1 component.
#include "step_kernels.hpp"

namespace mepgl {

template hipError_t launch_step_nc<float, 1>(const StepArgs& a, int which, hipStream_t st);

}  // namespace mepgl
