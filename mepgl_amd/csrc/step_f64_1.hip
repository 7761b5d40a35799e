// G, P and update kernels with double storage and 1 component.
#include "step_kernels.hpp"

namespace mepgl {

template hipError_t launch_step_nc<double, 1>(const StepArgs& a, int which, hipStream_t st);

}  // namespace mepgl
