// G, P kernels with double storage and 2 components.
#include "step_kernels.hpp"

namespace mepgl {

template hipError_t launch_step_nc<double, 2>(const StepArgs& a, int which, hipStream_t st);

}  // namespace mepgl
