// fwe_step.hip -- one instantiation unit of the stencil kernels: the
// launch templates of fwe_tiles.h for one law (FWE_MODE) and one damping
// form (FWE_PML), every tile height and wavefield count.  The build compiles
// this file once per supported (law, damping) pair, in parallel.
#include "fwe_tiles.h"

#if !defined(FWE_MODE) || !defined(FWE_PML)
#error "compile with -DFWE_MODE=<law> -DFWE_PML=<damping form>"
#endif

namespace fwe {
template hipError_t stress_mp<FWE_MODE, FWE_PML>(const StepArgs&, int, hipStream_t);
template hipError_t velocity_mp<FWE_MODE, FWE_PML>(const StepArgs&, hipStream_t);
#if FWE_MODE != 2 && FWE_MODE != 4 && FWE_PML != 4
// groups read the medium arrays (no material table) and take no
// convolutional PML
template hipError_t group_mp<FWE_MODE, FWE_PML>(const StepArgs*, const int32_t*, int, int, int, int, int, bool,
                                                hipStream_t);
#endif
}  // namespace fwe
