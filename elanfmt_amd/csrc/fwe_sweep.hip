// fwe_sweep.hip -- one instantiation unit of the single-pass sweep kernel
// (fwe_sweep.h) for one law (FWE_MODE) and one damping form (FWE_PML), both
// row counts and both top boundaries.  Compiled once per supported pair.
#include "fwe_sweep.h"

#if !defined(FWE_MODE) || !defined(FWE_PML)
#error "compile with -DFWE_MODE=<law> -DFWE_PML=<damping form>"
#endif

namespace fwe {
template hipError_t sweep_mp<FWE_MODE, FWE_PML>(const SweepArgs&, int, hipStream_t);
}  // namespace fwe
