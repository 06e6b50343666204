// Integrator instantiations: teams of 2 and 3 wavefronts (FP32 and FP64 Newton matrices), with the Krylov (SPGMR) linear solver.
#include "bdf_kernels.h"

namespace zrk {
template hipError_t launch_bdf_team<true>(const BdfLaunch&, bool, int, size_t, bool);
template int resident_bdf_team<true>(bool, int, size_t, bool);
}  // namespace zrk
