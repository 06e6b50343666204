// Integrator instantiations: one wavefront per reactor (FP64 and FP32 Newton matrices), with the Krylov (SPGMR) linear solver.
#include "bdf_kernels.h"

namespace zrk {
template hipError_t launch_bdf_wave<true>(const BdfLaunch&, bool, bool);
template int resident_bdf_wave<true>(bool, bool, size_t);
}  // namespace zrk
