// Integrator instantiations: one wavefront per reactor (FP64 and FP32 Newton matrices), with the direct linear solver.
#include "bdf_kernels.h"

namespace zrk {
template hipError_t launch_bdf_wave<false>(const BdfLaunch&, bool, bool);
template int resident_bdf_wave<false>(bool, bool, size_t);
}  // namespace zrk
