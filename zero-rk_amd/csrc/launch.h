// Host-side launchers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "dev_mech.h"

namespace zrk {

hipError_t launch_rates(const DevMech& m, int n, const double* T, const double* P,
                        const double* Y, int stride, double L, double* ydot, double* rop,
                        double* create, double* destroy, double* net, hipStream_t st);
hipError_t launch_jacobian(const DevMech& m, const BdfOpts& o, int n, const double* T,
                           const double* P, const double* Y, int stride, double* J,
                           double* jscratch, int nblocks, hipStream_t st);
// Wavefronts per reactor of the integrator: 1 (nsp <= 63), 2 or 3 (FP32
// preconditioner only), 0 if the mechanism is too large.
int bdf_team_width(int nsp);
// doubles of saved-Jacobian workspace per resident team
size_t bdf_slot_doubles(const DevMech& m);
size_t bdf_lds_bytes(const DevMech& m, bool f32);
// teams the persistent integrator keeps resident (0: does not fit)
int bdf_max_resident_teams(const DevMech& m, bool f32);
hipError_t launch_bdf(const DevMech& m, const BdfOpts& o, int n, double dt, double* T,
                      double* P, double* Y, int stride, int* nsteps, int* stats,
                      const int* order, unsigned int* counter, double* jscratch, int nwaves,
                      hipStream_t st);

}  // namespace zrk
