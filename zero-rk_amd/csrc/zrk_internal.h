// Internal handle types shared by the C-ABI layers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "../../include/zrk_device.h"
#include "ck_mech.h"
#include "dev_mech.h"
#include "sparse_plan.h"

namespace zrk {

struct MechHandle {
  Mechanism host;
  DevMech dev{};
  DevMech* dev_self = nullptr;  // device copy of dev (dev.self points here)
  void* dev_blob = nullptr;
  size_t dev_bytes = 0;
  int device = 0;
  // persistent grid size in teams, by [FP32 preconditioner][Krylov solver]
  // [wavefronts per team]
  int resident[2][2][4] = {};
  // wavefronts per team chosen by the autotuner for [FP32 preconditioner] (0: smallest)
  int team_pref[2] = {0, 0};
  // integrator workspace
  double* ws_jac = nullptr;            // saved-Jacobian slots, one per resident team
  unsigned int* ws_counter = nullptr;
  size_t ws_doubles = 0;
  // sparse path (mechanisms above 191 species, or forced): plan, its device
  // copy and the resident workgroups
  SparsePlan* plan = nullptr;
  DevSparse sp_dev{};
  void* sp_blob = nullptr;
  int sp_resident = 0;
  int force_sparse = 0;
  ~MechHandle();
};

int build_device_mech(MechHandle* h, std::string* err);
// team width the integrator uses for these options (0 if none fits)
int team_width_for(MechHandle* h, bool f32, bool iter = false);
int ensure_device(MechHandle* h);
int ensure_workspace(MechHandle* h, size_t doubles);
// builds and uploads the sparse plan once (0 on success)
int ensure_sparse(MechHandle* h);
BdfOpts to_bdf_opts(const zrk_bdf_options& o);
// Integrator scratch supplied by the caller instead of the handle's own
// allocations: `doubles` doubles of per-team slots and a 256-byte work counter.
// With it integrate_device allocates nothing: the device image must already
// exist (-6 otherwise), the dense paths only (-4 for the sparse path), and the
// persistent grid shrinks to the slots that fit (-7 if not even one does).
struct ExtWorkspace {
  double* jac;
  size_t doubles;
  unsigned int* counter;
};
// Step records of a one-reactor batch (the plugin's step callback): a device
// buffer of cap records of 2 nsp + 5 doubles and a device int that receives
// the number of successful CVode calls (see BdfOpts::trace).
struct StepTrace {
  double* buf;
  int cap;
  int* count;
};
// dpdt[n], esrc[n], ysrc[n*stride]: optional device arrays of source terms;
// ign_time[n]: optional ignition-time output (with o.delta_temperature_ignition > 0)
// tdelta[n] (optional, with always_T = 0): the manager's temperature change
// of each reactor's previous solve; a reactor whose entry is not above 0 is
// integrated at fixed temperature (zerork_reactor_manager.cpp:1109-1113)
// ws (optional): caller-supplied scratch, see ExtWorkspace
// trace (optional, n = 1 only): step records, see StepTrace
int integrate_device(MechHandle* h, const zrk_bdf_options& o, int n, double dt,
                     const double* dts, double* T, double* P, double* Y, int stride, int* nsteps, int* stats,
                     const int* order, hipStream_t st, const double* dpdt = nullptr,
                     const double* esrc = nullptr, const double* ysrc = nullptr,
                     double* ign_time = nullptr, const double* tdelta = nullptr,
                     int always_T = 1, const ExtWorkspace* ws = nullptr,
                     const StepTrace* trace = nullptr);

}  // namespace zrk
