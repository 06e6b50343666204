// Internal handle types shared by the C-ABI layers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "../../include/zrk_device.h"
#include "ck_mech.h"
#include "dev_mech.h"

namespace zrk {

struct MechHandle {
  Mechanism host;
  DevMech dev{};
  DevMech* dev_self = nullptr;  // device copy of dev (dev.self points here)
  void* dev_blob = nullptr;
  size_t dev_bytes = 0;
  int device = 0;
  int resident_waves[2] = {0, 0};  // persistent grid size: FP64, FP32 preconditioner
  // integrator workspace
  double* ws_jac = nullptr;            // saved-Jacobian slots, one per resident team
  unsigned int* ws_counter = nullptr;
  size_t ws_doubles = 0;
  ~MechHandle();
};

int build_device_mech(MechHandle* h, std::string* err);
int ensure_device(MechHandle* h);
int ensure_workspace(MechHandle* h, size_t doubles);
BdfOpts to_bdf_opts(const zrk_bdf_options& o);
int integrate_device(MechHandle* h, const zrk_bdf_options& o, int n, double dt,
                     const double* dts, double* T, double* P, double* Y, int stride, int* nsteps, int* stats,
                     const int* order, hipStream_t st);

}  // namespace zrk
