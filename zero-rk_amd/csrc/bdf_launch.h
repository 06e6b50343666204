// Host interface of the integrator kernels (bdf_kernels.h): one launch
// description and the per-family launch / occupancy entry points, whose
// explicit instantiations live in bdf_{wave,team}_{direct,krylov}.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_mech.h"

namespace zrk {

// doubles of the Krylov solver's uniform block (Givens rotations, triangular
// factor, least-squares right side) after the RHS arrays of a team's LDS
constexpr size_t kKryDoubles = 46;

// LDS bytes of the one-wavefront integrator's (nsp+1) x (nsp+1) matrix of
// element size mt, padded to 16 bytes.
__host__ __device__ constexpr size_t wave_matrix_bytes(int nsp, size_t mt) {
  return ((size_t)(nsp + 1) * (size_t)(nsp + 1) * mt + 15) & ~(size_t)15;
}

struct BdfLaunch {
  const DevMech& m;
  const BdfOpts& o;
  int n;
  double dt;
  const double* dts;
  double *T, *P, *Y;
  int stride;
  int *nsteps, *stats;
  const int* order;
  unsigned int* counter;
  double* jscratch;
  int grid;
  size_t lds;
  hipStream_t st;
};

template <bool kIter>
hipError_t launch_bdf_wave(const BdfLaunch& a, bool ext, bool f32);
template <bool kIter>
hipError_t launch_bdf_team(const BdfLaunch& a, bool ext, int W, size_t slot, bool f32);
template <bool kIter>
int resident_bdf_wave(bool ext, bool f32, size_t lds);
template <bool kIter>
int resident_bdf_team(bool ext, int W, size_t lds, bool f32);

inline size_t max_lds_per_block() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 0;
  return (size_t)v;
}

template <class K>
inline int resident_blocks(K kernel, int threads, size_t lds) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (lds > max_lds_per_block()) return 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess)
    return 0;
  return ncu * (per_cu > 0 ? per_cu : 1);
}


}  // namespace zrk
