// Host-side launchers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "dev_mech.h"

namespace zrk {

// Mechanisms whose per-reactor rate arrays exceed a wavefront's LDS share
// (rates_need_block) run one reactor per workgroup on a persistent grid of
// rates_block_grid workgroups, each with nstep + 1 doubles of scratch.
bool rates_need_block(const DevMech& m);
int rates_block_grid(const DevMech& m, int n);
hipError_t launch_rates(const DevMech& m, int n, const double* T, const double* P,
                        const double* Y, int stride, double L, double* ydot, double* rop,
                        double* create, double* destroy, double* net, hipStream_t st,
                        double* scratch = nullptr, int grid = 0);
hipError_t launch_jacobian(const DevMech& m, const BdfOpts& o, int n, const double* T,
                           const double* P, const double* Y, int stride, double* J,
                           double* jscratch, int nblocks, hipStream_t st);
// Smallest number of wavefronts per reactor of the integrator: 1 (nsp <= 63),
// 2 or 3 (nsp <= 191), 0 if the mechanism is too large.
int bdf_team_width(int nsp);
// whether a team of W wavefronts can integrate this mechanism
bool bdf_team_ok(int nsp, bool f32, int W);
// doubles of device workspace per resident team of W wavefronts (the saved
// Jacobian; for FP64 teams also the Newton matrix)
size_t bdf_slot_doubles(const DevMech& m, int W, bool f32);
// iter: the Krylov solver's kernels (their LDS holds its uniform block too)
size_t bdf_lds_bytes(const DevMech& m, bool f32, int W, bool iter = false);
// teams the persistent integrator keeps resident (0: does not fit)
int bdf_max_resident_teams(const DevMech& m, bool f32, int W, bool iter = false);
hipError_t launch_bdf(const DevMech& m, const BdfOpts& o, int n, double dt, const double* dts,
                      double* T, double* P, double* Y, int stride, int* nsteps, int* stats,
                      const int* order, unsigned int* counter, double* jscratch, int nwaves,
                      int W, hipStream_t st);

// sparse path (bdf_sparse.hip): one reactor per workgroup of kSparseWaves
// wavefronts, any mechanism size.  Four wavefronts with four workgroups per
// CU (128 VGPRs) gave the 2,000-species workload 123 reactors/s, against 98
// for two workgroups of eight wavefronts, 89 for eight of two and 60 for one
// workgroup of eight: its BDF steps are latency chains, so more reactors in
// flight per CU win over more threads per reactor
constexpr int kSparseWaves = 4;
struct DevSparse;
size_t sparse_slot_bytes(const DevMech& m, const DevSparse& S);
size_t sparse_lds_bytes(const DevMech& m);
int sparse_max_resident(const DevMech& m);
hipError_t launch_bdf_sparse(const DevMech& m, const DevSparse& S, const BdfOpts& o, int n,
                             double dt, const double* dts, double* T, double* P, double* Y,
                             int stride, int* nsteps, int* stats, const int* order,
                             unsigned int* counter, double* ws, int nblocks, hipStream_t st);

// doubles of per-team LDS used by the RHS (C, G, rop with its zero slot)
size_t rhs_lds_bytes(const DevMech& m);

// learned surrogate (mlp.hip)
hipError_t launch_mlp_features(int n, int nsp, const double* T, const double* P, const double* Y,
                               int stride, const double* dt, double dt_all, const float* mu,
                               const float* sig, unsigned short* X, int ldx, int nfeat,
                               hipStream_t st);
hipError_t launch_gemm_bf16(int M, int N, int K, const unsigned short* A, long sAm, long sAk,
                            const unsigned short* B, long sBk, long sBn, float* C, long sCm,
                            long sCn, int accumulate, hipStream_t st);
hipError_t launch_bias_act(int M, int N, const float* C, const float* b, float* Z,
                           unsigned short* Yb, float* Yf, int act, hipStream_t st);
hipError_t launch_act_backward(int M, int N, const float* dY, const float* Z, float* dZ,
                               unsigned short* dZb, int act, hipStream_t st);
hipError_t launch_col_sum(int M, int N, const float* dZ, float* db, hipStream_t st);
hipError_t launch_adam(long n, float* w, const float* g, float* m, float* v, unsigned short* wb,
                       float lr, float b1, float b2, float eps, float c1, float c2,
                       hipStream_t st);
hipError_t launch_to_bf16(long n, const float* x, unsigned short* y, hipStream_t st);
hipError_t launch_accept(const DevMech& m, const BdfOpts& o, int n, const double* dts,
                         double dt_all, const double* T0, const double* P, const double* Y0,
                         int nodes, const double* tau, double* Tp, double* Yp, int stride,
                         int max_newton, int* accept, double* res, double* jscratch,
                         int nblocks, hipStream_t st);

}  // namespace zrk
