// The persistent per-reactor BDF integrator kernels (one wavefront per
// reactor for nsp <= 63, teams of 2-3 wavefronts up to 191 species).
// Included by the four translation units that instantiate them
// (bdf_{wave,team}_{direct,krylov}.hip), so the variants compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "bdf_launch.h"
#include "cvode_wave.h"

namespace zrk {

static_assert(Cvode<double, TeamWave, true>::kKry == kKryDoubles,
              "the Krylov block size the host reserves in LDS");

// ---------------------------------------------------------------------------
// Persistent per-reactor BDF: each team (one wavefront, or W wavefronts of a
// workgroup) takes reactors from *counter until the batch is exhausted.
// T[n], Y[n*stride] updated in place (and P[n] at constant volume);
// nsteps[n] = plugin cost (CVode one-step calls, negative on failure);
// stats[n*8] (optional): nst nfe nje nsetups nni ncfn netf status.
// ---------------------------------------------------------------------------
template <typename MT, class Team, bool kIter>
__device__ __forceinline__ void bdf_driver(Cvode<MT, Team, kIter>& cv, const DevMech& m,
                                           const BdfOpts& o, int n, double dt_all,
                                           const double* __restrict__ dts,
                                           double* __restrict__ Tio, double* __restrict__ Pio,
                                           double* __restrict__ Yio, int stride,
                                           int* __restrict__ nsteps_out,
                                           int* __restrict__ stats_out,
                                           const int* __restrict__ order,
                                           unsigned int* __restrict__ counter) {
  const int lane = cv.team.tid();
  const int nsp = m.nsp;
  cv.mech = m;
  cv.opt = o;
  cv.has_src = o.dpdt || o.esrc || o.ysrc;
  cv.has_ysrc = o.ysrc != nullptr;
  cv.nlscoef = o.nlscoef;
  cv.hmax_inv = o.hmax_inv;
  for (;;) {
    unsigned int r = 0;
    if (lane == 0) r = atomicAdd(counter, 1u);
    r = (unsigned int)cv.team.from((double)r, 0);
    if (r >= (unsigned int)n) break;
    if (order) r = (unsigned int)order[r];
    const double dt = dts ? dts[r] : dt_all;
    double y0 = 0.0;
    if (lane < nsp) y0 = Yio[(size_t)r * stride + lane];
    if (lane == nsp) y0 = Tio[r] / o.Tref;
    // the manager's solve_temperature (zerork_reactor_manager.cpp:1109-1113):
    // held at its initial value unless always solved for or it moved by more
    // than the threshold in the previous solve; then the state is the nsp
    // mass fractions (the temperature lane is inactive)
    cv.solveT = o.always_T || (o.tdelta && o.tdelta[r] > 0.0);
    cv.nvar = nsp + (cv.solveT ? 1 : 0);
    cv.active = lane < nsp || (lane == nsp && cv.solveT);
    cv.inv_nvar = 1.0 / (double)cv.nvar;
    cv.P = Pio[r];
    cv.v0 = 0.0;
    if (o.constant_volume) {
      // ReactorConstantVolumeCPU::InitializeState: the volume of the initial
      // state (reactor_constant_volume_cpu.cpp:28-29)
      const double mws = cv.team.sum(lane < nsp ? y0 * m.inv_mw[lane] : 0.0);
      cv.v0 = 1.0 / (cv.P * (1.0 / mws) / (kRu * (cv.team.from(y0, nsp) * o.Tref)));
    }
    // source terms (InitializeState, reactor_constant_pressure_cpu.cpp:15-42,
    // reactor_constant_volume_cpu.cpp:15-42): the initial mass enthalpy or
    // internal energy and 1/density
    cv.dpdt = (o.dpdt ? o.dpdt[r] : 0.0);
    cv.esrc = (o.esrc ? o.esrc[r] : 0.0);
    cv.ysrc = (o.ysrc && lane < nsp) ? o.ysrc[(size_t)r * stride + lane] : 0.0;
    cv.e0 = cv.vinit = 0.0;
    if (cv.has_src) {
      const double T0 = cv.team.from(y0, nsp) * o.Tref;
      const double Yj = lane < nsp ? y0 : 0.0;
      const double mws = cv.team.sum(lane < nsp ? y0 * m.inv_mw[lane] : 0.0);
      cv.vinit = 1.0 / (cv.P * (1.0 / mws) / (kRu * T0));
      double e0, c;
      team_mass_energy(cv.team, m, o.constant_volume != 0, false, T0, Yj, e0, c);
      cv.e0 = e0;
    }
    // absolute tolerances, with abstol_dens scaled by 1 / (rho0 / MW_j 1e-3)
    // for the species, rho0 the initial density
    // (ReactorNVectorSerial::GetAbsoluteToleranceCorrection,
    // reactor_nvector_serial.cpp:107-120)
    double atl = o.atol;
    if (o.abstol_dens) {
      const double mws = cv.team.sum(lane < nsp ? y0 * m.inv_mw[lane] : 0.0);
      const double T0 = cv.team.from(y0, nsp) * o.Tref;
      const double rho0 = 1.0 / (1.0 / (cv.P * (1.0 / mws) / (kRu * T0)));
      if (lane < nsp) atl = o.atol * (1.0 / (rho0 * m.inv_mw[lane] * 1.0e-3));
    }
    cv.atolv = cv.active ? atl : 0.0;
    cv.rtol = o.rtol;
    cv.inv_rho_last = 0.0;
#ifdef ZRK_PHASE_CYCLES
    for (int k = 0; k < 12; ++k) cv.cyc[k] = 0;
    const long long tr0 = (long long)__builtin_amdgcn_s_memtime();
#endif
    cv.reinit(0.0, y0);
    // ignition root function (reactor_nvector_serial.cpp:774-784): the
    // crossing of T0 + delta_T, T0 the input temperature (compiled into the
    // kExt variants only)
    const bool ign = Team::kExt && o.ign_dT > 0.0 && cv.solveT;
    double troot_last = 0.0;
    bool irfnd = false;
    // root returns per reactor are bounded (an interpolant of degree <= 5 has
    // at most 5 roots per step; round-off flips of g near a double root must
    // not keep a wavefront here)
    constexpr int kMaxRootReturns = 64;
    int nroots = 0;
    if (ign) {
      cv.rt_Tign = Tio[r] + o.ign_dT;
      cv.root_init();
    }
    int nsteps = 0, num_tries = 0, flag = kCvSuccess;
    int tot[7] = {0, 0, 0, 0, 0, 0, 0};
    double tcurr = 0.0;
    // the step record of a one-reactor batch after each successful CVode
    // call: the state CVode returned (zn[0] at tn, else the interpolant) and
    // CVodeGetDky(t, 1); t_prev follows every call, failed ones included
    // (solver_cvode.cpp:247-265)
    const bool tr = o.trace != nullptr && r == 0;
    int ntr = 0;
    double tprev = 0.0;
    auto record = [&](double tc) {
      if (ntr < o.trace_cap) {
        double* rec = o.trace + (size_t)ntr * (2 * nsp + 5);
        const double yv = (tc == cv.tn) ? cv.zn[0] : cv.dky0(tc);
        const double yd = cv.dky1(tc);
        if (lane == 0) {
          rec[0] = (double)nsteps;
          rec[1] = tc;
          rec[2] = tc - tprev;
        }
        if (lane < cv.nvar) {
          rec[3 + lane] = yv;
          rec[4 + nsp + lane] = yd;
        }
      }
      ++ntr;
      tprev = tc;
    };
    // one CVode call per iteration, CV_ONE_STEP (solver_cvode.cpp:228-266)
    while (tcurr < dt) {
      if (ign && cv.nst > 0) {
        // a later call: after a root return, roots in the rest of the last
        // step, else the step's end tn, which was not returned yet
        if (irfnd) cv.root_check2();
        const double troundoff = 100.0 * DBL_EPSILON * (fabs(cv.tn) + fabs(cv.h));
        if (fabs(cv.tn - cv.rt_tret) > troundoff) {
          nsteps += 1;
          double troot;
          if (nroots < kMaxRootReturns && cv.root_check3(troot)) {
            ++nroots;
            irfnd = true;
            tcurr = cv.rt_tret = troot_last = troot;
            if (tr) record(tcurr);
            if (o.stop_after_ignition) break;
            continue;
          }
          irfnd = false;
          tcurr = cv.rt_tret = cv.tn;
          if (tr) record(tcurr);
          continue;
        }
      }
      flag = cv.one_step(dt, tcurr);
      nsteps += 1;
      if (cv.nst >= o.max_steps) flag = kCvTooMuchWork;
      if (flag < 0) {
        num_tries += 1;
        cv.atolv *= o.retry_atol_adj;
        cv.rtol *= o.retry_rtol_adj;
        tot[0] += cv.nst; tot[1] += cv.nfe; tot[2] += cv.nje; tot[3] += cv.nsetups;
        tot[4] += cv.nni; tot[5] += cv.ncfn; tot[6] += cv.netf;
        tprev = tcurr;
        if (num_tries == o.num_retries) break;
        cv.reinit(tcurr, cv.zn[0]);
        if (ign) {
          cv.root_init();
          irfnd = false;
        }
      } else {
        bool stop = false;
        if (ign) {
          cv.rt_tret = cv.tn;
          double troot;
          if (nroots < kMaxRootReturns && cv.root_check3(troot)) {  // returned before tn
            ++nroots;
            irfnd = true;
            tcurr = cv.rt_tret = troot_last = troot;
            stop = o.stop_after_ignition != 0;
          }
        }
        if (tr) record(tcurr);
        if (stop) break;
      }
    }
    if (tr && lane == 0) *o.trace_n = ntr;
    // failed reactors keep their input state (zerork_reactor_manager.cpp:1123-1139)
    if (flag == kCvSuccess) {
      if (ign && o.ign_time && lane == 0) o.ign_time[r] = troot_last;
      // the state at dt, or at the ignition time when stopped there
      const double yout = cv.dky0(tcurr < dt ? tcurr : dt);
      tot[0] += cv.nst; tot[1] += cv.nfe; tot[2] += cv.nje; tot[3] += cv.nsetups;
      tot[4] += cv.nni; tot[5] += cv.ncfn; tot[6] += cv.netf;
      if (lane < nsp) Yio[(size_t)r * stride + lane] = yout;
      if (lane == nsp) Tio[r] = yout * o.Tref;
      // GetState (reactor_constant_volume_cpu.cpp:44-63,
      // reactor_constant_pressure_cpu.cpp:44-66): constant volume P = Ru T /
      // (v Wmix) + dpdt t; constant pressure P + dpdt t, or with mass sources
      // Ru T / (v Wmix) at the initial v; t = dt (the manager passes the
      // solve's dt, also when the integration stopped at ignition)
      const double tend = dt;
      if (o.constant_volume || cv.has_ysrc) {
        const double mws = cv.team.sum(lane < nsp ? yout * m.inv_mw[lane] : 0.0);
        const double Tend = cv.team.from(yout, nsp) * o.Tref;
        const double v = o.constant_volume ? cv.v0 : cv.vinit;
        const double Pend = kRu * Tend / (v * (1.0 / mws)) + (o.constant_volume ? cv.dpdt * tend : 0.0);
        if (lane == 0) Pio[r] = Pend;
      } else if (cv.has_src) {
        const double Pend = cv.P + cv.dpdt * tend;
        if (lane == 0) Pio[r] = Pend;
      }
    } else {
      nsteps = (nsteps <= 0) ? -1 : -nsteps;
    }
#ifdef ZRK_PHASE_CYCLES
    cv.cyc[4] = (long long)__builtin_amdgcn_s_memtime() - tr0;
#endif
    if (lane == 0) {
      if (nsteps_out) nsteps_out[r] = nsteps;
      if (stats_out) {
#ifdef ZRK_PHASE_CYCLES
        // profiling layout: 20 ints per reactor, cycles / 16 in slots 8..19
#pragma unroll
        for (int k = 0; k < 7; ++k) stats_out[(size_t)r * 20 + k] = tot[k];
        stats_out[(size_t)r * 20 + 7] = flag;
#pragma unroll
        for (int k = 0; k < 12; ++k) stats_out[(size_t)r * 20 + 8 + k] = (int)(cv.cyc[k] >> 4);
#else
#pragma unroll
        for (int k = 0; k < 7; ++k) stats_out[(size_t)r * 8 + k] = tot[k];
        stats_out[(size_t)r * 8 + 7] = flag;
#endif
      }
    }
    cv.team.sync();
  }
}

// One wavefront per reactor (nsp <= 63).  MT = double: plain FP64 method;
// MT = float: FP32 preconditioner.  LDS (bdf_lds_bytes): CvUni, Mlds[N*N] MT
// (N = nsp + 1, padded to 16 bytes), t2l[kNP], then C[nsp+1], G[nsp+1],
// rop[nstep+1] doubles.
template <typename MT, bool kExt, bool kIter>
__device__ __forceinline__ void bdf_wave(DevMech m, BdfOpts o, int n, double dt,
                                         const double* __restrict__ dts, double* __restrict__ Tio,
                                         double* __restrict__ Pio, double* __restrict__ Yio,
                                         int stride, int* __restrict__ nsteps_out,
                                         int* __restrict__ stats_out,
                                         const int* __restrict__ order,
                                         unsigned int* __restrict__ counter,
                                         double* __restrict__ jscratch) {
  extern __shared__ double lds[];
  const int nsp = m.nsp;
  CvWave<MT, kExt, kIter> cv(reinterpret_cast<CvUni*>(lds));
  char* base = reinterpret_cast<char*>(lds) + sizeof(CvUni);
  cv.Mlds = reinterpret_cast<MT*>(base);
  cv.prow = nullptr;
  cv.rvec = nullptr;
  cv.t2l = reinterpret_cast<double*>(base + wave_matrix_bytes(nsp, sizeof(MT)));
  cv.hl = nullptr;
  double* rl = cv.t2l + kNP;
  cv.s = rhs_lds_at(rl, nsp);
  cv.kry = rl + 2 * (nsp + 1) + m.nstep + 1;  // after the RHS arrays (kIter kernels)
  cv.savedJ = reinterpret_cast<MT*>(jscratch + (size_t)blockIdx.x * kNP * kNP);
  bdf_driver(cv, m, o, n, dt, dts, Tio, Pio, Yio, stride, nsteps_out, stats_out, order, counter);
}

// Minimum waves per SIMD each instantiation's register allocation targets
// (build-time tuning knobs).  The FP32 kernel's LDS (N*N floats + the rate
// arrays, about 18 KB at 53 species) allows 2 waves per SIMD, and holding
// registers to that budget is faster than one unconstrained wave; the FP64
// kernel's matrix (about 23 KB) does not leave room for a second wave.
#ifndef ZRK_BDF_F64_WAVES_PER_EU
#define ZRK_BDF_F64_WAVES_PER_EU 1
#endif
#ifndef ZRK_BDF_F32_WAVES_PER_EU
#define ZRK_BDF_F32_WAVES_PER_EU 2
#endif
#define ZRK_BDF_KERNEL_ARGS                                                                     \
  DevMech m, BdfOpts o, int n, double dt, const double* __restrict__ dts,                      \
      double* __restrict__ Tio, double* __restrict__ Pio, double* __restrict__ Yio, int stride, \
      int* __restrict__ nsteps_out, int* __restrict__ stats_out, const int* __restrict__ order, \
      unsigned int* __restrict__ counter, double* __restrict__ jscratch
template <typename MT>
constexpr int bdf_waves_per_eu() {
  return sizeof(MT) == sizeof(float) ? ZRK_BDF_F32_WAVES_PER_EU : ZRK_BDF_F64_WAVES_PER_EU;
}
template <typename MT, bool kExt, bool kIter>
__global__ void __launch_bounds__(kWave, bdf_waves_per_eu<MT>())
bdf_kernel(ZRK_BDF_KERNEL_ARGS) {
  bdf_wave<MT, kExt, kIter>(m, o, n, dt, dts, Tio, Pio, Yio, stride, nsteps_out, stats_out,
                            order, counter, jscratch);
}
#undef ZRK_BDF_KERNEL_ARGS

// W wavefronts per reactor (64 <= nsp < 64W).  LDS (bdf_lds_bytes):
// CvUni[W] (one per wavefront), for an FP32 matrix Mlds[(nsp+1)*(64W+1)]
// float, then prow[64W], rvec[64W] of MT, t2l[64W], hl[64W], red[W+1]
// double, then C[nsp+1], G[nsp+1], rop[nstep+1] doubles.  Slot of
// slot_doubles doubles per workgroup: the saved Jacobian (MT, column stride
// 64W+1), then for an FP64 matrix the Newton matrix and its inverse (same
// layout).
template <typename MT, int W, bool kExt, bool kIter>
__global__ void __launch_bounds__(kWave * W)
bdf_team_kernel(DevMech m, BdfOpts o, int n, double dt, const double* __restrict__ dts,
                double* __restrict__ Tio,
                double* __restrict__ Pio, double* __restrict__ Yio, int stride,
                int* __restrict__ nsteps_out, int* __restrict__ stats_out,
                const int* __restrict__ order, unsigned int* __restrict__ counter,
                double* __restrict__ jscratch, size_t slot_doubles) {
  extern __shared__ double lds[];
  constexpr int L = kWave * W;
  using CV = Cvode<MT, TeamBlock<W, kExt>, kIter>;
  const int nsp = m.nsp;
  CV cv(reinterpret_cast<CvUni*>(lds) + threadIdx.x / kWave);
  char* p = reinterpret_cast<char*>(lds) + W * sizeof(CvUni);
  MT* slot = reinterpret_cast<MT*>(jscratch + (size_t)blockIdx.x * slot_doubles);
  if constexpr (std::is_same<MT, float>::value) {
    cv.Mlds = reinterpret_cast<MT*>(p);
    p += ((size_t)(nsp + 1) * CV::kJS * sizeof(float) + 15) & ~(size_t)15;
  } else {
    cv.Mlds = slot + (size_t)(nsp + 1) * CV::kJS;
  }
  cv.prow = reinterpret_cast<MT*>(p);
  cv.rvec = cv.prow + L;
  cv.t2l = reinterpret_cast<double*>(cv.rvec + L);
  cv.hl = cv.t2l + L;
  double* red = cv.hl + L;
  cv.team.red = red;
  double* rl = red + ((W + 2) & ~1);
  cv.s = rhs_lds_at(rl, nsp);
  cv.kry = rl + 2 * (nsp + 1) + m.nstep + 1;  // after the RHS arrays (kIter kernels)
  cv.savedJ = slot;
  bdf_driver(cv, m, o, n, dt, dts, Tio, Pio, Yio, stride, nsteps_out, stats_out, order, counter);
}

// ---------------------------------------------------------------------------
// host side: launch / occupancy of one family of instantiations
// ---------------------------------------------------------------------------
template <class F>
hipError_t with_ext(bool ext, F&& f) {
  if (ext) return f(std::true_type{});
  return f(std::false_type{});
}

template <bool kIter>
hipError_t launch_bdf_wave(const BdfLaunch& a, bool ext, bool f32) {
  return with_ext(ext, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    if (f32)
      hipLaunchKernelGGL((bdf_kernel<float, E, kIter>), dim3(a.grid), dim3(kWave), a.lds, a.st,
                         a.m, a.o, a.n, a.dt, a.dts, a.T, a.P, a.Y, a.stride, a.nsteps, a.stats,
                         a.order, a.counter, a.jscratch);
    else
      hipLaunchKernelGGL((bdf_kernel<double, E, kIter>), dim3(a.grid), dim3(kWave), a.lds, a.st,
                         a.m, a.o, a.n, a.dt, a.dts, a.T, a.P, a.Y, a.stride, a.nsteps, a.stats,
                         a.order, a.counter, a.jscratch);
    return hipGetLastError();
  });
}

template <typename MT, bool kIter>
hipError_t launch_bdf_team_mt(const BdfLaunch& a, bool ext, int W, size_t slot) {
  return with_ext(ext, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    if (W == 2)
      hipLaunchKernelGGL((bdf_team_kernel<MT, 2, E, kIter>), dim3(a.grid), dim3(2 * kWave),
                         a.lds, a.st, a.m, a.o, a.n, a.dt, a.dts, a.T, a.P, a.Y, a.stride,
                         a.nsteps, a.stats, a.order, a.counter, a.jscratch, slot);
    else
      hipLaunchKernelGGL((bdf_team_kernel<MT, 3, E, kIter>), dim3(a.grid), dim3(3 * kWave),
                         a.lds, a.st, a.m, a.o, a.n, a.dt, a.dts, a.T, a.P, a.Y, a.stride,
                         a.nsteps, a.stats, a.order, a.counter, a.jscratch, slot);
    return hipGetLastError();
  });
}

template <bool kIter>
hipError_t launch_bdf_team(const BdfLaunch& a, bool ext, int W, size_t slot, bool f32) {
  return f32 ? launch_bdf_team_mt<float, kIter>(a, ext, W, slot)
             : launch_bdf_team_mt<double, kIter>(a, ext, W, slot);
}

template <bool kIter>
int resident_bdf_wave(bool ext, bool f32, size_t lds) {
  int r = 0;
  (void)with_ext(ext, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    r = f32 ? resident_blocks(bdf_kernel<float, E, kIter>, kWave, lds)
            : resident_blocks(bdf_kernel<double, E, kIter>, kWave, lds);
    return hipSuccess;
  });
  return r;
}

template <typename MT, bool kIter>
int resident_bdf_team_mt(bool ext, int W, size_t lds) {
  int r = 0;
  (void)with_ext(ext, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    r = W == 2 ? resident_blocks(bdf_team_kernel<MT, 2, E, kIter>, 2 * kWave, lds)
               : resident_blocks(bdf_team_kernel<MT, 3, E, kIter>, 3 * kWave, lds);
    return hipSuccess;
  });
  return r;
}

template <bool kIter>
int resident_bdf_team(bool ext, int W, size_t lds, bool f32) {
  return f32 ? resident_bdf_team_mt<float, kIter>(ext, W, lds)
             : resident_bdf_team_mt<double, kIter>(ext, W, lds);
}

}  // namespace zrk
