// gfx950 kernels of the batched reactor-chemistry path and their launchers.
//
//   rates_kernel     FP64 rates of progress + species production + const-P RHS
//                    for a batch of reactor states (any mechanism size):
//                    one wavefront per reactor, 4 per workgroup.
//   jacobian_kernel  reference analytic Jacobian for nsp <= 63 (parity tests).
//   bdf_kernel       persistent per-reactor BDF integrator (nsp <= 63): each
//                    wavefront takes reactors from a device work counter until
//                    the batch is exhausted.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "cvode_wave.h"
#include "launch.h"

namespace zrk {

constexpr int kRatesWaves = 4;  // wavefronts per workgroup in rates_kernel

// LDS bytes of the one-wavefront integrator's (nsp+1) x (nsp+1) matrix of
// element size mt, padded to 16 bytes.
__host__ __device__ constexpr size_t wave_matrix_bytes(int nsp, size_t mt) {
  return ((size_t)(nsp + 1) * (size_t)(nsp + 1) * mt + 15) & ~(size_t)15;
}

// ---------------------------------------------------------------------------
// rates_kernel: inputs T[n], P[n], Y[n*stride] (reactor-major AoS, as the
// plugin API); outputs (each optional): ydot[n*(nsp+1)] ([Y.., T] order),
// rop[n*nstep], create/destroy/net[n*nsp].
// ---------------------------------------------------------------------------
template <bool kExt>
__global__ void __launch_bounds__(kWave* kRatesWaves)
rates_kernel(DevMech m, int n, const double* __restrict__ Tin, const double* __restrict__ Pin,
             const double* __restrict__ Yin, int stride, double L, double* __restrict__ ydot,
             double* __restrict__ rop_out, double* __restrict__ create_out,
             double* __restrict__ destroy_out, double* __restrict__ net_out) {
  extern __shared__ double lds[];
  const int wave = threadIdx.x / kWave;
  const int lane = lane_id();
  const int nsp = m.nsp, nstep = m.nstep;
  const int per_wave = 3 * nsp + 2 + nstep + 1;
  double* C = lds + wave * per_wave;
  double* G = C + nsp + 1;
  double* H = G + nsp + 1;
  double* rop = H + nsp;
  if (lane == 0) {  // entries read by absent reaction participants
    C[nsp] = 1.0;
    G[nsp] = 0.0;
  }
  for (int r = blockIdx.x * kRatesWaves + wave; r < n; r += gridDim.x * kRatesWaves) {
    const double* Y = Yin + (size_t)r * stride;
    const double T = fmin(Tin[r], kTlimit);
    const double P = Pin[r];
    const TPow p = make_tpow(T);
    double part = 0.0;
    for (int j = lane; j < nsp; j += kWave) part += Y[j] * m.inv_mw[j];
    const double wmix = 1.0 / wave_sum(part);
    const double inv_rho = 1.0 / (P * wmix / (kRu * T));
    const double rho_c = 1.0 / inv_rho;
    double cps = 0.0, cs = 0.0;
    for (int j = lane; j < nsp; j += kWave) {
      double cp, h, g;
      thermo_one(m.thermo + 16 * j, p, cp, h, g);
      const double Cj = rho_c * m.inv_mw[j] * Y[j];
      C[j] = Cj;
      G[j] = g;
      H[j] = h;
      cps += cp * m.ru_mw[j] * Y[j];
      cs += Cj;
    }
    const double cp_mix = wave_sum(cps);
    const double Csum = wave_sum(cs);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    team_rates<kWave, kExt>(lane, m, C, G, rop, p, Csum, L);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double e = 0.0;
    for (int j = lane; j < nsp; j += kWave) {
      double cr, de;
      species_cd<kExt>(m, rop, j, cr, de);
      const double net = cr - de;
      if (create_out) create_out[(size_t)r * nsp + j] = cr;
      if (destroy_out) destroy_out[(size_t)r * nsp + j] = de;
      if (net_out) net_out[(size_t)r * nsp + j] = net;
      if (ydot) ydot[(size_t)r * (nsp + 1) + j] = (net * m.mw[j]) * inv_rho;
      e += H[j] * net * inv_rho;
    }
    const double esum = wave_sum(e);
    if (ydot && lane == 0)
      ydot[(size_t)r * (nsp + 1) + nsp] = -esum * kRu * Tin[r] / cp_mix + 0.0 / cp_mix;
    if (rop_out)
      for (int s = lane; s < nstep; s += kWave) rop_out[(size_t)r * nstep + s] = rop[s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// ---------------------------------------------------------------------------
// jacobian_kernel (nsp <= 63): J[n][N][N] row-major, N = nsp+1, T last.
// ---------------------------------------------------------------------------
template <bool kExt>
__global__ void __launch_bounds__(kWave)
jacobian_kernel(DevMech m, BdfOpts o, int n, const double* __restrict__ Tin,
                const double* __restrict__ Pin, const double* __restrict__ Yin, int stride,
                double* __restrict__ Jout, double* __restrict__ jscratch) {
  extern __shared__ double lds[];
  const int lane = lane_id();
  const int nsp = m.nsp, N = nsp + 1;
  double* Jl = lds;
  double* t2l = lds + kNP * kNP;
  double* rl = t2l + kNP;
  RhsLds s = rhs_lds_at(rl, nsp);
  double* saved = jscratch + (size_t)blockIdx.x * kNP * kNP;
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    double y = 0.0;
    if (lane < nsp) y = Yin[(size_t)r * stride + lane];
    if (lane == nsp) y = Tin[r] / o.Tref;
    const double P = Pin[r];
    double fy;
    RhsAux aux;
    rhs64<kExt>(m, s, P, 0.0, o.Tref, o.step_limiter, y, fy, aux);
    jacobian_lds<double, kExt>(m, s, P, 0.0, o, y, fy, aux.inv_rho, Jl, kNP, saved, t2l);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane <= nsp)
      for (int c = 0; c < N; ++c) Jout[(size_t)r * N * N + (size_t)lane * N + c] = Jl[c * kNP + lane];
    __builtin_amdgcn_wave_barrier();
  }
}

// jacobian_team_kernel (64 <= nsp < 64W): the same Jacobian computed by a team
// of W wavefronts (the integrator's team code, FP64), through a global
// scratch J (column stride N) per workgroup.  LDS: t2l[64W], hl[64W],
// red[W+1] doubles, then C, G, rop.
template <int W, bool kExt>
__global__ void __launch_bounds__(kWave * W)
jacobian_team_kernel(DevMech m, BdfOpts o, int n, const double* __restrict__ Tin,
                     const double* __restrict__ Pin, const double* __restrict__ Yin, int stride,
                     double* __restrict__ Jout, double* __restrict__ jscratch) {
  extern __shared__ double lds[];
  constexpr int L = kWave * W;
  const int nsp = m.nsp, N = nsp + 1;
  TeamBlock<W, kExt> team;
  double* t2l = lds;
  double* hl = t2l + L;
  team.red = hl + L;
  double* rl = team.red + ((W + 2) & ~1);
  RhsLds s = rhs_lds_at(rl, nsp);
  double* J = jscratch + (size_t)blockIdx.x * 2 * N * N;
  double* saved = J + (size_t)N * N;
  const int tid = team.tid();
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    double y = 0.0;
    if (tid < nsp) y = Yin[(size_t)r * stride + tid];
    if (tid == nsp) y = Tin[r] / o.Tref;
    const double P = Pin[r];
    double fy;
    RhsAux aux;
    rhs_team(team, m, s, P, 0.0, o.Tref, o.step_limiter, y, fy, aux);
    jacobian_team<double>(team, m, s, P, 0.0, o, y, fy, aux.inv_rho, J, saved, N, t2l, hl);
    if (tid < N)
      for (int c = 0; c < N; ++c) Jout[(size_t)r * N * N + (size_t)tid * N + c] = J[c * N + tid];
    team.sync();
  }
}

// ---------------------------------------------------------------------------
// Persistent per-reactor BDF: each team (one wavefront, or W wavefronts of a
// workgroup) takes reactors from *counter until the batch is exhausted.
// T[n], Y[n*stride] updated in place (and P[n] at constant volume);
// nsteps[n] = plugin cost (CVode one-step calls, negative on failure);
// stats[n*8] (optional): nst nfe nje nsetups nni ncfn netf status.
// ---------------------------------------------------------------------------
template <typename MT, class Team>
__device__ __forceinline__ void bdf_driver(Cvode<MT, Team>& cv, const DevMech& m,
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
  cv.active = lane <= nsp;
  cv.has_src = o.dpdt || o.esrc || o.ysrc;
  cv.has_ysrc = o.ysrc != nullptr;
  cv.inv_nvar = 1.0 / (double)(nsp + 1);
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
    cv.atolv = cv.active ? o.atol : 0.0;
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
    const bool ign = Team::kExt && o.ign_dT > 0.0;
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
            if (o.stop_after_ignition) break;
            continue;
          }
          irfnd = false;
          tcurr = cv.rt_tret = cv.tn;
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
        if (num_tries == o.num_retries) break;
        cv.reinit(tcurr, cv.zn[0]);
        if (ign) {
          cv.root_init();
          irfnd = false;
        }
      } else if (ign) {
        cv.rt_tret = cv.tn;
        double troot;
        if (nroots < kMaxRootReturns && cv.root_check3(troot)) {  // returned before tn
          ++nroots;
          irfnd = true;
          tcurr = cv.rt_tret = troot_last = troot;
          if (o.stop_after_ignition) break;
        }
      }
    }
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
template <typename MT, bool kExt>
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
  CvWave<MT, kExt> cv(reinterpret_cast<CvUni*>(lds));
  char* base = reinterpret_cast<char*>(lds) + sizeof(CvUni);
  cv.Mlds = reinterpret_cast<MT*>(base);
  cv.prow = nullptr;
  cv.rvec = nullptr;
  cv.t2l = reinterpret_cast<double*>(base + wave_matrix_bytes(nsp, sizeof(MT)));
  cv.hl = nullptr;
  double* rl = cv.t2l + kNP;
  cv.s = rhs_lds_at(rl, nsp);
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
template <typename MT, bool kExt>
__global__ void __launch_bounds__(kWave, bdf_waves_per_eu<MT>())
bdf_kernel(ZRK_BDF_KERNEL_ARGS) {
  bdf_wave<MT, kExt>(m, o, n, dt, dts, Tio, Pio, Yio, stride, nsteps_out, stats_out, order,
                     counter, jscratch);
}
#undef ZRK_BDF_KERNEL_ARGS

// W wavefronts per reactor (64 <= nsp < 64W), FP32 preconditioner.  LDS
// (bdf_team_lds_bytes): CvUni[W] (one per wavefront), Mlds[(nsp+1)*(64W+1)]
// float, prow[64W], rvec[64W] float, t2l[64W], hl[64W], red[W+1] double,
// then C[nsp+1], G[nsp+1], rop[nstep+1] doubles.  savedJ slot: slot_doubles
// doubles per workgroup.
template <int W, bool kExt>
__global__ void __launch_bounds__(kWave * W)
bdf_team_kernel(DevMech m, BdfOpts o, int n, double dt, const double* __restrict__ dts,
                double* __restrict__ Tio,
                double* __restrict__ Pio, double* __restrict__ Yio, int stride,
                int* __restrict__ nsteps_out, int* __restrict__ stats_out,
                const int* __restrict__ order, unsigned int* __restrict__ counter,
                double* __restrict__ jscratch, size_t slot_doubles) {
  extern __shared__ double lds[];
  constexpr int L = kWave * W;
  using CV = Cvode<float, TeamBlock<W, kExt>>;
  const int nsp = m.nsp;
  CV cv(reinterpret_cast<CvUni*>(lds) + threadIdx.x / kWave);
  char* p = reinterpret_cast<char*>(lds) + W * sizeof(CvUni);
  cv.Mlds = reinterpret_cast<float*>(p);
  p += ((size_t)(nsp + 1) * CV::kJS * sizeof(float) + 15) & ~(size_t)15;
  cv.prow = reinterpret_cast<float*>(p);
  cv.rvec = cv.prow + L;
  cv.t2l = reinterpret_cast<double*>(cv.rvec + L);
  cv.hl = cv.t2l + L;
  double* red = cv.hl + L;
  cv.team.red = red;
  double* rl = red + ((W + 2) & ~1);
  cv.s = rhs_lds_at(rl, nsp);
  cv.savedJ = reinterpret_cast<float*>(jscratch + (size_t)blockIdx.x * slot_doubles);
  bdf_driver(cv, m, o, n, dt, dts, Tio, Pio, Yio, stride, nsteps_out, stats_out, order, counter);
}

__global__ void fill_u32_kernel(unsigned int* p, unsigned int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// launchers.  The kExt variants (with the general record evaluation and, in
// the integrator, the ignition root finding) are launched only for
// mechanisms that have PLOG, non-integer or SRI records, or integrations
// that ask for ignition times.
// ---------------------------------------------------------------------------
namespace {
template <class F>
hipError_t with_ext(bool ext, F&& f) {
  if (ext) return f(std::true_type{});
  return f(std::false_type{});
}
}  // namespace

hipError_t launch_rates(const DevMech& m, int n, const double* T, const double* P,
                        const double* Y, int stride, double L, double* ydot, double* rop,
                        double* create, double* destroy, double* net, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)kRatesWaves * (3 * m.nsp + 2 + m.nstep + 1);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  int blocks = (n + kRatesWaves - 1) / kRatesWaves;
  return with_ext(m.n_ext > 0, [&](auto e) {
    hipLaunchKernelGGL((rates_kernel<decltype(e)::value>), dim3(blocks),
                       dim3(kWave * kRatesWaves), lds, st, m, n, T, P, Y, stride, L, ydot, rop,
                       create, destroy, net);
    return hipGetLastError();
  });
}

size_t rhs_lds_bytes(const DevMech& m) {
  return sizeof(double) * (size_t)(2 * (m.nsp + 1) + m.nstep + 1);
}
hipError_t launch_jacobian(const DevMech& m, const BdfOpts& o, int n, const double* T,
                           const double* P, const double* Y, int stride, double* J,
                           double* jscratch, int nblocks, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int grid = nblocks < n ? nblocks : n;
  const int W = bdf_team_width(m.nsp);
  if (W == 0) return hipErrorInvalidValue;
  return with_ext(m.n_ext > 0, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    if (W > 1) {
      const size_t L = (size_t)kWave * W;
      const size_t lds = sizeof(double) * (2 * L + ((W + 2) & ~1)) + rhs_lds_bytes(m);
      if (W == 2)
        hipLaunchKernelGGL((jacobian_team_kernel<2, E>), dim3(grid), dim3(2 * kWave), lds, st, m,
                           o, n, T, P, Y, stride, J, jscratch);
      else
        hipLaunchKernelGGL((jacobian_team_kernel<3, E>), dim3(grid), dim3(3 * kWave), lds, st, m,
                           o, n, T, P, Y, stride, J, jscratch);
      return hipGetLastError();
    }
    const size_t lds = sizeof(double) * (size_t)(kNP * kNP + kNP) + rhs_lds_bytes(m);
    hipLaunchKernelGGL((jacobian_kernel<E>), dim3(grid), dim3(kWave), lds, st, m, o, n, T, P, Y,
                       stride, J, jscratch);
    return hipGetLastError();
  });
}

int bdf_team_width(int nsp) {
  if (nsp <= kNP - 1) return 1;
  if (nsp + 1 <= 2 * kWave) return 2;
  if (nsp + 1 <= 3 * kWave) return 3;
  return 0;
}

bool bdf_team_ok(int nsp, bool f32, int W) {
  if (W == 1) return nsp <= kNP - 1;
  return f32 && (W == 2 || W == 3) && nsp + 1 <= kWave * W;
}

size_t bdf_slot_doubles(const DevMech& m, int W) {
  if (W <= 1) return (size_t)kNP * kNP;
  const size_t floats = (size_t)(m.nsp + 1) * (kWave * W + 1);
  return ((floats + 1) / 2 + 31) & ~(size_t)31;  // 256-byte aligned slots
}

size_t bdf_lds_bytes(const DevMech& m, bool f32, int W) {
  if (W <= 1)
    return sizeof(CvUni) + wave_matrix_bytes(m.nsp, f32 ? sizeof(float) : sizeof(double)) +
           sizeof(double) * kNP + rhs_lds_bytes(m);
  const size_t L = (size_t)kWave * W;
  const size_t mat = ((size_t)(m.nsp + 1) * (L + 1) * sizeof(float) + 15) & ~(size_t)15;
  return W * sizeof(CvUni) + mat + 2 * L * sizeof(float) + 2 * L * sizeof(double) +
         (size_t)((W + 2) & ~1) * sizeof(double) + rhs_lds_bytes(m);
}

namespace {

size_t max_lds_per_block() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 0;
  return (size_t)v;
}

template <class K>
int resident_blocks(K kernel, int threads, size_t lds) {
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

}  // namespace

int bdf_max_resident_teams(const DevMech& m, bool f32, int W) {
  if (!bdf_team_ok(m.nsp, f32, W)) return 0;
  const size_t lds = bdf_lds_bytes(m, f32, W);
  int r = 0;
  (void)with_ext(m.n_ext > 0, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    switch (W) {
      case 1:
        r = f32 ? resident_blocks(bdf_kernel<float, E>, kWave, lds)
                : resident_blocks(bdf_kernel<double, E>, kWave, lds);
        break;
      case 2:
        r = f32 ? resident_blocks(bdf_team_kernel<2, E>, 2 * kWave, lds) : 0;
        break;
      case 3:
        r = f32 ? resident_blocks(bdf_team_kernel<3, E>, 3 * kWave, lds) : 0;
        break;
      default:
        r = 0;
    }
    return hipSuccess;
  });
  return r;
}

hipError_t launch_bdf(const DevMech& m, const BdfOpts& o, int n, double dt, const double* dts,
                      double* T, double* P, double* Y, int stride, int* nsteps, int* stats,
                      const int* order, unsigned int* counter, double* jscratch, int nwaves,
                      int W, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const bool f32 = o.fp32_preconditioner != 0;
  if (!bdf_team_ok(m.nsp, f32, W)) return hipErrorInvalidValue;
  const size_t lds = bdf_lds_bytes(m, f32, W);
  if (lds > max_lds_per_block()) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, counter, 0u);
  const int grid = nwaves < n ? nwaves : n;
  return with_ext(m.n_ext > 0 || o.ign_dT > 0.0, [&](auto e) {
    constexpr bool E = decltype(e)::value;
    if (W == 1 && f32)
      hipLaunchKernelGGL((bdf_kernel<float, E>), dim3(grid), dim3(kWave), lds, st, m, o, n, dt,
                         dts, T, P, Y, stride, nsteps, stats, order, counter, jscratch);
    else if (W == 1)
      hipLaunchKernelGGL((bdf_kernel<double, E>), dim3(grid), dim3(kWave), lds, st, m, o, n, dt,
                         dts, T, P, Y, stride, nsteps, stats, order, counter, jscratch);
    else if (W == 2)
      hipLaunchKernelGGL((bdf_team_kernel<2, E>), dim3(grid), dim3(2 * kWave), lds, st, m, o, n,
                         dt, dts, T, P, Y, stride, nsteps, stats, order, counter, jscratch,
                         bdf_slot_doubles(m, W));
    else
      hipLaunchKernelGGL((bdf_team_kernel<3, E>), dim3(grid), dim3(3 * kWave), lds, st, m, o, n,
                         dt, dts, T, P, Y, stride, nsteps, stats, order, counter, jscratch,
                         bdf_slot_doubles(m, W));
    return hipGetLastError();
  });
}

}  // namespace zrk
