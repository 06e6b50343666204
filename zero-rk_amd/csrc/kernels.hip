// gfx950 kernels of the batched reactor-chemistry path and their launchers.
//
//   rates_kernel     FP64 rates of progress + species production + const-P RHS
//                    for a batch of reactor states (any mechanism size):
//                    one wavefront per reactor, 4 per workgroup.
//   jacobian_kernel  reference analytic Jacobian for nsp <= 63 (parity tests).
//   launch_bdf       the persistent per-reactor BDF integrator (bdf_kernels.h;
//                    its instantiations are compiled in bdf_{wave,team}_
//                    {direct,krylov}.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "bdf_launch.h"
#include "cvode_wave.h"
#include "launch.h"

namespace zrk {

constexpr int kRatesWaves = 4;  // wavefronts per workgroup in rates_kernel

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
// rates_block_kernel: rates_kernel for mechanisms whose per-reactor arrays do
// not fit in one wavefront's share of LDS (the 874-species iso-octane and
// 2,000-species mechanisms): one reactor per workgroup of kRatesBlockWaves
// wavefronts at a time, C, g/RT and h/RT in LDS, the rates of progress in
// this workgroup's slot of scratch (nstep + 1 doubles: the last is the zero
// that absent participants read), copied to rop_out when it is requested.
// Sums are reduced per wavefront, then across wavefronts in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kRatesBlockWaves = 4;

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int wave = threadIdx.x / kWave;
  v = wave_sum(v);
  if (lane_id() == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kRatesBlockWaves; ++w) s += red[w];
  __syncthreads();
  return s;
}

template <bool kExt>
__global__ void __launch_bounds__(kWave* kRatesBlockWaves)
rates_block_kernel(DevMech m, int n, const double* __restrict__ Tin,
                   const double* __restrict__ Pin, const double* __restrict__ Yin, int stride,
                   double L, double* __restrict__ ydot, double* __restrict__ rop_out,
                   double* __restrict__ create_out, double* __restrict__ destroy_out,
                   double* __restrict__ net_out, double* __restrict__ scratch) {
  constexpr int S = kWave * kRatesBlockWaves;
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int nsp = m.nsp, nstep = m.nstep;
  double* red = lds;  // kRatesBlockWaves (padded to 8)
  double* C = lds + 8;
  double* G = C + nsp + 1;
  double* H = G + nsp + 1;
  double* rop = scratch + (size_t)blockIdx.x * (nstep + 1);
  if (tid == 0) {  // entries read by absent reaction participants
    C[nsp] = 1.0;
    G[nsp] = 0.0;
  }
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    const double* Y = Yin + (size_t)r * stride;
    const double T = fmin(Tin[r], kTlimit);
    const double P = Pin[r];
    const TPow p = make_tpow(T);
    double part = 0.0;
    for (int j = tid; j < nsp; j += S) part += Y[j] * m.inv_mw[j];
    const double wmix = 1.0 / block_sum(part, red);
    const double inv_rho = 1.0 / (P * wmix / (kRu * T));
    const double rho_c = 1.0 / inv_rho;
    double cps = 0.0, cs = 0.0;
    for (int j = tid; j < nsp; j += S) {
      double cp, h, g;
      thermo_one(m.thermo + 16 * j, p, cp, h, g);
      const double Cj = rho_c * m.inv_mw[j] * Y[j];
      C[j] = Cj;
      G[j] = g;
      H[j] = h;
      cps += cp * m.ru_mw[j] * Y[j];
      cs += Cj;
    }
    const double cp_mix = block_sum(cps, red);
    const double Csum = block_sum(cs, red);  // ends with a barrier: C, G visible
    team_rates<S, kExt>(tid, m, C, G, rop, p, Csum, L);
    __syncthreads();  // the rates of progress (this workgroup's slot) visible
    double e = 0.0;
    for (int j = tid; j < nsp; j += S) {
      double cr, de;
      species_cd<kExt>(m, rop, j, cr, de);
      const double net = cr - de;
      if (create_out) create_out[(size_t)r * nsp + j] = cr;
      if (destroy_out) destroy_out[(size_t)r * nsp + j] = de;
      if (net_out) net_out[(size_t)r * nsp + j] = net;
      if (ydot) ydot[(size_t)r * (nsp + 1) + j] = (net * m.mw[j]) * inv_rho;
      e += H[j] * net * inv_rho;
    }
    const double esum = block_sum(e, red);
    if (ydot && tid == 0)
      ydot[(size_t)r * (nsp + 1) + nsp] = -esum * kRu * Tin[r] / cp_mix + 0.0 / cp_mix;
    if (rop_out)
      for (int s2 = tid; s2 < nstep; s2 += S) rop_out[(size_t)r * nstep + s2] = rop[s2];
    __syncthreads();  // the slot and the LDS arrays are reused by the next reactor
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

bool rates_need_block(const DevMech& m) {
  return sizeof(double) * (size_t)kRatesWaves * (3 * m.nsp + 2 + m.nstep + 1) > 160 * 1024;
}

size_t rates_block_lds_bytes(const DevMech& m) {
  return sizeof(double) * (size_t)(8 + 3 * m.nsp + 2);
}

int rates_block_grid(const DevMech& m, int n) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const size_t lds = rates_block_lds_bytes(m);
  hipError_t e = m.n_ext > 0
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rates_block_kernel<true>,
                                                     kWave * kRatesBlockWaves, lds)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rates_block_kernel<false>,
                                                     kWave * kRatesBlockWaves, lds);
  if (e != hipSuccess || per_cu <= 0) return 0;
  return std::max(1, std::min(n, per_cu * ncu));
}

hipError_t launch_rates(const DevMech& m, int n, const double* T, const double* P,
                        const double* Y, int stride, double L, double* ydot, double* rop,
                        double* create, double* destroy, double* net, hipStream_t st,
                        double* scratch, int grid) {
  if (n <= 0) return hipSuccess;
  if (rates_need_block(m)) {
    const size_t lds = rates_block_lds_bytes(m);
    if (lds > 160 * 1024 || !scratch || grid <= 0) return hipErrorInvalidValue;
    return with_ext(m.n_ext > 0, [&](auto e) {
      hipLaunchKernelGGL((rates_block_kernel<decltype(e)::value>), dim3(std::min(grid, n)),
                         dim3(kWave * kRatesBlockWaves), lds, st, m, n, T, P, Y, stride, L, ydot,
                         rop, create, destroy, net, scratch);
      return hipGetLastError();
    });
  }
  const size_t lds = sizeof(double) * (size_t)kRatesWaves * (3 * m.nsp + 2 + m.nstep + 1);
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
  (void)f32;  // teams have both FP32 (LDS) and FP64 (device slot) Newton matrices
  if (W == 1) return nsp <= kNP - 1;
  return (W == 2 || W == 3) && nsp + 1 <= kWave * W;
}

size_t bdf_slot_doubles(const DevMech& m, int W, bool f32) {
  if (W <= 1) return (size_t)kNP * kNP;
  // the saved Jacobian, column stride 64W + 1; FP64 teams add the Newton
  // matrix (inverted in place) with the same layout
  const size_t elems = (size_t)(m.nsp + 1) * (kWave * W + 1);
  const size_t d = f32 ? (elems + 1) / 2 : 2 * elems;
  return (d + 31) & ~(size_t)31;  // 256-byte aligned slots
}

size_t bdf_lds_bytes(const DevMech& m, bool f32, int W, bool iter) {
  const size_t kry = iter ? sizeof(double) * kKryDoubles : 0;
  if (W <= 1)
    return sizeof(CvUni) + wave_matrix_bytes(m.nsp, f32 ? sizeof(float) : sizeof(double)) +
           sizeof(double) * kNP + rhs_lds_bytes(m) + kry;
  const size_t L = (size_t)kWave * W;
  const size_t mt = f32 ? sizeof(float) : sizeof(double);
  const size_t mat = f32 ? ((size_t)(m.nsp + 1) * (L + 1) * sizeof(float) + 15) & ~(size_t)15 : 0;
  return W * sizeof(CvUni) + mat + 2 * L * mt + 2 * L * sizeof(double) +
         (size_t)((W + 2) & ~1) * sizeof(double) + rhs_lds_bytes(m) + kry;
}


int bdf_max_resident_teams(const DevMech& m, bool f32, int W, bool iter) {
  if (!bdf_team_ok(m.nsp, f32, W)) return 0;
  const size_t lds = bdf_lds_bytes(m, f32, W, iter);
  const bool ext = m.n_ext > 0;
  if (W == 1) return iter ? resident_bdf_wave<true>(ext, f32, lds)
                          : resident_bdf_wave<false>(ext, f32, lds);
  return iter ? resident_bdf_team<true>(ext, W, lds, f32)
              : resident_bdf_team<false>(ext, W, lds, f32);
}

hipError_t launch_bdf(const DevMech& m, const BdfOpts& o, int n, double dt, const double* dts,
                      double* T, double* P, double* Y, int stride, int* nsteps, int* stats,
                      const int* order, unsigned int* counter, double* jscratch, int nwaves,
                      int W, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const bool f32 = o.fp32_preconditioner != 0;
  const bool iter = bdf_alt_modes(o);
  if (!bdf_team_ok(m.nsp, f32, W)) return hipErrorInvalidValue;
  const size_t lds = bdf_lds_bytes(m, f32, W, iter);
  if (lds > max_lds_per_block()) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, counter, 0u);
  const int grid = nwaves < n ? nwaves : n;
  const bool ext = m.n_ext > 0 || o.ign_dT > 0.0;
  const BdfLaunch a{m, o, n, dt, dts, T, P, Y, stride, nsteps, stats, order, counter, jscratch,
                    grid, lds, st};
  if (W == 1) return iter ? launch_bdf_wave<true>(a, ext, f32) : launch_bdf_wave<false>(a, ext, f32);
  const size_t slot = bdf_slot_doubles(m, W, f32);
  return iter ? launch_bdf_team<true>(a, ext, W, slot, f32)
              : launch_bdf_team<false>(a, ext, W, slot, f32);
}

}  // namespace zrk
