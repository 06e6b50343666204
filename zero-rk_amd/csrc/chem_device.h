// Wavefront-level FP64 chemistry: NASA-7 thermo, rate constants, rates of
// progress, species production and the constant-pressure right-hand side.
//
// One 64-lane wavefront owns one reactor.  Lanes sweep reaction records
// (records are kind-sorted so neighbouring lanes branch alike); rates of
// progress go to the wave's LDS slice; each species' creation/destruction is
// then a gather over its own step list, so every output has exactly one
// writer (no atomics, bitwise-reproducible).
//
// Formulas follow the reference: nasa_poly.cpp:50-172 (thermo),
// rate_const.cpp:539-1012 (updateK), perf_net.cpp:190-275 (limiter, ROP,
// create/destroy/net), reactor_constant_pressure_cpu.cpp:68-133 (RHS).
#pragma once

#include <hip/hip_runtime.h>

#include "dev_mech.h"

namespace zrk {

constexpr double kRu = 8.314462618e3;   // J/kmol/K
constexpr double kPatm = 1.01325e5;     // Pa
constexpr double kTlimit = 1.0e4;       // K, RHS temperature clamp
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Make a value known-uniform to the compiler (all lanes hold the same value).
__device__ __forceinline__ double uni(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readfirstlane(lo);
  hi = __builtin_amdgcn_readfirstlane(hi);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Butterfly sums: every lane ends with the same (bitwise) value.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return uni(v);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, kWave));
  return uni(v);
}

// Value of lane `src` (wave-uniform) broadcast to every lane.
__device__ __forceinline__ double bcast(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}

struct TPow {
  double T, invT, logT;
  double hm0, hm1, hm2, hm3;
  double gm0, gm1, gm2, gm3, gm4;
  double log_patm_rut;
};

__device__ __forceinline__ TPow make_tpow(double T) {
  TPow p;
  p.T = T;
  p.invT = 1.0 / T;
  p.logT = log(T);
  double h0 = T, h1 = T * h0, h2 = T * h1, h3 = T * h2;
  p.hm0 = h0 * 0.5;
  p.hm1 = h1 * 0.33333333333333333333;
  p.hm2 = h2 * 0.25;
  p.hm3 = h3 * 0.2;
  p.gm0 = 1.0 - p.logT;
  p.gm1 = T * -0.5;
  p.gm2 = (T * T) * -0.16666666666666666667;
  p.gm3 = (T * (T * T)) * -0.08333333333333333333;
  p.gm4 = (T * (T * (T * T))) * -0.05;
  p.log_patm_rut = log(kPatm / (kRu * T));
  return p;
}

// NASA-7 cp/R, h/RT, g/RT of one species (coefficient row c[16]).
__device__ __forceinline__ void thermo_one(const double* __restrict__ c, const TPow& p,
                                           double& cp, double& h, double& g) {
  const double* a = (p.T < c[0]) ? c + 1 : c + 8;
  const double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6];
  const double T = p.T;
  cp = a0 + T * (a1 + T * (a2 + T * (a3 + T * a4)));
  h = a0 + a1 * p.hm0 + a2 * p.hm1 + a3 * p.hm2 + a4 * p.hm3 + a5 * p.invT;
  g = a0 * p.gm0 + a1 * p.gm1 + a2 * p.gm2 + a3 * p.gm3 + a4 * p.gm4 + a5 * p.invT - a6;
}

// Rate constants, limiter and rates of progress of every step of one reactor.
// C, G: species concentrations and g/RT in the wave's LDS; rop: [nstep] LDS.
__device__ __forceinline__ void wave_rates(const DevMech& m, const double* __restrict__ C,
                                           const double* __restrict__ G, double* __restrict__ rop,
                                           const TPow& p, double Csum, double L) {
  const bool limit = (0.0 < L && L < 1.0e300);
  for (int r = lane_id(); r < m.nrec; r += kWave) {
    const int fl = m.r_flags[r];
    double kf = exp(m.r_fA[r] + m.r_fB[r] * p.logT - m.r_fT[r] * p.invT);
    double kr = 0.0;
    const int* re = m.r_reac + kMaxPart * r;
    const int* pr = m.r_prod + kMaxPart * r;
    if (fl & kRecRevArr) {
      kr = exp(m.r_rA[r] + m.r_rB[r] * p.logT - m.r_rT[r] * p.invT);
    } else if (fl & kRecRevKeq) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k)
        if (pr[k] >= 0) s += G[pr[k]];
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k)
        if (re[k] >= 0) s -= G[re[k]];
      kr = exp(s - m.r_rA[r] * p.log_patm_rut) * kf;
    }
    if (fl & (kRecThirdBody | kRecFalloff)) {
      double cm;
      const int fs = m.r_fspc[r];
      if (fs >= 0) {
        cm = C[fs];
      } else {
        cm = Csum;
        for (int k = m.r_eptr[r]; k < m.r_eptr[r + 1]; ++k) cm += C[m.e_sp[k]] * m.e_eff[k];
      }
      if (fl & kRecThirdBody) {
        kf *= cm;
        kr *= cm;
      } else {
        const double* q = m.r_fo + 7 * r;
        const double klow = exp(q[0] + q[1] * p.logT - q[2] * p.invT);
        double Pr = klow * cm / kf;
        if (Pr < 1.0e-300) Pr = 1.0e-300;
        double lpr = log10(Pr);
        double f = 1.0;
        if (fl & kRecTroe) {
          double fc = 0.0;
          if (q[4] != 0) fc += (1.0 - q[3]) * exp(-p.T / q[4]);
          if (q[5] != 0) fc += q[3] * exp(-p.T / q[5]);
          if (fl & kRecTroe4) fc += exp(-q[6] * p.invT);
          if (fc < 1.0e-300) fc = 1.0e-300;
          f = log10(fc);
          const double n = 0.75 - 1.27 * f;
          lpr -= (0.4 + 0.67 * f);
          lpr = lpr / (n - 0.14 * lpr);
          lpr *= lpr;
          f /= (1.0 + lpr);
          f = pow(10.0, f);
        }
        const double pc = f * Pr / (1.0 + Pr);
        kf *= pc;
        kr *= pc;
      }
    }
    if (limit) {
      kf *= L / (L + kf);
      kr *= L / (L + kr);
    }
#pragma unroll
    for (int k = 0; k < kMaxPart; ++k)
      if (re[k] >= 0) kf *= C[re[k]];
    rop[m.r_fstep[r]] = kf;
    const int rs = m.r_rstep[r];
    if (rs >= 0) {
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k)
        if (pr[k] >= 0) kr *= C[pr[k]];
      rop[rs] = kr;
    }
  }
}

// Creation and destruction rates of species j (gather, reference summation order).
__device__ __forceinline__ void species_cd(const DevMech& m, const double* __restrict__ rop, int j,
                                           double& create, double& destroy) {
  double d = 0.0, c = 0.0;
  for (int k = m.d_ptr[j]; k < m.d_ptr[j + 1]; ++k) d += rop[m.d_step[k]];
  for (int k = m.c_ptr[j]; k < m.c_ptr[j + 1]; ++k) c += rop[m.c_step[k]];
  create = c;
  destroy = d;
}

// ---------------------------------------------------------------------------
// Constant-pressure RHS for nsp <= 63: lane j < nsp holds species j, lane nsp
// holds the temperature variable y_T = T/Tref (padding lanes hold 0).
// Returns 1 when T <= 0 (recoverable for the integrator), else 0.
// Outputs per lane: ydot, and for species lanes h/RT and mass cp (J/kg/K).
// ---------------------------------------------------------------------------
struct RhsLds {
  double* C;    // [>= nsp]
  double* G;    // [>= nsp]
  double* rop;  // [>= nstep]
};

struct RhsAux {
  double inv_rho;  // 1/density of this evaluation
  double cp_mix;   // mixture cp (J/kg/K)
  double h;        // lane: h/RT of species lane
  double cpm;      // lane: cp (J/kg/K) of species lane
};

__device__ __forceinline__ int rhs64(const DevMech& m, const RhsLds& s, double P, double Tref,
                                     double L, double y, double& ydot, RhsAux& aux) {
  const int lane = lane_id();
  const int nsp = m.nsp;
  const bool sp = lane < nsp;
  const double yT = bcast(y, nsp);
  double T = yT * Tref;
  if (T <= 0.0) {
    ydot = 0.0;
    return 1;
  }
  T = fmin(T, kTlimit);
  const TPow p = make_tpow(T);
  const double Yj = sp ? y : 0.0;
  const double imw = sp ? m.inv_mw[lane] : 0.0;
  const double mws = wave_sum(Yj * imw);
  const double wmix = 1.0 / mws;
  const double inv_rho = 1.0 / (P * wmix / (kRu * T));
  const double rho_c = 1.0 / inv_rho;
  double cp = 0.0, h = 0.0, g = 0.0, Cj = 0.0, cpm = 0.0;
  if (sp) {
    thermo_one(m.thermo + 16 * lane, p, cp, h, g);
    cpm = cp * m.ru_mw[lane];
    Cj = rho_c * imw * Yj;
    s.C[lane] = Cj;
    s.G[lane] = g;
  }
  const double cp_mix = wave_sum(cpm * Yj);
  const double Csum = wave_sum(Cj);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  wave_rates(m, s.C, s.G, s.rop, p, Csum, L);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double e = 0.0;
  double yd = 0.0;
  if (sp) {
    double cr, de;
    species_cd(m, s.rop, lane, cr, de);
    const double net = cr - de;
    yd = (net * m.mw[lane]) * inv_rho;
    e = h * net * inv_rho;
  }
  const double esum = wave_sum(e);
  if (lane == nsp) yd = -esum * kRu * yT / cp_mix + 0.0 / (cp_mix * Tref);
  ydot = yd;
  aux.inv_rho = inv_rho;
  aux.cp_mix = cp_mix;
  aux.h = h;
  aux.cpm = cpm;
  __builtin_amdgcn_wave_barrier();
  return 0;
}

}  // namespace zrk
