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

// Lane values moved by a DPP pattern within rows of 16 lanes (both halves of
// the double); CTRL: 0xB1 quad xor 1, 0x4E quad xor 2, 0x141 half-row mirror,
// 0x140 row mirror.  All lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Wave sums: symmetric DPP exchanges inside each row of 16 lanes (every lane
// of a row ends with the same row sum, since a + b == b + a bitwise), then the
// four row sums added in row order from scalar registers.  Every lane gets
// the same bits.  All lanes must be active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return uni(((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48));
}

__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_move<0xB1>(v));
  v = fmax(v, dpp_move<0x4E>(v));
  v = fmax(v, dpp_move<0x141>(v));
  v = fmax(v, dpp_move<0x140>(v));
  return uni(fmax(fmax(lane_value(v, 0), lane_value(v, 16)),
                  fmax(lane_value(v, 32), lane_value(v, 48))));
}

// Value of lane `src` (wave-uniform) broadcast to every lane.
__device__ __forceinline__ double bcast(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// ---------------------------------------------------------------------------
// Teams: the threads that share one reactor.  A TeamWave is one wavefront
// (state components on lanes 0..63); a TeamBlock<W> is W wavefronts of one
// workgroup (components on threads 0..64W-1), whose reductions and
// broadcasts go through a few LDS slots (red[0..W]) and workgroup barriers.
// Reductions sum the per-wavefront butterfly results in wavefront order, so
// every thread gets the same bits.
// kExt: the mechanism has records that need the general evaluation (PLOG,
// non-integer, SRI); the code for them is compiled only into kExt variants,
// so the common kernels carry none of it.
// ---------------------------------------------------------------------------
template <bool E>
struct TeamWaveX {
  static constexpr int kThreads = kWave;
  static constexpr bool kExt = E;
  __device__ __forceinline__ int tid() const { return lane_id(); }
  __device__ __forceinline__ double sum(double v) const { return wave_sum(v); }
  __device__ __forceinline__ double max(double v) const { return wave_max(v); }
  __device__ __forceinline__ double from(double v, int src) const { return bcast(v, src); }
  __device__ __forceinline__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

using TeamWave = TeamWaveX<false>;

template <int W, bool E = false>
struct TeamBlock {
  static constexpr int kThreads = kWave * W;
  static constexpr bool kExt = E;
  double* red;  // LDS [W + 1]
  __device__ __forceinline__ int tid() const { return (int)threadIdx.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
  __device__ __forceinline__ double sum(double v) const {
    const double s = wave_sum(v);
    if (lane_id() == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < W; ++w) r += red[w];
    __syncthreads();
    return uni(r);
  }
  __device__ __forceinline__ double max(double v) const {
    const double s = wave_max(v);
    if (lane_id() == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < W; ++w) r = fmax(r, red[w]);
    __syncthreads();
    return uni(r);
  }
  __device__ __forceinline__ double from(double v, int src) const {
    if ((int)threadIdx.x == src) red[W] = v;
    __syncthreads();
    const double r = red[W];
    __syncthreads();
    return uni(r);
  }
};

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

// NASA-7 cp/R, h/RT, g/RT of one species (coefficient row c[16]).  Both
// ranges are loaded before the T < Tmid choice so the loads issue together.
__device__ __forceinline__ void thermo_one(const double* __restrict__ c, const TPow& p,
                                           double& cp, double& h, double& g) {
  double lo[7], hi[7];
  const double tmid = c[0];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    lo[k] = c[1 + k];
    hi[k] = c[8 + k];
  }
  const bool low = p.T < tmid;
  double a[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) a[k] = low ? lo[k] : hi[k];
  const double T = p.T;
  cp = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  h = a[0] + a[1] * p.hm0 + a[2] * p.hm1 + a[3] * p.hm2 + a[4] * p.hm3 + a[5] * p.invT;
  g = a[0] * p.gm0 + a[1] * p.gm1 + a[2] * p.gm2 + a[3] * p.gm3 + a[4] * p.gm4 + a[5] * p.invT -
      a[6];
}

__device__ __forceinline__ int lo16(int w) { return w & 0xFFFF; }
__device__ __forceinline__ int hi16(int w) { return (int)((unsigned)w >> 16); }

// One reaction record as loaded from HBM (all fields issued together).
struct RecLoad {
  int4 meta, sp;
  double2 c0, c1, c2;
  int e0, e1;
};

__device__ __forceinline__ RecLoad load_rec(const DevMech& m, int r) {
  RecLoad x;
  x.meta = m.r_meta[r];
  x.sp = m.r_sp[r];
  x.c0 = m.r_coef[3 * r];
  x.c1 = m.r_coef[3 * r + 1];
  x.c2 = m.r_coef[3 * r + 2];
  x.e0 = m.r_eptr[r];
  x.e1 = m.r_eptr[r + 1];
  return x;
}

// PLOG forward rate coefficient at temperature p.T and pressure P from the
// table at t (plog_reaction.cpp:163-320): the lines of one pressure are summed
// with the sign of their A; between two tabulated pressures the rate is
// interpolated linearly in log p (zero if either end is not positive);
// outside the table the end pressure's rate is used.
__device__ __forceinline__ double plog_at(const double* __restrict__ t, int np, int k,
                                          const TPow& p) {
  const double* __restrict__ lptr = t + 1 + 2 * np;
  const double* __restrict__ ln = lptr + np + 1;
  double r = 0.0;
  const int a = (int)lptr[k], b = (int)lptr[k + 1];
  for (int q = a; q < b; ++q) {
    const double* __restrict__ l = ln + 4 * q;
    r += l[1] * exp(l[0] + p.logT * l[2] - p.invT * l[3]);
  }
  return r;
}

__device__ __forceinline__ double plog_rate(const double* __restrict__ t, const TPow& p,
                                            double P) {
  const int np = (int)t[0];
  const double* __restrict__ pp = t + 1;
  const double* __restrict__ lp = pp + np;
  if (p.T <= 0.0 || P <= 0.0) return 0.0;
  if (np == 1 || P <= pp[0]) return plog_at(t, np, 0, p);
  if (P >= pp[np - 1]) return plog_at(t, np, np - 1, p);
  int r = 0;
  while (r + 2 < np && P >= pp[r + 1]) ++r;  // pp[r] <= P < pp[r + 1]
  const double k1 = plog_at(t, np, r, p);
  if (k1 <= 0.0) return 0.0;
  const double k2 = plog_at(t, np, r + 1, p);
  if (k2 <= 0.0) return 0.0;
  return k1 * pow(P / pp[r], log(k2 / k1) / (lp[r + 1] - lp[r]));
}

// k times prod sign(C_s) |C_s|^p_s over a non-integer step's exponent list
// (non_integer_reaction_network.cpp:270-293).
__device__ __forceinline__ double ni_rop(double k, const DevMech& m, const double* __restrict__ C,
                                         int off, int n) {
  for (int q = off; q < off + n; ++q) {
    const double c = C[m.ni_sp[q]];
    const double f = pow(fabs(c), m.ni_pow[q]);
    k *= (c >= 0.0) ? f : -f;
  }
  return k;
}

// Build-time knob: 1 loads the next record while the current one is
// evaluated (hides the load latency at the cost of about 22 vector registers;
// measured no faster in the integrator at two waves per SIMD, so off).
#ifndef ZRK_RATES_PREFETCH
#define ZRK_RATES_PREFETCH 0
#endif

// One record of any kind, plainly evaluated: the records with a PLOG forward
// rate, non-integer exponents or SRI blending (rare; kept out of the main
// loop so that its register use does not grow).  The order of the updates is
// the reference's (rate_const.cpp:539-582): Arrhenius or PLOG (precomputed
// into rop), reverse from REV or Keq, third body, fall-off; then the limiter
// (perf_net.cpp:204-216) and the rates of progress.
__device__ __forceinline__ void rates_general_record(const DevMech& m, int r,
                                                  const double* __restrict__ C,
                                                  const double* __restrict__ G,
                                                  double* __restrict__ rop, const TPow& p,
                                                  double Csum, double L) {
  const RecLoad x = load_rec(m, r);
  const int fl = x.meta.z;
  const int4 xe = m.r_ext[fl >> kRecExtShift];
  const int re[kMaxPart] = {lo16(x.sp.x), hi16(x.sp.x), lo16(x.sp.y), hi16(x.sp.y)};
  const int pr[kMaxPart] = {lo16(x.sp.z), hi16(x.sp.z), lo16(x.sp.w), hi16(x.sp.w)};
  const int nf = xe.w & 0xFFFF, nrv = xe.w >> 16;
  // a PLOG forward rate was stored in rop[forward step] by the pass before
  double kf = (fl & kRecPlog) ? rop[x.meta.x] : exp(x.c0.x + x.c0.y * p.logT - x.c1.x * p.invT);
  double kr = 0.0;
  if (fl & kRecRevArr) {
    kr = exp(x.c1.y + x.c2.x * p.logT - x.c2.y * p.invT);
  } else if (fl & kRecRevKeq) {
    double s = 0.0;
    if (fl & kRecNonInt) {
      // exponent-weighted change of g/RT: reverse list minus forward list
      // (non_integer_reaction_network.cpp:335-358)
      for (int k = xe.z; k < xe.z + nrv; ++k) s += m.ni_pow[k] * G[m.ni_sp[k]];
      for (int k = xe.y; k < xe.y + nf; ++k) s -= m.ni_pow[k] * G[m.ni_sp[k]];
    } else {
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k) s += G[pr[k]];
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k) s -= G[re[k]];
    }
    kr = exp(s - x.c1.y * p.log_patm_rut) * kf;
  }
  if (fl & (kRecThirdBody | kRecFalloff)) {
    double cm;
    if (x.meta.w >= 0) {
      cm = C[x.meta.w];
    } else {
      cm = Csum;
      for (int k = x.e0; k < x.e1; ++k) cm += C[m.e_sp[k]] * m.e_eff[k];
    }
    if (fl & kRecThirdBody) {
      kf *= cm;
      kr *= cm;
    } else {
      const double* __restrict__ q = m.r_fo + 8 * r;
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
      } else if (fl & kRecSri) {
        // d (a exp(-b/T) + exp(-T/c))^X T^e, X = 1/(1 + log10(Pr)^2)
        // (rate_const.cpp:975-1004; d = 1 and e = 0 when not given)
        const double xp = 1.0 / (1.0 + lpr * lpr);
        f = q[3] * exp(-q[4] * p.invT);
        if (q[5] > 0) f += exp(-p.T * q[5]);
        f = pow(f, xp);
        f *= q[6];
        f *= pow(p.T, q[7]);
      }
      const double pc = f * Pr / (1.0 + Pr);
      kf *= pc;
      kr *= pc;
    }
  }
  if (0.0 < L && L < 1.0e300) {
    kf *= L / (L + kf);
    kr *= L / (L + kr);
  }
  for (int k = 0; k < kMaxPart; ++k) kf *= C[re[k]];
  const int rs = x.meta.y;
  if (fl & kRecNonInt) {
    // a non-integer reaction has no expanded participants: its exponent
    // lists give the rates of progress
    kf = ni_rop(kf, m, C, xe.y, nf);
    if (rs >= 0) kr = ni_rop(kr, m, C, xe.z, nrv);
  } else if (rs >= 0) {
    for (int k = 0; k < kMaxPart; ++k) kr *= C[pr[k]];
  }
  rop[x.meta.x] = kf;
  if (rs >= 0) rop[rs] = kr;
}

// Rate constants, limiter and rates of progress of every step of one reactor.
// C, G: species concentrations and g/RT in the wave's LDS, with the entries
// C[nsp] = 1 and G[nsp] = 0 that absent participants (index nsp) read;
// rop: [nstep+1] LDS (slot nstep is the zero rate the padded gathers point at).
// Thread t of a team of S threads handles records t, t+S, ...; a record's
// third-body and fall-off data are requested before its rate constants are
// computed.
// Records come fall-off first, then third-body, then plain, so the costly
// fall-off branch is taken by as few wavefront iterations as possible.
template <int S, bool kExt>
__device__ __forceinline__ void team_rates(int tid, const DevMech& m, const double* __restrict__ C,
                                           const double* __restrict__ G, double* __restrict__ rop,
                                           const TPow& p, double Csum, double L) {
  const bool limit = (0.0 < L && L < 1.0e300);
  // PLOG, non-integer and SRI records come last and are evaluated by
  // rates_general_record (kExt variants only)
  const int nrec = kExt ? m.nrec - m.n_ext : m.nrec;
  if (tid == 0) rop[m.nstep] = 0.0;
  int r = tid;
#if ZRK_RATES_PREFETCH
  RecLoad nx;
  if (r < nrec) nx = load_rec(m, r);
#endif
  const int4* __restrict__ es = reinterpret_cast<const int4*>(m.e_sp);
  const double2* __restrict__ ee = reinterpret_cast<const double2*>(m.e_eff);
  for (; r < nrec; r += S) {
#if ZRK_RATES_PREFETCH
    const RecLoad x = nx;
#else
    const RecLoad x = load_rec(m, r);
#endif
    const int fl = x.meta.z;
    const bool enh = (fl & (kRecThirdBody | kRecFalloff)) && x.meta.w < 0 && x.e1 > x.e0;
    // The loads below are issued on every path (from valid dummy addresses
    // where they are not needed) and in this order: the vector-memory counter
    // drains in order, so with path-independent counts a wait for this
    // record's data does not also wait for the next record's.
    const int e0 = enh ? x.e0 : 0;
    const int4 es0 = es[e0 >> 2];
    const double2 ea0 = ee[e0 >> 1], eb0 = ee[(e0 >> 1) + 1];
    const double* __restrict__ qp = m.r_fo + 8 * ((fl & kRecFalloff) ? r : 0);
    double q[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = qp[k];
#if ZRK_RATES_PREFETCH
    nx = load_rec(m, (r + S < nrec) ? r + S : r);
#endif
    const int re[kMaxPart] = {lo16(x.sp.x), hi16(x.sp.x), lo16(x.sp.y), hi16(x.sp.y)};
    const int pr[kMaxPart] = {lo16(x.sp.z), hi16(x.sp.z), lo16(x.sp.w), hi16(x.sp.w)};
    double kf = exp(x.c0.x + x.c0.y * p.logT - x.c1.x * p.invT);
    double kr = 0.0;
    if (fl & kRecRevArr) {
      kr = exp(x.c1.y + x.c2.x * p.logT - x.c2.y * p.invT);
    } else if (fl & kRecRevKeq) {
      // absent participants (listed last) read the 0 in G[nsp]: s +- 0 == s
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k) s += G[pr[k]];
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k) s -= G[re[k]];
      kr = exp(s - x.c1.y * p.log_patm_rut) * kf;
    }
    if (fl & (kRecThirdBody | kRecFalloff)) {
      double cm;
      const int fs = x.meta.w;
      if (fs >= 0) {
        cm = C[fs];
      } else {
        // list order; padding entries (species 0, efficiency 0) add +-0
        cm = Csum;
        if (enh) {
          cm += C[es0.x] * ea0.x;
          cm += C[es0.y] * ea0.y;
          cm += C[es0.z] * eb0.x;
          cm += C[es0.w] * eb0.y;
          for (int k = (x.e0 >> 2) + 1; k < (x.e1 >> 2); ++k) {
            const int4 sk = es[k];
            const double2 a = ee[2 * k], b = ee[2 * k + 1];
            cm += C[sk.x] * a.x;
            cm += C[sk.y] * a.y;
            cm += C[sk.z] * b.x;
            cm += C[sk.w] * b.y;
          }
        }
      }
      if (fl & kRecThirdBody) {
        kf *= cm;
        kr *= cm;
      } else {
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
    // absent participants read the 1 in C[nsp]: kf * 1 == kf
#pragma unroll
    for (int k = 0; k < kMaxPart; ++k) kf *= C[re[k]];
    rop[x.meta.x] = kf;
    const int rs = x.meta.y;
    if (rs >= 0) {
#pragma unroll
      for (int k = 0; k < kMaxPart; ++k) kr *= C[pr[k]];
      rop[rs] = kr;
    }
  }
  if constexpr (kExt) {
    // PLOG forward rates at the pressure of the concentrations, Csum Ru T
    // (rate_const.cpp:551, 565-568), in a pass of their own (kept apart from
    // the record evaluation, which then reads them before the reverse rate
    // from Keq); a record is handled by the same thread in both passes
    for (int q = nrec + tid; q < m.nrec; q += S) {
      const int4 meta = m.r_meta[q];
      if (meta.z & kRecPlog)
        rop[meta.x] = plog_rate(m.plog + m.r_ext[meta.z >> kRecExtShift].x, p, Csum * kRu * p.T);
    }
    for (int q = nrec + tid; q < m.nrec; q += S) rates_general_record(m, q, C, G, rop, p, Csum, L);
  }
}

// Creation and destruction rates of species j (gathers in the reference's
// summation order; padding entries add the zero rate, which is exact).
template <bool kExt>
__device__ __forceinline__ void species_cd(const DevMech& m, const double* __restrict__ rop, int j,
                                           double& create, double& destroy) {
  const int jb = j >> 6, jl = j & 63;
  const uint4* __restrict__ gd = m.g_d + (size_t)jb * m.gd_blocks * kWave + jl;
  const uint4* __restrict__ gc = m.g_c + (size_t)jb * m.gc_blocks * kWave + jl;
  double d = 0.0, c = 0.0;
  if (kExt && m.has_nig) {
    // non-integer terms first (perf_net.cpp:250-253)
    for (int k = m.nig_dptr[j]; k < m.nig_dptr[j + 1]; ++k) {
      const double2 e = m.nig[k];
      d += e.x * rop[(int)e.y];
    }
    for (int k = m.nig_cptr[j]; k < m.nig_cptr[j + 1]; ++k) {
      const double2 e = m.nig[k];
      c += e.x * rop[(int)e.y];
    }
  }
  const int2 gn = m.g_nb[jb];  // this group's blocks in use (padding beyond)
  const int nd = gn.x, nc = gn.y;
  const int nb = nd > nc ? nd : nc;
  uint4 wd = nd > 0 ? gd[0] : make_uint4(0, 0, 0, 0);
  uint4 wc = nc > 0 ? gc[0] : make_uint4(0, 0, 0, 0);
  for (int b = 0; b < nb; ++b) {
    const uint4 xd = wd, xc = wc;
    if (b + 1 < nd) wd = gd[(b + 1) * kWave];
    if (b + 1 < nc) wc = gc[(b + 1) * kWave];
    if (b < nd) {
      d += rop[xd.x & 0xFFFF];
      d += rop[xd.x >> 16];
      d += rop[xd.y & 0xFFFF];
      d += rop[xd.y >> 16];
      d += rop[xd.z & 0xFFFF];
      d += rop[xd.z >> 16];
      d += rop[xd.w & 0xFFFF];
      d += rop[xd.w >> 16];
    }
    if (b < nc) {
      c += rop[xc.x & 0xFFFF];
      c += rop[xc.x >> 16];
      c += rop[xc.y & 0xFFFF];
      c += rop[xc.y >> 16];
      c += rop[xc.z & 0xFFFF];
      c += rop[xc.z >> 16];
      c += rop[xc.w & 0xFFFF];
      c += rop[xc.w >> 16];
    }
  }
  create = c;
  destroy = d;
}

// ---------------------------------------------------------------------------
// Reactor RHS for nsp <= 63: lane j < nsp holds species j, lane nsp holds the
// temperature variable y_T = T/Tref (padding lanes hold 0).
// v0 == 0: constant pressure (reactor_constant_pressure_cpu.cpp:68-133), the
// density follows from P; h/RT and cp.  v0 > 0: constant volume at specific
// volume v0 (reactor_constant_volume_cpu.cpp:66-130); u/RT = h/RT - 1 and
// cv_j = (cp_j/R - 1) R/MW_j (mechanism.cpp:381-386, 428-441).
// Returns 1 when T <= 0 (recoverable for the integrator), else 0.
// Outputs per lane: ydot, and for species lanes the energy (h/RT or u/RT) and
// mass heat capacity (cp or cv, J/kg/K) the Jacobian's temperature row uses.
// ---------------------------------------------------------------------------
struct RhsLds {
  double* C;    // [nsp + 1]: slot nsp is 1 (the factor of an absent participant)
  double* G;    // [nsp + 1]: slot nsp is 0 (the g/RT of an absent participant)
  double* rop;  // [nstep + 1]
};

// The three arrays laid out from one LDS base (2 (nsp + 1) + nstep + 1 doubles).
__device__ __forceinline__ RhsLds rhs_lds_at(double* base, int nsp) {
  return RhsLds{base, base + nsp + 1, base + 2 * (nsp + 1)};
}

struct RhsAux {
  double inv_rho;  // 1/density of this evaluation
  double cp_mix;   // mixture cp or cv (J/kg/K)
  double h;        // lane: h/RT or u/RT of species lane
  double cpm;      // lane: cp or cv (J/kg/K) of species lane
#ifdef ZRK_PHASE_CYCLES
  long long cyc[3];  // thermo and sums, rates of progress, gathers and T row
#endif
};

// The plugin's source terms at time t of a solve (reactor_constant_pressure_cpu
// .cpp:15-42 and reactor_constant_volume_cpu.cpp:15-42): pressure rate, energy
// source, this thread's mass-fraction source; e0 the initial mass enthalpy
// (constant pressure) or internal energy (constant volume), vinit the initial
// 1/density (kept at constant pressure when mass sources are present).
struct RhsSrc {
  bool on = false;        // any source term
  bool has_ysrc = false;  // mass-fraction sources
  double t = 0.0, dpdt = 0.0, esrc = 0.0, e0 = 0.0, vinit = 0.0, ysrc = 0.0;
};

// Mass enthalpy (cvol = false) or internal energy of the mixture at T
// (mechanism::getMassEnthalpyFromTY / getMassIntEnergyFromTY,
// mechanism.cpp:336-367), and its cp or cv (getMassCpFromTY /
// getMassCvFromTY, :408-441); Yj: this thread's mass fraction (0 off species).
template <class Team>
__device__ __forceinline__ void team_mass_energy(const Team& team, const DevMech& m, bool cvol,
                                                 bool cv_capacity, double T, double Yj,
                                                 double& e, double& cx) {
  const int tid = team.tid();
  double ej = 0.0, cj = 0.0;
  if (tid < m.nsp) {
    double cp, h, g;
    thermo_one(m.thermo + 16 * tid, make_tpow(T), cp, h, g);
    if (cvol) h -= 1.0;
    if (cv_capacity) cp -= 1.0;
    ej = h * (m.ru_mw[tid] * T) * Yj;
    cj = cp * m.ru_mw[tid] * Yj;
  }
  e = team.sum(ej);
  cx = team.sum(cj);
}

// Temperature at which the mixture's mass enthalpy (or internal energy) is E:
// mechanism::getTemperatureFromHY / FromEY (mechanism.cpp:477-547), with their
// heat-capacity choices (the enthalpy iteration and its upper extrapolation
// use cv).  Every quantity is a team sum, so the loop is uniform.
template <class Team>
__device__ __forceinline__ double team_temperature_from_energy(const Team& team,
                                                                const DevMech& m, bool cvol,
                                                                double E, double Yj,
                                                                double guess) {
  const double tmin = 100.0, tmax = 5000.0;
  double emin, emax, c;
  team_mass_energy(team, m, cvol, cvol, tmin, Yj, emin, c);
  if (E < emin) return tmin - (emin - E) / c;
  team_mass_energy(team, m, cvol, true, tmax, Yj, emax, c);
  if (E > emax) return tmax - (emax - E) / c;
  double T = guess;
  if (T < tmin || T > tmax) T = tmin + (tmax - tmin) / (emax - emin) * (E - emin);
  for (int i = 0; i < 400; ++i) {
    double Ei;
    team_mass_energy(team, m, cvol, true, T, Yj, Ei, c);
    const double d = fmin(fmax((E - Ei) / c, -100.0), 100.0);
    if (fabs(d) < 1.0e-6 || T + d == T) break;
    T += d;
  }
  return T;
}

template <class Team>
__device__ __forceinline__ int rhs_team(const Team& team, const DevMech& m, const RhsLds& s,
                                        double P, double v0, double Tref, double L, double y,
                                        double& ydot, RhsAux& aux, RhsSrc src = {}) {
  const int lane = team.tid();
  const int nsp = m.nsp;
  const bool sp = lane < nsp;
#ifdef ZRK_PHASE_CYCLES
  const long long tc0 = (long long)__builtin_amdgcn_s_memtime();
#endif
  const double yT = team.from(y, nsp);
  double T = yT * Tref;
  if (T <= 0.0) {
    ydot = 0.0;
    return 1;
  }
  T = fmin(T, kTlimit);
  const double Yj = sp ? y : 0.0;
  const double imw = sp ? m.inv_mw[lane] : 0.0;
  const bool cvol = v0 > 0.0;
  // inv_rho: the reference's carried 1/density member; vd: this evaluation's
  double inv_rho = v0;
  if (!cvol) {
    if (src.has_ysrc) {
      inv_rho = src.vinit;  // mass sources: the density keeps its initial value
    } else {
      const double Pc = src.on ? P + src.dpdt * src.t : P;  // :91-94
      const double mws = team.sum(Yj * imw);
      const double wmix = 1.0 / mws;
      inv_rho = 1.0 / (Pc * wmix / (kRu * T));
    }
  }
  if (src.on && src.esrc != 0.0)  // reactor_constant_pressure_cpu.cpp:96-99 (volume :87-90)
    T = team_temperature_from_energy(team, m, cvol, src.e0 + src.esrc * src.t, Yj, T);
  double vd = inv_rho;
  if (cvol && src.on && src.dpdt != 0.0) {  // reactor_constant_volume_cpu.cpp:94-98
    const double wmix = 1.0 / team.sum(Yj * imw);
    const double Pc = kRu * T / (v0 * wmix) + src.dpdt * src.t;
    vd = 1.0 / (Pc * wmix / (kRu * T));
  }
  const TPow p = make_tpow(T);
  const double rho_c = 1.0 / vd;
  double cp = 0.0, h = 0.0, g = 0.0, Cj = 0.0, cpm = 0.0;
  if (sp) {
    thermo_one(m.thermo + 16 * lane, p, cp, h, g);
    if (cvol) {
      h -= 1.0;
      cp -= 1.0;
    }
    cpm = cp * m.ru_mw[lane];
    Cj = rho_c * imw * Yj;
    s.C[lane] = Cj;
    s.G[lane] = g;
  }
  if (lane == nsp) {
    s.C[nsp] = 1.0;
    s.G[nsp] = 0.0;
  }
  const double cp_mix = team.sum(cpm * Yj);
  const double Csum = team.sum(Cj);
  team.sync();
#ifdef ZRK_PHASE_CYCLES
  const long long tc1 = (long long)__builtin_amdgcn_s_memtime();
#endif
  team_rates<Team::kThreads, Team::kExt>(lane, m, s.C, s.G, s.rop, p, Csum, L);
  team.sync();
#ifdef ZRK_PHASE_CYCLES
  const long long tc2 = (long long)__builtin_amdgcn_s_memtime();
#endif
  double e = 0.0;
  double yd = 0.0;
  if (sp) {
    double cr, de;
    species_cd<Team::kExt>(m, s.rop, lane, cr, de);
    const double net = cr - de;
    if (src.has_ysrc) {
      yd = (net * m.mw[lane] * vd + src.ysrc);
      e = h * (net * vd + src.ysrc * imw);
    } else {
      yd = (net * m.mw[lane]) * vd;
      e = h * net * vd;
    }
  }
  const double esum = team.sum(e);
  if (lane == nsp)
    yd = -esum * kRu * yT / cp_mix +
         (src.on ? src.esrc + src.dpdt * vd : 0.0) / (cp_mix * Tref);
  ydot = yd;
  aux.inv_rho = inv_rho;
  aux.cp_mix = cp_mix;
  aux.h = h;
  aux.cpm = cpm;
  team.sync();
#ifdef ZRK_PHASE_CYCLES
  const long long tc3 = (long long)__builtin_amdgcn_s_memtime();
  aux.cyc[0] = tc1 - tc0;
  aux.cyc[1] = tc2 - tc1;
  aux.cyc[2] = tc3 - tc2;
#endif
  return 0;
}

// One reactor per wavefront (nsp <= 63).
template <bool kExt = false>
__device__ __forceinline__ int rhs64(const DevMech& m, const RhsLds& s, double P, double v0,
                                     double Tref, double L, double y, double& ydot,
                                     RhsAux& aux, RhsSrc src = {}) {
  return rhs_team(TeamWaveX<kExt>{}, m, s, P, v0, Tref, L, y, ydot, aux, src);
}

}  // namespace zrk
