// Per-reactor variable-order BDF (CVODE v5.8.0 algorithm) on one wavefront.
//
// Layout: lane j < nsp holds species j, lane nsp holds y_T = T/Tref (the
// reference state order), lanes above are padding (identically zero, zero
// error weight).  Every CVODE vector (Nordsieck history zn[0..5], ewt,
// acor, ...) is one register per lane; CVODE's scalars are wave-uniform.
//
// Linear algebra: the analytic Jacobian
// (applications/cfd_plugin/reactor_nvector_serial.cpp:610-771) is assembled
// column by column into the wave's LDS and saved to HBM for reuse; the Newton
// matrix M = I - gamma*J is inverted in registers by Gauss-Jordan elimination
// without pivoting (row per lane, columns rotated one register per step so
// every register index is a compile-time constant), and each Newton solve is
// then an LDS matrix-vector product.  This is the explicit-inverse scheme of
// the reference GPU path (interfaces/cublas_manager/cublas_manager.cu:73-91,
// getrf/getri without pivoting + block-diagonal matvec); a zero pivot is a
// recoverable setup failure, like a singular dense factorisation in CVODE.
//
// CVODE structure restated (names kept): cvStep, cvPredict, cvSet/cvSetBDF,
// cvSetTqBDF, cvNls + SUNNonlinSol_Newton, cvLsSetup/cvLsLinSys, cvLsSolve
// (with the 2/(1+gamrat) BDF correction), cvHandleNFlag, cvDoErrorTest,
// cvCompleteStep, cvPrepareNextStep, cvChooseEta, cvSetEta, cvAdjustOrder,
// cvRescale, cvRestore, cvHin, CVodeGetDky; the plugin's driver loop is
// applications/cfd_plugin/solver_cvode.cpp:220-299.
#pragma once

#include <cfloat>

#include "chem_device.h"

namespace zrk {

constexpr int kQmax = 5;

// ---------------------------------------------------------------------------
// Reference analytic Jacobian into LDS, column-major J[c*kNP + row], and a
// copy to HBM (savedJ).  inv_rho_prev is the inverse density of the last RHS
// call (at y), which the reference reuses for the 1/C factors
// (reactor_nvector_serial.cpp:625-637).  Returns the inverse density of the
// last RHS evaluated here (the reference's carried value).  t2l: kNP doubles
// of LDS scratch.  Jl is column-major with column stride ld >= nsp + 1 and
// only rows <= nsp are written (the term table holds no terms for lanes at
// or beyond nsp); saved is column-major with stride kNP.
//
// Species block: lane i walks row i's term list (sorted by column), so each
// entry has one writer and the loads of the list are coalesced across lanes;
// the temperature row and the MW_i/MW_j scaling follow column by column.
// Every entry is accumulated in FP64; MT = float stores the finished entries
// in FP32 (the FP32 preconditioner keeps its Jacobian in FP32).
// ---------------------------------------------------------------------------
//
// solveT = false (temperature held, not a state variable): the temperature
// row and column are zero and the finite-difference evaluation is skipped,
// as the reference's Jacobian without temperature has neither
// (reactor_nvector_serial.cpp:641-643, 691-750).
template <typename MT, bool kExt = false>
__device__ __forceinline__ double jacobian_lds(const DevMech& m, const RhsLds& s, double P,
                                               double v0, const BdfOpts& o, double y, double fy,
                                               double inv_rho_prev, MT* __restrict__ Jl, int ld,
                                               MT* __restrict__ saved,
                                               double* __restrict__ t2l, RhsSrc src = {},
                                               bool solveT = true) {
  const int lane = lane_id();
  const int nsp = m.nsp;
  const bool sp = lane < nsp;
  const bool act = lane <= nsp;
  const double mmf = o.min_mass_fraction;
  double t1 = y;
  if (sp && fabs(t1) < mmf) t1 = (t1 < 0.0) ? -mmf : mmf;
  const double t2 = sp ? m.mw[lane] / t1 * inv_rho_prev : 0.0;
  t2l[lane] = t2;
  // d(ydot)/dT by forward difference at the unfloored state.  Evaluated
  // first: it does not depend on the floored evaluation, and its inverse
  // density is the one carried on (it is the reference's last call).
  const double yT = bcast(y, nsp);
  double dT = yT * sqrt(DBL_EPSILON);
  const double Tp = yT + dT;
  dT = Tp - yT;
  const double t4in = (lane == nsp) ? Tp : (sp ? y : 0.0);
  double colT = 0.0;
  RhsAux aux2;
  if (solveT) {
    double t4;
    rhs64<kExt>(m, s, P, v0, o.Tref, o.step_limiter, t4in, t4, aux2, src);
    colT = act ? (t4 - fy) * (1.0 / dT) : 0.0;
  }
  // floored evaluation: rates of progress for every species term
  double t3;
  RhsAux aux;
  rhs64<kExt>(m, s, P, v0, o.Tref, o.step_limiter, t1, t3, aux, src);
  if (!solveT) aux2.inv_rho = aux.inv_rho;  // the last evaluation's
  const double Tdot = bcast(t3, nsp);
  const double RuTemp = kRu * yT;
  const double mult = -1.0 / aux.cp_mix;
  const double mwi = sp ? m.mw[lane] : 0.0;
  if (act)
    for (int c = 0; c < nsp; ++c) Jl[c * ld + lane] = MT(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  {
    // each product ROP*invC is rounded once, then added or subtracted as the
    // reference's J -= ROP*invC / J += ROP*invC, so balanced terms cancel exactly
    const uint4* __restrict__ jr = m.jr + lane;
    const int nb = m.jr_blocks;
    int cur = -1;
    double v = 0.0;
    auto term = [&](unsigned t) {
      const int col = (int)((t >> 16) & 0xFFu);
      if (col != 0xFF) {
        if (col != cur) {
          cur = col;
          v = 0.0;
        }
        if (kExt && (t & kJtNonInt)) {
          // (multiplier * ROP) / C_col (non_integer_reaction_network.cpp:493-506)
          const double2 e = m.ni_jt[t & 0xFFFFu];
          v = __dadd_rn(v, __dmul_rn(__dmul_rn(e.x, s.rop[(int)e.y]), t2l[col]));
        } else {
          const double p = __dmul_rn(s.rop[t & 0xFFFFu], t2l[col]);
          v = (t & 0x80000000u) ? __dsub_rn(v, p) : __dadd_rn(v, p);
        }
        Jl[col * ld + lane] = (MT)v;
      }
    };
    uint4 w = nb > 0 ? jr[0] : make_uint4(0, 0, 0, 0);
    for (int b = 0; b < nb; ++b) {
      const uint4 x = w;
      if (b + 1 < nb) w = jr[(b + 1) * kWave];
      term(x.x);
      term(x.y);
      term(x.z);
      term(x.w);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int c = 0; c < nsp; ++c) {
    const double v = sp ? (double)Jl[c * ld + lane] : 0.0;
    // d(Tdot)/dy_c from the concentration-space column
    double val = 0.0;
    if (solveT) {
      const double sc = wave_sum(v * aux.h);
      val = sc * RuTemp;
      val += Tdot * bcast(aux.cpm, c) * m.mw[c];
      val *= m.inv_mw[c] * mult;
    }
    const double out = sp ? v * (mwi * m.inv_mw[c]) : (lane == nsp ? val : 0.0);
    if (act) Jl[c * ld + lane] = (MT)out;
    saved[c * kNP + lane] = (MT)out;
  }
  if (act) Jl[nsp * ld + lane] = (MT)colT;
  saved[nsp * kNP + lane] = (MT)colT;
  return aux2.inv_rho;
}

// ---------------------------------------------------------------------------
// The same Jacobian for a team of several wavefronts (nsp < 64W): thread i
// owns row i; J and saved are column-major with column stride js; t2l and hl
// are kThreads doubles of LDS scratch.  Only rows <= nsp are written (rows
// beyond are left as they are), so js >= nsp + 1 suffices.  The temperature row is a
// matrix-vector product read from LDS (column c by thread c, conflict-free for
// an odd js) instead of one reduction per column.
// ---------------------------------------------------------------------------
template <typename MT, class Team>
__device__ __forceinline__ double jacobian_team(const Team& team, const DevMech& m,
                                                const RhsLds& s, double P, double v0,
                                                const BdfOpts& o, double y, double fy,
                                                double inv_rho_prev, MT* __restrict__ J,
                                                MT* __restrict__ saved, int js,
                                                double* __restrict__ t2l,
                                                double* __restrict__ hl, RhsSrc src = {},
                                                bool solveT = true) {
  const int tid = team.tid();
  const int nsp = m.nsp;
  const bool sp = tid < nsp;
  const bool act = tid <= nsp;
  const double mmf = o.min_mass_fraction;
  double t1 = y;
  if (sp && fabs(t1) < mmf) t1 = (t1 < 0.0) ? -mmf : mmf;
  t2l[tid] = sp ? m.mw[tid] / t1 * inv_rho_prev : 0.0;
  const double yT = team.from(y, nsp);
  double dT = yT * sqrt(DBL_EPSILON);
  const double Tp = yT + dT;
  dT = Tp - yT;
  double colT = 0.0;
  RhsAux aux2;
  if (solveT) {
    double t4;
    rhs_team(team, m, s, P, v0, o.Tref, o.step_limiter, (tid == nsp) ? Tp : (sp ? y : 0.0), t4,
             aux2, src);
    colT = act ? (t4 - fy) * (1.0 / dT) : 0.0;
  }
  double t3;
  RhsAux aux;
  rhs_team(team, m, s, P, v0, o.Tref, o.step_limiter, t1, t3, aux, src);
  if (!solveT) aux2.inv_rho = aux.inv_rho;  // the last evaluation's
  const double Tdot = team.from(t3, nsp);
  hl[tid] = sp ? aux.h : 0.0;
  if (tid < js)
    for (int c = 0; c < nsp; ++c) J[c * js + tid] = MT(0);
  team.sync();
  if (tid < m.jr_stride) {
    // rows beyond the table's stride have no terms (team larger than needed)
    const uint4* __restrict__ jr = m.jr + tid;
    const int nb = m.jr_blocks;
    const int js_tab = m.jr_stride;
    int cur = -1;
    double v = 0.0;
    auto term = [&](unsigned t) {
      const int col = (int)((t >> 16) & 0xFFu);
      if (col != 0xFF) {
        if (col != cur) {
          cur = col;
          v = 0.0;
        }
        if (Team::kExt && (t & kJtNonInt)) {
          const double2 e = m.ni_jt[t & 0xFFFFu];
          v = __dadd_rn(v, __dmul_rn(__dmul_rn(e.x, s.rop[(int)e.y]), t2l[col]));
        } else {
          const double p = __dmul_rn(s.rop[t & 0xFFFFu], t2l[col]);
          v = (t & 0x80000000u) ? __dsub_rn(v, p) : __dadd_rn(v, p);
        }
        J[col * js + tid] = (MT)v;
      }
    };
    uint4 w = nb > 0 ? jr[0] : make_uint4(0, 0, 0, 0);
    for (int b = 0; b < nb; ++b) {
      const uint4 x = w;
      if (b + 1 < nb) w = jr[(b + 1) * js_tab];
      term(x.x);
      term(x.y);
      term(x.z);
      term(x.w);
    }
  }
  team.sync();
  // temperature row: thread c forms d(Tdot)/dy_c from column c
  double trow = 0.0;
  if (sp && solveT) {
    double acc = 0.0;
    for (int i = 0; i < nsp; ++i) acc += (double)J[tid * js + i] * hl[i];
    trow = acc * (kRu * yT);
    trow += Tdot * aux.cpm * m.mw[tid];
    trow *= m.inv_mw[tid] * (-1.0 / aux.cp_mix);
  }
  team.sync();
  // row i of the species block: J_ic *= MW_i / MW_c; then the temperature row
  // (column c's entry at row nsp) and the finite-difference column
  if (sp) {
    const double mwi = m.mw[tid];
    for (int c = 0; c < nsp; ++c) {
      const MT out = (MT)((double)J[c * js + tid] * (mwi * m.inv_mw[c]));
      J[c * js + tid] = out;
      saved[c * js + tid] = out;
    }
    J[tid * js + nsp] = (MT)trow;
    saved[tid * js + nsp] = (MT)trow;
  }
  if (act) {
    J[nsp * js + tid] = (MT)colT;
    saved[nsp * js + tid] = (MT)colT;
  }
  team.sync();
  return aux2.inv_rho;
}

// Gauss-Jordan inverse (no pivoting) of the N x N leading block of a matrix
// held one row per lane in R[0..63]; padding rows/columns must be identity.
// MT is the working precision (double, or float for the FP32 preconditioner).
// Returns false on a zero pivot.
//
// The pivot row is broadcast from its lane with readlane (no LDS round trip)
// and is not normalised when it is used: row k is kept unscaled with a
// per-row factor s_k = 1/a_kk set at its pivot step, and every other row i
// subtracts g_i * (row k) with g_i = a_ik / a_kk, which does not depend on
// how row k is scaled.  The pivot column becomes -g_i (1 on the pivot row,
// whose factor carries 1/a_kk), and every row is multiplied by its factor at
// the end.  Per column and pivot that is one broadcast and one fma.
// Register indices must be compile-time constants, so pivots are taken in
// blocks of kGjBlock whose pivot columns sit in registers 0..kGjBlock-1, and
// the registers are rotated by kGjBlock after each block: on return register
// c holds logical column (c + gj_rotation(N)) mod 64 of the inverse.
constexpr int kGjBlock = 8;
__host__ __device__ constexpr int gj_rotation(int N) {
  return (((N + kGjBlock - 1) / kGjBlock) * kGjBlock) & (kNP - 1);
}

template <typename MT>
__device__ __forceinline__ bool gj_invert(MT (&R)[kNP], int N) {
  const int lane = lane_id();
  bool ok = true;
  MT s = MT(1);
  for (int k0 = 0; k0 < N; k0 += kGjBlock) {
#pragma unroll
    for (int kk = 0; kk < kGjBlock; ++kk) {
      const int k = k0 + kk;
      if (k < N) {
        const MT f = R[kk];
        const MT piv = bcast(f, k);
        if (piv == MT(0)) ok = false;
        const MT rp = MT(1) / piv;
        const bool me = (lane == k);
        const MT g = me ? MT(0) : f * rp;
        if (me) s = rp;
#pragma unroll
        for (int c = 0; c < kNP; ++c)
          if (c != kk) R[c] = fma(-g, bcast(R[c], k), R[c]);
        R[kk] = me ? MT(1) : -g;
      }
    }
    MT t[kGjBlock];
#pragma unroll
    for (int c = 0; c < kGjBlock; ++c) t[c] = R[c];
#pragma unroll
    for (int c = 0; c < kNP - kGjBlock; ++c) R[c] = R[c + kGjBlock];
#pragma unroll
    for (int c = 0; c < kGjBlock; ++c) R[kNP - kGjBlock + c] = t[c];
  }
#pragma unroll
  for (int c = 0; c < kNP; ++c) R[c] *= s;
  return ok;
}

// In-place Gauss-Jordan inverse (no pivoting) of the N x N matrix A held in
// LDS column-major with column stride js (thread i owns row i), for a team of
// several wavefronts.  Step k: every thread c scales entry (k, c) of the
// pivot row into prow[c]; then every row i != k subtracts f_i * prow from its
// row (f_i its column-k entry, read before the step) and sets its column-k
// entry to -f_i / pivot, while row k takes prow.  Returns false on a zero pivot.
template <typename MT, class Team>
__device__ __forceinline__ bool gj_invert_lds(const Team& team, MT* __restrict__ A, int js, int N,
                                              MT* __restrict__ prow) {
  const int tid = team.tid();
  const bool row = tid < N;
  bool ok = true;
  team.sync();
  for (int k = 0; k < N; ++k) {
    const MT piv = A[k * js + k];
    if (piv == MT(0)) ok = false;
    const MT rp = MT(1) / piv;
    const MT f = row ? A[k * js + tid] : MT(0);
    if (row) prow[tid] = (tid == k) ? rp : A[tid * js + k] * rp;
    team.sync();
    if (row) {
      if (tid == k) {
        for (int c = 0; c < N; ++c) A[c * js + tid] = prow[c];
      } else {
#pragma unroll 4
        for (int c = 0; c < N; ++c) {
          const MT a = A[c * js + tid];
          A[c * js + tid] = (c == k) ? -f * rp : fma(-f, prow[c], a);
        }
      }
    }
    team.sync();
  }
  return ok;
}

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inverse (no pivoting) on MFMA for a team of W
// wavefronts (FP32 tiles for the FP32 Newton matrix, FP64 tiles for the FP64
// one): the N x N matrix A (in LDS, or for FP64 teams in the team's slot of
// device memory, which stays in the last-level cache), column-major with
// column stride js
// (element (r, c) at A[c*js + r]), is inverted in place in 16 x 16 tiles.  The
// padding to a multiple of 16 is the identity and is never stored (tile
// loads make it up, stores beyond N are dropped).  Block step kb with the
// diagonal tile D = A[K,K]:
//   wavefront 0 inverts D in registers (16 rows on lanes 0..15) -> A[K,K];
//   row tiles      A[K,J] <- D^-1 A[K,J]                 (J != K)
//   other tiles    A[I,J] <- A[I,J] - A[I,K] A[K,J]      (I != K, J != K)
//                  A[I,K] <- -A[I,K] D^-1                (I != K)
// with the tiles of each phase spread over the wavefronts and a barrier
// between phases.  v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64 operand
// layouts (one 16 x 16 block):
//   A[i][k] (16 x 4): lane l holds i = l % 16, k = l / 16
//   B[k][j] (4 x 16): lane l holds k = l / 16, j = l % 16
//   C[i][j] (16 x 16): register v of lane l holds j = l % 16 and i = 4 (l / 16) + v
//                      (FP32) or i = l / 16 + 4 v (FP64)
// flag: one element of LDS for the zero-pivot result.  Returns false on a zero
// pivot of a diagonal-tile inversion.
// ---------------------------------------------------------------------------
// 16 x 16 x 4 MFMA by element type: v_mfma_f32_16x16x4f32 for the FP32
// Newton matrices, v_mfma_f64_16x16x4f64 for the FP64 ones.  Both take one A
// element (row lane % 16, k lane / 16) and one B element (k lane / 16, column
// lane % 16) per lane and accumulate the 16 x 16 tile as 4 elements per lane
// in column lane % 16, but in different rows (row(lk, v), lk = lane / 16):
// the FP32 instruction holds rows 4 lk + v, the FP64 one rows lk + 4 v
// (checked on gfx950 against a host product).
template <typename T>
struct Mfma16;
template <>
struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};
template <>
struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};

template <typename T>
__device__ __forceinline__ T tile_elem(const T* __restrict__ A, int js, int N, int r, int c) {
  return (r < N && c < N) ? A[c * js + r] : (r == c ? T(1) : T(0));
}

// reciprocal of a pivot: the FP32 inverse only preconditions the Newton
// iteration (its residual and convergence test stay in FP64), so the
// hardware reciprocal (1 ulp) serves; the FP64 inverse divides exactly
__device__ __forceinline__ float pivot_recip(float p) { return __builtin_amdgcn_rcpf(p); }
__device__ __forceinline__ double pivot_recip(double p) { return 1.0 / p; }

// In-register Gauss-Jordan inverse of a 16 x 16 matrix held one row per lane
// (lanes 0..15; other lanes compute values that are not used), with the
// deferred row scaling of gj_invert.
template <typename T>
__device__ __forceinline__ bool gj_invert16(T (&d)[16]) {
  const int lane = lane_id();
  bool ok = true;
  T s = T(1);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const T piv = bcast(d[k], k);
    if (piv == T(0)) ok = false;
    const T rp = pivot_recip(piv);
    const bool me = (lane == k);
    const T g = me ? T(0) : d[k] * rp;
    if (me) s = rp;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (c != k) d[c] = fma(-g, bcast(d[c], k), d[c]);
    d[k] = me ? T(1) : -g;
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) d[c] *= s;
  return ok;
}

// C (16 x 16 tile of A at rows I0.., columns J0..) += a * (tile at rows K0..,
// columns J0..): every access of a lane is in its column J0 + (lane % 16).
// full: the tiles lie inside the N x N matrix (uniform); otherwise elements
// beyond N read as the identity and are not stored.
template <bool kFull, typename T>
__device__ __forceinline__ void mfma_tile_update(T* __restrict__ A, int js, int N, int I0,
                                                 int J0, int K0, const T (&a)[4]) {
  const int lane = lane_id();
  const int li = lane & 15, lk = lane >> 4;
  const int c = J0 + li;
  T* __restrict__ col = A + c * js;
  typename Mfma16<T>::v4 acc;
  T b[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = I0 + Mfma16<T>::row(lk, v);
    acc[v] = kFull ? col[r] : tile_elem(A, js, N, r, c);
  }
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4)
    b[s4] = kFull ? col[K0 + 4 * s4 + lk] : tile_elem(A, js, N, K0 + 4 * s4 + lk, c);
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) acc = Mfma16<T>::mac(a[s4], b[s4], acc);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = I0 + Mfma16<T>::row(lk, v);
    if (kFull || (r < N && c < N)) col[r] = acc[v];
  }
}

// Two full-tile updates (columns J0.. and J1..) interleaved: all loads first,
// then the two independent MFMA chains, then the stores.
template <typename T>
__device__ __forceinline__ void mfma_tile_update2(T* __restrict__ A, int js, int I0, int J0,
                                                  int J1, int K0, const T (&a)[4]) {
  const int lane = lane_id();
  const int li = lane & 15, lk = lane >> 4;
  T* __restrict__ c0 = A + (J0 + li) * js;
  T* __restrict__ c1 = A + (J1 + li) * js;
  typename Mfma16<T>::v4 x0, x1;
  T b0[4], b1[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    x0[v] = c0[I0 + Mfma16<T>::row(lk, v)];
    x1[v] = c1[I0 + Mfma16<T>::row(lk, v)];
  }
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) {
    b0[s4] = c0[K0 + 4 * s4 + lk];
    b1[s4] = c1[K0 + 4 * s4 + lk];
  }
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) {
    x0 = Mfma16<T>::mac(a[s4], b0[s4], x0);
    x1 = Mfma16<T>::mac(a[s4], b1[s4], x1);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    c0[I0 + Mfma16<T>::row(lk, v)] = x0[v];
    c1[I0 + Mfma16<T>::row(lk, v)] = x1[v];
  }
}

// C (tile at rows R0.., columns J0..) = a * (tile at rows K0.., columns J0..)
template <bool kFull, typename T>
__device__ __forceinline__ void mfma_tile_product(T* __restrict__ A, int js, int N, int R0,
                                                  int J0, int K0, const T (&a)[4]) {
  const int lane = lane_id();
  const int li = lane & 15, lk = lane >> 4;
  const int c = J0 + li;
  T* __restrict__ col = A + c * js;
  T b[4];
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4)
    b[s4] = kFull ? col[K0 + 4 * s4 + lk] : tile_elem(A, js, N, K0 + 4 * s4 + lk, c);
  typename Mfma16<T>::v4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) acc = Mfma16<T>::mac(a[s4], b[s4], acc);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = R0 + Mfma16<T>::row(lk, v);
    if (kFull || (r < N && c < N)) col[r] = acc[v];
  }
}

// the A operand of tile (rows I0.., columns K0..), times sign
template <typename T>
__device__ __forceinline__ void mfma_a_operand(const T* __restrict__ A, int js, int N, int I0,
                                               int K0, T sign, T (&a)[4]) {
  const int lane = lane_id();
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) a[s4] = sign * tile_elem(A, js, N, I0 + li, K0 + 4 * s4 + lk);
}

template <class Team, typename T>
__device__ __forceinline__ bool gj_invert_mfma(const Team& team, T* __restrict__ A, int js,
                                               int N, T* __restrict__ flag) {
  constexpr int kWaves = Team::kThreads / kWave;
  const int wave = team.tid() / kWave;
  const int lane = lane_id();
  const int li = lane & 15;
  const int nb = (N + 15) >> 4;
  const int nfull = N >> 4;  // tiles 0..nfull-1 lie inside the matrix
  if (team.tid() == 0) *flag = T(1);
  team.sync();
  for (int kb = 0; kb < nb; ++kb) {
    const int K0 = kb * 16;
    if (wave == 0) {
      T d[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) d[c] = tile_elem(A, js, N, K0 + li, K0 + c);
      const bool ok = gj_invert16(d);
      if (lane < 16 && K0 + li < N) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
          if (K0 + c < N) A[(K0 + c) * js + K0 + li] = d[c];
      }
      if (!ok && lane == 0) *flag = T(0);
    }
    team.sync();
    // row tiles: A[K,J] <- D^-1 A[K,J]
    T dk[4];
    mfma_a_operand(A, js, N, K0, K0, T(1), dk);
    for (int jb = wave; jb < nb; jb += kWaves) {
      if (jb == kb) continue;
      if (kb < nfull && jb < nfull)
        mfma_tile_product<true>(A, js, N, K0, jb * 16, K0, dk);
      else
        mfma_tile_product<false>(A, js, N, K0, jb * 16, K0, dk);
    }
    team.sync();
    // other row blocks: A[I,J] -= A[I,K] A[K,J], then A[I,K] = -A[I,K] D^-1
    for (int ib = wave; ib < nb; ib += kWaves) {
      if (ib == kb) continue;
      const int I0 = ib * 16;
      T ak[4];
      mfma_a_operand(A, js, N, I0, K0, T(-1), ak);
      const bool rin = ib < nfull && kb < nfull;
      // full tiles in pairs (skipping column block kb), then the rest one by one
      int jb = 0;
      if (rin) {
        for (;;) {
          while (jb == kb) ++jb;
          int jb1 = jb + 1;
          while (jb1 == kb) ++jb1;
          if (jb1 >= nfull) break;
          mfma_tile_update2(A, js, I0, jb * 16, jb1 * 16, K0, ak);
          jb = jb1 + 1;
        }
      }
      for (; jb < nb; ++jb) {
        if (jb == kb) continue;
        if (rin && jb < nfull)
          mfma_tile_update<true>(A, js, N, I0, jb * 16, K0, ak);
        else
          mfma_tile_update<false>(A, js, N, I0, jb * 16, K0, ak);
      }
      if (rin)
        mfma_tile_product<true>(A, js, N, I0, K0, K0, ak);
      else
        mfma_tile_product<false>(A, js, N, I0, K0, K0, ak);
    }
    team.sync();
  }
  return *flag != T(0);
}

}  // namespace zrk
