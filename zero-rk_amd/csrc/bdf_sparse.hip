// Sparse-path integrator for large mechanisms (any size; the dense paths stop
// at 191 species): CVODE v5.8.0's BDF (cvode.c, cvode_ls.c,
// sunnonlinsol_newton.c, as restated for one reactor per wavefront in
// cvode_wave.h) for one reactor per workgroup of kSpWaves wavefronts, with
//   * the reactor's vectors (Nordsieck history, weights, corrections, the
//     RHS scratch and the rates of progress) in its own slot of device
//     memory: a thread owns components i = tid (mod threads) of every vector,
//     so vector updates need no barrier;
//   * the analytic Jacobian (reactor_nvector_serial.cpp:610-771) assembled at
//     the LU positions of a static sparse plan (sparse_plan.h) from per-
//     position term lists, the temperature row from its column sums and the
//     temperature column by a forward difference, FP64;
//   * the Newton matrix I - gamma J rounded to FP32 and factorised in place
//     without pivoting, row by row (up-looking) over the plan's level sets,
//     one wavefront per row, each row scattered into a dense FP32 row of LDS;
//   * the two triangular solves over the same level sets, FP32 factors with
//     FP64 vectors and sums.
// Constant pressure and constant volume reactors; no source terms, ignition
// root or extended (PLOG / non-integer / SRI) records on this path.
#include <hip/hip_runtime.h>

#include "cvode_wave.h"
#include "launch.h"
#include "sparse_plan.h"

namespace zrk {

constexpr int kSpWaves = kSparseWaves;
constexpr int kSpThreads = kSpWaves * kWave;
static_assert(kSpWaves >= 2, "the solves' N doubles live in the kSpWaves LDS rows of N floats");
constexpr int kSpQmax = 5;
using SpTeam = TeamBlock<kSpWaves, false>;

// doubles of one reactor's slot: 16 vectors of N, C, G, h, cp of N, the rates
// of progress (nstep + 1), the Jacobian (nnz), the FP32 LU (nnz floats) and
// the FP32 dense trailing block (nd x nd floats)
__host__ __device__ inline size_t sparse_slot_doubles(int nsp, int nstep, int nnz, int nd) {
  const size_t N = (size_t)nsp + 1;
  const size_t d = 20 * N + (size_t)nstep + 1 + (size_t)nnz + ((size_t)nnz + 1) / 2 +
                   ((size_t)nd * nd + 1) / 2 + 1;
  return (d + 31) & ~(size_t)31;
}

// -DZRK_PHASE_CYCLES: per-reactor cycle counts of the phases (RHS, Jacobian,
// Newton-matrix setup and factorisation, solves, whole reactor; then the
// factorisation's leading levels, trailing rows and dense block) in stats
// slots 8..15 of a 20-int layout, as the dense integrators report theirs
#ifdef ZRK_PHASE_CYCLES
#define SP_TIC(t) const long long t = (long long)__builtin_amdgcn_s_memtime()
#define SP_TOC(t, i) cyc[i] += (long long)__builtin_amdgcn_s_memtime() - t
#else
#define SP_TIC(t)
#define SP_TOC(t, i)
#endif

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SpCv {
  SpTeam team;
  DevMech m;
  DevSparse S;
  BdfOpts o;
  int nsp, N;
  // per-lane and per-wavefront indices (the object itself is shared by the
  // wavefront's lanes in LDS)
  __device__ __forceinline__ int tid() const { return (int)threadIdx.x; }
  __device__ __forceinline__ int wave() const { return (int)threadIdx.x / kWave; }
  __device__ __forceinline__ int lane() const { return lane_id(); }
  float* w;  // this wavefront's dense LU row (LDS, N floats, zero between rows)
  double* xs;  // the solves' work vector (N doubles): the LDS of the rows w, which
               // only the factorisation uses; it zeroes its row on entry
  double *zn[kSpQmax + 1], *ewt, *acor, *tempv, *ftemp, *y, *delta, *t1, *t2, *t3, *x;
  double *C, *G, *H, *CPM, *rop, *J;
  float* lu;
  float* D;  // dense trailing block, row-major nd x nd
  // reactor
  double P, v0, inv_rho_last, cp_last;
  // CVODE (names as in cvode.c)
  double rtol, atol, tn, h, hscale, hprime, next_h, eta, etamax, hmin, hmax_inv, hu, h0u;
  double tau[kSpQmax + 2], tq[6], l[kSpQmax + 1];
  double rl1, gamma, gammap, gamrat, crate, delp, acnrm, nlscoef, saved_tq5, tolsf;
  double etaq, etaqm1, etaqp1;
  int q, qprime, next_q, qwait, L, qu, indx_acor;
  int nst, nstlp, nstlj, nfe, nje, nsetups, nni, ncfn, netf, nscon;
  int convfail, jcur, acnrmcur;
#ifdef ZRK_PHASE_CYCLES
  long long cyc[8];
#endif

  __device__ void bind(double* slot) {
    double* p = slot;
    for (int j = 0; j <= kSpQmax; ++j, p += N) zn[j] = p;
    ewt = p; p += N;
    acor = p; p += N;
    tempv = p; p += N;
    ftemp = p; p += N;
    y = p; p += N;
    delta = p; p += N;
    t1 = p; p += N;
    t2 = p; p += N;
    t3 = p; p += N;
    x = p; p += N;
    C = p; p += N;
    G = p; p += N;
    H = p; p += N;
    CPM = p; p += N;
    rop = p; p += m.nstep + 1;
    J = p; p += S.nnz;
    lu = reinterpret_cast<float*>(p);
    p += ((size_t)S.nnz + 1) / 2;
    D = reinterpret_cast<float*>(p);
  }

  __device__ __forceinline__ double wrms(const double* v) {
    double s = 0.0;
    for (int i = tid(); i < N; i += kSpThreads) {
      const double p = v[i] * ewt[i];
      s += p * p;
    }
    return sqrt(team.sum(s) / N);
  }

  // ReactorConstantPressureCPU / ReactorConstantVolumeCPU::GetTimeDerivative
  // (reactor_constant_pressure_cpu.cpp:68-133) without source terms
  __device__ int rhs(const double* yy, double* yd) {
    team.sync();
    const double yT = yy[nsp];
    double T = yT * o.Tref;
    if (T <= 0.0) return 1;
    T = fmin(T, kTlimit);
    double inv_rho = v0;
    if (!(v0 > 0.0)) {
      double part = 0.0;
      for (int j = tid(); j < nsp; j += kSpThreads) part += yy[j] * m.inv_mw[j];
      const double wmix = 1.0 / team.sum(part);
      inv_rho = 1.0 / (P * wmix / (kRu * T));
    }
    const bool cvol = v0 > 0.0;
    const double vd = inv_rho, rho_c = 1.0 / vd;
    const TPow p = make_tpow(T);
    double cps = 0.0, cs = 0.0;
    for (int j = tid(); j < nsp; j += kSpThreads) {
      double cp, hh, g;
      thermo_one(m.thermo + 16 * j, p, cp, hh, g);
      if (cvol) {
        hh -= 1.0;
        cp -= 1.0;
      }
      const double cpm = cp * m.ru_mw[j];
      const double Cj = rho_c * m.inv_mw[j] * yy[j];
      C[j] = Cj;
      G[j] = g;
      H[j] = hh;
      CPM[j] = cpm;
      cps += cpm * yy[j];
      cs += Cj;
    }
    if (tid() == 0) {
      C[nsp] = 1.0;
      G[nsp] = 0.0;
    }
    const double cp_mix = team.sum(cps);
    const double Csum = team.sum(cs);  // the sums end with a barrier: C, G visible
    team_rates<kSpThreads, false>(tid(), m, C, G, rop, p, Csum, o.step_limiter);
    team.sync();
    double e = 0.0;
    for (int j = tid(); j < nsp; j += kSpThreads) {
      double cr, de;
      species_cd<false>(m, rop, j, cr, de);
      const double net = cr - de;
      yd[j] = (net * m.mw[j]) * vd;
      e += H[j] * net * vd;
    }
    const double esum = team.sum(e);
    if (tid() == nsp % kSpThreads) yd[nsp] = -esum * kRu * yT / cp_mix;
    inv_rho_last = inv_rho;
    cp_last = cp_mix;
    return 0;
  }

  __device__ __forceinline__ int f(const double* yy, double* yd) {
    nfe++;
    SP_TIC(t0);
    const int rv = rhs(yy, yd);
    SP_TOC(t0, 0);
    return rv;
  }

  // the reference's Jacobian at (yp, fp) into J at the plan's positions
  __device__ void jacobian(const double* yp, const double* fp) {
    const double mmf = o.min_mass_fraction;
    team.sync();
    for (int i = tid(); i < nsp; i += kSpThreads) {
      double v = yp[i];
      if (fabs(v) < mmf) v = (v < 0.0) ? -mmf : mmf;
      t1[i] = v;
      t2[i] = m.mw[i] / v * inv_rho_last;
    }
    if (tid() == nsp % kSpThreads) t1[nsp] = yp[nsp];
    (void)rhs(t1, t3);
    team.sync();
    for (int p = tid(); p < S.nnz; p += kSpThreads) {
      const int r = S.pos_row[p], c = S.pos_col[p];
      double acc = 0.0;
      if (r < nsp && c < nsp) {
        const double tc = t2[c];
        for (int k = S.jt_ptr[p]; k < S.jt_ptr[p + 1]; ++k) {
          const uint32_t t = S.jt[k];
          const double v = rop[t & 0x7FFFFFFFu] * tc;
          acc = (t >> 31) ? acc - v : acc + v;
        }
      }
      J[p] = acc;
    }
    team.sync();
    // d(Tdot)/dy_j from the species columns, then the heat-capacity term
    const double RuTemp = kRu * yp[nsp];
    const double mult = -1.0 / cp_last;
    const double tdot = t3[nsp];
    for (int j = tid(); j < nsp; j += kSpThreads) {
      double acc = 0.0;
      for (int k = S.cp_ptr[j]; k < S.cp_ptr[j + 1]; ++k) {
        const int p = S.cp_pos[k];
        acc += J[p] * H[S.pos_row[p]];
      }
      double v = acc * RuTemp;
      v += tdot * CPM[j] * m.mw[j];
      v *= (1.0 / m.mw[j]) * mult;
      J[S.trow_pos[j]] = v;
    }
    // d(ydot)/dT by a forward difference
    const double yT = yp[nsp];
    double dT = yT * 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)
    const double tT = yT + dT;
    dT = tT - yT;
    const double inv = 1.0 / dT;
    team.sync();
    for (int i = tid(); i < nsp; i += kSpThreads) t1[i] = yp[i];
    if (tid() == nsp % kSpThreads) t1[nsp] = tT;
    (void)rhs(t1, t3);
    team.sync();
    for (int k = tid(); k < N; k += kSpThreads) J[S.tcol_pos[k]] = (t3[k] - fp[k]) * inv;
    // concentration -> mass-fraction columns
    for (int p = tid(); p < S.nnz; p += kSpThreads) {
      const int r = S.pos_row[p], c = S.pos_col[p];
      if (r < nsp && c < nsp) J[p] *= m.mw[r] * (1.0 / m.mw[c]);
    }
    team.sync();
  }

  // row i of the LU by one wavefront (up-looking): scatter into w, eliminate
  // the lower-part columns below kend in ascending order.  The eliminations
  // form a chain (each multiplier reads w after the previous updates), but
  // what they read of the finished rows k (pivot, and each lane's first
  // entry of row k's upper part) does not depend on the chain: it is loaded
  // kPre columns at a time before the chain runs through them, so the chain
  // waits for memory once per kPre columns instead of once per column.
  static constexpr int kPre = 8;
  __device__ __forceinline__ void eliminate_row(int i, int kend) {
    const int b = S.rp[i], e = S.rp[i + 1];
    {
      // a hub row's hundreds of entries: four chunks of 64 loaded together
      int p = b + lane();
      for (; p + 3 * kWave < e; p += 4 * kWave) {
        int c[4];
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          c[u] = S.ci[p + u * kWave];
          v[u] = lu[p + u * kWave];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) w[c[u]] = v[u];
      }
      for (; p < e; p += kWave) w[S.ci[p]] = lu[p];
    }
    wave_sync();
    for (int p0 = b; p0 < e; p0 += kPre) {
      int kk[kPre], dkk[kPre], ekk[kPre], uc[kPre];
      float pv[kPre], uv[kPre];
#pragma unroll
      for (int j = 0; j < kPre; ++j) {
        const int p = p0 + j;
        const int k = p < e ? S.ci[p] : kend;
        kk[j] = k < kend ? k : -1;
        dkk[j] = ekk[j] = uc[j] = 0;
        pv[j] = uv[j] = 0.0f;
        if (k < kend) {
          const int dk = S.dpos[k], ek = S.rp[k + 1], q2 = dk + 1 + lane();
          dkk[j] = dk;
          ekk[j] = ek;
          pv[j] = lu[dk];
          if (q2 < ek) {
            uv[j] = lu[q2];
            uc[j] = S.ci[q2];
          }
        }
      }
      bool done = false;
#pragma unroll
      for (int j = 0; j < kPre; ++j) {
        if (!done) {
          const int k = kk[j];
          if (k < 0) {
            done = true;
          } else {
            const float lk = w[k] / pv[j];
            wave_sync();
            if (lane() == 0) w[k] = lk;
            if (dkk[j] + 1 + lane() < ekk[j]) w[uc[j]] -= lk * uv[j];
            for (int q2 = dkk[j] + 1 + kWave + lane(); q2 < ekk[j]; q2 += kWave)
              w[S.ci[q2]] -= lk * lu[q2];
            wave_sync();
          }
        }
      }
      if (done) break;
    }
  }

  // in-place LU of the FP32 matrix in lu; false on a zero pivot.  Leading
  // rows (< split) row by row over their level sets; trailing rows against
  // the leading columns (independent of each other) into the dense block;
  // the block by a right-looking dense LU with the whole workgroup; the
  // block back into the CSR positions (its other entries stay exactly zero)
  __device__ bool factor() {
    const int sp = S.split, nd = N - sp;
    bool ok = true;
    for (int i = lane(); i < N; i += kWave) w[i] = 0.0f;  // the solves reuse this LDS
    wave_sync();
    SP_TIC(f0);
    for (int lev = 0; lev < S.nlev_l; ++lev) {
      for (int idx = S.ll_ptr[lev] + wave(); idx < S.ll_ptr[lev + 1]; idx += kSpWaves) {
        const int i = S.ll_rows[idx];
        eliminate_row(i, i);
        if (w[i] == 0.0f) ok = false;
        for (int p = S.rp[i] + lane(); p < S.rp[i + 1]; p += kWave) {
          const int c = S.ci[p];
          lu[p] = w[c];
          w[c] = 0.0f;
        }
        wave_sync();
      }
      team.sync();
    }
    SP_TOC(f0, 5);
    SP_TIC(f1);
    for (int i = sp + wave(); i < N; i += kSpWaves) {
      eliminate_row(i, sp);
      {
        const int e = S.rp[i + 1];
        int p = S.rp[i] + lane();
        for (; p + 3 * kWave < e; p += 4 * kWave) {
          int c[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = S.ci[p + u * kWave];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (c[u] < sp) lu[p + u * kWave] = w[c[u]];
        }
        for (; p < e; p += kWave) {
          const int c = S.ci[p];
          if (c < sp) lu[p] = w[c];
        }
      }
      float* Di = D + (size_t)(i - sp) * nd;
      for (int c = sp + lane(); c < N; c += kWave) Di[c - sp] = w[c];
      wave_sync();
      {
        const int e = S.rp[i + 1];
        int p = S.rp[i] + lane();
        for (; p + 3 * kWave < e; p += 4 * kWave) {
          int c[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = S.ci[p + u * kWave];
#pragma unroll
          for (int u = 0; u < 4; ++u) w[c[u]] = 0.0f;
        }
        for (; p < e; p += kWave) w[S.ci[p]] = 0.0f;
      }
      wave_sync();
    }
    team.sync();
    SP_TOC(f1, 6);
    SP_TIC(f2);
    for (int k = 0; k < nd; ++k) {
      const float* Dk = D + (size_t)k * nd;
      const float piv = Dk[k];
      if (piv == 0.0f) ok = false;
      for (int i = k + 1 + wave(); i < nd; i += kSpWaves) {
        float* Di = D + (size_t)i * nd;
        const float a = Di[k];
        if (a == 0.0f) continue;
        const float li = a / piv;
        for (int j = k + 1 + lane(); j < nd; j += kWave) Di[j] -= li * Dk[j];
        if (lane() == 0) Di[k] = li;
      }
      team.sync();
    }
    for (int i = sp + wave(); i < N; i += kSpWaves) {
      const float* Di = D + (size_t)(i - sp) * nd;
      for (int p = S.rp[i] + lane(); p < S.rp[i + 1]; p += kWave) {
        const int c = S.ci[p];
        if (c >= sp) lu[p] = Di[c - sp];
      }
    }
    team.sync();
    SP_TOC(f2, 7);
    return team.max(ok ? 0.0 : 1.0) == 0.0;
  }

  // b <- (LU)^-1 b in the plan's order.  Leading rows over their level sets
  // (one wavefront per row); the trailing block's rows against the leading
  // columns; the block densely in column panels of 64 (the panel's triangle
  // by one wavefront with shuffles, the rows beyond it by all wavefronts)
  __device__ void solve(double* b) {
    const int sp = S.split, nd = N - sp;
    team.sync();
    for (int i = tid(); i < N; i += kSpThreads) xs[i] = b[S.perm[i]];
    // the level-set sweeps are chains of levels; the factor entries a
    // wavefront needs at the next level do not depend on this one, so the
    // first 64 of its first row there are loaded before this level's
    // barrier (the sums keep their order: the same bits)
    int nx_i = -1, nx_c = 0;
    float nx_v = 0.0f;
    auto next_lower = [&](int lev) {
      nx_i = -1;
      if (lev >= S.nlev_l) return;
      const int idx = S.ll_ptr[lev] + wave();
      if (idx >= S.ll_ptr[lev + 1]) return;
      nx_i = S.ll_rows[idx];
      const int p = S.rp[nx_i] + lane();
      nx_v = 0.0f;
      nx_c = 0;
      if (p < S.dpos[nx_i]) {
        nx_v = lu[p];
        nx_c = S.ci[p];
      }
    };
    next_lower(0);
    team.sync();
    for (int lev = 0; lev < S.nlev_l; ++lev) {
      const int i0 = nx_i, c0 = nx_c;
      const float v0 = nx_v;
      next_lower(lev + 1);
      for (int idx = S.ll_ptr[lev] + wave(); idx < S.ll_ptr[lev + 1]; idx += kSpWaves) {
        const bool first = idx == S.ll_ptr[lev] + wave();
        const int i = first ? i0 : S.ll_rows[idx];
        double s = 0.0;
        int p = S.rp[i] + lane();
        if (first) {
          if (p < S.dpos[i]) s += (double)v0 * xs[c0];
          p += kWave;
        }
        for (; p < S.dpos[i]; p += kWave) s += (double)lu[p] * xs[S.ci[p]];
        s = wave_sum(s);
        if (lane() == 0) xs[i] -= s;
      }
      team.sync();
    }
    for (int i = sp + wave(); i < N; i += kSpWaves) {
      // the hub rows hold hundreds of entries: four chunks of 64 are loaded
      // together before they are summed (in the same order: the same bits)
      double s = 0.0;
      const int e = S.rp[i + 1];
      int p = S.rp[i] + lane();
      for (; p + 3 * kWave < e; p += 4 * kWave) {
        int c[4];
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          c[u] = S.ci[p + u * kWave];
          v[u] = lu[p + u * kWave];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (c[u] < sp) s += (double)v[u] * xs[c[u]];
      }
      for (; p < e; p += kWave)
        if (S.ci[p] < sp) s += (double)lu[p] * xs[S.ci[p]];
      s = wave_sum(s);
      if (lane() == 0) xs[i] -= s;
    }
    team.sync();
    double* xb = xs + sp;
    for (int kb = 0; kb < nd; kb += kWave) {  // forward, unit lower
      const int kn = min(kWave, nd - kb);
      if (wave() == 0) {
        double xv = lane() < kn ? xb[kb + lane()] : 0.0;
        for (int c = 0; c < kn; ++c) {
          const double xc = __shfl(xv, c);
          if (lane() > c && lane() < kn) xv -= (double)D[(size_t)(kb + lane()) * nd + kb + c] * xc;
        }
        if (lane() < kn) xb[kb + lane()] = xv;
      }
      team.sync();
      const double xl = lane() < kn ? xb[kb + lane()] : 0.0;
      int j = kb + kn + wave();
      for (; j + kSpWaves < nd; j += 2 * kSpWaves) {  // two rows' loads together
        const float d0 = lane() < kn ? D[(size_t)j * nd + kb + lane()] : 0.0f;
        const float d1 = lane() < kn ? D[(size_t)(j + kSpWaves) * nd + kb + lane()] : 0.0f;
        double s0 = lane() < kn ? (double)d0 * xl : 0.0;
        double s1 = lane() < kn ? (double)d1 * xl : 0.0;
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        if (lane() == 0) {
          xb[j] -= s0;
          xb[j + kSpWaves] -= s1;
        }
      }
      for (; j < nd; j += kSpWaves) {
        double s = lane() < kn ? (double)D[(size_t)j * nd + kb + lane()] * xl : 0.0;
        s = wave_sum(s);
        if (lane() == 0) xb[j] -= s;
      }
      team.sync();
    }
    for (int kb = ((nd - 1) / kWave) * kWave; nd > 0 && kb >= 0; kb -= kWave) {  // backward
      const int kn = min(kWave, nd - kb);
      if (wave() == 0) {
        double xv = lane() < kn ? xb[kb + lane()] : 0.0;
        for (int c = kn - 1; c >= 0; --c) {
          if (lane() == c) xv /= (double)D[(size_t)(kb + c) * nd + kb + c];
          const double xc = __shfl(xv, c);
          if (lane() < c) xv -= (double)D[(size_t)(kb + lane()) * nd + kb + c] * xc;
        }
        if (lane() < kn) xb[kb + lane()] = xv;
      }
      team.sync();
      const double xl = lane() < kn ? xb[kb + lane()] : 0.0;
      int j = wave();
      for (; j + kSpWaves < kb; j += 2 * kSpWaves) {  // two rows' loads together
        const float d0 = lane() < kn ? D[(size_t)j * nd + kb + lane()] : 0.0f;
        const float d1 = lane() < kn ? D[(size_t)(j + kSpWaves) * nd + kb + lane()] : 0.0f;
        double s0 = lane() < kn ? (double)d0 * xl : 0.0;
        double s1 = lane() < kn ? (double)d1 * xl : 0.0;
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        if (lane() == 0) {
          xb[j] -= s0;
          xb[j + kSpWaves] -= s1;
        }
      }
      for (; j < kb; j += kSpWaves) {
        double s = lane() < kn ? (double)D[(size_t)j * nd + kb + lane()] * xl : 0.0;
        s = wave_sum(s);
        if (lane() == 0) xb[j] -= s;
      }
      team.sync();
    }
    // backward over the leading rows, with the next level's first row of
    // each wavefront loaded ahead as in the forward sweep
    float nx_d = 1.0f;
    auto next_upper = [&](int lev) {
      nx_i = -1;
      if (lev >= S.nlev_lb) return;
      const int idx = S.lbl_ptr[lev] + wave();
      if (idx >= S.lbl_ptr[lev + 1]) return;
      nx_i = S.lbl_rows[idx];
      const int dp = S.dpos[nx_i];
      const int p = dp + 1 + lane();
      nx_v = 0.0f;
      nx_c = 0;
      nx_d = lu[dp];
      if (p < S.rp[nx_i + 1]) {
        nx_v = lu[p];
        nx_c = S.ci[p];
      }
    };
    next_upper(0);
    for (int lev = 0; lev < S.nlev_lb; ++lev) {
      const int i0 = nx_i, c0 = nx_c;
      const float v0 = nx_v, d0 = nx_d;
      next_upper(lev + 1);
      for (int idx = S.lbl_ptr[lev] + wave(); idx < S.lbl_ptr[lev + 1]; idx += kSpWaves) {
        const bool first = idx == S.lbl_ptr[lev] + wave();
        const int i = first ? i0 : S.lbl_rows[idx];
        double s = 0.0;
        int p = S.dpos[i] + 1 + lane();
        if (first) {
          if (p < S.rp[i + 1]) s += (double)v0 * xs[c0];
          p += kWave;
        }
        for (; p < S.rp[i + 1]; p += kWave) s += (double)lu[p] * xs[S.ci[p]];
        s = wave_sum(s);
        const float piv = first ? d0 : lu[S.dpos[i]];
        if (lane() == 0) xs[i] = (xs[i] - s) / (double)piv;
      }
      team.sync();
    }
    for (int i = tid(); i < N; i += kSpThreads) b[S.perm[i]] = xs[i];
    team.sync();
  }

  // ---- CVODE (cvode.c) ----
  __device__ bool set_ewt(const double* yy) {
    double bad = 0.0;
    for (int i = tid(); i < N; i += kSpThreads) {
      const double v = rtol * fabs(yy[i]) + atol;
      if (v <= 0.0) bad = 1.0;
      else ewt[i] = 1.0 / v;
    }
    return team.max(bad) == 0.0;
  }

  __device__ void reinit(double t0, const double* y0) {
    if (y0 != zn[0])
      for (int i = tid(); i < N; i += kSpThreads) zn[0][i] = y0[i];
    tn = t0;
    q = 1;
    L = 2;
    qwait = L;
    etamax = kEtamx1;
    qu = 0;
    hu = 0.0;
    tolsf = 1.0;
    nst = nfe = ncfn = netf = nni = nsetups = 0;
    nstlp = nstlj = nje = nscon = 0;
    hmin = 0.0;
    h0u = 0.0;
    next_h = 0.0;
    next_q = 0;
    indx_acor = kSpQmax;
    saved_tq5 = 0.0;
    crate = 1.0;
    delp = 0.0;
    gammap = 0.0;
    gamrat = 1.0;
    for (int i = 0; i < kSpQmax + 2; ++i) tau[i] = 0.0;
    for (int i = 0; i < 6; ++i) tq[i] = 0.0;
    for (int i = 0; i <= kSpQmax; ++i) l[i] = 0.0;
  }

  __device__ double upper_bound_h0(double tdist) {
    double hub_inv = 0.0;
    for (int i = tid(); i < N; i += kSpThreads) {
      const double a = kHubFactor * fabs(zn[0][i]) + 1.0 / ewt[i];
      const double b = fabs(zn[1][i]) / a;
      if (b > hub_inv) hub_inv = b;
    }
    hub_inv = team.max(hub_inv);
    double hub = kHubFactor * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    return hub;
  }

  __device__ int ydd_norm(double hg, double& yddnrm) {
    for (int i = tid(); i < N; i += kSpThreads) y[i] = hg * zn[1][i] + zn[0][i];
    if (f(y, tempv) > 0) return 1;
    for (int i = tid(); i < N; i += kSpThreads) tempv[i] = (tempv[i] - zn[1][i]) * (1.0 / hg);
    yddnrm = wrms(tempv);
    return 0;
  }

  __device__ int hin(double tout) {
    const double tdiff = tout - tn;
    if (tdiff == 0.0) return -1;
    const double tdist = fabs(tdiff);
    const double tround = DBL_EPSILON * fmax(fabs(tn), fabs(tout));
    if (tdist < 2.0 * tround) return -1;
    const double hlb = kHlbFactor * tround;
    const double hub = upper_bound_h0(tdist);
    double hg = sqrt(hlb * hub);
    if (hub < hlb) {
      h = hg;
      return 0;
    }
    double hnew = hg, hs = hg, yddnrm = 0.0;
    for (int count1 = 1; count1 <= kMaxIters; ++count1) {
      bool hgOK = false;
      for (int count2 = 1; count2 <= kMaxIters; ++count2) {
        if (ydd_norm(hg, yddnrm) == 0) {
          hgOK = true;
          break;
        }
        hg *= 0.2;
      }
      if (!hgOK) {
        if (count1 <= 2) return -2;
        hnew = hs;
        break;
      }
      hs = hg;
      hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
      if (count1 == kMaxIters) break;
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) break;
      if (count1 > 1 && hrat > 2.0) {
        hnew = hg;
        break;
      }
      hg = hnew;
    }
    double h0 = kHBias * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    h = h0;
    return 0;
  }

  __device__ void rescale() {
    double fac = eta;
    for (int j = 1; j <= q; ++j) {
      for (int i = tid(); i < N; i += kSpThreads) zn[j][i] *= fac;
      fac *= eta;
    }
    h = hscale * eta;
    next_h = h;
    hscale = h;
    nscon = 0;
  }

  __device__ void increase_bdf() {
    for (int i = 0; i <= kSpQmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
    double hsum = hscale;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
    double* zL = zn[L];
    const double* za = zn[indx_acor];
    for (int i = tid(); i < N; i += kSpThreads) zL[i] = A1 * za[i];
    for (int j = 2; j <= q; ++j)
      for (int i = tid(); i < N; i += kSpThreads) zn[j][i] += l[j] * zL[i];
  }

  __device__ void decrease_bdf() {
    for (int i = 0; i <= kSpQmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; ++j)
      for (int i = tid(); i < N; i += kSpThreads) zn[j][i] = -l[j] * zn[q][i] + zn[j][i];
  }

  __device__ void adjust_order(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1) increase_bdf();
    else decrease_bdf();
  }

  __device__ void adjust_params() {
    if (qprime != q) {
      adjust_order(qprime - q);
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    rescale();
  }

  __device__ void predict() {
    tn += h;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
        for (int i = tid(); i < N; i += kSpThreads) zn[j - 1][i] += zn[j][i];
  }

  __device__ void restore(double saved_t) {
    tn = saved_t;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
        for (int i = tid(); i < N; i += kSpThreads) zn[j - 1][i] -= zn[j][i];
  }

  __device__ void set_tq(double hsum, double alpha0, double alpha0_hat, double xi_inv,
                         double xistar_inv) {
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double Cc = xistar_inv / l[q];
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(Cc * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = alpha0 - (1.0 / (q + 1));
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = nlscoef / tq[2];
  }

  __device__ void set() {
    double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
    l[0] = l[1] = xi_inv = xistar_inv = 1.0;
    for (int i = 2; i <= q; ++i) l[i] = 0.0;
    alpha0 = alpha0_hat = -1.0;
    hsum = h;
    if (q > 1) {
      for (int j = 2; j < q; ++j) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        alpha0 -= 1.0 / j;
        for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
      for (int i = q; i >= 1; --i) l[i] += l[i - 1] * xistar_inv;
    }
    set_tq(hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  // cvLsSetup: Jacobian (or the saved one), M = I - gamma J, factorisation
  __device__ int lsetup(int cf) {
    const double dgamma = fabs(gamma / gammap - 1.0);
    const bool jbad = (nst == 0) || (nst >= nstlj + kCvlsMsbj) ||
                      ((cf == kFailBadJ) && (dgamma < kCvlsDgmax)) || (cf == kFailOther);
    if (jbad) {
      jcur = 1;
      nje++;
      nstlj = nst;
      SP_TIC(t0);
      jacobian(y, ftemp);
      SP_TOC(t0, 1);
    } else {
      jcur = 0;
    }
    SP_TIC(t1);
    for (int p = tid(); p < S.nnz; p += kSpThreads) {
      double v = J[p] * (-gamma);
      if (S.pos_row[p] == S.pos_col[p]) v += 1.0;
      lu[p] = (float)v;
    }
    team.sync();
    const bool ok = factor();
    SP_TOC(t1, 2);
    return ok ? 0 : 1;
  }

  __device__ int nls_lsetup(bool jbad) {
    if (jbad) convfail = kFailBadJ;
    const int rv = lsetup(convfail);
    nsetups++;
    gamrat = 1.0;
    gammap = gamma;
    crate = 1.0;
    nstlp = nst;
    return rv > 0 ? kNlsConvRecvr : kNlsSuccess;
  }

  __device__ int nls_residual(const double* ycor, double* res) {
    for (int i = tid(); i < N; i += kSpThreads) y[i] = zn[0][i] + ycor[i];
    if (f(y, ftemp) > 0) return kRhsRecvr;
    for (int i = tid(); i < N; i += kSpThreads) {
      double r = rl1 * zn[1][i] + ycor[i];
      res[i] = -gamma * ftemp[i] + r;
    }
    return kNlsSuccess;
  }

  __device__ int nls_convtest(int mi, const double* ycor, const double* dl, double tol) {
    const double del = wrms(dl);
    if (mi > 0) crate = fmax(kCrdown * crate, del / delp);
    const double dcon = del * fmin(1.0, crate) / tol;
    if (dcon <= 1.0) {
      acnrm = (mi == 0) ? del : wrms(ycor);
      acnrmcur = 1;
      return kNlsSuccess;
    }
    if (mi >= 1 && del > kRdiv * delp) return kNlsConvRecvr;
    delp = del;
    return kNlsContinue;
  }

  __device__ int newton(double* ycor, double tol, bool callLSetup) {
    bool jbad = false;
    int rv;
    for (;;) {
      rv = nls_residual(ycor, delta);
      if (rv == kNlsSuccess && callLSetup) rv = nls_lsetup(jbad);
      if (rv == kNlsSuccess) {
        int curiter = 0;
        for (;;) {
          nni++;
          for (int i = tid(); i < N; i += kSpThreads) delta[i] = -delta[i];
          SP_TIC(t0);
          solve(delta);
          SP_TOC(t0, 3);
          if (gamrat != 1.0) {
            const double s = 2.0 / (1.0 + gamrat);
            for (int i = tid(); i < N; i += kSpThreads) delta[i] *= s;
          }
          for (int i = tid(); i < N; i += kSpThreads) ycor[i] += delta[i];
          rv = nls_convtest(curiter, ycor, delta, tol);
          if (rv == kNlsSuccess) {
            jcur = 0;
            return kNlsSuccess;
          }
          if (rv != kNlsContinue) break;
          curiter++;
          if (curiter >= kNlsMaxcor) {
            rv = kNlsConvRecvr;
            break;
          }
          rv = nls_residual(ycor, delta);
          if (rv != kNlsSuccess) break;
        }
      }
      if (rv > 0 && !jcur) {
        callLSetup = true;
        jbad = true;
        for (int i = tid(); i < N; i += kSpThreads) ycor[i] = 0.0;
        continue;
      }
      return rv;
    }
  }

  __device__ int nls(int nflag) {
    convfail = (nflag == kFirstCall || nflag == kPrevErrFail) ? kNoFailures : kFailOther;
    const bool callSetup = (nflag == kPrevConvFail) || (nflag == kPrevErrFail) || (nst == 0) ||
                           (nst >= nstlp + kMsbp) || (fabs(gamrat - 1.0) > kDgmaxLsetup);
    for (int i = tid(); i < N; i += kSpThreads) acor[i] = 0.0;
    acnrmcur = 0;
    const int rv = newton(acor, tq[4], callSetup);
    if (rv != kNlsSuccess) return rv;
    for (int i = tid(); i < N; i += kSpThreads) y[i] = zn[0][i] + acor[i];
    if (!acnrmcur) acnrm = wrms(acor);
    jcur = 0;
    return kCvSuccess;
  }

  __device__ int handle_nflag(int& nflag, double saved_t, int& ncf) {
    if (nflag == kCvSuccess) return kDoErrorTest;
    ncfn++;
    restore(saved_t);
    ncf++;
    etamax = 1.0;
    if (fabs(h) <= hmin * kOnepsm || ncf == kMxncf) {
      if (nflag == kNlsConvRecvr) return kCvConvFailure;
      return kCvReptdRhsErr;
    }
    eta = fmax(kEtacf, hmin / fabs(h));
    nflag = kPrevConvFail;
    rescale();
    return kPredictAgain;
  }

  __device__ int do_error_test(int& nflag, double saved_t, int& nef, double& dsm) {
    dsm = acnrm * tq[2];
    if (dsm <= 1.0) return kCvSuccess;
    nef++;
    netf++;
    nflag = kPrevErrFail;
    restore(saved_t);
    if (fabs(h) <= hmin * kOnepsm || nef == kMxnef) return kCvErrFailure;
    etamax = 1.0;
    if (nef <= kMxnef1) {
      eta = 1.0 / (pow(kBias2 * dsm, 1.0 / L) + kAddon);
      eta = fmax(kEtamin, fmax(eta, hmin / fabs(h)));
      if (nef >= kSmallNef) eta = fmin(eta, kEtamxf);
      rescale();
      return kTryAgain;
    }
    if (q > 1) {
      eta = fmax(kEtamin, hmin / fabs(h));
      adjust_order(-1);
      L = q;
      q--;
      qwait = L;
      rescale();
      return kTryAgain;
    }
    eta = fmax(kEtamin, hmin / fabs(h));
    h *= eta;
    next_h = h;
    hscale = h;
    qwait = kLongWait;
    nscon = 0;
    if (f(zn[0], tempv) > 0) return kCvReptdRhsErr;
    for (int i = tid(); i < N; i += kSpThreads) zn[1][i] = h * tempv[i];
    return kTryAgain;
  }

  __device__ void complete_step() {
    nst++;
    nscon++;
    hu = h;
    qu = q;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
    for (int j = 0; j <= q; ++j)
      for (int i = tid(); i < N; i += kSpThreads) zn[j][i] += l[j] * acor[i];
    qwait--;
    if (qwait == 1 && q != kSpQmax) {
      for (int i = tid(); i < N; i += kSpThreads) zn[kSpQmax][i] = acor[i];
      saved_tq5 = tq[5];
      indx_acor = kSpQmax;
    }
  }

  __device__ void set_eta() {
    if (eta < kThresh) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = fmin(eta, etamax);
      eta /= fmax(1.0, fabs(h) * hmax_inv * eta);
      hprime = h * eta;
      if (qprime < q) nscon = 0;
    }
  }

  __device__ double comp_etaqm1() {
    etaqm1 = 0.0;
    if (q > 1) {
      const double ddn = wrms(zn[q]) * tq[1];
      etaqm1 = 1.0 / (pow(kBias1 * ddn, 1.0 / q) + kAddon);
    }
    return etaqm1;
  }

  __device__ double comp_etaqp1() {
    etaqp1 = 0.0;
    if (q != kSpQmax) {
      if (saved_tq5 == 0.0) return etaqp1;
      double hr = 1.0;
      for (int i = 0; i < L; ++i) hr *= h / tau[2];
      const double cquot = (tq[5] / saved_tq5) * hr;
      for (int i = tid(); i < N; i += kSpThreads) tempv[i] = -cquot * zn[kSpQmax][i] + acor[i];
      const double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (pow(kBias3 * dup, 1.0 / (L + 1)) + kAddon);
    }
    return etaqp1;
  }

  __device__ void choose_eta() {
    const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam < kThresh) {
      eta = 1.0;
      qprime = q;
      return;
    }
    if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else {
      eta = etaqp1;
      qprime = q + 1;
      for (int i = tid(); i < N; i += kSpThreads) zn[kSpQmax][i] = acor[i];
    }
  }

  __device__ void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = (qwait > 2) ? qwait : 2;
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (pow(kBias2 * dsm, 1.0 / L) + kAddon);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    etaqm1 = comp_etaqm1();
    etaqp1 = comp_etaqp1();
    choose_eta();
    set_eta();
  }

  __device__ int step() {
    const double saved_t = tn;
    int ncf = 0, nef = 0, nflag = kFirstCall;
    double dsm = 0.0;
    if (nst > 0 && hprime != h) adjust_params();
    for (;;) {
      predict();
      set();
      nflag = nls(nflag);
      const int kflag = handle_nflag(nflag, saved_t, ncf);
      if (kflag == kPredictAgain) continue;
      if (kflag != kDoErrorTest) return kflag;
      const int eflag = do_error_test(nflag, saved_t, nef, dsm);
      if (eflag == kTryAgain) continue;
      if (eflag != kCvSuccess) return eflag;
      break;
    }
    complete_step();
    prepare_next_step(dsm);
    etamax = (nst <= kSmallNst) ? kEtamx2 : kEtamx3;
    for (int i = tid(); i < N; i += kSpThreads) acor[i] *= tq[2];
    return kCvSuccess;
  }

  __device__ int one_step(double tout, double& tret) {
    if (nst == 0) {
      if (!set_ewt(zn[0])) return kCvIllInput;
      if (f(zn[0], zn[1]) > 0) return kCvReptdRhsErr;
      if (hin(tout) != 0) return kCvIllInput;
      const double rh = fabs(h) * hmax_inv;
      if (rh > 1.0) h /= rh;
      hscale = h;
      h0u = h;
      hprime = h;
      for (int i = tid(); i < N; i += kSpThreads) zn[1][i] *= h;
    } else {
      if (!set_ewt(zn[0])) return kCvIllInput;
    }
    const double nrm = wrms(zn[0]);
    tolsf = DBL_EPSILON * nrm;
    if (tolsf > 1.0) return kCvTooMuchAcc;
    tolsf = 1.0;
    const int kflag = step();
    tret = tn;
    if (kflag != kCvSuccess) return kflag;
    next_q = qprime;
    next_h = hprime;
    return kCvSuccess;
  }

  // CVodeGetDky(t, 0) into out
  __device__ void dky0(double t, double* out) {
    const double s = (t - tn) / h;
    for (int i = tid(); i < N; i += kSpThreads) out[i] = 0.0;
    for (int j = q; j >= 0; --j) {
      double c = 1.0;
      for (int i = 0; i < j; ++i) c *= s;
      for (int i = tid(); i < N; i += kSpThreads) out[i] += c * zn[j][i];
    }
  }
};

// doubles of LDS holding one SpCv per wavefront (16-byte multiple)
__host__ __device__ constexpr size_t sp_object_doubles() {
  return ((sizeof(SpCv) * kSpWaves + 15) & ~(size_t)15) / sizeof(double);
}

__global__ void sp_counter_clear(unsigned int* p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

// at least four wavefronts per SIMD (128 VGPRs): four workgroups per CU,
// which the LDS rows (four of N floats per workgroup) also allow
__global__ void __launch_bounds__(kSpThreads, 4)
bdf_sparse_kernel(DevMech m, DevSparse S, BdfOpts o, int n, double dt_all,
                  const double* __restrict__ dts, double* __restrict__ Tio,
                  double* __restrict__ Pio, double* __restrict__ Yio, int stride,
                  int* __restrict__ nsteps_out, int* __restrict__ stats_out,
                  const int* __restrict__ order, unsigned int* __restrict__ counter,
                  double* __restrict__ ws, size_t slot_doubles) {
  extern __shared__ double lds[];
  // LDS: this wavefront's integrator object (every member is uniform: the
  // lanes compute and store the same values; kept here rather than in
  // registers or private memory, which its size would otherwise spill to),
  // the team's reduction slots, then a dense FP32 row of N per wavefront
  SpCv& c = reinterpret_cast<SpCv*>(lds)[threadIdx.x / kWave];
  double* tail = lds + sp_object_doubles();
  c.team.red = tail;
  c.m = m;
  c.S = S;
  c.o = o;
  c.nsp = m.nsp;
  c.N = m.nsp + 1;
  c.w = reinterpret_cast<float*>(tail + ((kSpWaves + 2) & ~1)) + (size_t)c.wave() * c.N;
  c.xs = tail + ((kSpWaves + 2) & ~1);  // kSpWaves rows of N floats hold N doubles
  wave_sync();
  for (int i = c.lane(); i < c.N; i += kWave) c.w[i] = 0.0f;
  wave_sync();
  c.bind(ws + (size_t)blockIdx.x * slot_doubles);
  c.nlscoef = o.nlscoef;
  c.hmax_inv = o.hmax_inv;
  wave_sync();
  const int nsp = m.nsp, N = c.N, tid = c.tid();
  for (;;) {
    unsigned int r = 0;
    if (tid == 0) r = atomicAdd(counter, 1u);
    r = (unsigned int)c.team.from((double)r, 0);
    if (r >= (unsigned int)n) break;
    if (order) r = (unsigned int)order[r];
    const double dt = dts ? dts[r] : dt_all;
    double* y0 = c.t1;
    for (int i = tid; i < nsp; i += kSpThreads) y0[i] = Yio[(size_t)r * stride + i];
    if (tid == nsp % kSpThreads) y0[nsp] = Tio[r] / o.Tref;
    c.P = Pio[r];
    c.v0 = 0.0;
    if (o.constant_volume) {
      // ReactorConstantVolumeCPU::InitializeState (reactor_constant_volume_cpu.cpp:28-29)
      double part = 0.0;
      for (int i = tid; i < nsp; i += kSpThreads) part += y0[i] * m.inv_mw[i];
      const double mws = c.team.sum(part);
      const double T0 = (Tio[r] / o.Tref) * o.Tref;
      c.v0 = 1.0 / (c.P * (1.0 / mws) / (kRu * T0));
    }
    c.atol = o.atol;
    c.rtol = o.rtol;
    c.inv_rho_last = 0.0;
    c.team.sync();
#ifdef ZRK_PHASE_CYCLES
    for (int k = 0; k < 8; ++k) c.cyc[k] = 0;
    const long long tr0 = (long long)__builtin_amdgcn_s_memtime();
#endif
    c.reinit(0.0, y0);
    int nsteps = 0, num_tries = 0, flag = kCvSuccess;
    int tot[7] = {0, 0, 0, 0, 0, 0, 0};
    double tcurr = 0.0;
    // the step record of a one-reactor batch after each successful CVode
    // call (solver_cvode.cpp:247-265): zn[0] at tn and CVodeGetDky(tn, 1) =
    // zn[1] / h; t_prev follows every call, failed ones included
    const bool tr = o.trace != nullptr && r == 0;
    int ntr = 0;
    double tprev = 0.0;
    // one CVode call per iteration, CV_ONE_STEP (solver_cvode.cpp:228-266)
    while (tcurr < dt) {
      flag = c.one_step(dt, tcurr);
      nsteps += 1;
      if (c.nst >= o.max_steps) flag = kCvTooMuchWork;
      if (flag < 0) {
        num_tries += 1;
        c.atol *= o.retry_atol_adj;
        c.rtol *= o.retry_rtol_adj;
        tot[0] += c.nst; tot[1] += c.nfe; tot[2] += c.nje; tot[3] += c.nsetups;
        tot[4] += c.nni; tot[5] += c.ncfn; tot[6] += c.netf;
        tprev = tcurr;
        if (num_tries == o.num_retries) break;
        c.team.sync();
        c.reinit(tcurr, c.zn[0]);
      } else if (tr) {
        if (ntr < o.trace_cap) {
          double* rec = o.trace + (size_t)ntr * (2 * nsp + 5);
          if (tid == 0) {
            rec[0] = (double)nsteps;
            rec[1] = tcurr;
            rec[2] = tcurr - tprev;
          }
          const double hinv = 1.0 / c.h;
          for (int i = tid; i < N; i += kSpThreads) {
            rec[3 + i] = c.zn[0][i];
            rec[4 + nsp + i] = c.zn[1][i] * hinv;
          }
        }
        ++ntr;
        tprev = tcurr;
      }
    }
    if (tr && tid == 0) *o.trace_n = ntr;
    if (flag == kCvSuccess) {
      double* yout = c.tempv;
      c.dky0(tcurr < dt ? tcurr : dt, yout);
      tot[0] += c.nst; tot[1] += c.nfe; tot[2] += c.nje; tot[3] += c.nsetups;
      tot[4] += c.nni; tot[5] += c.ncfn; tot[6] += c.netf;
      c.team.sync();
      for (int i = tid; i < nsp; i += kSpThreads) Yio[(size_t)r * stride + i] = yout[i];
      if (tid == 0) Tio[r] = yout[nsp] * o.Tref;
      if (o.constant_volume) {
        // GetState: P = Ru T / (v Wmix) (reactor_constant_volume_cpu.cpp:44-63)
        double part = 0.0;
        for (int i = tid; i < nsp; i += kSpThreads) part += yout[i] * m.inv_mw[i];
        const double mws = c.team.sum(part);
        const double Tend = yout[nsp] * o.Tref;
        if (tid == 0) Pio[r] = kRu * Tend / (c.v0 * (1.0 / mws));
      }
    } else {
      nsteps = (nsteps <= 0) ? -1 : -nsteps;
    }
#ifdef ZRK_PHASE_CYCLES
    c.cyc[4] = (long long)__builtin_amdgcn_s_memtime() - tr0;
#endif
    if (tid == 0) {
      if (nsteps_out) nsteps_out[r] = nsteps;
      if (stats_out) {
#ifdef ZRK_PHASE_CYCLES
        for (int k = 0; k < 7; ++k) stats_out[(size_t)r * 20 + k] = tot[k];
        stats_out[(size_t)r * 20 + 7] = flag;
        for (int k = 0; k < 8; ++k) stats_out[(size_t)r * 20 + 8 + k] = (int)(c.cyc[k] >> 4);
#else
        for (int k = 0; k < 7; ++k) stats_out[(size_t)r * 8 + k] = tot[k];
        stats_out[(size_t)r * 8 + 7] = flag;
#endif
      }
    }
    c.team.sync();
  }
  (void)N;
}

size_t sparse_slot_bytes(const DevMech& m, const DevSparse& S) {
  return sizeof(double) * sparse_slot_doubles(m.nsp, m.nstep, S.nnz, S.N - S.split);
}

size_t sparse_lds_bytes(const DevMech& m) {
  return sizeof(double) * (sp_object_doubles() + (size_t)((kSpWaves + 2) & ~1)) +
         ((sizeof(float) * (size_t)kSpWaves * (m.nsp + 1) + 15) & ~(size_t)15);
}

int sparse_max_resident(const DevMech& m) {
  const size_t lds = sparse_lds_bytes(m);
  int dev = 0, ncu = 0, per_cu = 0, maxlds = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
      hipSuccess || lds > (size_t)maxlds)
    return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(bdf_sparse_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bdf_sparse_kernel, kSpThreads,
                                                   lds) != hipSuccess)
    return 0;
  return ncu * (per_cu > 0 ? per_cu : 1);
}

hipError_t launch_bdf_sparse(const DevMech& m, const DevSparse& S, const BdfOpts& o, int n,
                             double dt, const double* dts, double* T, double* P, double* Y,
                             int stride, int* nsteps, int* stats, const int* order,
                             unsigned int* counter, double* ws, int nblocks, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (S.N != m.nsp + 1 || m.n_ext > 0 || nblocks <= 0) return hipErrorInvalidValue;
  const size_t lds = sparse_lds_bytes(m);
  hipLaunchKernelGGL(sp_counter_clear, dim3(1), dim3(64), 0, st, counter);
  const int grid = nblocks < n ? nblocks : n;
  hipLaunchKernelGGL(bdf_sparse_kernel, dim3(grid), dim3(kSpThreads), lds, st, m, S, o, n, dt,
                     dts, T, P, Y, stride, nsteps, stats, order, counter, ws,
                     sparse_slot_doubles(m.nsp, m.nstep, S.nnz, S.N - S.split));
  return hipGetLastError();
}

}  // namespace zrk
