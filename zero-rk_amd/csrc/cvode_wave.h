// CVODE v5.8.0 BDF restated for one reactor per wavefront (see bdf_device.h).
//
// Small CVODE arrays (zn, tau, tq, l) are indexed only with compile-time
// indices inside unrolled loops guarded by the wave-uniform order q, so they
// stay in registers.  Return codes follow CVODE's.
#pragma once
#include <type_traits>

#include "bdf_device.h"

namespace zrk {

// CVODE constants (cvode.c / cvode_impl.h / cvode_ls_impl.h / sunnonlinsol_newton)
constexpr double kHlbFactor = 100.0, kHubFactor = 0.1, kHBias = 0.5;
constexpr int kMaxIters = 4;
constexpr double kThresh = 1.5, kEtamx1 = 10000.0, kEtamx2 = 10.0, kEtamx3 = 10.0;
constexpr double kEtamxf = 0.2, kEtamin = 0.1, kEtacf = 0.25, kAddon = 1.0e-6;
constexpr double kBias1 = 6.0, kBias2 = 6.0, kBias3 = 10.0, kOnepsm = 1.000001;
constexpr int kSmallNst = 10, kMxncf = 10, kMxnef = 7, kMxnef1 = 3, kSmallNef = 2, kLongWait = 10;
constexpr int kNlsMaxcor = 3;
constexpr double kCrdown = 0.3, kDgmaxLsetup = 0.3, kRdiv = 2.0, kCvlsDgmax = 0.2;
constexpr int kMsbp = 20, kCvlsMsbj = 51;

enum { kFirstCall = 0, kPrevConvFail, kPrevErrFail };
enum { kNoFailures = 0, kFailBadJ, kFailOther };
enum { kDoErrorTest = 100, kPredictAgain, kTryAgain };
enum { kNlsSuccess = 0, kNlsContinue = 901, kNlsConvRecvr = 902, kRhsRecvr = 903 };
enum {
  kCvSuccess = 0, kCvTooMuchWork = -1, kCvTooMuchAcc = -2, kCvErrFailure = -3,
  kCvConvFailure = -4, kCvReptdRhsErr = -10, kCvIllInput = -22
};

// Optional per-phase cycle counters (build with -DZRK_PHASE_CYCLES): RHS,
// Jacobian, Newton-matrix inverse and Newton solves, in shader clock cycles.
#ifdef ZRK_PHASE_CYCLES
#define ZRK_TIC(t) const long long t = (long long)__builtin_amdgcn_s_memtime()
#define ZRK_TOC(t, i) cyc[i] += (long long)__builtin_amdgcn_s_memtime() - t
#else
#define ZRK_TIC(t)
#define ZRK_TOC(t, i)
#endif

// The integrator's uniform FP64 state, one block per wavefront in LDS.
struct CvUni {
  double P, v0, rtol, tn, h, hscale, hprime, eta, etamax, hmin, hmax_inv, hu;
  double tau[kQmax + 2], tq[6], l[kQmax + 1];
  double rl1, gamma, gammap, gamrat, crate, delp, acnrm, nlscoef, saved_tq5;
  double etaq, etaqm1, etaqp1, inv_rho_last, inv_nvar;
  double dpdt, esrc, e0, vinit;  // source terms of this reactor (RhsSrc)
  // root finding (ignition): interval start tlo and g(tlo), the ignition
  // temperature, the last time returned by a one-step call
  double rt_tlo, rt_glo, rt_Tign, rt_tret;
  double pad_;
};
static_assert(sizeof(CvUni) % 16 == 0, "CvUni keeps the LDS layout 16-byte aligned");

// MT: precision of the Newton matrix, its inverse and the saved Jacobian
// (double: the plain FP64 method; float: the FP32 preconditioner, with the
// residual, error test and state update still in FP64).
// Team: TeamWave (one wavefront per reactor, nsp <= 63) or TeamBlock<W> (W
// wavefronts per reactor, nsp < 64W; an FP32 matrix lives in LDS, an FP64
// one in the team's slot of device memory).
// kIter: the alternate solver modes are compiled in -- the Krylov solver
// (SPGMR, the plugin's iterative option) and the difference-quotient
// Jacobians (analytic = 0); these instantiations are used exactly when
// BdfOpts::krylov is set or the Jacobian is not the analytic one, so the
// default direct analytic kernels carry none of that code (their register
// allocation, and their speed, stay those of the direct solver alone).
template <typename MT, class Team = TeamWave, bool kIter = false>
struct Cvode {
  static constexpr int kL = Team::kThreads;        // state slots = matrix row registers
  static constexpr bool kWaveTeam = (kL == kWave);
  // teams invert their Newton matrix with the blocked MFMA Gauss-Jordan, FP32
  // or FP64 tiles (build with -DZRK_TEAM_MFMA_INVERSE=0 for the element-wise
  // version)
#ifndef ZRK_TEAM_MFMA_INVERSE
#define ZRK_TEAM_MFMA_INVERSE 1
#endif
  static constexpr bool kUseMfmaInverse = !kWaveTeam && ZRK_TEAM_MFMA_INVERSE;
  // one wavefront, FP32 matrix: the same blocked MFMA inverse on the compact
  // N x N matrix in LDS (-DZRK_WAVE_MFMA_INVERSE=0: the in-register inverse)
#ifndef ZRK_WAVE_MFMA_INVERSE
#define ZRK_WAVE_MFMA_INVERSE 1
#endif
  static constexpr bool kWaveMfmaInverse =
      kWaveTeam && ZRK_WAVE_MFMA_INVERSE && std::is_same<MT, float>::value;
  static constexpr int kJS = kWaveTeam ? kNP : kL + 1;  // column stride of savedJ (and of
                                                       // Mlds for teams)
  Team team;
#ifdef ZRK_PHASE_CYCLES
  long long cyc[12];  // rhs, jacobian, inverse, solve, whole reactor; rhs parts (3);
                      // predict+set, nonlinear solve, error test, step completion
#endif
  DevMech mech;
  RhsLds s;
  BdfOpts opt;
  MT* savedJ;      // this team's saved Jacobian in HBM (column stride kJS)
  MT* Mlds;        // column-major: Jacobian workspace, then the inverse of the
                   // Newton matrix; stride nsp + 1 for one wavefront, kJS for teams;
                   // in LDS, except for FP64 teams (the team's device slot)
  MT* prow;        // LDS [kL]: published pivot row (teams)
  MT* rvec;        // LDS [kL]: right-hand side of a solve (teams)
  double* t2l;     // LDS [kL]: Jacobian column factors
  double* hl;      // LDS [kL]: energies for the Jacobian's temperature row (teams of
                   // several wavefronts)
  double* kry;     // LDS [kKry] per wavefront (kIter): the Krylov solver's Givens
                   // rotations, triangular factor and least-squares right side
  // lane vectors
  double zn[kQmax + 1];
  double ewt, acor, tempv, ftemp, y, atolv, delta;
  double ysrc;  // this lane's mass-fraction source (when has_ysrc)
  bool active;  // species lane, or the temperature lane when it is solved for
  bool has_src, has_ysrc;  // uniform: any source term / mass-fraction sources
  bool solveT;  // uniform: the temperature is a state variable (else held fixed)
  int nvar;     // uniform: state variables (nsp + 1, or nsp when T is held)
  // uniform FP64 scalars: kept in this wavefront's CvUni block in LDS (every
  // lane holds the same values), so they occupy no vector registers across
  // the rate evaluation and the matrix inversion
  double &P;
  double &v0;  // constant volume: fixed specific volume; 0 at constant pressure
  double &rtol, &tn, &h, &hscale, &hprime, &eta, &etamax, &hmin, &hmax_inv, &hu;
  double (&tau)[kQmax + 2];
  double (&tq)[6];
  double (&l)[kQmax + 1];
  double &rl1, &gamma, &gammap, &gamrat, &crate, &delp, &acnrm, &nlscoef, &saved_tq5;
  double &etaq, &etaqm1, &etaqp1, &inv_rho_last, &inv_nvar;
  double &dpdt, &esrc, &e0, &vinit;
  double &rt_tlo, &rt_glo, &rt_Tign, &rt_tret;
  int q, qprime, qwait, L;
  int nst, nstlp, nstlj, nfe, nje, nsetups, nni, ncfn, netf;
  int convfail, jcur, acnrmcur;

  __device__ explicit Cvode(CvUni* u)
      : P(u->P), v0(u->v0), rtol(u->rtol), tn(u->tn), h(u->h), hscale(u->hscale),
        hprime(u->hprime), eta(u->eta), etamax(u->etamax), hmin(u->hmin),
        hmax_inv(u->hmax_inv), hu(u->hu), tau(u->tau), tq(u->tq), l(u->l), rl1(u->rl1),
        gamma(u->gamma), gammap(u->gammap), gamrat(u->gamrat), crate(u->crate),
        delp(u->delp), acnrm(u->acnrm), nlscoef(u->nlscoef), saved_tq5(u->saved_tq5),
        etaq(u->etaq), etaqm1(u->etaqm1), etaqp1(u->etaqp1), inv_rho_last(u->inv_rho_last),
        inv_nvar(u->inv_nvar), dpdt(u->dpdt), esrc(u->esrc), e0(u->e0), vinit(u->vinit),
        rt_tlo(u->rt_tlo), rt_glo(u->rt_glo), rt_Tign(u->rt_Tign), rt_tret(u->rt_tret) {}

  // the source terms at time tt (src.on false when the reactor has none)
  __device__ __forceinline__ RhsSrc sources(double tt) const {
    RhsSrc rs;
    if (has_src) {
      rs.on = true;
      rs.has_ysrc = has_ysrc;
      rs.t = tt;
      rs.dpdt = dpdt;
      rs.esrc = esrc;
      rs.e0 = e0;
      rs.vinit = vinit;
      rs.ysrc = ysrc;
    }
    return rs;
  }

  __device__ __forceinline__ double wrms(double x) const {
    const double p = active ? x * ewt : 0.0;
    return sqrt(team.sum(p * p) * inv_nvar);
  }

  // f(tt, yy): the reactor's RHS at time tt of the solve
  __device__ __forceinline__ int f(double tt, double yy, double& yd) {
    nfe++;
    RhsAux aux;
    const RhsSrc src = sources(tt);
    ZRK_TIC(t0);
    int fl;
    if constexpr (kWaveTeam)
      fl = rhs64<Team::kExt>(mech, s, P, v0, opt.Tref, opt.step_limiter, yy, yd, aux, src);
    else
      fl = rhs_team(team, mech, s, P, v0, opt.Tref, opt.step_limiter, yy, yd, aux, src);
    ZRK_TOC(t0, 0);
#ifdef ZRK_PHASE_CYCLES
    if (!fl)
      for (int k = 0; k < 3; ++k) cyc[5 + k] += aux.cyc[k];
#endif
    if (!fl) inv_rho_last = aux.inv_rho;
    if (!active) yd = 0.0;
    return fl;
  }

  __device__ __forceinline__ int set_ewt(double yy) {
    const double v = rtol * fabs(yy) + atolv;
    const double bad = team.max((active && v <= 0.0) ? 1.0 : 0.0);
    ewt = active ? 1.0 / v : 0.0;
    return bad > 0.0 ? -1 : 0;
  }

  __device__ __forceinline__ void reinit(double t0, double y0) {
    zn[0] = y0;
#pragma unroll
    for (int j = 1; j <= kQmax; ++j) zn[j] = 0.0;
    tn = t0;
    q = 1;
    L = 2;
    qwait = L;
    etamax = kEtamx1;
    hu = 0.0;
    nst = nfe = ncfn = netf = nni = nsetups = 0;
    nstlp = nstlj = nje = 0;
    hmin = 0.0;
    saved_tq5 = 0.0;
    crate = 1.0;
    delp = 0.0;
    gammap = 0.0;
    gamrat = 1.0;
    jcur = 0;
#pragma unroll
    for (int i = 0; i < kQmax + 2; ++i) tau[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) tq[i] = 0.0;
#pragma unroll
    for (int i = 0; i <= kQmax; ++i) l[i] = 0.0;
  }

  // cvUpperBoundH0
  __device__ __forceinline__ double upper_bound_h0(double tdist) {
    const double t1 = kHubFactor * fabs(zn[0]) + (active ? 1.0 / ewt : 1.0);
    const double t2 = active ? fabs(zn[1]) / t1 : 0.0;
    const double hub_inv = team.max(t2);
    double hub = kHubFactor * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    return hub;
  }

  // cvYddNorm
  __device__ __forceinline__ int ydd_norm(double hg, double& yddnrm) {
    y = hg * zn[1] + zn[0];
    double fy;
    if (f(tn + hg, y, fy) > 0) return 1;
    tempv = (fy - zn[1]) * (1.0 / hg);
    yddnrm = wrms(tempv);
    return 0;
  }

  // cvHin
  __device__ __forceinline__ int hin(double tout) {
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

  // cvRescale
  __device__ __forceinline__ void rescale() {
    double fct = eta;
#pragma unroll
    for (int j = 1; j <= kQmax; ++j) {
      if (j <= q) {
        zn[j] *= fct;
        fct *= eta;
      }
    }
    h = hscale * eta;
    hscale = h;
  }

  // cvIncreaseBDF
  __device__ __forceinline__ void increase_bdf() {
#pragma unroll
    for (int i = 0; i <= kQmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
    double hsum = hscale;
#pragma unroll
    for (int j = 1; j < kQmax; ++j) {
      if (j < q) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
#pragma unroll
        for (int i = j + 2; i >= 2; --i)
          if (i <= kQmax) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
    const double zL = A1 * zn[kQmax];  // zn[indx_acor], indx_acor == qmax
#pragma unroll
    for (int j = 2; j <= kQmax; ++j)
      if (j <= q) zn[j] += l[j] * zL;
    // zn[L] = zL with L = q + 1 (q < qmax here)
#pragma unroll
    for (int j = 2; j <= kQmax; ++j)
      if (j == q + 1) zn[j] = zL;
  }

  // cvDecreaseBDF
  __device__ __forceinline__ void decrease_bdf() {
#pragma unroll
    for (int i = 0; i <= kQmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
#pragma unroll
    for (int j = 1; j <= kQmax - 2; ++j) {
      if (j <= q - 2) {
        hsum += tau[j];
        const double xi = hsum / hscale;
#pragma unroll
        for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
      }
    }
    double zq = 0.0;
#pragma unroll
    for (int j = 0; j <= kQmax; ++j)
      if (j == q) zq = zn[j];
#pragma unroll
    for (int j = 2; j < kQmax; ++j)
      if (j < q) zn[j] = -l[j] * zq + zn[j];
  }

  __device__ __forceinline__ void adjust_order(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1) increase_bdf();
    else decrease_bdf();
  }

  __device__ __forceinline__ void adjust_params() {
    if (qprime != q) {
      adjust_order(qprime - q);
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    rescale();
  }

  // cvPredict
  __device__ __forceinline__ void predict() {
    tn += h;
#pragma unroll
    for (int k = 1; k <= kQmax; ++k)
#pragma unroll
      for (int j = kQmax; j >= k; --j)
        if (k <= q && j <= q) zn[j - 1] += zn[j];
  }

  // cvRestore
  __device__ __forceinline__ void restore(double saved_t) {
    tn = saved_t;
#pragma unroll
    for (int k = 1; k <= kQmax; ++k)
#pragma unroll
      for (int j = kQmax; j >= k; --j)
        if (k <= q && j <= q) zn[j - 1] -= zn[j];
  }

  // cvSetBDF + cvSetTqBDF + gamma bookkeeping (cvSet)
  __device__ __forceinline__ void set() {
    double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
    l[0] = l[1] = xi_inv = xistar_inv = 1.0;
#pragma unroll
    for (int i = 2; i <= kQmax; ++i) l[i] = 0.0;
    alpha0 = alpha0_hat = -1.0;
    hsum = h;
    if (q > 1) {
#pragma unroll
      for (int j = 2; j < kQmax; ++j) {
        if (j < q) {
          hsum += tau[j - 1];
          xi_inv = h / hsum;
          alpha0 -= 1.0 / j;
#pragma unroll
          for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
        }
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      double tq1m = 0.0;
#pragma unroll
      for (int j = 1; j <= kQmax; ++j)
        if (j == q - 1) tq1m = tau[j];
      hsum += tq1m;
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = kQmax; i >= 1; --i)
        if (i <= q) l[i] += l[i - 1] * xistar_inv;
    }
    // cvSetTqBDF
    double lq = 0.0, tauq = 0.0;
#pragma unroll
    for (int j = 1; j <= kQmax; ++j)
      if (j == q) {
        lq = l[j];
        tauq = tau[j];
      }
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / lq;
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tauq;
      xi_inv = h / hsum;
      const double A5 = alpha0 - (1.0 / (q + 1));
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = nlscoef / tq[2];
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  // cvLsSetup/cvLsLinSys + inversion of M = I - gamma*J; >0 recoverable failure
  __device__ __forceinline__ int lsetup(int cf, double ypred, double fpred, int& jc) {
    if constexpr (kWaveTeam)
      return lsetup_wave(cf, ypred, fpred, jc);
    else
      return lsetup_team(cf, ypred, fpred, jc);
  }

  __device__ __forceinline__ bool jacobian_is_bad(int cf) const {
    const double dgamma = fabs(gamma / gammap - 1.0);
    return (nst == 0) || (nst >= nstlj + kCvlsMsbj) ||
           ((cf == kFailBadJ) && (dgamma < kCvlsDgmax)) || (cf == kFailOther);
  }

  // one wavefront: the matrix is inverted in registers (row per lane); the
  // inverse is kept in LDS as the N x N block, column-major with stride N
  __device__ __forceinline__ int lsetup_wave(int cf, double ypred, double fpred, int& jc) {
    const int nsp = mech.nsp;
    const int N = nsp + 1;
    const int lane = team.tid();
    if (!jacobian_is_bad(cf)) {
      jc = 0;
    } else {
      jc = 1;
      nje++;
      nstlj = nst;
      ZRK_TIC(t0);
      if (!kIter || opt.jac == kJacAnalytic) {
        inv_rho_last = jacobian_lds<MT, Team::kExt>(mech, s, P, v0, opt, ypred, fpred,
                                                    inv_rho_last, Mlds, N, savedJ, t2l,
                                                    sources(tn), solveT);
      } else if (dq_jacobian(ypred, fpred)) {
        return 1;  // the RHS failed at a perturbed state: recoverable
      }
      ZRK_TOC(t0, 1);
    }
    if constexpr (kWaveMfmaInverse) {
      ZRK_TIC(t1);
      // M = I - gamma J into the compact LDS matrix (rows/columns < N), from
      // the saved Jacobian (rewritten above by this lane when re-evaluated)
      const double ng = -gamma;
      team.sync();  // every lane is done with the Jacobian workspace
      if (lane < N) {
        for (int c = 0; c < N; ++c) {
          const double v = newton_entry(ng * (double)savedJ[c * kNP + lane], lane, c);
          Mlds[c * N + lane] = (MT)v;
        }
      }
      team.sync();
      const bool ok = gj_invert_mfma(team, reinterpret_cast<float*>(Mlds), N, N,
                                     reinterpret_cast<float*>(t2l));
      ZRK_TOC(t1, 2);
      return ok ? 0 : 1;
    }
    // the saved Jacobian (rewritten above by this lane when it was re-evaluated)
    MT R[kNP];
#pragma unroll
    for (int c = 0; c < kNP; ++c) R[c] = savedJ[c * kNP + lane];
    ZRK_TIC(t1);
    const double ng = -gamma;
#pragma unroll
    for (int c = 0; c < kNP; ++c) {
      double v = (c <= nsp && lane <= nsp) ? ng * (double)R[c] : 0.0;
      if (c <= nsp && lane <= nsp) v = newton_entry(v, lane, c);
      else if (c == lane) v += 1.0;
      R[c] = (MT)v;
    }
    const bool ok = gj_invert<MT>(R, N);
    team.sync();  // every lane is done with the Jacobian workspace
    // register c holds logical column (c + rot) mod 64: keep logical columns < N
    const int rot = gj_rotation(N);
    if (lane < N) {
#pragma unroll
      for (int c = 0; c < kNP; ++c) {
        const int j = (c + rot) & (kNP - 1);
        if (j < N) Mlds[j * N + lane] = R[c];
      }
    }
    team.sync();
    ZRK_TOC(t1, 2);
    return ok ? 0 : 1;
  }

  // several wavefronts: the matrix is inverted in place in LDS (thread i owns
  // row i; column-major with odd stride kJS, so every access is conflict-free)
  __device__ __forceinline__ int lsetup_team(int cf, double ypred, double fpred, int& jc) {
    const int nsp = mech.nsp;
    const int N = nsp + 1;
    const int tid = team.tid();
    bool copy = true;
    if (!jacobian_is_bad(cf)) {
      jc = 0;
    } else {
      jc = 1;
      nje++;
      nstlj = nst;
      ZRK_TIC(t0);
      if (!kIter || opt.jac == kJacAnalytic) {
        inv_rho_last = jacobian_team<MT>(team, mech, s, P, v0, opt, ypred, fpred, inv_rho_last,
                                         Mlds, savedJ, kJS, t2l, hl, sources(tn), solveT);
        copy = false;
      } else if (dq_jacobian(ypred, fpred)) {
        return 1;  // the RHS failed at a perturbed state: recoverable
      }
      ZRK_TOC(t0, 1);
    }
    if (copy) {
      team.sync();
      if (tid < N)
        for (int c = 0; c < N; ++c) Mlds[c * kJS + tid] = savedJ[c * kJS + tid];
    }
    ZRK_TIC(t1);
    const double ng = -gamma;
    if (tid < N)
      for (int c = 0; c < N; ++c) {
        const double v = newton_entry(ng * (double)Mlds[c * kJS + tid], tid, c);
        Mlds[c * kJS + tid] = (MT)v;
      }
    bool ok;
    if constexpr (kUseMfmaInverse)
      ok = gj_invert_mfma(team, Mlds, kJS, N, prow);
    else
      ok = gj_invert_lds<MT>(team, Mlds, kJS, N, prow);
    ZRK_TOC(t1, 2);
    return ok ? 0 : 1;
  }

  // Entry (row, col) of the Newton matrix from v = -gamma J_row,col: the
  // direct solver's I - gamma J, or the Krylov solver's preconditioner
  // (FormPreconditioner, reactor_nvector_serial.cpp:414-493): the diagonal
  // and, unless every entry is kept, only entries of the Jacobian's pattern
  // with |v| > threshold.  Rows and columns < nsp + 1.
  __device__ __forceinline__ double newton_entry(double v, int row, int col) const {
    if (row == col) return v + 1.0;
    if (!kIter || opt.pfull) return v;
    const bool in = (mech.jpat[row * 3 + (col >> 6)] >> (col & 63)) & 1ull;
    return (in && fabs(v) > opt.pthr) ? v : 0.0;
  }

  // Jacobian by difference quotients into savedJ (column c at savedJ[c * kJS
  // + row]), one RHS per column: CVODE's dense quotients (cvLsDenseDQJac:
  // inc_c = max(sqrt(uround) |y_c|, minInc / ewt_c), minInc = 1000 |h|
  // uround N wrms(f), column (f(y + inc e_c) - f) / inc as (1/inc)(f' - f)),
  // or the reactor's divided differences (reactor_nvector_serial.cpp:
  // 816-835: a fresh f(y), delta_c = sqrt(1e-16 max(1e-5, |y_c|)), column
  // (f' - f) / delta).  Columns of held variables are zero.  Returns nonzero
  // when the RHS fails.
  __device__ __noinline__ int dq_jacobian(double ypred, double fpred) {
    const int lane = team.tid();
    const int N = mech.nsp + 1;
    double f0 = fpred;
    double minInc = 1.0;
    if (opt.jac == kJacCvodeDQ) {
      const double fnorm = wrms(fpred);
      if (fnorm != 0.0) minInc = 1000.0 * fabs(h) * DBL_EPSILON * (double)nvar * fnorm;
    } else {
      if (f(tn, ypred, f0) > 0) return 1;
    }
    const double srur = sqrt(DBL_EPSILON);
    for (int c = 0; c < N; ++c) {
      double col = 0.0;
      if (c < nvar) {
        double incl = 0.0;
        if (lane == c)
          incl = (opt.jac == kJacCvodeDQ) ? fmax(srur * fabs(ypred), minInc / ewt)
                                          : sqrt(1.0e-16 * fmax(1.0e-5, fabs(ypred)));
        const double inc = team.from(incl, c);
        double fp;
        if (f(tn, lane == c ? ypred + inc : ypred, fp) > 0) return 1;
        if (active)
          col = (opt.jac == kJacCvodeDQ) ? (1.0 / inc) * (fp - f0) : (fp - f0) / inc;
      }
      if (lane < N) savedJ[c * kJS + lane] = (MT)col;
    }
    team.sync();
    return 0;
  }

  // x = M^{-1} b with the inverse in LDS (for one wavefront the entries of b
  // are broadcast from their lanes; teams publish b in LDS)
  __device__ __forceinline__ double lsolve(double b) {
    ZRK_TIC(t0);
    const int lane = team.tid();
    const int N = mech.nsp + 1;
    MT acc = MT(0);
    if constexpr (kWaveTeam) {
      const MT bm = (MT)b;
      if (lane < N) {
        const MT* __restrict__ col = Mlds + lane;
        int j = 0;
        for (; j + 4 <= N; j += 4) {  // same order as one j at a time
          acc = fma(col[j * N], bcast(bm, j), acc);
          acc = fma(col[(j + 1) * N], bcast(bm, j + 1), acc);
          acc = fma(col[(j + 2) * N], bcast(bm, j + 2), acc);
          acc = fma(col[(j + 3) * N], bcast(bm, j + 3), acc);
        }
        for (; j < N; ++j) acc = fma(col[j * N], bcast(bm, j), acc);
      }
    } else {
      rvec[lane] = (MT)b;
      team.sync();
#pragma unroll 8
      for (int j = 0; j < N; ++j) acc = fma(Mlds[j * kJS + lane], rvec[j], acc);
      if (lane >= N) acc = MT(0);
    }
    team.sync();
    ZRK_TOC(t0, 3);
    return (double)acc;
  }

  // ---- the Krylov solver (SUNLinSol_SPGMR as the plugin configures it:
  // maxl 5, left preconditioning, modified Gram-Schmidt, no restarts, scaling
  // by ewt; cvLsSolve's small-residual shortcut and tolerance; cvLsATimes
  // with cvLsDQJtimes).  The Krylov vectors are lane registers; the Givens
  // rotations, the triangular factor and the least-squares right side are
  // uniform and kept in this wavefront's kry block of LDS.
  static constexpr int kMaxl = 5;
  static constexpr int kKry = 2 * kMaxl + (kMaxl + 1) * kMaxl + kMaxl + 1;

  // z = v - gamma J v, J v = (f(y + sig v) - f(y)) / sig at the Newton
  // iterate (y, ftemp), sig = 1 / wrms(v) shrunk by 4 on a failed RHS (3
  // tries); nonzero on failure
  __device__ __forceinline__ int atimes(double v, double& z) {
    double sig = 1.0 / wrms(v);
    double fv = 0.0;
    int rv = 1;
    for (int it = 0; it < 3; ++it) {
      rv = f(tn, sig * v + y, fv);
      if (rv == 0) break;
      sig *= 0.25;
    }
    if (rv != 0) return 1;
    const double jv = (1.0 / sig) * (fv - ftemp);
    z = active ? -gamma * jv + v : 0.0;
    return 0;
  }

  // b <- the Newton correction; lflag: 0, or 1 for a recoverable failure
  // (no convergence, or a reduced residual after the first Newton iteration)
  __device__ double lsolve_krylov(double b, int curiter, int& lflag) {
    const int lane = team.tid();
    double* G = kry;                        // rotations (c, s) [2 kMaxl]
    double* R = kry + 2 * kMaxl;            // R[i * kMaxl + j]
    double* Yg = R + (kMaxl + 1) * kMaxl;   // [kMaxl + 1]
    lflag = 0;
    const double deltar = opt.eps_lin * tq[4];
    if (wrms(b) <= deltar) return curiter > 0 ? 0.0 : b;
    const double delta = deltar * sqrt((double)nvar);
    const double w = active ? ewt : 0.0;
    double V[kMaxl + 1] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    V[0] = w * lsolve(b);
    const double beta = sqrt(team.sum(V[0] * V[0]));
    if (beta <= delta) return 0.0;
    V[0] *= 1.0 / beta;
    double rot = 1.0, rho = beta;
    int krydim = 0;
    bool conv = false, fail = false;
    // the iterations run as a loop (one copy of the inlined RHS); every
    // register array is indexed at compile time through predicated loops
    for (int l = 0; l < kMaxl && !conv && !fail; ++l) {
      krydim = l + 1;
      double vl = 0.0;
#pragma unroll
      for (int k = 0; k < kMaxl; ++k)
        if (k == l) vl = V[k];
      double z = 0.0;
      fail = atimes(active ? vl / ewt : 0.0, z) != 0;
      double vk = fail ? 0.0 : w * lsolve(z);
      // SUNModifiedGS: column l of the Hessenberg matrix
      const double vk_norm = sqrt(team.sum(vk * vk));
      double hc[kMaxl + 1];
#pragma unroll
      for (int i = 0; i <= kMaxl; ++i) {
        hc[i] = 0.0;
        if (i <= l) {
          hc[i] = team.sum(V[i] * vk);
          vk = vk - hc[i] * V[i];
        }
      }
      double nn = sqrt(team.sum(vk * vk));
      double t = 1000.0 * vk_norm;
      if ((t + nn) == t) {  // reorthogonalise
        double nn2 = 0.0;
#pragma unroll
        for (int i = 0; i < kMaxl; ++i) {
          if (i <= l) {
            const double np = team.sum(V[i] * vk);
            t = 1000.0 * hc[i];
            if ((t + np) != t) {
              hc[i] += np;
              vk = vk - np * V[i];
              nn2 += np * np;
            }
          }
        }
        if (nn2 != 0.0) {
          const double np = nn * nn - nn2;
          nn = np > 0.0 ? sqrt(np) : 0.0;
        }
      }
      // SUNQRfact: the earlier rotations on this column, then a new one
      // annihilating its subdiagonal entry nn
#pragma unroll
      for (int k = 0; k < kMaxl - 1; ++k) {
        if (k < l) {
          const double c = G[2 * k], sn = G[2 * k + 1];
          const double t1 = hc[k], t2 = hc[k + 1];
          hc[k] = c * t1 - sn * t2;
          hc[k + 1] = sn * t1 + c * t2;
        }
      }
      double t1 = 0.0;
#pragma unroll
      for (int k = 0; k < kMaxl; ++k)
        if (k == l) t1 = hc[k];
      const double t2 = nn;
      double c, sn;
      if (t2 == 0.0) {
        c = 1.0;
        sn = 0.0;
      } else if (fabs(t2) >= fabs(t1)) {
        const double q = t1 / t2;
        sn = -1.0 / sqrt(1.0 + q * q);
        c = -sn * q;
      } else {
        const double q = t2 / t1;
        c = 1.0 / sqrt(1.0 + q * q);
        sn = -c * q;
      }
      const double rll = c * t1 - sn * t2;
      team.sync();
      if (lane == 0) {
        G[2 * l] = c;
        G[2 * l + 1] = sn;
#pragma unroll
        for (int i = 0; i < kMaxl; ++i)
          if (i < l) R[i * kMaxl + l] = hc[i];
        R[l * kMaxl + l] = rll;
      }
      team.sync();
      fail = fail || rll == 0.0;  // SUNLS_QRFACT_FAIL: recoverable
      rot *= sn;
      rho = fabs(rot * beta);
      if (rho <= delta) {
        conv = true;
      } else if (!fail) {
        vk *= 1.0 / nn;
      }
#pragma unroll
      for (int k = 1; k <= kMaxl; ++k)
        if (k == l + 1) V[k] = vk;
    }
    if (fail) {
      lflag = 1;
      return 0.0;
    }
    // SUNQRsol: y = R^{-1} Q (beta e_0), uniform work on one lane, then shared
    if (lane == 0) {
      Yg[0] = beta;
      for (int i = 1; i <= krydim; ++i) Yg[i] = 0.0;
      for (int k = 0; k < krydim; ++k) {
        const double c = G[2 * k], sn = G[2 * k + 1];
        const double t1 = Yg[k], t2 = Yg[k + 1];
        Yg[k] = c * t1 - sn * t2;
        Yg[k + 1] = sn * t1 + c * t2;
      }
      for (int k = krydim - 1; k >= 0; --k) {
        Yg[k] /= R[k * kMaxl + k];
        for (int i = 0; i < k; ++i) Yg[i] -= Yg[k] * R[i * kMaxl + k];
      }
    }
    team.sync();
    double xc = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxl; ++k)
      if (k < krydim) xc += Yg[k] * V[k];
    team.sync();
    if (conv || rho < beta) {
      lflag = conv ? 0 : (curiter == 0 ? 0 : 1);
      return active ? xc / ewt : 0.0;
    }
    lflag = 1;
    return 0.0;
  }

  // cvNlsLSetup
  __device__ __forceinline__ int nls_lsetup(bool jbad, int& jc) {
    if (jbad) convfail = kFailBadJ;
    const int rv = lsetup(convfail, y, ftemp, jc);
    nsetups++;
    gamrat = 1.0;
    gammap = gamma;
    crate = 1.0;
    nstlp = nst;
    return rv > 0 ? kNlsConvRecvr : kNlsSuccess;
  }

  // cvNlsResidual
  __device__ __forceinline__ int nls_residual(double ycor, double& res) {
    y = zn[0] + ycor;
    if (f(tn, y, ftemp) > 0) return kRhsRecvr;
    res = rl1 * zn[1] + ycor;
    res = -gamma * ftemp + res;
    return kNlsSuccess;
  }

  // cvNlsConvTest
  __device__ __forceinline__ int nls_convtest(int mi, double ycor, double dlt, double tol) {
    const double del = wrms(dlt);
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

  // SUNNonlinSolSolve_Newton
  __device__ __forceinline__ int newton(double& ycor, double tol, bool callLSetup) {
    bool jbad = false;
    int jc = 0, rv;
    for (;;) {
      rv = nls_residual(ycor, delta);
      if (rv == kNlsSuccess && callLSetup) {
        rv = nls_lsetup(jbad, jc);
        jcur = jc;
      }
      if (rv == kNlsSuccess) {
        int curiter = 0;
        for (;;) {
          nni++;
          double b = -delta;
          if (kIter && opt.krylov) {
            // cvLsSolve with SPGMR: no rescaling for a lagged gamma
            int lflag;
            b = lsolve_krylov(b, curiter, lflag);
            if (lflag) {
              rv = kNlsConvRecvr;
              break;
            }
          } else {
            b = lsolve(b);
            if (gamrat != 1.0) b *= 2.0 / (1.0 + gamrat);
          }
          if (!active) b = 0.0;
          delta = b;
          ycor += delta;
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
        ycor = 0.0;
        continue;
      }
      return rv;
    }
  }

  // cvNls
  __device__ __forceinline__ int nls(int nflag) {
    convfail = (nflag == kFirstCall || nflag == kPrevErrFail) ? kNoFailures : kFailOther;
    const bool callSetup = (nflag == kPrevConvFail) || (nflag == kPrevErrFail) || (nst == 0) ||
                           (nst >= nstlp + kMsbp) || (fabs(gamrat - 1.0) > kDgmaxLsetup);
    acor = 0.0;
    acnrmcur = 0;
    const int rv = newton(acor, tq[4], callSetup);
    if (rv != kNlsSuccess) return rv;
    y = zn[0] + acor;
    if (!acnrmcur) acnrm = wrms(acor);
    jcur = 0;
    return kCvSuccess;
  }

  __device__ __forceinline__ int handle_nflag(int& nflag, double saved_t, int& ncf) {
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

  __device__ __forceinline__ int do_error_test(int& nflag, double saved_t, int& nef, double& dsm) {
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
    hscale = h;
    qwait = kLongWait;
    double fz;
    if (f(tn, zn[0], fz) > 0) return kCvReptdRhsErr;
    zn[1] = h * fz;
    return kTryAgain;
  }

  __device__ __forceinline__ void complete_step() {
    nst++;
    hu = h;
#pragma unroll
    for (int i = kQmax; i >= 2; --i)
      if (i <= q) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
#pragma unroll
    for (int j = 0; j <= kQmax; ++j)
      if (j <= q) zn[j] += l[j] * acor;
    qwait--;
    if (qwait == 1 && q != kQmax) {
      zn[kQmax] = acor;
      saved_tq5 = tq[5];
    }
  }

  __device__ __forceinline__ void set_eta() {
    if (eta < kThresh) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = fmin(eta, etamax);
      eta /= fmax(1.0, fabs(h) * hmax_inv * eta);
      hprime = h * eta;
    }
  }

  __device__ __forceinline__ double comp_etaqm1() {
    etaqm1 = 0.0;
    if (q > 1) {
      double zq = 0.0;
#pragma unroll
      for (int j = 0; j <= kQmax; ++j)
        if (j == q) zq = zn[j];
      const double ddn = wrms(zq) * tq[1];
      etaqm1 = 1.0 / (pow(kBias1 * ddn, 1.0 / q) + kAddon);
    }
    return etaqm1;
  }

  __device__ __forceinline__ double comp_etaqp1() {
    etaqp1 = 0.0;
    if (q != kQmax) {
      if (saved_tq5 == 0.0) return etaqp1;
      double hr = 1.0;  // SUNRpowerI(h/tau[2], L): repeated product
      const double base = h / tau[2];
      for (int i = 0; i < L; ++i) hr *= base;
      const double cquot = (tq[5] / saved_tq5) * hr;
      tempv = -cquot * zn[kQmax] + acor;
      const double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (pow(kBias3 * dup, 1.0 / (L + 1)) + kAddon);
    }
    return etaqp1;
  }

  __device__ __forceinline__ void choose_eta() {
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
      zn[kQmax] = acor;
    }
  }

  __device__ __forceinline__ void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = qwait > 2 ? qwait : 2;
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

  // cvStep
  __device__ __forceinline__ int step() {
    const double saved_t = tn;
    int ncf = 0, nef = 0, nflag = kFirstCall;
    double dsm = 0.0;
    if (nst > 0 && hprime != h) adjust_params();
    for (;;) {
      ZRK_TIC(p0);
      predict();
      set();
      ZRK_TOC(p0, 8);
      ZRK_TIC(p1);
      nflag = nls(nflag);
      ZRK_TOC(p1, 9);
      const int kflag = handle_nflag(nflag, saved_t, ncf);
      if (kflag == kPredictAgain) continue;
      if (kflag != kDoErrorTest) return kflag;
      ZRK_TIC(p2);
      const int eflag = do_error_test(nflag, saved_t, nef, dsm);
      ZRK_TOC(p2, 10);
      if (eflag == kTryAgain) continue;
      if (eflag != kCvSuccess) return eflag;
      break;
    }
    ZRK_TIC(p3);
    complete_step();
    prepare_next_step(dsm);
    ZRK_TOC(p3, 11);
    etamax = (nst <= kSmallNst) ? kEtamx2 : kEtamx3;
    acor *= tq[2];
    return kCvSuccess;
  }

  // CVode(..., CV_ONE_STEP)
  __device__ __forceinline__ int one_step(double tout, double& tret) {
    if (nst == 0) {
      if (set_ewt(zn[0]) != 0) return kCvIllInput;
      double f0;
      if (f(tn, zn[0], f0) > 0) return kCvReptdRhsErr;
      zn[1] = f0;
      if (hin(tout) != 0) return kCvIllInput;
      const double rh = fabs(h) * hmax_inv;
      if (rh > 1.0) h /= rh;
      hscale = h;
      hprime = h;
      zn[1] *= h;
    } else {
      if (set_ewt(zn[0]) != 0) return kCvIllInput;
    }
    const double tolsf = DBL_EPSILON * wrms(zn[0]);
    if (tolsf > 1.0) return kCvTooMuchAcc;
    const int kflag = step();
    tret = tn;
    return kflag;
  }

  // ---- root finding for the ignition time: SUNDIALS CVODE 5.8.0's
  // cvRcheck1-3 and cvRootfind for the one root function
  // g = T0 + delta_T - T (reactor_nvector_serial.cpp:774-784), both
  // directions, in the one-step mode of solver_cvode.cpp:228-266.  All values
  // are uniform over the team.
  static constexpr int kRootMaxIter = 200;  // bound on the secant iterations

  __device__ __forceinline__ double root_g_of(double yv) const {
    return rt_Tign - team.from(yv, mech.nsp) * opt.Tref;
  }
  __device__ __forceinline__ double root_ttol() const {
    return (fabs(tn) + fabs(h)) * DBL_EPSILON * 100.0;
  }

  // cvRcheck1: the interval starts at the (re)initialisation time
  __device__ __forceinline__ void root_init() {
    rt_tlo = tn;
    rt_glo = root_g_of(zn[0]);
    rt_tret = tn;
  }

  // cvRootfind: a sign change of g over (tlo, thi] is narrowed to within
  // ttol by the Illinois-weighted secant; true when (tlo, thi] holds a root,
  // which is then thi (ghi = g there)
  __device__ __forceinline__ bool rootfind(double& tlo, double& glo, double& thi, double& ghi,
                           double ttol) const {
    const bool zroot0 = ghi == 0.0;
    if (zroot0 || !(glo * ghi < 0.0)) return zroot0;
    double alph = 1.0;
    int side = 0, sideprev = -1;
    for (int it = 0; it < kRootMaxIter; ++it) {
      alph = (sideprev == side) ? ((side == 2) ? alph * 2.0 : alph * 0.5) : 1.0;
      double tmid = thi - (thi - tlo) * ghi / (ghi - alph * glo);
      if (fabs(tmid - tlo) < 0.5 * ttol) {
        const double fracint = fabs(thi - tlo) / ttol;
        const double fracsub = (fracint > 5.0) ? 0.1 : 0.5 / fracint;
        tmid = tlo + fracsub * (thi - tlo);
      }
      if (fabs(thi - tmid) < 0.5 * ttol) {
        const double fracint = fabs(thi - tlo) / ttol;
        const double fracsub = (fracint > 5.0) ? 0.1 : 0.5 / fracint;
        tmid = thi - fracsub * (thi - tlo);
      }
      const double gmid = root_g_of(dky0(tmid));
      sideprev = side;
      if (gmid != 0.0 && glo * gmid < 0.0) {  // sign change in (tlo, tmid)
        thi = tmid;
        ghi = gmid;
        side = 1;
        if (fabs(thi - tlo) <= ttol) break;
        continue;
      }
      if (gmid == 0.0) {  // a zero at tmid
        thi = tmid;
        ghi = gmid;
        break;
      }
      tlo = tmid;  // the sign change is in (tmid, thi)
      glo = gmid;
      side = 2;
      if (fabs(thi - tlo) <= ttol) break;
    }
    return true;
  }

  // cvRcheck3 after a step (thi = tn, y = zn[0]): true with the root time in
  // troot when (tlo, tn] holds a root; the interval start moves to the root or
  // to tn
  __device__ __forceinline__ bool root_check3(double& troot) {
    double tlo = rt_tlo, glo = rt_glo, thi = tn;
    double ghi = root_g_of(zn[0]);
    const bool found = rootfind(tlo, glo, thi, ghi, root_ttol());
    rt_tlo = thi;
    rt_glo = ghi;
    troot = thi;
    return found;
  }

  // cvRcheck2 at the start of the call after a root return: g at the root
  // again; if it is exactly zero, its value a roundoff interval later
  __device__ __forceinline__ void root_check2() {
    const double ylo = dky0(rt_tlo);
    double g = root_g_of(ylo);
    if (g == 0.0) {
      const double ttol = root_ttol();
      const double smallh = h > 0.0 ? ttol : -ttol;
      const double tplus = rt_tlo + smallh;
      const double yp = ((tplus - tn) * h >= 0.0) ? ylo + (smallh / h) * zn[1] : dky0(tplus);
      const double gp = root_g_of(yp);
      if (gp != 0.0) g = gp;
    }
    rt_glo = g;
  }

  // CVodeGetDky(t, 1): sum over j = q..1 of j s^(j-1) zn[j], times 1/h
  __device__ __forceinline__ double dky1(double t) const {
    const double sv = (t - tn) / h;
    double out = 0.0;
#pragma unroll
    for (int j = kQmax; j >= 1; --j) {
      if (j <= q) {
        double c = (double)j;
#pragma unroll
        for (int i = 0; i < j - 1; ++i) c *= sv;
        out += c * zn[j];
      }
    }
    return out * (1.0 / h);
  }

  // CVodeGetDky(t, 0)
  __device__ __forceinline__ double dky0(double t) const {
    const double sv = (t - tn) / h;
    double out = 0.0;
#pragma unroll
    for (int j = kQmax; j >= 0; --j) {
      if (j <= q) {
        double c = 1.0;
#pragma unroll
        for (int i = 0; i < j; ++i) c *= sv;
        out += c * zn[j];
      }
    }
    return out;
  }
};

template <typename MT, bool E = false, bool I = false>
using CvWave = Cvode<MT, TeamWaveX<E>, I>;

}  // namespace zrk
