/* CPU FP64 oracle: plain-C restatement of the reference cfd_plugin path.
 * TEST INFRASTRUCTURE ONLY (checker + CPU baseline); see zrk_oracle.h. */
#include "zrk_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NIST_RU_DEFAULT 8.314462618e3 /* src/zerork/constants.h:10 */
/* the mechanism's gas constant (zrko_mech.ru; 0 selects the default) */
#define NIST_RU (m->ru > 0.0 ? m->ru : NIST_RU_DEFAULT)
#define P_ATM 1.01325e5       /* src/zerork/constants.h:12 */
#define TLIMIT 1.0e4          /* reactor_constant_pressure_cpu.cpp:88 */

void zrko_default_opts(zrko_opts* o) {
  /* zerork_reactor_manager.cpp:50-101 defaults, BDF dense direct */
  o->rtol = 1.0e-8;
  o->atol = 1.0e-20;
  o->max_dt = 0.05;
  o->nlscoef = 0.05;
  o->max_steps = 5000;
  o->num_retries = 5;
  o->retry_atol_adj = 0.1;
  o->retry_rtol_adj = 1.0;
  o->min_mass_fraction = 1.0e-30;
  o->step_limiter = 1.0e22;
  o->Tref = 1.0;
  o->constant_volume = 0;
  o->delta_temperature_ignition = 0.0;
  o->stop_after_ignition = 0;
  o->abstol_dens = 0;
  o->solve_temperature = 1;
  o->dense = 1;
  o->analytic = 1;
  o->iterative = 0;
  o->eps_lin = 1.0e-3;
  o->preconditioner_threshold = 1.0e-3;
}

/* ------------------------------------------------------------------------- */
/* NASA-7 thermo: src/zerork/nasa_poly.cpp:50-172                              */
/* ------------------------------------------------------------------------- */
static void thermo_cp(const zrko_mech* m, double T, double* cp) {
  for (int j = 0; j < m->nsp; ++j) {
    const double* c = m->thermo + 16 * j;
    const double* a = (T < c[0]) ? c + 1 : c + 8;
    cp[j] = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  }
}

static void thermo_h(const zrko_mech* m, double T, double* h) {
  const double invT = 1.0 / T;
  double hm[4];
  hm[0] = T;
  hm[1] = T * hm[0];
  hm[2] = T * hm[1];
  hm[3] = T * hm[2];
  hm[0] *= 0.50000000000000000000;
  hm[1] *= 0.33333333333333333333;
  hm[2] *= 0.25000000000000000000;
  hm[3] *= 0.20000000000000000000;
  for (int j = 0; j < m->nsp; ++j) {
    const double* c = m->thermo + 16 * j;
    const double* a = (T < c[0]) ? c + 1 : c + 8;
    h[j] = a[0] + a[1] * hm[0] + a[2] * hm[1] + a[3] * hm[2] + a[4] * hm[3] + a[5] * invT;
  }
}

static void thermo_g(const zrko_mech* m, double T, double* g) {
  const double invT = 1.0 / T;
  double gm[5];
  gm[0] = 1.0 - log(T);
  gm[1] = T;
  gm[2] = T * gm[1];
  gm[3] = T * gm[2];
  gm[4] = T * gm[3];
  gm[1] *= -0.50000000000000000000;
  gm[2] *= -0.16666666666666666667;
  gm[3] *= -0.08333333333333333333;
  gm[4] *= -0.05000000000000000000;
  for (int j = 0; j < m->nsp; ++j) {
    const double* c = m->thermo + 16 * j;
    const double* a = (T < c[0]) ? c + 1 : c + 8;
    g[j] = a[0] * gm[0] + a[1] * gm[1] + a[2] * gm[2] + a[3] * gm[3] + a[4] * gm[4] +
           a[5] * invT - a[6];
  }
}

void zrko_thermo(const zrko_mech* m, double T, double* cp_R, double* h_RT, double* g_RT) {
  if (cp_R) thermo_cp(m, T, cp_R);
  if (h_RT) thermo_h(m, T, h_RT);
  if (g_RT) thermo_g(m, T, g_RT);
}

/* PLOG rate at pressure-table point k of PLOG step i: the lines of that
 * pressure summed with the sign of their A (plog_reaction.cpp:163-184) */
static double plog_at(const zrko_mech* m, int i, int k, double logT, double invT) {
  const int g = m->plog_ptr[i] + k;
  double r = 0.0;
  for (int q = m->plog_lptr[g]; q < m->plog_lptr[g + 1]; ++q)
    r += m->plog_sign[q] * exp(m->plog_lnA[q] + logT * m->plog_beta[q] - invT * m->plog_Ta[q]);
  return r;
}

/* PLogReaction::GetRateCoefficientFromTP (plog_reaction.cpp:186-320): no
 * extrapolation; log-p interpolation between the bracketing pressures, zero
 * when either end is not positive */
static double plog_rate(const zrko_mech* m, int i, double T, double logT, double invT,
                        double P) {
  const int np = m->plog_ptr[i + 1] - m->plog_ptr[i];
  const double* pp = m->plog_p + m->plog_ptr[i];
  const double* lp = m->plog_lnp + m->plog_ptr[i];
  if (T <= 0.0 || P <= 0.0) return 0.0;
  if (np == 1 || P <= pp[0]) return plog_at(m, i, 0, logT, invT);
  if (P >= pp[np - 1]) return plog_at(m, i, np - 1, logT, invT);
  int r = 0;
  for (int j = 0; j < np - 1; ++j)
    if (pp[j] <= P && P < pp[j + 1]) {
      r = j;
      break;
    }
  const double k1 = plog_at(m, i, r, logT, invT);
  if (k1 <= 0.0) return 0.0;
  const double k2 = plog_at(m, i, r + 1, logT, invT);
  if (k2 <= 0.0) return 0.0;
  return k1 * pow(P / pp[r], log(k2 / k1) / (lp[r + 1] - lp[r]));
}

/* ------------------------------------------------------------------------- */
/* Rate constants: rate_const::updateK (rate_const.cpp:539-582) and its parts  */
/* ------------------------------------------------------------------------- */
static void rate_constants_ws(const zrko_mech* m, double T, const double* C, double* K,
                              double* G) {
  const int nstep = m->nstep;
  for (int s = 0; s < nstep; ++s) K[s] = 0.0;
  double Csum = 0.0;
  for (int j = 0; j < m->nsp; ++j) Csum += C[j];
  const double logT = log(T);
  const double invT = 1.0 / T;
  const double log_patm_rut = log(P_ATM / (NIST_RU * T));
  /* Arrhenius (rate_const.cpp:669-687) */
  for (int i = 0; i < m->n_arr; ++i)
    K[m->arr_step[i]] = exp(m->arr_lnA[i] + m->arr_beta[i] * logT - m->arr_Ta[i] * invT);
  /* PLOG at the pressure Csum Ru T, before Keq (rate_const.cpp:551, 565-568,
   * 1188-1200; plog_reaction.cpp:163-320) */
  if (m->n_plog > 0) {
    const double P = Csum * NIST_RU * T;
    for (int i = 0; i < m->n_plog; ++i)
      K[m->plog_step[i]] = plog_rate(m, i, T, logT, invT, P);
  }
  /* reverse from equilibrium (rate_const.cpp:690-731); a non-integer step
   * weighs g/RT by its rate-of-progress exponents, reverse list minus forward
   * list (non_integer_reaction_network.cpp:335-358) */
  if (m->n_keq > 0) {
    thermo_g(m, T, G);
    for (int i = 0; i < m->n_keq; ++i) {
      const int f = m->keq_fwd[i];
      double sum = 0.0;
      if (m->n_ni > 0 && m->step_ni[f] >= 0) {
        const int a = m->step_ni[f], b = m->step_ni[m->keq_step[i]];
        for (int k = m->ni_rop_ptr[b]; k < m->ni_rop_ptr[b + 1]; ++k)
          sum += m->ni_rop_pow[k] * G[m->ni_rop_sp[k]];
        for (int k = m->ni_rop_ptr[a]; k < m->ni_rop_ptr[a + 1]; ++k)
          sum -= m->ni_rop_pow[k] * G[m->ni_rop_sp[k]];
      } else {
        for (int k = m->prod_ptr[f]; k < m->prod_ptr[f + 1]; ++k) sum += G[m->prod_sp[k]];
        for (int k = m->reac_ptr[f]; k < m->reac_ptr[f + 1]; ++k) sum -= G[m->reac_sp[k]];
      }
      const double arg = sum - m->keq_ndelta[i] * log_patm_rut;
      K[m->keq_step[i]] = exp(arg) * K[f];
    }
  }
  /* third body (rate_const.cpp:862-888) */
  for (int i = 0; i < m->n_tb; ++i) {
    double cm = Csum;
    for (int k = m->tb_ptr[i]; k < m->tb_ptr[i + 1]; ++k) cm += C[m->tb_sp[k]] * m->tb_eff[k];
    K[m->tb_fwd[i]] *= cm;
    if (m->tb_rev[i] >= 0) K[m->tb_rev[i]] *= cm;
  }
  /* falloff (rate_const.cpp:890-1012; Lindemann, Troe and SRI) */
  for (int i = 0; i < m->n_fo; ++i) {
    const double* p = m->fo_param + 8 * i;
    const double klow = exp(p[0] + p[1] * logT - p[2] * invT);
    double cm;
    if (m->fo_spc[i] >= 0) {
      cm = C[m->fo_spc[i]];
    } else {
      cm = Csum;
      for (int k = m->fo_ptr[i]; k < m->fo_ptr[i + 1]; ++k) cm += C[m->fo_sp[k]] * m->fo_eff[k];
    }
    double Pr = klow * cm / K[m->fo_fwd[i]];
    if (Pr < 1.0e-300) Pr = 1.0e-300;
    double lpr = log10(Pr);
    double f = 1.0;
    if (m->fo_type[i] == 1 || m->fo_type[i] == 2) {
      double fc = 0.0;
      if (p[4] != 0) fc += (1.0 - p[3]) * exp(-T / p[4]);
      if (p[5] != 0) fc += p[3] * exp(-T / p[5]);
      if (m->fo_type[i] == 2) fc += exp(-p[6] * invT);
      if (fc < 1.0e-300) fc = 1.0e-300;
      f = log10(fc);
      const double n = 0.75 - 1.27 * f;
      lpr -= (0.4 + 0.67 * f);
      lpr = lpr / (n - 0.14 * lpr);
      lpr *= lpr;
      f /= (1.0 + lpr);
      f = pow(10.0, f);
    } else if (m->fo_type[i] == 3) {
      /* d (a exp(-b/T) + exp(-T/c))^X T^e with X = 1/(1 + log10(Pr)^2);
       * d = 1 and e = 0 multiply exactly when they were not given */
      const double xp = 1.0 / (1.0 + lpr * lpr);
      f = p[3] * exp(-p[4] * invT);
      if (p[5] > 0) f += exp(-T * p[5]);
      f = pow(f, xp);
      f *= p[6];
      f *= pow(T, p[7]);
    }
    const double pc = f * Pr / (1.0 + Pr);
    K[m->fo_fwd[i]] *= pc;
    if (m->fo_rev[i] >= 0) K[m->fo_rev[i]] *= pc;
  }
}

void zrko_rate_constants(const zrko_mech* m, double T, const double* C, double* K) {
  double* G = (double*)malloc(sizeof(double) * (size_t)m->nsp);
  rate_constants_ws(m, T, C, K, G);
  free(G);
}

/* perf_net::calcRatesFromTC_StepLimiter (perf_net.cpp:190-275) */
static void rates_ws(const zrko_mech* m, double T, const double* C, double L, double* rop,
                     double* create, double* destroy, double* net, double* G) {
  rate_constants_ws(m, T, C, rop, G);
  if (0.0 < L && L < 1.0e300)
    for (int s = 0; s < m->nstep; ++s) rop[s] *= L / (L + rop[s]);
  for (int s = 0; s < m->nstep; ++s)
    for (int k = m->reac_ptr[s]; k < m->reac_ptr[s + 1]; ++k) rop[s] *= C[m->reac_sp[k]];
  memset(create, 0, sizeof(double) * (size_t)m->nsp);
  memset(destroy, 0, sizeof(double) * (size_t)m->nsp);
  /* non-integer steps first (perf_net.cpp:250-253;
   * non_integer_reaction_network.cpp:270-333): the rate of progress takes
   * sign(C) |C|^p per exponent, then creation and destruction by coefficient */
  for (int q = 0; q < m->n_ni; ++q) {
    const int s = m->ni_step[q];
    for (int k = m->ni_rop_ptr[q]; k < m->ni_rop_ptr[q + 1]; ++k) {
      const double c = C[m->ni_rop_sp[k]];
      const double f = pow(fabs(c), m->ni_rop_pow[k]);
      rop[s] *= (c >= 0.0) ? f : -f;
    }
  }
  for (int q = 0; q < m->n_ni; ++q)
    for (int k = m->ni_pr_ptr[q]; k < m->ni_pr_ptr[q + 1]; ++k)
      create[m->ni_pr_sp[k]] += m->ni_pr_nu[k] * rop[m->ni_step[q]];
  for (int q = 0; q < m->n_ni; ++q)
    for (int k = m->ni_re_ptr[q]; k < m->ni_re_ptr[q + 1]; ++k)
      destroy[m->ni_re_sp[k]] += m->ni_re_nu[k] * rop[m->ni_step[q]];
  for (int s = 0; s < m->nstep; ++s)
    for (int k = m->reac_ptr[s]; k < m->reac_ptr[s + 1]; ++k) destroy[m->reac_sp[k]] += rop[s];
  for (int s = 0; s < m->nstep; ++s)
    for (int k = m->prod_ptr[s]; k < m->prod_ptr[s + 1]; ++k) create[m->prod_sp[k]] += rop[s];
  for (int j = 0; j < m->nsp; ++j) net[j] = create[j] - destroy[j];
}

void zrko_rates(const zrko_mech* m, double T, const double* C, double L, double* rop,
                double* create, double* destroy, double* net) {
  double* G = (double*)malloc(sizeof(double) * (size_t)m->nsp);
  rates_ws(m, T, C, L, rop, create, destroy, net, G);
  free(G);
}

/* ------------------------------------------------------------------------- */
/* Constant-pressure reactor (reactor_constant_pressure_cpu.cpp) and           */
/* constant-volume reactor (reactor_constant_volume_cpu.cpp)                   */
/* ------------------------------------------------------------------------- */
typedef struct {
  const zrko_mech* m;
  const zrko_opts* o;
  double P;
  double inv_density; /* carried between calls like the reference member */
  double v0;          /* constant volume: the fixed specific volume (1/density) */
  double cp_mix;      /* mass cp (constant pressure) or cv (constant volume) */
  /* source terms (InitializeState, reactor_constant_pressure_cpu.cpp:15-42 and
   * reactor_constant_volume_cpu.cpp:15-42): pressure rate, energy source,
   * mass-fraction sources; e0 the initial mass enthalpy (constant pressure) or
   * internal energy (constant volume), vinit the initial 1/density */
  double dpdt, esrc, e0, vinit;
  const double* ysrc;
  /* > 0: the temperature is held at fixT * Tref and is not a state variable
   * (SetSolveTemperature(false)); the state is then the nsp mass fractions */
  double fixT;
  double *C, *h, *cpm, *G, *rop, *cr, *de, *net, *esc, *tmp1, *tmp2, *tmp3;
} reactor_t;

static void reactor_init(reactor_t* r, const zrko_mech* m, const zrko_opts* o, double P) {
  const int nsp = m->nsp, N = nsp + 1;
  r->m = m;
  r->o = o;
  r->P = P;
  r->inv_density = 0.0;
  r->v0 = 0.0;
  r->cp_mix = 0.0;
  r->dpdt = r->esrc = r->e0 = r->vinit = 0.0;
  r->ysrc = NULL;
  r->fixT = 0.0;
  double* w = (double*)calloc((size_t)(8 * nsp + m->nstep + 3 * N), sizeof(double));
  r->C = w;
  r->h = w + nsp;
  r->cpm = w + 2 * nsp;
  r->G = w + 3 * nsp;
  r->cr = w + 4 * nsp;
  r->de = w + 5 * nsp;
  r->net = w + 6 * nsp;
  r->esc = w + 7 * nsp; /* scratch of the energy -> temperature solve */
  r->rop = w + 8 * nsp;
  r->tmp1 = r->rop + m->nstep;
  r->tmp2 = r->tmp1 + N;
  r->tmp3 = r->tmp2 + N;
}

static void reactor_free(reactor_t* r) { free(r->C); }

/* 1/density at (T, P, Y) (mechanism::getDensityFromTPY, mechanism.cpp:279-284) */
static double inv_density_tpy(const zrko_mech* m, double T, double P, const double* y) {
  double mws = 0.0;
  for (int j = 0; j < m->nsp; ++j) mws += y[j] * (1.0 / m->mw[j]);
  return 1.0 / (P * (1.0 / mws) / (NIST_RU * T));
}

/* Mass enthalpy (mechanism::getMassEnthalpyFromTY, mechanism.cpp:336-347) or,
 * with cvol, internal energy (getMassIntEnergyFromTY, :355-367) of (T, y);
 * w: nsp doubles of scratch. */
static double mass_energy(const zrko_mech* m, int cvol, double T, const double* y, double* w) {
  thermo_h(m, T, w);
  double e = 0.0;
  for (int j = 0; j < m->nsp; ++j) {
    if (cvol) w[j] -= 1.0;
    w[j] *= (NIST_RU / m->mw[j]) * T;
    e += w[j] * y[j];
  }
  return e;
}

/* Mass cp, or with cv the mass cv (getMassCpFromTY / getMassCvFromTY,
 * mechanism.cpp:408-441). */
static double mass_heat_capacity(const zrko_mech* m, int cv, double T, const double* y,
                                 double* w) {
  thermo_cp(m, T, w);
  double c = 0.0;
  for (int j = 0; j < m->nsp; ++j) {
    if (cv) w[j] -= 1.0;
    w[j] *= NIST_RU / m->mw[j];
    c += w[j] * y[j];
  }
  return c;
}

/* Temperature at which the mass enthalpy (cvol = 0) or internal energy
 * (cvol = 1) of y equals E: mechanism::getTemperatureFromHY / FromEY
 * (mechanism.cpp:477-547), including their choices of heat capacity (the HY
 * iteration and its upper extrapolation use cv). */
static double temperature_from_energy(const zrko_mech* m, int cvol, double E, const double* y,
                                      double guess, double* w) {
  const int max_it = 400;
  const double tol = 1.0e-6, tmin = 100.0, tmax = 5000.0;
  const double emin = mass_energy(m, cvol, tmin, y, w);
  const double emax = mass_energy(m, cvol, tmax, y, w);
  if (E < emin) return tmin - (emin - E) / mass_heat_capacity(m, cvol, tmin, y, w);
  if (E > emax) return tmax - (emax - E) / mass_heat_capacity(m, 1, tmax, y, w);
  double T = guess;
  if (T < tmin || T > tmax) T = tmin + (tmax - tmin) / (emax - emin) * (E - emin);
  for (int i = 0; i < max_it; ++i) {
    const double Ei = mass_energy(m, cvol, T, y, w);
    const double c = mass_heat_capacity(m, 1, T, y, w);
    double d = (E - Ei) / c;
    d = d < -100.0 ? -100.0 : (d > 100.0 ? 100.0 : d);
    if (fabs(d) < tol || T + d == T) break;
    T += d;
  }
  return T;
}

/* Mixture pressure at (T, v, y) (mechanism::getPressureFromTVY, mechanism.cpp:292-297) */
static double pressure_tvy(const zrko_mech* m, double T, double v, const double* y) {
  double mws = 0.0;
  for (int j = 0; j < m->nsp; ++j) mws += y[j] * (1.0 / m->mw[j]);
  return NIST_RU * T / (v * (1.0 / mws));
}

/* InitializeState's source-term part: y0 the initial state, P0 its pressure */
static void reactor_set_sources(reactor_t* r, const double* y0, double P0, double dpdt,
                                double esrc, const double* ysrc) {
  const zrko_mech* m = r->m;
  const int cvol = r->o->constant_volume;
  const double T0 = y0[m->nsp] * r->o->Tref;
  r->dpdt = dpdt;
  r->esrc = esrc;
  r->ysrc = ysrc;
  r->vinit = inv_density_tpy(m, T0, P0, y0);
  r->e0 = mass_energy(m, cvol, T0, y0, r->esc);
  if (!cvol) r->inv_density = r->vinit;
}

/* GetTimeDerivative: constant pressure
 * (reactor_constant_pressure_cpu.cpp:68-133) recomputes the density from P
 * and uses h/RT and cp; constant volume (reactor_constant_volume_cpu.cpp:
 * 66-130) keeps the initial density and uses u/RT = h/RT - 1 and
 * cv_j = (cp_j/R - 1) R/MW_j (mechanism.cpp:381-386, 428-441). */
static int reactor_rhs(reactor_t* r, double t, const double* y, double* ydot) {
  const zrko_mech* m = r->m;
  const int nsp = m->nsp;
  const int cvol = r->o->constant_volume;
  /* the initial temperature when it is not solved for (GetTimeDerivative's
   * initial_temperature_, reactor_constant_volume_cpu.cpp:79-82) */
  const double yT = r->fixT > 0.0 ? r->fixT : y[nsp];
  double T = yT * r->o->Tref;
  if (T <= 0.0) return 1;
  if (T > TLIMIT) T = TLIMIT;
  if (!cvol && !r->ysrc) {
    /* P + dpdt t (reactor_constant_pressure_cpu.cpp:91-94); with mass sources
     * the density keeps its initial value */
    const double Pc = r->P + r->dpdt * t;
    double mws = 0.0;
    for (int j = 0; j < nsp; ++j) mws += y[j] * (1.0 / m->mw[j]);
    const double wmix = 1.0 / mws;
    r->inv_density = 1.0 / (Pc * wmix / (NIST_RU * T));
  }
  if (r->esrc != 0.0)  /* :96-99 (constant volume :87-90) */
    T = temperature_from_energy(m, cvol, r->e0 + r->esrc * t, y, T, r->esc);
  /* vd: this evaluation's 1/density; r->inv_density is the reference's
   * carried member (constant volume: the fixed v0 even when dpdt moves vd) */
  double vd = r->inv_density;
  if (cvol) {
    r->inv_density = vd = r->v0;
    if (r->dpdt != 0.0) /* reactor_constant_volume_cpu.cpp:94-98 */
      vd = inv_density_tpy(m, T, pressure_tvy(m, T, r->v0, y) + r->dpdt * t, y);
  }
  thermo_h(m, T, r->h);
  thermo_cp(m, T, r->cpm);
  double cp = 0.0;
  for (int j = 0; j < nsp; ++j) {
    if (cvol) {
      r->h[j] -= 1.0;
      r->cpm[j] -= 1.0;
    }
    r->cpm[j] *= NIST_RU / m->mw[j];
    cp += r->cpm[j] * y[j];
  }
  r->cp_mix = cp;
  const double dens = 1.0 / vd;
  for (int j = 0; j < nsp; ++j) r->C[j] = dens * (1.0 / m->mw[j]) * y[j];
  rates_ws(m, T, r->C, r->o->step_limiter, r->rop, r->cr, r->de, r->net, r->G);
  double esum = 0.0;
  if (r->ysrc) {
    for (int j = 0; j < nsp; ++j) {
      ydot[j] = (r->net[j] * m->mw[j] * vd + r->ysrc[j]);
      esum += r->h[j] * (r->net[j] * vd + r->ysrc[j] * (1.0 / m->mw[j]));
    }
  } else {
    for (int j = 0; j < nsp; ++j) {
      ydot[j] = (r->net[j] * m->mw[j]) * vd;
      esum += r->h[j] * r->net[j] * vd;
    }
  }
  if (r->fixT > 0.0) return 0;  /* no temperature equation */
  const double dTdt = -esum * NIST_RU * y[nsp] / cp;
  ydot[nsp] = dTdt + (r->esrc + r->dpdt * vd) / (cp * r->o->Tref);
  return 0;
}

int zrko_rhs_src(const zrko_mech* m, const zrko_opts* o, double P0, const double* y0,
                 double t, const double* y, double dpdt, double esrc, const double* ysrc,
                 double* ydot) {
  reactor_t r;
  reactor_init(&r, m, o, P0);
  if (o->constant_volume) r.v0 = inv_density_tpy(m, y0[m->nsp] * o->Tref, P0, y0);
  reactor_set_sources(&r, y0, P0, dpdt, esrc, ysrc);
  int flag = reactor_rhs(&r, t, y, ydot);
  reactor_free(&r);
  return flag;
}

int zrko_rhs(const zrko_mech* m, const zrko_opts* o, double P, double t, const double* y,
             double* ydot) {
  reactor_t r;
  reactor_init(&r, m, o, P);
  if (o->constant_volume) r.v0 = inv_density_tpy(m, y[m->nsp] * o->Tref, P, y);
  int flag = reactor_rhs(&r, t, y, ydot);
  reactor_free(&r);
  return flag;
}

/* SetupJacobianSparse (reactor_nvector_serial.cpp:610-771) in dense column-major
 * storage.  r->inv_density must hold the value of the last RHS call (at y). */
static void reactor_jacobian(reactor_t* r, double t, const double* y, const double* fy,
                             double* J) {
  const zrko_mech* m = r->m;
  const int nsp = m->nsp, N = nsp + 1;
  const double mmf = r->o->min_mass_fraction;
  double* t1 = r->tmp1;
  double* t2 = r->tmp2;
  double* t3 = r->tmp3;
  for (int j = 0; j < nsp; ++j) {
    t1[j] = y[j];
    if (fabs(t1[j]) < mmf) t1[j] = (t1[j] < 0.0) ? -mmf : mmf;
    t2[j] = m->mw[j] / t1[j] * r->inv_density;
  }
  t1[nsp] = y[nsp];
  reactor_rhs(r, t, t1, t3);
  memset(J, 0, sizeof(double) * (size_t)N * N);
  /* every destruction term first, then every creation term, each list in
   * (step, perturbed reactant, row) order: reactor_nvector_serial.cpp:146-172
   * builds the lists in that order and :655-669 applies them one after the
   * other, so an entry fed by both lists sums all its destructions first */
  for (int s = 0; s < m->nstep; ++s)
    for (int k = m->reac_ptr[s]; k < m->reac_ptr[s + 1]; ++k) {
      const int col = m->reac_sp[k];
      for (int q = m->reac_ptr[s]; q < m->reac_ptr[s + 1]; ++q)
        J[col * N + m->reac_sp[q]] -= r->rop[s] * t2[col];
    }
  for (int s = 0; s < m->nstep; ++s)
    for (int k = m->reac_ptr[s]; k < m->reac_ptr[s + 1]; ++k) {
      const int col = m->reac_sp[k];
      for (int q = m->prod_ptr[s]; q < m->prod_ptr[s + 1]; ++q)
        J[col * N + m->prod_sp[q]] += r->rop[s] * t2[col];
    }
  /* non-integer terms (non_integer_reaction_network.cpp:493-506, 774-905):
   * summed per entry in process order (step, exponent species with a non-zero
   * exponent, reactants then products), then added to the entry */
  if (m->n_ni > 0) {
    double* nj = (double*)calloc((size_t)N * N, sizeof(double));
    for (int q = 0; q < m->n_ni; ++q) {
      const double rp = r->rop[m->ni_step[q]];
      for (int k = m->ni_rop_ptr[q]; k < m->ni_rop_ptr[q + 1]; ++k) {
        const double pw = m->ni_rop_pow[k];
        if (pw == 0.0) continue;
        const int col = m->ni_rop_sp[k];
        for (int e = m->ni_re_ptr[q]; e < m->ni_re_ptr[q + 1]; ++e)
          nj[col * N + m->ni_re_sp[e]] += (-m->ni_re_nu[e] * pw) * rp * t2[col];
        for (int e = m->ni_pr_ptr[q]; e < m->ni_pr_ptr[q + 1]; ++e)
          nj[col * N + m->ni_pr_sp[e]] += (m->ni_pr_nu[e] * pw) * rp * t2[col];
      }
    }
    for (int i = 0; i < N * N; ++i) J[i] += nj[i];
    free(nj);
  }
  if (r->fixT > 0.0) {
    /* no temperature row or column: the species block only
     * (reactor_nvector_serial.cpp:641-643, 691-750 under solve_temperature_) */
    for (int j = 0; j < nsp; ++j)
      for (int k = 0; k < nsp; ++k) J[j * N + k] *= m->mw[k] * (1.0 / m->mw[j]);
    return;
  }
  /* d(Tdot)/dy_j */
  const double RuTemp = NIST_RU * y[nsp];
  for (int j = 0; j < nsp; ++j) {
    double acc = 0.0;
    for (int k = 0; k < nsp; ++k) acc += J[j * N + k] * r->h[k];
    J[j * N + nsp] = acc * RuTemp;
  }
  for (int j = 0; j < nsp; ++j) J[j * N + nsp] += t3[nsp] * r->cpm[j] * m->mw[j];
  const double mult = -1.0 / r->cp_mix;
  for (int j = 0; j < nsp; ++j) J[j * N + nsp] *= (1.0 / m->mw[j]) * mult;
  /* d(ydot)/dT by forward difference */
  double dT = y[nsp] * sqrt(DBL_EPSILON);
  t1[nsp] += dT;
  dT = t1[nsp] - y[nsp];
  const double inv = 1.0 / dT;
  for (int j = 0; j < nsp; ++j) t1[j] = y[j];
  reactor_rhs(r, t, t1, t3);
  for (int k = 0; k < N; ++k) J[nsp * N + k] = (t3[k] - fy[k]) * inv;
  /* concentration -> mass fraction columns */
  for (int j = 0; j < nsp; ++j)
    for (int k = 0; k < nsp; ++k) J[j * N + k] *= m->mw[k] * (1.0 / m->mw[j]);
}

void zrko_jacobian(const zrko_mech* m, const zrko_opts* o, double P, double t,
                   const double* y, const double* fy, double* J) {
  reactor_t r;
  reactor_init(&r, m, o, P);
  if (o->constant_volume) r.v0 = inv_density_tpy(m, y[m->nsp] * o->Tref, P, y);
  double* f = (double*)malloc(sizeof(double) * (size_t)(m->nsp + 1));
  reactor_rhs(&r, t, y, f); /* sets the carried inverse density at y */
  reactor_jacobian(&r, t, y, fy, J);
  free(f);
  reactor_free(&r);
}

/* ------------------------------------------------------------------------- */
/* Dense LU with partial pivoting (SUNDIALS sundials_dense.c denseGETRF/RS)     */
/* ------------------------------------------------------------------------- */
int zrko_dense_getrf(int n, double* a, int* p) {
  for (int k = 0; k < n; ++k) {
    double* ck = a + (size_t)k * n;
    int l = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(ck[i]) > fabs(ck[l])) l = i;
    p[k] = l;
    if (ck[l] == 0.0) return k + 1;
    if (l != k)
      for (int i = 0; i < n; ++i) {
        double* ci = a + (size_t)i * n;
        double tmp = ci[l];
        ci[l] = ci[k];
        ci[k] = tmp;
      }
    const double mult = 1.0 / ck[k];
    for (int i = k + 1; i < n; ++i) ck[i] *= mult;
    for (int j = k + 1; j < n; ++j) {
      double* cj = a + (size_t)j * n;
      const double akj = cj[k];
      if (akj != 0.0)
        for (int i = k + 1; i < n; ++i) cj[i] -= akj * ck[i];
    }
  }
  return 0;
}

void zrko_dense_getrs(int n, const double* a, const int* p, double* b) {
  for (int k = 0; k < n; ++k) {
    const int pk = p[k];
    if (pk != k) {
      double tmp = b[k];
      b[k] = b[pk];
      b[pk] = tmp;
    }
  }
  for (int k = 0; k < n - 1; ++k) {
    const double* ck = a + (size_t)k * n;
    const double bk = b[k];
    for (int i = k + 1; i < n; ++i) b[i] -= ck[i] * bk;
  }
  for (int k = n - 1; k > 0; --k) {
    const double* ck = a + (size_t)k * n;
    b[k] /= ck[k];
    const double bk = b[k];
    for (int i = 0; i < k; ++i) b[i] -= ck[i] * bk;
  }
  b[0] /= a[0];
}

/* ------------------------------------------------------------------------- */
/* CVODE v5.8.0 BDF restatement (cvode.c, cvode_ls.c, sunnonlinsol_newton.c)   */
/* with a dense direct linear solver.  Names follow CVODE's.                    */
/* ------------------------------------------------------------------------- */
#define QMAX 5
#define L_MAX (QMAX + 1)
#define HLB_FACTOR 100.0
#define HUB_FACTOR 0.1
#define H_BIAS 0.5
#define MAX_ITERS 4
#define THRESH 1.5
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 1.0e-6
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define ONEPSM 1.000001
#define SMALL_NST 10
#define MXNCF 10
#define MXNEF 7
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define NLS_MAXCOR 3
#define CRDOWN 0.3
#define DGMAX_LSETUP 0.3
#define RDIV 2.0
#define MSBP 20
#define CVLS_MSBJ 51
#define CVLS_DGMAX 0.2

enum { FIRST_CALL = 0, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J, CV_FAIL_OTHER };
enum { DO_ERROR_TEST = 100, PREDICT_AGAIN, TRY_AGAIN };
enum { NLS_SUCCESS = 0, NLS_CONTINUE = 901, NLS_CONV_RECVR = 902, RHSFUNC_RECVR = 903 };
#define CV_SUCCESS 0
#define CV_TOO_MUCH_WORK (-1)
#define CV_TOO_MUCH_ACC (-2)
#define CV_ERR_FAILURE (-3)
#define CV_CONV_FAILURE (-4)
#define CV_REPTD_RHSFUNC_ERR (-10)
#define CV_ILL_INPUT (-22)

#define SPGMR_MAXL 5 /* SUNLinSol_SPGMR(state, PREC_LEFT, 0): maxl 0 -> 5 */

typedef struct {
  reactor_t* r;
  const zrko_opts* o;
  int N;      /* state variables: nsp + 1, or nsp at fixed temperature */
  int nsp;
  int direct; /* CVODE's dense direct solver (dense = 1, iterative = 0) */
  double* zn[L_MAX];
  double *ewt, *acor, *tempv, *ftemp, *y, *atolv, *delta;
  double *savedJ, *M;
  int* piv;
  const char* pat; /* [N*N] the analytic Jacobian's pattern (preconditioner) */
  /* SPGMR work: Krylov basis V[0..maxl], its Hessenberg matrix and rotations */
  double* V[SPGMR_MAXL + 1];
  double *xcor, *vtemp, *work, *jv;
  double hes[SPGMR_MAXL + 1][SPGMR_MAXL];
  double givens[2 * SPGMR_MAXL];
  double rtol;
  double tn, h, hscale, hprime, next_h, eta, etamax, hmin, hmax_inv, hu, h0u;
  double tau[L_MAX + 1], tq[6], l[L_MAX];
  double rl1, gamma, gammap, gamrat, crate, delp, acnrm, nlscoef, saved_tq5, tolsf;
  double etaq, etaqm1, etaqp1;
  int q, qprime, next_q, qwait, L, qu, indx_acor;
  long nst, nstlp, nstlj, nfe, nje, nsetups, nni, ncfn, netf, nscon;
  int convfail, jcur, acnrmcur;
} cvmem;

static double wrms(const double* x, const double* w, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    const double p = x[i] * w[i];
    s += p * p;
  }
  return sqrt(s / n);
}

static int cv_ewt(cvmem* cv, const double* y) {
  for (int i = 0; i < cv->N; ++i) {
    const double v = cv->rtol * fabs(y[i]) + cv->atolv[i];
    if (v <= 0.0) return -1;
    cv->ewt[i] = 1.0 / v;
  }
  return 0;
}

static int cv_f(cvmem* cv, double t, const double* y, double* yd) {
  cv->nfe++;
  return reactor_rhs(cv->r, t, y, yd);
}

static void cv_reinit(cvmem* cv, double t0, const double* y0) {
  memcpy(cv->zn[0], y0, sizeof(double) * (size_t)cv->N);
  cv->tn = t0;
  cv->q = 1;
  cv->L = 2;
  cv->qwait = cv->L;
  cv->etamax = ETAMX1;
  cv->qu = 0;
  cv->hu = 0.0;
  cv->tolsf = 1.0;
  cv->nst = cv->nfe = cv->ncfn = cv->netf = cv->nni = cv->nsetups = 0;
  cv->nstlp = 0;
  cv->nstlj = 0;
  cv->nje = 0;
  cv->nscon = 0;
  cv->hmin = 0.0;
  cv->h0u = 0.0;
  cv->next_h = 0.0;
  cv->next_q = 0;
  cv->indx_acor = QMAX;
  cv->saved_tq5 = 0.0;
  cv->crate = 1.0;
  cv->delp = 0.0;
  cv->gammap = 0.0;
  cv->gamrat = 1.0;
  for (int i = 0; i <= L_MAX; ++i) cv->tau[i] = 0.0;
  for (int i = 0; i < 6; ++i) cv->tq[i] = 0.0;
  for (int i = 0; i < L_MAX; ++i) cv->l[i] = 0.0;
}

static double cv_upper_bound_h0(cvmem* cv, double tdist) {
  double hub_inv = 0.0;
  for (int i = 0; i < cv->N; ++i) {
    const double t1 = HUB_FACTOR * fabs(cv->zn[0][i]) + 1.0 / cv->ewt[i];
    const double t2 = fabs(cv->zn[1][i]) / t1;
    if (t2 > hub_inv) hub_inv = t2;
  }
  double hub = HUB_FACTOR * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  return hub;
}

static int cv_ydd_norm(cvmem* cv, double hg, double* yddnrm) {
  for (int i = 0; i < cv->N; ++i) cv->y[i] = hg * cv->zn[1][i] + cv->zn[0][i];
  int rv = cv_f(cv, cv->tn + hg, cv->y, cv->tempv);
  if (rv > 0) return 1;
  for (int i = 0; i < cv->N; ++i) cv->tempv[i] = (cv->tempv[i] - cv->zn[1][i]) * (1.0 / hg);
  *yddnrm = wrms(cv->tempv, cv->ewt, cv->N);
  return 0;
}

static int cv_hin(cvmem* cv, double tout) {
  const double tdiff = tout - cv->tn;
  if (tdiff == 0.0) return -1;
  const double tdist = fabs(tdiff);
  const double tround = DBL_EPSILON * fmax(fabs(cv->tn), fabs(tout));
  if (tdist < 2.0 * tround) return -1;
  const double hlb = HLB_FACTOR * tround;
  const double hub = cv_upper_bound_h0(cv, tdist);
  double hg = sqrt(hlb * hub);
  if (hub < hlb) {
    cv->h = hg;
    return 0;
  }
  double hnew = hg, hs = hg, yddnrm = 0.0;
  for (int count1 = 1; count1 <= MAX_ITERS; ++count1) {
    int hgOK = 0;
    for (int count2 = 1; count2 <= MAX_ITERS; ++count2) {
      if (cv_ydd_norm(cv, hg, &yddnrm) == 0) {
        hgOK = 1;
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
    if (count1 == MAX_ITERS) break;
    const double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) break;
    if (count1 > 1 && hrat > 2.0) {
      hnew = hg;
      break;
    }
    hg = hnew;
  }
  double h0 = H_BIAS * hnew;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  cv->h = h0;
  return 0;
}

static void cv_rescale(cvmem* cv) {
  double f = cv->eta;
  for (int j = 1; j <= cv->q; ++j) {
    for (int i = 0; i < cv->N; ++i) cv->zn[j][i] *= f;
    f *= cv->eta;
  }
  cv->h = cv->hscale * cv->eta;
  cv->next_h = cv->h;
  cv->hscale = cv->h;
  cv->nscon = 0;
}

static void cv_increase_bdf(cvmem* cv) {
  for (int i = 0; i <= QMAX; ++i) cv->l[i] = 0.0;
  cv->l[2] = 1.0;
  double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
  double hsum = cv->hscale;
  if (cv->q > 1) {
    for (int j = 1; j < cv->q; ++j) {
      hsum += cv->tau[j + 1];
      const double xi = hsum / cv->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; --i) cv->l[i] = cv->l[i] * xiold + cv->l[i - 1];
      xiold = xi;
    }
  }
  const double A1 = (-alpha0 - alpha1) / prod;
  double* zL = cv->zn[cv->L];
  const double* za = cv->zn[cv->indx_acor];
  for (int i = 0; i < cv->N; ++i) zL[i] = A1 * za[i];
  for (int j = 2; j <= cv->q; ++j)
    for (int i = 0; i < cv->N; ++i) cv->zn[j][i] += cv->l[j] * zL[i];
}

static void cv_decrease_bdf(cvmem* cv) {
  for (int i = 0; i <= QMAX; ++i) cv->l[i] = 0.0;
  cv->l[2] = 1.0;
  double hsum = 0.0;
  for (int j = 1; j <= cv->q - 2; ++j) {
    hsum += cv->tau[j];
    const double xi = hsum / cv->hscale;
    for (int i = j + 2; i >= 2; --i) cv->l[i] = cv->l[i] * xi + cv->l[i - 1];
  }
  for (int j = 2; j < cv->q; ++j)
    for (int i = 0; i < cv->N; ++i) cv->zn[j][i] = -cv->l[j] * cv->zn[cv->q][i] + cv->zn[j][i];
}

static void cv_adjust_order(cvmem* cv, int dq) {
  if (cv->q == 2 && dq != 1) return;
  if (dq == 1)
    cv_increase_bdf(cv);
  else
    cv_decrease_bdf(cv);
}

static void cv_adjust_params(cvmem* cv) {
  if (cv->qprime != cv->q) {
    cv_adjust_order(cv, cv->qprime - cv->q);
    cv->q = cv->qprime;
    cv->L = cv->q + 1;
    cv->qwait = cv->L;
  }
  cv_rescale(cv);
}

static void cv_predict(cvmem* cv) {
  cv->tn += cv->h;
  for (int k = 1; k <= cv->q; ++k)
    for (int j = cv->q; j >= k; --j)
      for (int i = 0; i < cv->N; ++i) cv->zn[j - 1][i] += cv->zn[j][i];
}

static void cv_restore(cvmem* cv, double saved_t) {
  cv->tn = saved_t;
  for (int k = 1; k <= cv->q; ++k)
    for (int j = cv->q; j >= k; --j)
      for (int i = 0; i < cv->N; ++i) cv->zn[j - 1][i] -= cv->zn[j][i];
}

static void cv_set_tq_bdf(cvmem* cv, double hsum, double alpha0, double alpha0_hat,
                          double xi_inv, double xistar_inv) {
  const int q = cv->q;
  const double A1 = 1.0 - alpha0_hat + alpha0;
  const double A2 = 1.0 + q * A1;
  cv->tq[2] = fabs(A1 / (alpha0 * A2));
  cv->tq[5] = fabs(A2 * xistar_inv / (cv->l[q] * xi_inv));
  if (cv->qwait == 1) {
    if (q > 1) {
      const double C = xistar_inv / cv->l[q];
      const double A3 = alpha0 + 1.0 / q;
      const double A4 = alpha0_hat + xi_inv;
      const double Cpinv = (1.0 - A4 + A3) / A3;
      cv->tq[1] = fabs(C * Cpinv);
    } else {
      cv->tq[1] = 1.0;
    }
    hsum += cv->tau[q];
    xi_inv = cv->h / hsum;
    const double A5 = alpha0 - (1.0 / (q + 1));
    const double A6 = alpha0_hat - xi_inv;
    const double Cppinv = (1.0 - A6 + A5) / A2;
    cv->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  cv->tq[4] = cv->nlscoef / cv->tq[2];
}

static void cv_set(cvmem* cv) {
  const int q = cv->q;
  double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  cv->l[0] = cv->l[1] = xi_inv = xistar_inv = 1.0;
  for (int i = 2; i <= q; ++i) cv->l[i] = 0.0;
  alpha0 = alpha0_hat = -1.0;
  hsum = cv->h;
  if (q > 1) {
    for (int j = 2; j < q; ++j) {
      hsum += cv->tau[j - 1];
      xi_inv = cv->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; --i) cv->l[i] += cv->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -cv->l[1] - alpha0;
    hsum += cv->tau[q - 1];
    xi_inv = cv->h / hsum;
    alpha0_hat = -cv->l[1] - xi_inv;
    for (int i = q; i >= 1; --i) cv->l[i] += cv->l[i - 1] * xistar_inv;
  }
  cv_set_tq_bdf(cv, hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
  cv->rl1 = 1.0 / cv->l[1];
  cv->gamma = cv->h * cv->rl1;
  if (cv->nst == 0) cv->gammap = cv->gamma;
  cv->gamrat = (cv->nst > 0) ? cv->gamma / cv->gammap : 1.0;
}

/* cvLsDenseDQJac (SUNDIALS 5.8.0 cvode_ls.c): column j is
 * (f(y + inc e_j) - f(y)) / inc with inc = max(sqrt(uround) |y_j|,
 * minInc / ewt_j), minInc = 1000 |h| uround N wrms(f(y)) (1 when that norm is
 * 0); y is restored exactly.  J column-major with stride N. */
static int cv_dq_jac(cvmem* cv, double* y, const double* fy, double* J) {
  const int N = cv->N;
  const double srur = sqrt(DBL_EPSILON);
  const double fnorm = wrms(fy, cv->ewt, N);
  const double minInc = (fnorm != 0.0) ? (1000.0 * fabs(cv->h) * DBL_EPSILON * N * fnorm) : 1.0;
  double* ftemp = cv->work;
  for (int j = 0; j < N; ++j) {
    const double yjsaved = y[j];
    const double inc = fmax(srur * fabs(yjsaved), minInc / cv->ewt[j]);
    y[j] += inc;
    const int rv = cv_f(cv, cv->tn, y, ftemp);
    y[j] = yjsaved;
    if (rv != 0) return rv;
    const double inc_inv = 1.0 / inc;
    for (int i = 0; i < N; ++i) J[(size_t)j * N + i] = inc_inv * (ftemp[i] - fy[i]);
  }
  return 0;
}

/* ReactorNVectorSerial::DividedDifferenceJacobian
 * (reactor_nvector_serial.cpp:816-835): a fresh f(y), then column i is
 * (f(y + delta e_i) - f(y)) / delta with delta = sqrt(1e-16 max(1e-5, |y_i|)). */
static void reactor_dq_jac(cvmem* cv, double* y, double* J) {
  const int N = cv->N;
  double* f0 = cv->work;
  double* f1 = cv->jv;
  reactor_rhs(cv->r, cv->tn, y, f0);
  for (int i = 0; i < N; ++i) {
    const double ysafe = y[i];
    const double delta = sqrt(1.0e-16 * fmax(1.0e-5, fabs(ysafe)));
    y[i] = ysafe + delta;
    reactor_rhs(cv->r, cv->tn, y, f1);
    for (int j = 0; j < N; ++j) J[(size_t)i * N + j] = (f1[j] - f0[j]) / delta;
    y[i] = ysafe;
  }
}

/* The analytic Jacobian (reactor_jacobian) in CVODE's N x N layout: at fixed
 * temperature the species block of the (nsp+1)-strided matrix. */
static void cv_analytic_jac(cvmem* cv, const double* y, const double* fy, double* J) {
  const int nsp = cv->nsp;
  if (cv->N == nsp + 1) {
    reactor_jacobian(cv->r, cv->tn, y, fy, J);
    return;
  }
  const int Nf = nsp + 1;
  double* full = (double*)calloc((size_t)Nf * Nf, sizeof(double));
  double* yf = (double*)malloc(sizeof(double) * (size_t)Nf);
  memcpy(yf, y, sizeof(double) * (size_t)nsp);
  yf[nsp] = cv->r->fixT;
  reactor_jacobian(cv->r, cv->tn, yf, NULL, full);
  for (int j = 0; j < nsp; ++j)
    for (int i = 0; i < nsp; ++i) J[(size_t)j * nsp + i] = full[(size_t)j * Nf + i];
  free(yf);
  free(full);
}

/* cvLsSetup with the dense direct solver (cvLsLinSys + SUNLinSolSetup_Dense)
 * or, for SPGMR, the preconditioner setup ReactorJacobianSetupAndFactor
 * (reactor_base.cpp:96-109: Jacobian when CVODE does not vouch for the saved
 * one, then JacobianFactor / FormPreconditioner, reactor_nvector_serial.cpp:
 * 373-458).  The jbad test and its bookkeeping are CVODE's for both.
 * Returns 0, >0 recoverable. */
static int cv_lsetup(cvmem* cv, int convfail, double* ypred, const double* fpred, int* jcur) {
  const int N = cv->N;
  const zrko_opts* o = cv->o;
  const double dgamma = fabs(cv->gamma / cv->gammap - 1.0);
  const int jbad = (cv->nst == 0) || (cv->nst >= cv->nstlj + CVLS_MSBJ) ||
                   ((convfail == CV_FAIL_BAD_J) && (dgamma < CVLS_DGMAX)) ||
                   (convfail == CV_FAIL_OTHER);
  if (!jbad) {
    *jcur = 0;
  } else {
    *jcur = 1;
    cv->nje++;
    cv->nstlj = cv->nst;
    if (o->analytic) {
      cv_analytic_jac(cv, ypred, fpred, cv->savedJ);
    } else if (cv->direct) {
      if (cv_dq_jac(cv, ypred, fpred, cv->savedJ) != 0) return 1;
    } else {
      reactor_dq_jac(cv, ypred, cv->savedJ);
    }
  }
  const double g = cv->gamma;
  if (cv->direct || (o->dense && !o->analytic)) {
    /* M = I - gamma J, every entry */
    for (int j = 0; j < N; ++j) {
      for (int i = 0; i < N; ++i) cv->M[j * N + i] = -g * cv->savedJ[j * N + i];
      cv->M[j * N + j] += 1.0;
    }
  } else {
    /* FormPreconditioner: the Jacobian's pattern, diagonal kept, other
     * entries only when |-gamma J_ij| > threshold (0 when iterative = 0) */
    const double thr = o->iterative ? o->preconditioner_threshold : 0.0;
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double v = 0.0;
        if (cv->pat[j * N + i]) {
          v = -g * cv->savedJ[j * N + i];
          if (i == j) v += 1.0;
          else if (!(fabs(v) > thr)) v = 0.0;
        }
        cv->M[j * N + i] = v;
      }
  }
  return zrko_dense_getrf(N, cv->M, cv->piv) != 0 ? 1 : 0;
}

/* cvLsATimes with cvLsDQJtimes: z = v - gamma J v, J v by the difference
 * quotient (f(y + sig v) - f(y)) / sig at the Newton iterate, sig = 1 /
 * wrms(v), shrunk by 4 on a recoverable RHS failure (3 tries). */
static int cv_atimes(cvmem* cv, const double* v, double* z) {
  const int N = cv->N;
  double sig = 1.0 / wrms(v, cv->ewt, N);
  int rv = 1;
  for (int it = 0; it < 3; ++it) {
    for (int i = 0; i < N; ++i) cv->work[i] = sig * v[i] + cv->y[i];
    rv = cv_f(cv, cv->tn, cv->work, cv->jv);
    if (rv == 0) break;
    if (rv < 0) return -1;
    sig *= 0.25;
  }
  if (rv > 0) return 1;
  const double siginv = 1.0 / sig;
  for (int i = 0; i < N; ++i) cv->jv[i] = siginv * (cv->jv[i] - cv->ftemp[i]);
  for (int i = 0; i < N; ++i) z[i] = -cv->gamma * cv->jv[i] + v[i];
  return 0;
}

static double dotp(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* SUNModifiedGS: orthogonalise V[k] against V[0..k-1] into column k-1 of hes,
 * reorthogonalising when the new norm lost more than 1000 x roundoff. */
static void cv_mgs(cvmem* cv, int k, double* new_norm) {
  const int N = cv->N;
  double* vk = cv->V[k];
  const double vk_norm = sqrt(dotp(vk, vk, N));
  for (int i = 0; i < k; ++i) {
    cv->hes[i][k - 1] = dotp(cv->V[i], vk, N);
    const double h = cv->hes[i][k - 1];
    for (int e = 0; e < N; ++e) vk[e] = vk[e] - h * cv->V[i][e];
  }
  *new_norm = sqrt(dotp(vk, vk, N));
  double temp = 1000.0 * vk_norm;
  if ((temp + *new_norm) != temp) return;
  double new_norm_2 = 0.0;
  for (int i = 0; i < k; ++i) {
    const double np = dotp(cv->V[i], vk, N);
    temp = 1000.0 * cv->hes[i][k - 1];
    if ((temp + np) == temp) continue;
    cv->hes[i][k - 1] += np;
    for (int e = 0; e < N; ++e) vk[e] = vk[e] - np * cv->V[i][e];
    new_norm_2 += np * np;
  }
  if (new_norm_2 != 0.0) {
    const double np = (*new_norm) * (*new_norm) - new_norm_2;
    *new_norm = (np > 0.0) ? sqrt(np) : 0.0;
  }
}

/* one Givens rotation annihilating b against a: (c, s) with c a - s b = r */
static void givens_of(double a, double b, double* c, double* s) {
  if (b == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else if (fabs(b) >= fabs(a)) {
    const double t = a / b;
    *s = -1.0 / sqrt(1.0 + t * t);
    *c = -*s * t;
  } else {
    const double t = b / a;
    *c = 1.0 / sqrt(1.0 + t * t);
    *s = -*c * t;
  }
}

/* SUNQRfact for column l of the Hessenberg matrix (the earlier columns are
 * already factored): apply the previous rotations, then a new one; nonzero
 * when the new diagonal entry vanishes. */
static int cv_qrfact(cvmem* cv, int l) {
  for (int k = 0; k < l; ++k) {
    const double c = cv->givens[2 * k], s = cv->givens[2 * k + 1];
    const double t1 = cv->hes[k][l], t2 = cv->hes[k + 1][l];
    cv->hes[k][l] = c * t1 - s * t2;
    cv->hes[k + 1][l] = s * t1 + c * t2;
  }
  double c, s;
  const double t1 = cv->hes[l][l], t2 = cv->hes[l + 1][l];
  givens_of(t1, t2, &c, &s);
  cv->givens[2 * l] = c;
  cv->givens[2 * l + 1] = s;
  cv->hes[l][l] = c * t1 - s * t2;
  return cv->hes[l][l] == 0.0 ? l + 1 : 0;
}

/* SUNQRsol: y = R^{-1} Q b for the factored n-column Hessenberg matrix */
static int cv_qrsol(cvmem* cv, int n, double* b) {
  for (int k = 0; k < n; ++k) {
    const double c = cv->givens[2 * k], s = cv->givens[2 * k + 1];
    const double t1 = b[k], t2 = b[k + 1];
    b[k] = c * t1 - s * t2;
    b[k + 1] = s * t1 + c * t2;
  }
  for (int k = n - 1; k >= 0; --k) {
    if (cv->hes[k][k] == 0.0) return k + 1;
    b[k] /= cv->hes[k][k];
    for (int i = 0; i < k; ++i) b[i] -= b[k] * cv->hes[i][k];
  }
  return 0;
}

/* ReactorJacobianSolveCVODE: z = P^{-1} r with the factored preconditioner */
static void cv_psolve(cvmem* cv, const double* r, double* z) {
  memcpy(z, r, sizeof(double) * (size_t)cv->N);
  zrko_dense_getrs(cv->N, cv->M, cv->piv, z);
}

enum { LS_SUCCESS = 0, LS_RES_REDUCED, LS_CONV_FAIL, LS_REC_FAIL, LS_UNREC_FAIL };

/* SUNLinSolSolve_SPGMR with left preconditioning, scaling s1 = s2 = ewt, no
 * restarts, initial guess 0: x (returned in b) for A x = b to within delta in
 * the scaled preconditioned 2-norm. */
static int cv_spgmr(cvmem* cv, double* b, double delta) {
  const int N = cv->N;
  const double* s = cv->ewt;
  /* V[0] = s1 P^{-1} b (the initial guess is 0, so the residual is b) */
  cv_psolve(cv, b, cv->vtemp);
  for (int i = 0; i < N; ++i) cv->V[0][i] = s[i] * cv->vtemp[i];
  const double beta = sqrt(dotp(cv->V[0], cv->V[0], N));
  double rho = beta;
  if (beta <= delta) {  /* the initial guess 0 is returned */
    for (int i = 0; i < N; ++i) b[i] = 0.0;
    return LS_SUCCESS;
  }
  for (int i = 0; i <= SPGMR_MAXL; ++i)
    for (int j = 0; j < SPGMR_MAXL; ++j) cv->hes[i][j] = 0.0;
  double rotation_product = 1.0;
  const double rinv = 1.0 / beta;
  for (int i = 0; i < N; ++i) cv->V[0][i] *= rinv;
  int krydim = 0, converged = 0;
  for (int l = 0; l < SPGMR_MAXL; ++l) {
    krydim = l + 1;
    for (int i = 0; i < N; ++i) cv->vtemp[i] = cv->V[l][i] / s[i];
    const int ar = cv_atimes(cv, cv->vtemp, cv->V[l + 1]);
    if (ar != 0) return ar < 0 ? LS_UNREC_FAIL : LS_REC_FAIL;
    cv_psolve(cv, cv->V[l + 1], cv->vtemp);
    for (int i = 0; i < N; ++i) cv->V[l + 1][i] = s[i] * cv->vtemp[i];
    cv_mgs(cv, l + 1, &cv->hes[l + 1][l]);
    if (cv_qrfact(cv, l) != 0) return LS_REC_FAIL;
    rotation_product *= cv->givens[2 * l + 1];
    rho = fabs(rotation_product * beta);
    if (rho <= delta) {
      converged = 1;
      break;
    }
    const double hinv = 1.0 / cv->hes[l + 1][l];
    for (int i = 0; i < N; ++i) cv->V[l + 1][i] *= hinv;
  }
  double yg[SPGMR_MAXL + 1];
  yg[0] = beta;
  for (int i = 1; i <= krydim; ++i) yg[i] = 0.0;
  if (cv_qrsol(cv, krydim, yg) != 0) return LS_UNREC_FAIL;
  double* xc = cv->xcor;
  for (int i = 0; i < N; ++i) xc[i] = 0.0;
  for (int k = 0; k < krydim; ++k)
    for (int i = 0; i < N; ++i) xc[i] += yg[k] * cv->V[k][i];
  if (converged || rho < beta) {
    for (int i = 0; i < N; ++i) b[i] = xc[i] / s[i];
    return converged ? LS_SUCCESS : LS_RES_REDUCED;
  }
  for (int i = 0; i < N; ++i) b[i] = 0.0;
  return LS_CONV_FAIL;
}

/* cvLsSolve: b <- the Newton correction.  Dense direct: the LU solve and the
 * BDF rescaling 2/(1 + gamrat) for a lagged gamma; SPGMR: the small-residual
 * shortcut, tolerance eps_lin tq[4] sqrt(N), and the solver's outcome as
 * CVODE reads it (a reduced residual is accepted on the first Newton
 * iteration only).  Returns 0, >0 recoverable, <0 unrecoverable. */
static int cv_lsolve(cvmem* cv, double* b, int curiter) {
  const int N = cv->N;
  if (cv->direct) {
    zrko_dense_getrs(N, cv->M, cv->piv, b);
    if (cv->gamrat != 1.0) {
      const double sc = 2.0 / (1.0 + cv->gamrat);
      for (int i = 0; i < N; ++i) b[i] *= sc;
    }
    return 0;
  }
  const double deltar = cv->o->eps_lin * cv->tq[4];
  const double bnorm = wrms(b, cv->ewt, N);
  if (bnorm <= deltar) {
    if (curiter > 0)
      for (int i = 0; i < N; ++i) b[i] = 0.0;
    return 0;
  }
  const int rv = cv_spgmr(cv, b, deltar * sqrt((double)N));
  switch (rv) {
    case LS_SUCCESS: return 0;
    case LS_RES_REDUCED: return curiter == 0 ? 0 : 1;
    case LS_UNREC_FAIL: return -1;
    default: return 1;
  }
}

/* cvNlsLSetup */
static int cv_nls_lsetup(cvmem* cv, int jbad, int* jcur) {
  if (jbad) cv->convfail = CV_FAIL_BAD_J;
  int rv = cv_lsetup(cv, cv->convfail, cv->y, cv->ftemp, jcur);
  cv->nsetups++;
  cv->gamrat = 1.0;
  cv->gammap = cv->gamma;
  cv->crate = 1.0;
  cv->nstlp = cv->nst;
  return rv > 0 ? NLS_CONV_RECVR : NLS_SUCCESS;
}

/* cvNlsResidual */
static int cv_nls_residual(cvmem* cv, const double* ycor, double* res) {
  const int N = cv->N;
  for (int i = 0; i < N; ++i) cv->y[i] = cv->zn[0][i] + ycor[i];
  if (cv_f(cv, cv->tn, cv->y, cv->ftemp) > 0) return RHSFUNC_RECVR;
  for (int i = 0; i < N; ++i) {
    res[i] = cv->rl1 * cv->zn[1][i] + ycor[i];
    res[i] = -cv->gamma * cv->ftemp[i] + res[i];
  }
  return NLS_SUCCESS;
}

/* cvNlsConvTest */
static int cv_nls_convtest(cvmem* cv, int m, const double* ycor, const double* delta,
                           double tol) {
  const double del = wrms(delta, cv->ewt, cv->N);
  if (m > 0) cv->crate = fmax(CRDOWN * cv->crate, del / cv->delp);
  const double dcon = del * fmin(1.0, cv->crate) / tol;
  if (dcon <= 1.0) {
    cv->acnrm = (m == 0) ? del : wrms(ycor, cv->ewt, cv->N);
    cv->acnrmcur = 1;
    return NLS_SUCCESS;
  }
  if (m >= 1 && del > RDIV * cv->delp) return NLS_CONV_RECVR;
  cv->delp = del;
  return NLS_CONTINUE;
}

/* SUNNonlinSolSolve_Newton (sunnonlinsol_newton.c) with CVODE's callbacks */
static int cv_newton(cvmem* cv, double* ycor, double tol, int callLSetup) {
  const int N = cv->N;
  double* delta = cv->delta;
  int jbad = 0, jcur = 0, rv;
  for (;;) {
    rv = cv_nls_residual(cv, ycor, delta);
    if (rv != NLS_SUCCESS) goto fail;
    if (callLSetup) {
      rv = cv_nls_lsetup(cv, jbad, &jcur);
      cv->jcur = jcur;
      if (rv != NLS_SUCCESS) goto fail;
    }
    {
      int curiter = 0;
      for (;;) {
        cv->nni++;
        for (int i = 0; i < N; ++i) delta[i] = -delta[i];
        const int ls = cv_lsolve(cv, delta, curiter);
        if (ls != 0) {  /* (an unrecoverable solver failure is treated alike) */
          rv = NLS_CONV_RECVR;
          break;
        }
        for (int i = 0; i < N; ++i) ycor[i] += delta[i];
        rv = cv_nls_convtest(cv, curiter, ycor, delta, tol);
        if (rv == NLS_SUCCESS) {
          cv->jcur = 0;
          return NLS_SUCCESS;
        }
        if (rv != NLS_CONTINUE) break;
        curiter++;
        if (curiter >= NLS_MAXCOR) {
          rv = NLS_CONV_RECVR;
          break;
        }
        rv = cv_nls_residual(cv, ycor, delta);
        if (rv != NLS_SUCCESS) break;
      }
    }
  fail:
    if (rv > 0 && !cv->jcur) {
      callLSetup = 1;
      jbad = 1;
      for (int i = 0; i < N; ++i) ycor[i] = 0.0;
      continue;
    }
    return rv;
  }
}

/* cvNls */
static int cv_nls(cvmem* cv, int nflag) {
  cv->convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? CV_NO_FAILURES : CV_FAIL_OTHER;
  const int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (cv->nst == 0) ||
                        (cv->nst >= cv->nstlp + MSBP) ||
                        (fabs(cv->gamrat - 1.0) > DGMAX_LSETUP);
  for (int i = 0; i < cv->N; ++i) cv->acor[i] = 0.0;
  cv->acnrmcur = 0;
  const int rv = cv_newton(cv, cv->acor, cv->tq[4], callSetup);
  if (rv != NLS_SUCCESS) return rv;
  for (int i = 0; i < cv->N; ++i) cv->y[i] = cv->zn[0][i] + cv->acor[i];
  if (!cv->acnrmcur) cv->acnrm = wrms(cv->acor, cv->ewt, cv->N);
  cv->jcur = 0;
  return CV_SUCCESS;
}

static int cv_handle_nflag(cvmem* cv, int* nflagPtr, double saved_t, int* ncfPtr) {
  const int nflag = *nflagPtr;
  if (nflag == CV_SUCCESS) return DO_ERROR_TEST;
  cv->ncfn++;
  cv_restore(cv, saved_t);
  (*ncfPtr)++;
  cv->etamax = 1.0;
  if (fabs(cv->h) <= cv->hmin * ONEPSM || *ncfPtr == MXNCF) {
    if (nflag == NLS_CONV_RECVR) return CV_CONV_FAILURE;
    return CV_REPTD_RHSFUNC_ERR;
  }
  cv->eta = fmax(ETACF, cv->hmin / fabs(cv->h));
  *nflagPtr = PREV_CONV_FAIL;
  cv_rescale(cv);
  return PREDICT_AGAIN;
}

static int cv_do_error_test(cvmem* cv, int* nflagPtr, double saved_t, int* nefPtr,
                            double* dsmPtr) {
  const double dsm = cv->acnrm * cv->tq[2];
  *dsmPtr = dsm;
  if (dsm <= 1.0) return CV_SUCCESS;
  (*nefPtr)++;
  cv->netf++;
  *nflagPtr = PREV_ERR_FAIL;
  cv_restore(cv, saved_t);
  if (fabs(cv->h) <= cv->hmin * ONEPSM || *nefPtr == MXNEF) return CV_ERR_FAILURE;
  cv->etamax = 1.0;
  if (*nefPtr <= MXNEF1) {
    cv->eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / cv->L) + ADDON);
    cv->eta = fmax(ETAMIN, fmax(cv->eta, cv->hmin / fabs(cv->h)));
    if (*nefPtr >= SMALL_NEF) cv->eta = fmin(cv->eta, ETAMXF);
    cv_rescale(cv);
    return TRY_AGAIN;
  }
  if (cv->q > 1) {
    cv->eta = fmax(ETAMIN, cv->hmin / fabs(cv->h));
    cv_adjust_order(cv, -1);
    cv->L = cv->q;
    cv->q--;
    cv->qwait = cv->L;
    cv_rescale(cv);
    return TRY_AGAIN;
  }
  cv->eta = fmax(ETAMIN, cv->hmin / fabs(cv->h));
  cv->h *= cv->eta;
  cv->next_h = cv->h;
  cv->hscale = cv->h;
  cv->qwait = LONG_WAIT;
  cv->nscon = 0;
  if (cv_f(cv, cv->tn, cv->zn[0], cv->tempv) > 0) return CV_REPTD_RHSFUNC_ERR;
  for (int i = 0; i < cv->N; ++i) cv->zn[1][i] = cv->h * cv->tempv[i];
  return TRY_AGAIN;
}

static void cv_complete_step(cvmem* cv) {
  cv->nst++;
  cv->nscon++;
  cv->hu = cv->h;
  cv->qu = cv->q;
  for (int i = cv->q; i >= 2; --i) cv->tau[i] = cv->tau[i - 1];
  if (cv->q == 1 && cv->nst > 1) cv->tau[2] = cv->tau[1];
  cv->tau[1] = cv->h;
  for (int j = 0; j <= cv->q; ++j)
    for (int i = 0; i < cv->N; ++i) cv->zn[j][i] += cv->l[j] * cv->acor[i];
  cv->qwait--;
  if (cv->qwait == 1 && cv->q != QMAX) {
    memcpy(cv->zn[QMAX], cv->acor, sizeof(double) * (size_t)cv->N);
    cv->saved_tq5 = cv->tq[5];
    cv->indx_acor = QMAX;
  }
}

static void cv_set_eta(cvmem* cv) {
  if (cv->eta < THRESH) {
    cv->eta = 1.0;
    cv->hprime = cv->h;
  } else {
    cv->eta = fmin(cv->eta, cv->etamax);
    cv->eta /= fmax(1.0, fabs(cv->h) * cv->hmax_inv * cv->eta);
    cv->hprime = cv->h * cv->eta;
    if (cv->qprime < cv->q) cv->nscon = 0;
  }
}

static double cv_etaqm1(cvmem* cv) {
  cv->etaqm1 = 0.0;
  if (cv->q > 1) {
    const double ddn = wrms(cv->zn[cv->q], cv->ewt, cv->N) * cv->tq[1];
    cv->etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / cv->q) + ADDON);
  }
  return cv->etaqm1;
}

static double cv_etaqp1(cvmem* cv) {
  cv->etaqp1 = 0.0;
  if (cv->q != QMAX) {
    if (cv->saved_tq5 == 0.0) return cv->etaqp1;
    double hr = 1.0; /* SUNRpowerI: repeated product */
    for (int i = 0; i < cv->L; ++i) hr *= cv->h / cv->tau[2];
    const double cquot = (cv->tq[5] / cv->saved_tq5) * hr;
    for (int i = 0; i < cv->N; ++i) cv->tempv[i] = -cquot * cv->zn[QMAX][i] + cv->acor[i];
    const double dup = wrms(cv->tempv, cv->ewt, cv->N) * cv->tq[3];
    cv->etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (cv->L + 1)) + ADDON);
  }
  return cv->etaqp1;
}

static void cv_choose_eta(cvmem* cv) {
  const double etam = fmax(cv->etaqm1, fmax(cv->etaq, cv->etaqp1));
  if (etam < THRESH) {
    cv->eta = 1.0;
    cv->qprime = cv->q;
    return;
  }
  if (etam == cv->etaqm1) {
    cv->eta = cv->etaqm1;
    cv->qprime = cv->q - 1;
  } else if (etam == cv->etaq) {
    cv->eta = cv->etaq;
    cv->qprime = cv->q;
  } else {
    cv->eta = cv->etaqp1;
    cv->qprime = cv->q + 1;
    memcpy(cv->zn[QMAX], cv->acor, sizeof(double) * (size_t)cv->N);
  }
}

static void cv_prepare_next_step(cvmem* cv, double dsm) {
  if (cv->etamax == 1.0) {
    cv->qwait = (cv->qwait > 2) ? cv->qwait : 2;
    cv->qprime = cv->q;
    cv->hprime = cv->h;
    cv->eta = 1.0;
    return;
  }
  cv->etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / cv->L) + ADDON);
  if (cv->qwait != 0) {
    cv->eta = cv->etaq;
    cv->qprime = cv->q;
    cv_set_eta(cv);
    return;
  }
  cv->qwait = 2;
  cv->etaqm1 = cv_etaqm1(cv);
  cv->etaqp1 = cv_etaqp1(cv);
  cv_choose_eta(cv);
  cv_set_eta(cv);
}

static int cv_step(cvmem* cv) {
  const double saved_t = cv->tn;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  double dsm = 0.0;
  if (cv->nst > 0 && cv->hprime != cv->h) cv_adjust_params(cv);
  for (;;) {
    cv_predict(cv);
    cv_set(cv);
    nflag = cv_nls(cv, nflag);
    const int kflag = cv_handle_nflag(cv, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN) continue;
    if (kflag != DO_ERROR_TEST) return kflag;
    const int eflag = cv_do_error_test(cv, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != CV_SUCCESS) return eflag;
    break;
  }
  cv_complete_step(cv);
  cv_prepare_next_step(cv, dsm);
  cv->etamax = (cv->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  for (int i = 0; i < cv->N; ++i) cv->acor[i] *= cv->tq[2];
  return CV_SUCCESS;
}

/* CVode(..., CV_ONE_STEP): one internal step, initial setup on the first call. */
static int cv_one_step(cvmem* cv, double tout, double* tret) {
  if (cv->nst == 0) {
    if (cv_ewt(cv, cv->zn[0]) != 0) return CV_ILL_INPUT;
    if (cv_f(cv, cv->tn, cv->zn[0], cv->zn[1]) > 0) return CV_REPTD_RHSFUNC_ERR;
    if (cv_hin(cv, tout) != 0) return CV_ILL_INPUT;
    const double rh = fabs(cv->h) * cv->hmax_inv;
    if (rh > 1.0) cv->h /= rh;
    cv->hscale = cv->h;
    cv->h0u = cv->h;
    cv->hprime = cv->h;
    for (int i = 0; i < cv->N; ++i) cv->zn[1][i] *= cv->h;
  } else {
    if (cv_ewt(cv, cv->zn[0]) != 0) return CV_ILL_INPUT;
  }
  const double nrm = wrms(cv->zn[0], cv->ewt, cv->N);
  cv->tolsf = DBL_EPSILON * nrm;
  if (cv->tolsf > 1.0) return CV_TOO_MUCH_ACC;
  cv->tolsf = 1.0;
  const int kflag = cv_step(cv);
  *tret = cv->tn;
  if (kflag != CV_SUCCESS) return kflag;
  cv->next_q = cv->qprime;
  cv->next_h = cv->hprime;
  return CV_SUCCESS;
}

/* CVodeGetDky(t, k=0) */
static void cv_dky0(cvmem* cv, double t, double* out) {
  const double s = (t - cv->tn) / cv->h;
  for (int i = 0; i < cv->N; ++i) out[i] = 0.0;
  for (int j = cv->q; j >= 0; --j) {
    double c = 1.0;
    for (int i = 0; i < j; ++i) c *= s;
    for (int i = 0; i < cv->N; ++i) out[i] += c * cv->zn[j][i];
  }
}

/* CVodeGetDky(t, 1): the sum over j = q..1 of j s^(j-1) zn[j], times 1/h */
static void cv_dky1(cvmem* cv, double t, double* out) {
  const double s = (t - cv->tn) / cv->h;
  for (int i = 0; i < cv->N; ++i) out[i] = 0.0;
  for (int j = cv->q; j >= 1; --j) {
    double c = (double)j;
    for (int i = 0; i < j - 1; ++i) c *= s;
    for (int i = 0; i < cv->N; ++i) out[i] += c * cv->zn[j][i];
  }
  const double hinv = 1.0 / cv->h;
  for (int i = 0; i < cv->N; ++i) out[i] *= hinv;
}

/* The plugin's step callback after a successful CVode call
 * (solver_cvode.cpp:247-257): the state CVode returned (zn[0] at tn, else the
 * interpolant at tc), CVodeGetDky(tc, 1), the time since the previous call;
 * returns the callback's value (non-zero ends the integration). */
typedef struct {
  zrko_step_cb fn;
  void* data;
  double tprev;
  double* y;
  double* ydot;
} stepcb;

static int step_callback(cvmem* cv, stepcb* cb, long nsteps, double tc) {
  if (tc == cv->tn)
    memcpy(cb->y, cv->zn[0], sizeof(double) * (size_t)cv->N);
  else
    cv_dky0(cv, tc, cb->y);
  cv_dky1(cv, tc, cb->ydot);
  const int r = cb->fn(0, (int)nsteps, tc, tc - cb->tprev, cb->y, cb->ydot, cb->data);
  cb->tprev = tc;
  return r;
}

/* CvodeSolver::Integrate (solver_cvode.cpp:51-299): one-step loop, retries,
 * Dky interpolation at min(tcurr, end_time). */
/* ------------------------------------------------------------------------- */
/* Root finding for the ignition time: SUNDIALS CVODE 5.8.0 cvRcheck1-3 and    */
/* cvRootfind for the one function g = T0 + delta_T - T                        */
/* (reactor_nvector_serial.cpp:774-784), both directions, in the one-step loop  */
/* of solver_cvode.cpp:228-266                                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
  double tlo, glo, Tign, tret, Tref;
  int iT; /* index of the temperature component */
} rootst;

#define ROOT_MAX_ITER 200
#define ROOT_MAX_RETURNS 64 /* root returns per reactor (the device's bound) */

/* temperature component of CVodeGetDky(t, 0), summed as cv_dky0 does */
static double dky0_T(const cvmem* cv, int iT, double t) {
  const double s = (t - cv->tn) / cv->h;
  double out = 0.0;
  for (int j = cv->q; j >= 0; --j) {
    double c = 1.0;
    for (int i = 0; i < j; ++i) c *= s;
    out += c * cv->zn[j][iT];
  }
  return out;
}
static double root_g(const rootst* rs, double yT) { return rs->Tign - yT * rs->Tref; }
static double root_ttol(const cvmem* cv) {
  return (fabs(cv->tn) + fabs(cv->h)) * DBL_EPSILON * 100.0;
}
/* cvRcheck1: the interval starts at the (re)initialisation time */
static void root_init(const cvmem* cv, rootst* rs) {
  rs->tlo = cv->tn;
  rs->glo = root_g(rs, cv->zn[0][rs->iT]);
  rs->tret = cv->tn;
}
/* cvRootfind: Illinois-weighted secant on g over (tlo, thi]; returns 1 when
 * the interval holds a root, which is then thi (ghi = g there) */
static int rootfind(const cvmem* cv, const rootst* rs, double* tlo, double* glo, double* thi,
                    double* ghi, double ttol) {
  const int zroot0 = *ghi == 0.0;
  if (zroot0 || !(*glo * *ghi < 0.0)) return zroot0;
  double alph = 1.0;
  int side = 0, sideprev = -1;
  for (int it = 0; it < ROOT_MAX_ITER; ++it) {
    alph = (sideprev == side) ? ((side == 2) ? alph * 2.0 : alph * 0.5) : 1.0;
    double tmid = *thi - (*thi - *tlo) * *ghi / (*ghi - alph * *glo);
    if (fabs(tmid - *tlo) < 0.5 * ttol) {
      const double fracint = fabs(*thi - *tlo) / ttol;
      const double fracsub = (fracint > 5.0) ? 0.1 : 0.5 / fracint;
      tmid = *tlo + fracsub * (*thi - *tlo);
    }
    if (fabs(*thi - tmid) < 0.5 * ttol) {
      const double fracint = fabs(*thi - *tlo) / ttol;
      const double fracsub = (fracint > 5.0) ? 0.1 : 0.5 / fracint;
      tmid = *thi - fracsub * (*thi - *tlo);
    }
    const double gmid = root_g(rs, dky0_T(cv, rs->iT, tmid));
    sideprev = side;
    if (gmid != 0.0 && *glo * gmid < 0.0) { /* sign change in (tlo, tmid) */
      *thi = tmid;
      *ghi = gmid;
      side = 1;
      if (fabs(*thi - *tlo) <= ttol) break;
      continue;
    }
    if (gmid == 0.0) { /* a zero at tmid */
      *thi = tmid;
      *ghi = gmid;
      break;
    }
    *tlo = tmid; /* the sign change is in (tmid, thi) */
    *glo = gmid;
    side = 2;
    if (fabs(*thi - *tlo) <= ttol) break;
  }
  return 1;
}
/* cvRcheck3 after a step (thi = tn, y = zn[0]) */
static int root_check3(const cvmem* cv, rootst* rs, double* troot) {
  double tlo = rs->tlo, glo = rs->glo, thi = cv->tn;
  double ghi = root_g(rs, cv->zn[0][rs->iT]);
  const int found = rootfind(cv, rs, &tlo, &glo, &thi, &ghi, root_ttol(cv));
  rs->tlo = thi;
  rs->glo = ghi;
  *troot = thi;
  return found;
}
/* cvRcheck2 at the start of the call after a root return */
static void root_check2(const cvmem* cv, rootst* rs) {
  const double ylo = dky0_T(cv, rs->iT, rs->tlo);
  double g = root_g(rs, ylo);
  if (g == 0.0) {
    const double ttol = root_ttol(cv);
    const double smallh = cv->h > 0.0 ? ttol : -ttol;
    const double tplus = rs->tlo + smallh;
    const double yp = ((tplus - cv->tn) * cv->h >= 0.0)
                          ? ylo + (smallh / cv->h) * cv->zn[1][rs->iT]
                          : dky0_T(cv, rs->iT, tplus);
    const double gp = root_g(rs, yp);
    if (gp != 0.0) g = gp;
  }
  rs->glo = g;
}

/* The analytic Jacobian's pattern (ReactorNVectorSerial::
 * SetupSparseJacobianArrays, reactor_nvector_serial.cpp:122-281): diagonal,
 * temperature row and column, every step's reactant columns against its
 * reactant and product rows, the non-integer terms; pat[col*N + row], N = nsp
 * (no temperature) or nsp + 1. */
static char* jacobian_pattern(const zrko_mech* m, int N) {
  const int nsp = m->nsp;
  char* pat = (char*)calloc((size_t)N * N, 1);
  for (int j = 0; j < N; ++j) pat[j * N + j] = 1;
  if (N == nsp + 1)
    for (int k = 0; k < N; ++k) pat[k * N + nsp] = pat[nsp * N + k] = 1;
  for (int s = 0; s < m->nstep; ++s)
    for (int k = m->reac_ptr[s]; k < m->reac_ptr[s + 1]; ++k) {
      const int col = m->reac_sp[k];
      for (int q = m->reac_ptr[s]; q < m->reac_ptr[s + 1]; ++q) pat[col * N + m->reac_sp[q]] = 1;
      for (int q = m->prod_ptr[s]; q < m->prod_ptr[s + 1]; ++q) pat[col * N + m->prod_sp[q]] = 1;
    }
  for (int q = 0; q < m->n_ni; ++q)
    for (int k = m->ni_rop_ptr[q]; k < m->ni_rop_ptr[q + 1]; ++k) {
      const int col = m->ni_rop_sp[k];
      for (int e = m->ni_re_ptr[q]; e < m->ni_re_ptr[q + 1]; ++e) pat[col * N + m->ni_re_sp[e]] = 1;
      for (int e = m->ni_pr_ptr[q]; e < m->ni_pr_ptr[q + 1]; ++e) pat[col * N + m->ni_pr_sp[e]] = 1;
    }
  return pat;
}

static int integrate_one(const zrko_mech* m, const zrko_opts* o, double* P, double dt,
                         double* y, zrko_stats* st, double dpdt, double esrc,
                         const double* ysrc, double T0, double* ign_time, stepcb* cb) {
  const int nsp = m->nsp;
  const int Nf = nsp + 1;                            /* storage of every vector */
  const int N = o->solve_temperature ? nsp + 1 : nsp; /* CVODE's state size */
  reactor_t r;
  reactor_init(&r, m, o, *P);
  /* ReactorConstantVolumeCPU::InitializeState (reactor_constant_volume_cpu.cpp:15-42) */
  if (o->constant_volume) r.v0 = inv_density_tpy(m, y[m->nsp] * o->Tref, *P, y);
  reactor_set_sources(&r, y, *P, dpdt, esrc, ysrc);
  if (!o->solve_temperature) r.fixT = y[nsp];
  cvmem cv;
  memset(&cv, 0, sizeof(cv));
  cv.r = &r;
  cv.o = o;
  cv.N = N;
  cv.nsp = nsp;
  cv.direct = o->dense == 1 && o->iterative == 0;
  double* w = (double*)calloc((size_t)(L_MAX + 8 + SPGMR_MAXL + 1 + 4) * Nf +
                                  2 * (size_t)Nf * Nf, sizeof(double));
  for (int j = 0; j < L_MAX; ++j) cv.zn[j] = w + (size_t)j * Nf;
  cv.ewt = w + (size_t)L_MAX * Nf;
  cv.acor = cv.ewt + Nf;
  cv.tempv = cv.acor + Nf;
  cv.ftemp = cv.tempv + Nf;
  cv.y = cv.ftemp + Nf;
  cv.atolv = cv.y + Nf;
  cv.delta = cv.atolv + Nf;
  for (int j = 0; j <= SPGMR_MAXL; ++j) cv.V[j] = cv.delta + (size_t)(2 + j) * Nf;
  cv.xcor = cv.V[SPGMR_MAXL] + Nf;
  cv.vtemp = cv.xcor + Nf;
  cv.work = cv.vtemp + Nf;
  cv.jv = cv.work + Nf;
  cv.savedJ = cv.jv + Nf;
  cv.M = cv.savedJ + (size_t)Nf * Nf;
  cv.piv = (int*)calloc((size_t)Nf, sizeof(int));
  char* pat = jacobian_pattern(m, N);
  cv.pat = pat;
  /* absolute tolerances (CvodeSolver::Integrate, solver_cvode.cpp:74-80) times
   * GetAbsoluteToleranceCorrection: with abstol_dens 1 / (rho0 / MW_j 1e-3)
   * for the species, rho0 the density of the initial state */
  for (int i = 0; i < N; ++i) cv.atolv[i] = o->atol;
  if (o->abstol_dens) {
    const double rho0 = 1.0 / inv_density_tpy(m, y[nsp] * o->Tref, *P, y);
    for (int j = 0; j < nsp; ++j) {
      const double molar_density = rho0 * (1.0 / m->mw[j]) * 1.0e-3;
      cv.atolv[j] = o->atol * (1.0 / molar_density);
    }
  }
  cv.rtol = o->rtol;
  cv.nlscoef = o->nlscoef;
  cv.hmax_inv = (o->max_dt > 0.0) ? 1.0 / o->max_dt : 0.0;
  cv_reinit(&cv, 0.0, y);
  /* ignition root function: the crossing of T0 + delta_T, T0 the input T
   * (none when the temperature is held) */
  const int ign = o->delta_temperature_ignition > 0.0 && o->solve_temperature;
  rootst rs;
  rs.Tign = T0 + o->delta_temperature_ignition;
  rs.Tref = o->Tref;
  rs.iT = m->nsp;
  if (ign) root_init(&cv, &rs);
  int irfnd = 0, nroots = 0;
  double troot_last = 0.0;

  long nsteps = 0, nst_total = 0, nfe_total = 0, nje_total = 0, nset_total = 0;
  long nni_total = 0, ncfn_total = 0, netf_total = 0;
  int num_tries = 0, flag = CV_SUCCESS;
  double tcurr = 0.0;
  /* one CVode call per iteration, CV_ONE_STEP (solver_cvode.cpp:228-266) */
  while (tcurr < dt) {
    if (ign && cv.nst > 0) {
      /* a later call: after a root return, roots in the rest of the last
       * step, else the step's end tn, which was not returned yet (CVode's
       * stop tests at the start of a call) */
      if (irfnd) root_check2(&cv, &rs);
      const double troundoff = 100.0 * DBL_EPSILON * (fabs(cv.tn) + fabs(cv.h));
      if (fabs(cv.tn - rs.tret) > troundoff) {
        nsteps += 1;
        double troot;
        if (nroots < ROOT_MAX_RETURNS && root_check3(&cv, &rs, &troot)) {
          ++nroots;
          irfnd = 1;
          tcurr = rs.tret = troot;
          if (cb && step_callback(&cv, cb, nsteps, tcurr) != 0) break;
          troot_last = troot;
          if (o->stop_after_ignition) break;
          continue;
        }
        irfnd = 0;
        tcurr = rs.tret = cv.tn;
        if (cb && step_callback(&cv, cb, nsteps, tcurr) != 0) break;
        continue;
      }
    }
    flag = cv_one_step(&cv, dt, &tcurr);
    nsteps += 1;
    if (cv.nst >= o->max_steps) flag = CV_TOO_MUCH_WORK;
    if (flag < 0) {
      num_tries += 1;
      for (int i = 0; i < N; ++i) cv.atolv[i] *= o->retry_atol_adj;
      cv.rtol *= o->retry_rtol_adj;
      nst_total += cv.nst; nfe_total += cv.nfe; nje_total += cv.nje; nset_total += cv.nsetups;
      nni_total += cv.nni; ncfn_total += cv.ncfn; netf_total += cv.netf;
      if (cb) cb->tprev = tcurr;
      if (num_tries == o->num_retries) break;
      double* y0 = cv.tempv; /* yout = zn[0] at tn */
      memcpy(y0, cv.zn[0], sizeof(double) * (size_t)N);
      double* ysave = (double*)malloc(sizeof(double) * (size_t)N);
      memcpy(ysave, y0, sizeof(double) * (size_t)N);
      cv_reinit(&cv, tcurr, ysave);
      free(ysave);
      if (ign) {
        root_init(&cv, &rs);
        irfnd = 0;
      }
    } else {
      int rootret = 0;
      if (ign) {
        rs.tret = cv.tn;
        double troot;
        if (nroots < ROOT_MAX_RETURNS && root_check3(&cv, &rs, &troot)) { /* before tn */
          ++nroots;
          irfnd = 1;
          tcurr = rs.tret = troot;
          rootret = 1;
        }
      }
      /* the callback comes before the root time is kept (solver_cvode.cpp:247-264) */
      if (cb && step_callback(&cv, cb, nsteps, tcurr) != 0) break;
      if (rootret) {
        troot_last = tcurr;
        if (o->stop_after_ignition) break;
      }
    }
  }
  if (flag == CV_SUCCESS) {
    if (ign && ign_time) *ign_time = troot_last;
    /* the state at dt, or at the ignition time when stopped there; GetState's
     * pressure uses the solve's dt (the manager passes dt_calc_) */
    cv_dky0(&cv, tcurr < dt ? tcurr : dt, y);
    const double tend = dt;
    /* GetState: constant volume reports P = Ru T / (v Wmix) + dpdt t
     * (reactor_constant_volume_cpu.cpp:44-63, mechanism.cpp:292-297);
     * constant pressure P + dpdt t, or with mass sources Ru T / (v Wmix) at
     * the initial v (reactor_constant_pressure_cpu.cpp:44-66) */
    const double Tend = y[m->nsp] * o->Tref;
    if (o->constant_volume)
      *P = pressure_tvy(m, Tend, r.v0, y) + dpdt * tend;
    else if (ysrc)
      *P = pressure_tvy(m, Tend, r.vinit, y);
    else
      *P = *P + dpdt * tend;
    nst_total += cv.nst; nfe_total += cv.nfe; nje_total += cv.nje; nset_total += cv.nsetups;
    nni_total += cv.nni; ncfn_total += cv.ncfn; netf_total += cv.netf;
  }
  /* a failed reactor keeps its input state (zerork_reactor_manager.cpp:1123-1139) */
  if (flag < 0) nsteps = (nsteps <= 0) ? -1 : -nsteps;
  if (st) {
    st->nsteps = (int)nsteps;
    st->nst = (int)nst_total;
    st->nfe = (int)nfe_total;
    st->nje = (int)nje_total;
    st->nsetups = (int)nset_total;
    st->nni = (int)nni_total;
    st->ncfn = (int)ncfn_total;
    st->netf = (int)netf_total;
    st->status = flag;
  }
  free(cv.piv);
  free(pat);
  free(w);
  reactor_free(&r);
  return flag;
}

int zrko_integrate_batch(const zrko_mech* m, const zrko_opts* o, int n, double dt, double* T,
                         double* P, double* mf, int stride, zrko_stats* stats,
                         int nthreads) {
  return zrko_integrate_batch_aux(m, o, n, dt, T, P, mf, stride, stats, nthreads, NULL, NULL,
                                  NULL, NULL);
}

int zrko_integrate_batch_aux(const zrko_mech* m, const zrko_opts* o, int n, double dt,
                             double* T, double* P, double* mf, int stride, zrko_stats* stats,
                             int nthreads, const double* dpdt, const double* esrc,
                             const double* ysrc, double* ign_time) {
  const int nsp = m->nsp;
  int nfail = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nfail)
  for (int k = 0; k < n; ++k) {
    double* y = (double*)malloc(sizeof(double) * (size_t)(nsp + 1));
    for (int j = 0; j < nsp; ++j) y[j] = mf[(size_t)k * stride + j];
    y[nsp] = T[k] / o->Tref;
    int flag = integrate_one(m, o, &P[k], dt, y, stats ? stats + k : NULL,
                             dpdt ? dpdt[k] : 0.0, esrc ? esrc[k] : 0.0,
                             ysrc ? ysrc + (size_t)k * stride : NULL, T[k],
                             ign_time ? ign_time + k : NULL, NULL);
    for (int j = 0; j < nsp; ++j) mf[(size_t)k * stride + j] = y[j];
    T[k] = y[nsp] * o->Tref;
    if (flag < 0) nfail++;
    free(y);
  }
  return nfail;
}

int zrko_integrate_one_cb(const zrko_mech* m, const zrko_opts* o, double dt, double* T,
                          double* P, double* mf, zrko_stats* stats, zrko_step_cb fn,
                          void* data) {
  const int nsp = m->nsp;
  double* y = (double*)malloc(sizeof(double) * (size_t)(3 * (nsp + 1)));
  for (int j = 0; j < nsp; ++j) y[j] = mf[j];
  y[nsp] = *T / o->Tref;
  stepcb cb;
  cb.fn = fn;
  cb.data = data;
  cb.tprev = 0.0;
  cb.y = y + (nsp + 1);
  cb.ydot = y + 2 * (nsp + 1);
  const int flag = integrate_one(m, o, P, dt, y, stats, 0.0, 0.0, NULL, *T, NULL, fn ? &cb : NULL);
  for (int j = 0; j < nsp; ++j) mf[j] = y[j];
  *T = y[nsp] * o->Tref;
  free(y);
  return flag;
}
