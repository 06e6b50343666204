/* CPU FP64 oracle: a plain-C restatement of the reference cfd_plugin path.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/ (as the checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * (zero-rk_amd/) never links, loads or calls it.
 *
 * Each function cites the reference code it restates:
 *   zrko_thermo            src/zerork/nasa_poly.cpp:50-172
 *   zrko_rate_constants    src/zerork/rate_const.cpp:539-582 (updateK), 669-731,
 *                          862-1012, 1188-1200; plog_reaction.cpp:163-320;
 *                          non_integer_reaction_network.cpp:335-358 (Keq)
 *   zrko_rates             src/zerork/perf_net.cpp:190-275
 *                          (calcRatesFromTC_StepLimiter),
 *                          non_integer_reaction_network.cpp:270-333
 *   zrko_rhs               applications/cfd_plugin/reactor_constant_pressure_cpu.cpp:68-133
 *   zrko_jacobian          applications/cfd_plugin/reactor_nvector_serial.cpp:610-771,
 *                          non_integer_reaction_network.cpp:493-506, 774-905
 *   zrko_integrate         applications/cfd_plugin/solver_cvode.cpp:51-299 driving a
 *                          restatement of SUNDIALS CVODE v5.8.0 BDF (dense direct
 *                          Newton, the plugin's dense=1/iterative=0 mode), pinned
 *                          by cmake/sundials5.CMakeLists.txt:8-9 in the reference.
 */
#ifndef ZRK_ORACLE_H
#define ZRK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nsp, nrxn, nstep;
  const double* mw;        /* [nsp] kg/kmol */
  const double* thermo;    /* [nsp*16] Tmid, low a1..a7, high a1..a7 */
  const int* reac_ptr;     /* [nstep+1] */
  const int* reac_sp;
  const int* prod_ptr;     /* [nstep+1] */
  const int* prod_sp;
  int n_arr;
  const int* arr_step;
  const double* arr_lnA;
  const double* arr_beta;
  const double* arr_Ta;
  int n_keq;
  const int* keq_step;
  const int* keq_fwd;
  const double* keq_ndelta;
  int n_tb;
  const int* tb_fwd;
  const int* tb_rev;
  const int* tb_ptr;
  const int* tb_sp;
  const double* tb_eff;    /* efficiency - 1 */
  int n_fo;
  const int* fo_fwd;
  const int* fo_rev;
  const int* fo_type;      /* 0 Lindemann, 1 Troe-3, 2 Troe-4, 3 SRI */
  const int* fo_spc;       /* -1: mixture M */
  const double* fo_param;  /* [n_fo*8] lnAlow, beta, Ta, then alpha, T***, T*, T** (Troe)
                              or a, b, 1/c (-1 if c <= 0), d, e (SRI) */
  const int* fo_ptr;
  const int* fo_sp;
  const double* fo_eff;
  /* non-integer steps (no expanded participants), in step order */
  int n_ni;
  const int* step_ni;      /* [nstep] index into the ni lists or -1 */
  const int* ni_step;
  const int* ni_rop_ptr;   /* CSR of rate-of-progress exponents */
  const int* ni_rop_sp;
  const double* ni_rop_pow;
  const int* ni_re_ptr;    /* CSR of destroyed species and coefficients */
  const int* ni_re_sp;
  const double* ni_re_nu;
  const int* ni_pr_ptr;    /* CSR of created species and coefficients */
  const int* ni_pr_sp;
  const double* ni_pr_nu;
  /* PLOG forward steps */
  int n_plog;
  const int* plog_step;
  const int* plog_ptr;     /* CSR over distinct pressures */
  const double* plog_p;    /* Pa */
  const double* plog_lnp;
  const int* plog_lptr;    /* CSR (over all pressures) of Arrhenius lines */
  const double* plog_lnA;
  const double* plog_sign;
  const double* plog_beta;
  const double* plog_Ta;
  /* gas constant J/kmol/K; 0: src/zerork/constants.h:10's 8.314462618e3 (the
   * reference's earlier releases used CODATA 2014's 8.3144598e3) */
  double ru;
} zrko_mech;

typedef struct {
  double rtol, atol;
  double max_dt;
  double nlscoef;          /* nonlinear_convergence_coeff */
  int max_steps;
  int num_retries;         /* cvode_num_retries */
  double retry_atol_adj;   /* cvode_retry_absolute_tolerance_adjustment */
  double retry_rtol_adj;   /* cvode_retry_relative_tolerance_adjustment */
  double min_mass_fraction;
  double step_limiter;
  double Tref;             /* reference_temperature */
  int constant_volume;     /* 0: constant pressure, 1: constant volume */
  double delta_temperature_ignition; /* > 0: locate T crossing T0 + this */
  int stop_after_ignition; /* 1: end the integration at the crossing */
  /* 1: absolute tolerances scaled by the initial molar density of each
   * species, atol MW_j 1e3 / rho0 (reactor_nvector_serial.cpp:107-120) */
  int abstol_dens;
  /* 0: the temperature is held at its initial value and is not a state
   * variable (SetSolveTemperature(false), reactor_nvector_serial.cpp:69-101) */
  int solve_temperature;
  /* linear solver (solver_cvode.cpp:93-214, reactor_nvector_serial.cpp:
   * 363-507): dense = 1 with iterative = 0 is CVODE's dense direct solver with
   * the analytic Jacobian (analytic = 1) or CVODE's difference-quotient one;
   * otherwise SPGMR (maxl 5, left preconditioning, modified Gram-Schmidt,
   * tolerance factor eps_lin) preconditioned by the LU of I - gamma J, with
   * off-diagonal entries |.| <= preconditioner_threshold dropped when
   * iterative = 1, J analytic or the reactor's own divided differences */
  int dense, analytic, iterative;
  double eps_lin, preconditioner_threshold;
} zrko_opts;

typedef struct {
  int nsteps;      /* CVode one-step calls (the plugin's cost output), <0 on failure */
  int nst, nfe, nje, nsetups, nni, ncfn, netf;
  int status;      /* 0 ok, <0 CVODE-style failure flag */
} zrko_stats;

void zrko_default_opts(zrko_opts* o);

void zrko_thermo(const zrko_mech* m, double T, double* cp_R, double* h_RT, double* g_RT);
/* K per step after Arrhenius, Keq, third-body and falloff (no limiter). */
void zrko_rate_constants(const zrko_mech* m, double T, const double* C, double* K);
void zrko_rates(const zrko_mech* m, double T, const double* C, double step_limiter,
                double* rop, double* create, double* destroy, double* net);

/* Reactor state y = [Y_0..Y_{nsp-1}, T/Tref] (solve_temperature).  Constant
 * pressure, or with o->constant_volume the constant-volume reactor whose
 * volume is fixed by (P, y) itself. */
int zrko_rhs(const zrko_mech* m, const zrko_opts* o, double P, double t, const double* y,
             double* ydot);
/* Reference analytic Jacobian (dense, column-major J[col*N+row]) at y with
 * fy = f(y); reproduces the inverse-density carry-over of the reference. */
void zrko_jacobian(const zrko_mech* m, const zrko_opts* o, double P, double t,
                   const double* y, const double* fy, double* J);

/* Integrate n reactors over [0, dt] (AoS: mf[k*stride+j]), updating T and mf
 * in place (and P for constant volume); nthreads <= 0 uses OpenMP's default. */
int zrko_integrate_batch(const zrko_mech* m, const zrko_opts* o, int n, double dt,
                         double* T, double* P, double* mf, int stride,
                         zrko_stats* stats, int nthreads);
/* The same with the plugin's aux fields (each array optional): pressure rate
 * dpdt[n] (Pa/s), energy source esrc[n] (J/kg/s), mass-fraction sources
 * ysrc[n*stride] (1/s), ignition-time output ign_time[n] (with
 * o->delta_temperature_ignition > 0: the last crossing time, 0 if none); P is
 * updated to the end pressure. */
int zrko_integrate_batch_aux(const zrko_mech* m, const zrko_opts* o, int n, double dt,
                             double* T, double* P, double* mf, int stride,
                             zrko_stats* stats, int nthreads, const double* dpdt,
                             const double* esrc, const double* ysrc, double* ign_time);
/* The plugin's step callback (zerork_cfd_plugin.h:44): reactor id, CVode
 * calls so far, t, t - t_prev, state, its time derivative; non-zero ends the
 * integration. */
typedef int (*zrko_step_cb)(int id, int nsteps, double t, double dt, const double* y,
                            const double* ydot, void* data);
/* One reactor (T, P, mf[nsp] updated in place) with fn called after every
 * successful CVode call as CvodeSolver::Integrate calls it
 * (solver_cvode.cpp:228-272); a non-zero return ends the integration there
 * with the state CVodeGetDky(min(t, dt), 0). */
int zrko_integrate_one_cb(const zrko_mech* m, const zrko_opts* o, double dt, double* T,
                          double* P, double* mf, zrko_stats* stats, zrko_step_cb fn,
                          void* data);
/* RHS at time t of the reactor initialised at (P0, y0) with sources. */
int zrko_rhs_src(const zrko_mech* m, const zrko_opts* o, double P0, const double* y0,
                 double t, const double* y, double dpdt, double esrc, const double* ysrc,
                 double* ydot);

/* Dense LU with partial pivoting (SUNDIALS denseGETRF/denseGETRS semantics),
 * column-major; exposed for tests. Returns 0 or k+1 at a zero pivot. */
int zrko_dense_getrf(int n, double* a, int* piv);
void zrko_dense_getrs(int n, const double* a, const int* piv, double* b);

#ifdef __cplusplus
}
#endif
#endif
