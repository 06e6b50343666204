/* zrk_device.h - device-level C ABI of the batched reactor-chemistry path.
 *
 * The plugin API (zerork_cfd_plugin.h) takes host arrays; these entry points
 * take device pointers that are already resident in HBM and a HIP stream, so
 * a caller (or a test/benchmark) can run each stage of the path on its own.
 * All arrays are FP64; T [K], P [Pa], Y mass fractions reactor-major
 * Y[r*stride + j].  Return value: 0 on success, a negative code on invalid
 * arguments, a positive hipError_t value on a HIP failure.
 *
 * Entry point            replaces (reference file:line)
 *   zrk_mech_load        zerork::mechanism ctor (src/zerork/mechanism.cpp:17-66)
 *                        + mechanism_cuda device mirror (mechanism_cuda.cu:14-144)
 *   zrk_mech_info,       mechanism::getNumSpecies/Reactions/Steps, getMolWtSpc,
 *   zrk_mech_species,    getSpeciesName, step participant maps, reaction kinds
 *   zrk_mech_molwt,      (src/zerork/mechanism.h:33-56, 205-220)
 *   zrk_mech_steps,
 *   zrk_mech_reactions
 *   zrk_rates_device     mechanism_cuda::getReactionRatesLimiter_CUDA_mr_dev
 *                        (src/zerork/mechanism_cuda.cu:158-165) together with
 *                        ReactorConstantPressureGPU::GetTimeDerivative
 *                        (applications/cfd_plugin/reactor_constant_pressure_gpu.cu:181-244)
 *   zrk_jacobian_device  ReactorNVectorSerialCuda analytic Jacobian
 *                        (reactor_nvector_serial_cuda.cu:1455-1612; CPU twin
 *                        reactor_nvector_serial.cpp:610-771)
 *   zrk_integrate_device CvodeSolver::Integrate over a reactor batch
 *                        (applications/cfd_plugin/solver_cvode.cpp:51-299,
 *                        zerork_reactor_manager.cpp:909-1053), one BDF per reactor
 */
#ifndef ZRK_DEVICE_H
#define ZRK_DEVICE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ZRK_EXPORT __attribute__((visibility("default")))

typedef struct zrk_mech_s* zrk_mech;

/* Integrator options; names follow the plugin options of the same meaning
 * (zerork_reactor_manager.cpp:50-101). */
typedef struct {
  double rel_tol;
  double abs_tol;
  double max_dt;
  double nonlinear_convergence_coeff;
  double step_limiter;
  double reference_temperature;
  double min_mass_fraction;
  double cvode_retry_absolute_tolerance_adjustment;
  double cvode_retry_relative_tolerance_adjustment;
  int max_steps;
  int cvode_num_retries;
  /* 0: constant-pressure reactor (reactor_constant_pressure_cpu.cpp);
   * 1: constant-volume reactor (reactor_constant_volume_cpu.cpp), the
   * plugin's constant_volume option */
  int constant_volume;
  /* 0: Newton matrix, its inverse and the saved Jacobian in FP64 (the
   * reference's arithmetic); 1: in FP32, with rates, residual, error test and
   * state update still in FP64.  No reference counterpart (the reference
   * preconditions in FP64 only). */
  int fp32_preconditioner;
  /* > 0: report the time at which T crosses T0 + delta_temperature_ignition
   * (the plugin option of that name, zerork_reactor_manager.cpp:87; root
   * function reactor_nvector_serial.cpp:774-784, located by CVODE's root
   * finding as solver_cvode.cpp:90-92, 259-264 do); 0: off */
  double delta_temperature_ignition;
  /* 1: end a reactor's integration at its ignition time (the plugin option
   * stop_after_ignition, solver_cvode.cpp:262-264) */
  int stop_after_ignition;
  /* 1: absolute tolerance of species j scaled by 1 / (rho0 / MW_j 1e-3), rho0
   * the initial density (the plugin option abstol_dens,
   * reactor_nvector_serial.cpp:107-120) */
  int abstol_dens;
  /* The plugin's linear-solver options (solver_cvode.cpp:93-214,
   * reactor_nvector_serial.cpp:363-507).  dense = 1 with iterative = 0: the
   * direct solve with M = I - gamma J, J analytic (analytic = 1) or CVODE's
   * difference quotients.  Otherwise SPGMR (5 Krylov vectors, left
   * preconditioning, modified Gram-Schmidt, tolerance eps_lin * the Newton
   * tolerance) with Jacobian-vector products by difference quotients,
   * preconditioned by I - gamma J (J analytic, or the reactor's own
   * divided differences with analytic = 0) keeping the diagonal and, with
   * iterative = 1, only the off-diagonal entries of the Jacobian's pattern
   * whose magnitude exceeds preconditioner_threshold.  Device defaults:
   * dense 1, analytic 1, iterative 0 (the plugin's defaults are the
   * reference's: dense 0, analytic 1, iterative 1). */
  int dense;
  int analytic;
  int iterative;
  double eps_lin;
  double preconditioner_threshold;
} zrk_bdf_options;

ZRK_EXPORT void zrk_bdf_default_options(zrk_bdf_options* o);

ZRK_EXPORT int zrk_mech_load(const char* mech_file, const char* therm_file, zrk_mech* out,
                             char* err, int errlen);
ZRK_EXPORT void zrk_mech_free(zrk_mech m);
ZRK_EXPORT int zrk_mech_info(zrk_mech m, int* nsp, int* nrxn, int* nstep);
ZRK_EXPORT int zrk_mech_species(zrk_mech m, int j, char* name, int len);
ZRK_EXPORT int zrk_mech_molwt(zrk_mech m, double* mw);
/* step_rxn/step_dir [nstep]; reac_ptr/prod_ptr [nstep+1]; reac_sp/prod_sp sized
 * by the last pointer entries (pass NULL to query sizes via the ptr arrays). */
ZRK_EXPORT int zrk_mech_steps(zrk_mech m, int* step_rxn, int* step_dir, int* reac_ptr,
                              int* reac_sp, int* prod_ptr, int* prod_sp);

/* Per-reaction structure [nrxn]: forward/reverse step index (-1 if none),
 * reverse type (0 none, 1 explicit REV Arrhenius, 2 from Keq), number of
 * enhanced third-body species (-1 if not a third-body/fall-off reaction),
 * fall-off type (-1 none, 0 Lindemann, 1 Troe-3, 2 Troe-4).  NULL skips. */
ZRK_EXPORT int zrk_mech_reactions(zrk_mech m, int* fwd_step, int* rev_step, int* rev_type,
                                  int* n_enhanced, int* falloff_type);

/* Rates of progress (limiter step_limiter, as perf_net.cpp:190-275), species
 * creation/destruction/net rates and the constant-pressure derivative
 * ydot[r*(nsp+1) + k] ([dY_0/dt .. dY_{nsp-1}/dt, dT/dt]).  Any output may be
 * NULL.  Any mechanism size: one wavefront per reactor while a reactor's rate
 * arrays fit in its share of LDS, else one workgroup per reactor with the
 * rates of progress in device scratch of the handle. */
ZRK_EXPORT int zrk_rates_device(zrk_mech m, int n, const double* T, const double* P,
                                const double* Y, int stride, double step_limiter, double* ydot,
                                double* rop, double* create, double* destroy, double* net,
                                void* stream);

/* Reference analytic Jacobian J[r][row][col] (row-major, N = nsp+1, temperature
 * last) at the states, in FP64.  nsp <= 191 (above 63 species computed by
 * teams of 2-3 wavefronts, as in the integrator). */
ZRK_EXPORT int zrk_jacobian_device(zrk_mech m, const zrk_bdf_options* o, int n, const double* T,
                                   const double* P, const double* Y, int stride, double* J,
                                   void* stream);

/* Integrate n reactors over [0, dt]; T and Y updated in place, and P at
 * constant volume (P = Ru T / (v Wmix) of the end state).  nsteps[r] (optional) = CVode one-step calls, negative on failure
 * (the plugin's COST output); stats[r*8] (optional) = nst nfe nje nsetups nni
 * ncfn netf status.  Up to 63 species one wavefront per reactor, up to 191
 * a workgroup of 2-3 wavefronts (the FP32 Newton matrix in LDS, the FP64 one
 * in device memory), above that the sparse path (fp32_preconditioner = 1
 * only, -3 otherwise); -2 if the mechanism does not fit.  Returns 0 or an error
 * code; the number of failed reactors is available from nsteps. */
ZRK_EXPORT int zrk_integrate_device(zrk_mech m, const zrk_bdf_options* o, int n, double dt,
                                    double* T, double* P, double* Y, int stride,
                                    int* nsteps, int* stats, void* stream);

/* zrk_integrate_device with the plugin's aux fields (each device array
 * optional, null = none): source terms dpdt[n] (Pa/s), esrc[n] (J/kg/s),
 * ysrc[n*stride] (mass-fraction sources, 1/s, same layout as Y), and the
 * ignition-time output ignition_time[n].  Replaces the source-term arguments
 * of ReactorConstantPressureCPU / ReactorConstantVolumeCPU::InitializeState
 * and GetTimeDerivative (reactor_constant_pressure_cpu.cpp:15-133,
 * reactor_constant_volume_cpu.cpp:15-130), which the manager feeds from the
 * DPDT, E_SRC and Y_SRC aux fields, and the IGNITION_TIME field the manager
 * fills from the reactor's root time (zerork_reactor_manager.cpp:529-559,
 * 909-1053, 1131).  P is updated to the end pressure (constant pressure:
 * P + dpdt dt, or Ru T / (v Wmix) at the initial v with mass sources;
 * constant volume: Ru T / (v Wmix) + dpdt dt).  With
 * o->delta_temperature_ignition > 0, ignition_time[r] receives the last time
 * in [0, dt] (or just beyond dt, inside the last step) at which T crossed
 * T0 + delta_temperature_ignition, or 0 if it did not, and each crossing
 * counts one more CVode call in nsteps (the reference's one-step loop); with
 * o->stop_after_ignition = 1 the state returned is the state at the crossing. */
ZRK_EXPORT int zrk_integrate_device_aux(zrk_mech m, const zrk_bdf_options* o, int n, double dt,
                                        double* T, double* P, double* Y, int stride,
                                        const double* dpdt, const double* esrc,
                                        const double* ysrc, double* ignition_time, int* nsteps,
                                        int* stats, void* stream);

/* zrk_integrate_device without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments, kernels and results as zrk_integrate_device.  The integrator's
 * scratch (work counter and saved-Jacobian slots) is a 64 MiB device array
 * of the library, one per GPU, held until the process ends; the persistent
 * grid is capped at the slots that fit in it.  Every call on a GPU shares
 * that array, so calls that overlap on one GPU (other streams, other
 * threads) race: serialise them.  Dense paths only (up to 191 species): -4
 * for a mechanism on the sparse path; -6 if the handle's device image does
 * not exist yet (it is built by the first call of any other device entry
 * point, e.g. zrk_integrator_resident_waves); -7 if one reactor's slot does
 * not fit in the array. */
ZRK_EXPORT int zrk_integrate_device_noalloc(zrk_mech m, const zrk_bdf_options* o, int n,
                                            double dt, double* T, double* P, double* Y,
                                            int stride, int* nsteps, int* stats, void* stream);

/* Autotuner (no counterpart in the reference): times the integrator on the
 * given device batch (copies of T, P, Y; inputs unchanged) with each team
 * size that fits the mechanism and these options (1 wavefront per reactor,
 * or 2 / 3 with fp32_preconditioner) and keeps the fastest for later calls
 * with the same precision on this handle.  *choice = wavefronts per reactor
 * chosen; ms[1..3] (optional, 4 floats) = measured times, -1 if not tried. */
ZRK_EXPORT int zrk_autotune_integrator(zrk_mech m, const zrk_bdf_options* o, int n, double dt,
                                       const double* T, const double* P, const double* Y,
                                       int stride, int* choice, float* ms, void* stream);

/* Sets the team size (wavefronts per reactor) of later integrations with the
 * given precision on this handle, as the autotuner does; 0 restores the
 * smallest size that fits.  -1 if the size cannot integrate this mechanism. */
ZRK_EXPORT int zrk_integrator_set_team_width(zrk_mech m, int fp32_preconditioner, int W);

/* Sparse path: one reactor per workgroup with a static sparse FP32 LU of the
 * Newton matrix (the reference's sparse direct solve of I - gamma J,
 * reactor_nvector_serial.cpp:414-507 with superlu_manager.cpp:13-192).  Taken
 * by zrk_integrate_device / zrk_integrate_device_aux for mechanisms above 191
 * species, or for any mechanism after zrk_integrator_set_sparse(m, 1) (0
 * restores the dense paths).  Needs fp32_preconditioner = 1 (-3 otherwise);
 * source terms, the ignition root and PLOG / non-integer / SRI reactions
 * return -4.  *nnz, *levels (optional): LU entries and forward level count
 * of the plan (built on the first call). */
ZRK_EXPORT int zrk_integrator_set_sparse(zrk_mech m, int on);
ZRK_EXPORT int zrk_sparse_plan_info(zrk_mech m, int* nnz, int* levels, long long* factor_flops);

/* Number of wavefronts the persistent integrator keeps resident with these
 * options (NULL: defaults), for roofline bookkeeping. */
ZRK_EXPORT int zrk_integrator_resident_waves(zrk_mech m, const zrk_bdf_options* o);

#ifdef __cplusplus
}
#endif

#endif
