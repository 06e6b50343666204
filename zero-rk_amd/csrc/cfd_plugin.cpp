// The reference plugin API (include/zerork_cfd_plugin.h) over the HIP path.
//
// Option names, defaults and status codes follow
// applications/cfd_plugin/zerork_reactor_manager.cpp:30-191 and
// zerork_cfd_plugin.cpp:21-216.  A solve copies T/P/mf to HBM once, runs the
// persistent BDF kernel over the whole batch (reactors handed out in
// descending previous-cost order when sort_reactors is set, the reference's
// cost sort of :522-524), copies back, and fills the aux outputs
// (COST = CVode one-step calls, GPU = seconds per reactor of the batched
// solve, TEMPERATURE_DELTA = T change above solve_temperature_threshold).
// The linear solver follows the dense / analytic / iterative options (direct
// or SPGMR with the thresholded preconditioner, analytic or difference-
// quotient Jacobians), abstol_dens scales the absolute tolerances, and with
// always_solve_temperature = 0 reactors whose temperature changed by less
// than the threshold in their previous solve are integrated at fixed
// temperature (:1109-1113).
// The DPDT, E_SRC and Y_SRC aux fields are passed to the integrator as
// per-reactor source terms (P is then returned as the end pressure); with
// delta_temperature_ignition > 0 the IGNITION_TIME field receives each
// reactor's time of crossing T0 + delta_temperature_ignition (0 if none), and
// stop_after_ignition ends a reactor's integration there.
// Reactors whose batched integration fails are integrated again on their own
// with the FP64 Newton matrix when the batch used the FP32 one (the manager's
// CPU re-solve of a failed GPU batch, :1021-1034, 1090-1147); reactors that
// still fail keep their input state and report a negative cost, and the call
// returns ZERORK_STATUS_FAILED_SOLVE (:1123-1139).  Each solve appends a line
// to the reactor_timing_log when output_performance_log is set.  A step
// callback is called, as the manager sets it on the CPU solver, for a
// one-reactor solve with load_balance = 0 (:1082-1084): the integrator
// records every successful CVode call of that reactor on the device, and the
// callback is called for each record in order after the solve; a non-zero
// return ends the integration at that record, whose state is then the
// result (solver_cvode.cpp:247-257).  No C++ exception or exit() crosses the
// ABI.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/zerork_cfd_plugin.h"
#include "zrk_internal.h"

struct zerork_handle_impl {
  std::map<std::string, int> iopt;
  std::map<std::string, double> dopt;
  std::map<std::string, std::string> sopt;
  std::unique_ptr<zrk::MechHandle> mech;
  bool tried_init = false;
  double* dpdt = nullptr;
  double* cost = nullptr;
  double* gpu = nullptr;
  double* y_src = nullptr;
  double* e_src = nullptr;
  double* ign = nullptr;
  double* tdelta = nullptr;
  zerork_callback_fn cb = nullptr;
  void* cb_data = nullptr;
  int* ids = nullptr;
  hipStream_t stream = nullptr;
  // device buffers
  double* dT = nullptr;
  double* dP = nullptr;
  double* dY = nullptr;
  int* dsteps = nullptr;
  int* dorder = nullptr;
  double* dsrc_n = nullptr;  // dpdt, e_src, ignition times (3n)
  double* dysrc = nullptr;   // y_src (n*stride)
  double* dtrace = nullptr;  // step records of a one-reactor solve with a callback
  int* dtrace_n = nullptr;
  size_t cap_trace = 0;      // records
  size_t cap_n = 0, cap_y = 0, cap_src_n = 0, cap_src_y = 0;
  int n_calls = 0;           // solves so far (names the reactor dump files)
  // the manager's own cost and temperature-change arrays when the caller
  // supplies none (zerork_reactor_manager.cpp:503-518: kept across calls,
  // new entries cost 1 and change 0)
  std::vector<double> rc_own, tdelta_own;
  double* dtdelta = nullptr;  // device copy of the temperature changes
  size_t cap_td = 0;
  std::ofstream timing_log;   // reactor_timing_log (zerork_reactor_manager.cpp:614-636)
  bool timing_open = false;

  zerork_handle_impl() {
    sopt["mech_filename"] = "mech.dat";
    sopt["therm_filename"] = "therm.dat";
    iopt["verbosity"] = 4;
    iopt["sort_reactors"] = 1;
    iopt["load_balance"] = 0;
    iopt["load_balance_noise"] = 0;
    iopt["load_balance_mem"] = 1;
    iopt["reactor_weight_mult"] = 1;
    iopt["dump_reactors"] = 0;
    iopt["dump_failed_reactors"] = 0;
    iopt["output_performance_log"] = 1;
    iopt["max_steps"] = 5000;
    iopt["dense"] = 0;
    iopt["analytic"] = 1;
    iopt["iterative"] = 1;
    iopt["integrator"] = 0;
    iopt["abstol_dens"] = 0;
    iopt["cvode_num_retries"] = 5;
    dopt["rel_tol"] = 1.0e-8;
    dopt["abs_tol"] = 1.0e-20;
    dopt["eps_lin"] = 1.0e-3;
    dopt["nonlinear_convergence_coeff"] = 0.05;
    dopt["preconditioner_threshold"] = 1.0e-3;
    dopt["max_dt"] = 0.05;
    dopt["cvode_retry_absolute_tolerance_adjustment"] = 0.1;
    dopt["cvode_retry_relative_tolerance_adjustment"] = 1.0;
    iopt["constant_volume"] = 1;
    dopt["reference_temperature"] = 1.0;
    dopt["delta_temperature_ignition"] = 0.0;
    iopt["stop_after_ignition"] = 0;
    dopt["min_mass_fraction"] = 1.0e-30;
    iopt["always_solve_temperature"] = 1;
    dopt["solve_temperature_threshold"] = 2.0;
    dopt["step_limiter"] = 1.0e22;
    iopt["gpu"] = 1;
    // not a reference option: FP32 Newton matrix/inverse (FP64 everywhere else)
    iopt["fp32_preconditioner"] = 0;
    iopt["initial_gpu_multiplier"] = 8;
    iopt["n_reactors_min"] = 128;
    iopt["n_reactors_max"] = 1024;
    sopt["reactor_timing_log_filename"] = "/dev/null";
    sopt["mechanism_parsing_log_filename"] = "/dev/null";
  }
  ~zerork_handle_impl() {
    if (dT) (void)hipFree(dT);
    if (dP) (void)hipFree(dP);
    if (dY) (void)hipFree(dY);
    if (dsteps) (void)hipFree(dsteps);
    if (dorder) (void)hipFree(dorder);
    if (dsrc_n) (void)hipFree(dsrc_n);
    if (dysrc) (void)hipFree(dysrc);
    if (dtdelta) (void)hipFree(dtdelta);
    if (dtrace) (void)hipFree(dtrace);
    if (dtrace_n) (void)hipFree(dtrace_n);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// the universal gas constant, J/kmol/K (src/zerork/constants.h:10)
constexpr double kRuJ = 8.314462618e3;

zerork_status_t load_mechanism(zerork_handle h) {
  auto m = std::make_unique<zrk::MechHandle>();
  std::string err;
  const std::string& mf = h->sopt["mech_filename"];
  const std::string& tf = h->sopt["therm_filename"];
  {
    std::ifstream a(mf), b(tf);
    if (!a || (!tf.empty() && !b)) return ZERORK_STATUS_INVALID_FILENAME;
  }
  if (!zrk::load_mechanism(mf, tf, &m->host, &err)) {
    const std::string& logf = h->sopt["mechanism_parsing_log_filename"];
    std::ofstream log(logf);
    if (log) log << "mechanism parse failed: " << err << "\n";
    return ZERORK_STATUS_FAILED_MECHANISM_PARSE;
  }
  h->mech = std::move(m);
  return ZERORK_STATUS_SUCCESS;
}

// FinishInit (zerork_reactor_manager.cpp:573-651): lazy mechanism load, once.
zerork_status_t finish_init(zerork_handle h) {
  if (!h->tried_init) {
    h->tried_init = true;
    if (!h->mech) {
      zerork_status_t f = load_mechanism(h);
      if (f != ZERORK_STATUS_SUCCESS) return f;
    }
    if (h->iopt["output_performance_log"] != 0) {
      // the timing log's header (zerork_reactor_manager.cpp:614-636)
      h->timing_log.open(h->sopt["reactor_timing_log_filename"]);
      h->timing_open = h->timing_log.good();
      if (h->timing_open) {
        static const char* cols[] = {
            "reactors_solved", "n_cpu", "n_cpu_notemp", "n_gpu", "n_gpu_notemp",
            "n_gpu_groups", "n_steps_avg", "n_steps_avg_cpu", "n_steps_avg_gpu",
            "max_time_cpu", "max_time_gpu", "step_time_cpu", "step_time_gpu",
            "avg_time_total", "max_time_total"};
        h->timing_log << "#" << std::setw(11) << "solve_number";
        for (const char* c : cols) h->timing_log << std::setw(17) << c;
        h->timing_log << std::endl;
      }
    }
  } else if (!h->mech) {
    return ZERORK_STATUS_FAILED_MECHANISM_PARSE;
  }
  return ZERORK_STATUS_SUCCESS;
}

template <class T>
bool grow(T** p, size_t n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)) == hipSuccess;
}

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n\"'");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n\"'");
  return s.substr(a, b - a + 1);
}

// One key of the options file: the schema of ZeroRKCFDPluginIFP
// (applications/cfd_plugin/generateZeroRKCFDPluginIFP.py), i.e. the YAML key,
// the option it sets (zerork_reactor_manager.cpp:135-191), its type, default
// and allowed values.  A key without a default is required.
struct IfpKey {
  const char* yaml;
  const char* option;
  char type;  // 'i' int, 'd' double, 's' string
  bool has_default;
  double def;
  const char* sdef;
  double lo, hi;        // bounds (inclusive); NaN: none
  const char* allowed;  // discrete int values, "" for any
};

const double kNone = std::nan("");

const IfpKey kIfp[] = {
    {"mechanism_parsing_log", "mechanism_parsing_log_filename", 's', true, 0, "/dev/null", kNone, kNone, ""},
    {"reactor_timing_log", "reactor_timing_log_filename", 's', true, 0, "/dev/null", kNone, kNone, ""},
    {"n_reactors_max", "n_reactors_max", 'i', true, 2048, nullptr, 1, kNone, ""},
    {"n_reactors_min", "n_reactors_min", 'i', true, 256, nullptr, 1, kNone, ""},
    {"max_dt", "max_dt", 'd', true, 0.05, nullptr, kNone, kNone, ""},
    {"max_steps", "max_steps", 'i', true, 5000, nullptr, kNone, kNone, ""},
    {"relative_tolerance", "rel_tol", 'd', true, 1.0e-8, nullptr, kNone, kNone, ""},
    {"absolute_tolerance", "abs_tol", 'd', true, 1.0e-20, nullptr, kNone, kNone, ""},
    {"abstol_dens", "abstol_dens", 'i', true, 0, nullptr, kNone, kNone, ""},
    {"reference_temperature", "reference_temperature", 'd', true, 1.0, nullptr, kNone, kNone, ""},
    {"preconditioner_threshold", "preconditioner_threshold", 'd', false, 0, nullptr, kNone, kNone, ""},
    {"eps_lin", "eps_lin", 'd', true, 0.1, nullptr, kNone, kNone, ""},
    {"nonlinear_convergence_coeff", "nonlinear_convergence_coeff", 'd', true, 0.05, nullptr, kNone, kNone, ""},
    {"integrator", "integrator", 'i', true, 0, nullptr, kNone, kNone, "0123"},
    {"dense", "dense", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"analytic", "analytic", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"iterative", "iterative", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"gpu", "gpu", 'i', true, 0, nullptr, kNone, kNone, "01"},
    {"initial_gpu_multiplier", "initial_gpu_multiplier", 'd', true, 20.0, nullptr, 1.0, kNone, ""},
    {"load_balance", "load_balance", 'i', true, 1, nullptr, kNone, kNone, "012"},
    {"load_balance_noise", "load_balance_noise", 'i', true, 0, nullptr, 0, kNone, ""},
    {"load_balance_mem", "load_balance_mem", 'i', true, 1, nullptr, 0, kNone, ""},
    {"reactor_weight_mult", "reactor_weight_mult", 'i', true, 1, nullptr, 1, kNone, ""},
    {"sort_reactors", "sort_reactors", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"verbosity", "verbosity", 'i', true, 4, nullptr, kNone, kNone, "01234"},
    {"constant_volume", "constant_volume", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"delta_temperature_ignition", "delta_temperature_ignition", 'd', true, 0.0, nullptr, 0.0, 100000.0, ""},
    {"min_mass_fraction", "min_mass_fraction", 'd', true, 1.0e-30, nullptr, kNone, kNone, ""},
    {"step_limiter", "step_limiter", 'd', true, 1.0e22, nullptr, kNone, kNone, ""},
    {"always_solve_temperature", "always_solve_temperature", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"solve_temperature_threshold", "solve_temperature_threshold", 'd', true, 2.0, nullptr, 0.0, kNone, ""},
    {"dump_reactors", "dump_reactors", 'i', true, 0, nullptr, kNone, kNone, "01"},
    {"dump_failed_reactors", "dump_failed_reactors", 'i', true, 0, nullptr, kNone, kNone, "01"},
    {"output_performance_log", "output_performance_log", 'i', true, 1, nullptr, kNone, kNone, "01"},
    {"cvode_num_retries", "cvode_num_retries", 'i', true, 3, nullptr, 0, 20, ""},
    {"cvode_retry_absolute_tolerance_adjustment", "cvode_retry_absolute_tolerance_adjustment", 'd', true, 0.1, nullptr, 1.0e-6, 1.0, ""},
    {"cvode_retry_relative_tolerance_adjustment", "cvode_retry_relative_tolerance_adjustment", 'd', true, 1.0, nullptr, 1.0e-6, 1.0, ""},
};

// Options file (ZeroRKReactorManager::ReadOptionsFile,
// zerork_reactor_manager.cpp:135-191): a flat "key: value" YAML document
// validated against the schema above.  Every schema key sets its option,
// from the file or from the schema's default (which differs from the
// manager's constructor default for dense, eps_lin, cvode_num_retries,
// n_reactors_min/max and initial_gpu_multiplier); preconditioner_threshold
// has no default and must be present; a value outside its type, bounds or
// allowed set fails the parse and leaves the handle's options unchanged.
// Keys outside the schema are kept: an int option of this library's own
// (fp32_preconditioner, stop_after_ignition, num_species_stride) takes an
// integer, anything else is stored as a string.
zerork_status_t read_options(zerork_handle h, const char* fn) {
  std::ifstream f(fn);
  if (!f) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(f, line)) {
    auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    if (trim(line).empty() || trim(line) == "---" || trim(line) == "...") continue;
    auto colon = line.find(':');
    if (colon == std::string::npos) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
    kv[trim(line.substr(0, colon))] = trim(line.substr(colon + 1));
  }
  auto parse_int = [](const std::string& v, long* out) {
    if (v == "true" || v == "True") { *out = 1; return true; }
    if (v == "false" || v == "False") { *out = 0; return true; }
    char* end = nullptr;
    *out = std::strtol(v.c_str(), &end, 10);
    return !v.empty() && !*end;
  };
  auto parse_dbl = [](const std::string& v, double* out) {
    char* end = nullptr;
    *out = std::strtod(v.c_str(), &end);
    return !v.empty() && !*end;
  };
  std::map<std::string, int> iset;
  std::map<std::string, double> dset;
  std::map<std::string, std::string> sset;
  for (const IfpKey& k : kIfp) {
    auto it = kv.find(k.yaml);
    if (it == kv.end() && !k.has_default) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
    if (k.type == 's') {
      sset[k.option] = it == kv.end() ? std::string(k.sdef) : it->second;
      continue;
    }
    double v = k.def;
    if (it != kv.end()) {
      if (k.type == 'i') {
        long iv;
        if (!parse_int(it->second, &iv)) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
        v = (double)iv;
      } else if (!parse_dbl(it->second, &v)) {
        return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
      }
    }
    if ((!std::isnan(k.lo) && v < k.lo) || (!std::isnan(k.hi) && v > k.hi))
      return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
    if (k.allowed[0]) {
      bool ok = false;
      for (const char* c = k.allowed; *c; ++c) ok = ok || v == (double)(*c - '0');
      if (!ok) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
    }
    // initial_gpu_multiplier is a double in the schema and an int option of
    // the manager (the value is truncated, as the manager's assignment does)
    if (k.type == 'i' || h->iopt.count(k.option)) iset[k.option] = (int)v;
    else dset[k.option] = v;
  }
  for (const auto& e : kv) {
    bool schema = false;
    for (const IfpKey& k : kIfp) schema = schema || e.first == k.yaml;
    if (schema) continue;
    if (h->iopt.count(e.first)) {
      long iv;
      if (!parse_int(e.second, &iv)) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
      iset[e.first] = (int)iv;
    } else if (h->dopt.count(e.first)) {
      double dv;
      if (!parse_dbl(e.second, &dv)) return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
      dset[e.first] = dv;
    } else {
      sset[e.first] = e.second;
    }
  }
  for (auto& e : iset) h->iopt[e.first] = e.second;
  for (auto& e : dset) h->dopt[e.first] = e.second;
  for (auto& e : sset) h->sopt[e.first] = e.second;
  return ZERORK_STATUS_SUCCESS;
}

// ZeroRKReactorManager::DumpReactor (zerork_reactor_manager.cpp:1477-1497):
// one tab-separated line "id T P cost gpu mf_0 .. mf_{nsp-1}" (20 significant
// digits) appended to dumpfile_<tag>_<call, 6 digits>_<rank, 6 digits>.
void dump_reactor(const zerork_handle_impl* h, const char* tag, int id, double T, double P,
                  double rc, double rg, const double* mf, int nsp) {
  char name[96];
  std::snprintf(name, sizeof name, "dumpfile_%s_%06d_%06d", tag, h->n_calls, 0);
  std::FILE* f = std::fopen(name, "a");
  if (!f) return;
  std::fprintf(f, "%d\t%.20g\t%.20g\t%.20g\t%.20g\t", id, T, P, rc, rg);
  for (int k = 0; k < nsp; ++k) std::fprintf(f, "%.20g\t", mf[k]);
  std::fprintf(f, "\n");
  std::fclose(f);
}

// The device options of a solve from the handle's plugin options.
zrk_bdf_options solve_options(zerork_handle h, int nsp) {
  zrk_bdf_options o;
  zrk_bdf_default_options(&o);
  o.rel_tol = h->dopt["rel_tol"];
  o.abs_tol = h->dopt["abs_tol"];
  o.max_dt = h->dopt["max_dt"];
  o.nonlinear_convergence_coeff = h->dopt["nonlinear_convergence_coeff"];
  o.step_limiter = h->dopt["step_limiter"];
  o.reference_temperature = h->dopt["reference_temperature"];
  o.min_mass_fraction = h->dopt["min_mass_fraction"];
  o.cvode_retry_absolute_tolerance_adjustment = h->dopt["cvode_retry_absolute_tolerance_adjustment"];
  o.cvode_retry_relative_tolerance_adjustment = h->dopt["cvode_retry_relative_tolerance_adjustment"];
  o.max_steps = h->iopt["max_steps"];
  o.cvode_num_retries = h->iopt["cvode_num_retries"];
  o.constant_volume = h->iopt["constant_volume"] != 0;
  // mechanisms above 191 species take the sparse path, whose LU is FP32 only
  o.fp32_preconditioner = h->iopt["fp32_preconditioner"] != 0 || nsp > 191;
  o.delta_temperature_ignition = h->dopt["delta_temperature_ignition"];
  o.stop_after_ignition = h->iopt["stop_after_ignition"] != 0;
  // the linear solver (solver_cvode.cpp:93-214) and tolerances
  // (reactor_nvector_serial.cpp:107-120, 414-493)
  o.abstol_dens = h->iopt["abstol_dens"] != 0;
  o.dense = h->iopt["dense"];
  o.analytic = h->iopt["analytic"];
  o.iterative = h->iopt["iterative"];
  o.eps_lin = h->dopt["eps_lin"];
  o.preconditioner_threshold = h->dopt["preconditioner_threshold"];
  if (nsp > 191) {
    // the sparse path solves the Newton system directly with the static
    // sparse LU of I - gamma J and the analytic Jacobian (the reference's
    // sparse direct solve); the Krylov and difference-quotient options apply
    // to the dense paths only
    o.dense = 1;
    o.iterative = 0;
    o.analytic = 1;
  }
  return o;
}

// The step callback over the records of a one-reactor solve
// (solver_cvode.cpp:247-257): cb(reactor id, CVode calls, t, t - t_prev, y,
// ydot) for each successful CVode call in order; a non-zero return ends the
// integration there, as a success: the reactor's state is that record's and
// its cost that record's call count, the pressure the end pressure of that
// state (GetState at the solve's dt, reactor_constant_volume_cpu.cpp:44-63,
// reactor_constant_pressure_cpu.cpp:44-66); a stop on the call that passed dt
// leaves the solve's end state, CVodeGetDky at dt (solver_cvode.cpp:268-270).
bool replay_callback(zerork_handle h, const zrk_bdf_options& o, int nsp, int rec_len, double dt,
                     double T_in, double P_in, const double* Y_in, bool solve_T, bool srcs,
                     double* T, double* P, double* mf, int* steps) {
  int nrec = 0;
  if (hipMemcpy(&nrec, h->dtrace_n, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return false;
  // the solve was run until every record had room (it is deterministic)
  if (nrec < 0 || (size_t)nrec > h->cap_trace) return false;
  std::vector<double> rec((size_t)nrec * rec_len);
  if (nrec > 0 && hipMemcpy(rec.data(), h->dtrace, sizeof(double) * rec.size(),
                            hipMemcpyDeviceToHost) != hipSuccess)
    return false;
  const int id = h->ids ? h->ids[0] : 0;
  for (int k = 0; k < nrec; ++k) {
    const double* r = rec.data() + (size_t)k * rec_len;
    if (h->cb(id, (int)r[0], r[1], r[2], r + 3, r + 4 + nsp, h->cb_data) == 0) continue;
    // stopped on the call that passed dt: the state is the interpolant at
    // dt, the solve's own end state
    if (r[1] >= dt) break;
    const double* y = r + 3;
    std::copy(y, y + nsp, mf);
    const double Tref = o.reference_temperature;
    *T = solve_T ? y[nsp] * Tref : (T_in / Tref) * Tref;
    *steps = (int)r[0];
    const double dpdt = h->dpdt ? h->dpdt[0] : 0.0;
    if (o.constant_volume || (srcs && h->y_src)) {
      // the volume of the input state, P = Ru T / (v Wmix) (+ dpdt dt at
      // constant volume)
      const std::vector<double>& mw = h->mech->host.mw;
      double mws0 = 0.0, mws = 0.0;
      for (int j = 0; j < nsp; ++j) {
        mws0 += Y_in[j] * (1.0 / mw[j]);
        mws += y[j] * (1.0 / mw[j]);
      }
      const double v = 1.0 / (P_in * (1.0 / mws0) / (kRuJ * T_in));
      *P = kRuJ * *T / (v * (1.0 / mws)) + (o.constant_volume ? dpdt * dt : 0.0);
    } else if (srcs) {
      *P = P_in + dpdt * dt;
    }
    break;
  }
  return true;
}

zerork_status_t solve(zerork_handle h, int n_cycle, double dt, int n, double* T, double* P,
                      double* mf) {
  if (n < 0 || !T || !P || (n > 0 && !mf)) return ZERORK_STATUS_INVALID_POINTER;
  if (n == 0) return ZERORK_STATUS_SUCCESS;
  // only CVODE (integrator 0) is restated; SEULEX, SODEX and RADAU are not
  if (h->iopt["integrator"] != 0) {
    std::fprintf(stderr, "zerork-hip: integrator %d is not available (only 0, CVODE)\n",
                 h->iopt["integrator"]);
    return ZERORK_STATUS_UNKNOWN_ERROR;
  }
  zrk::MechHandle* m = h->mech.get();
  const int nsp = m->host.nsp;
  const int stride = h->iopt.count("num_species_stride") ? h->iopt["num_species_stride"] : nsp;
  if (stride < nsp) return ZERORK_STATUS_INVALID_POINTER;
  if (!h->stream && hipStreamCreate(&h->stream) != hipSuccess) return ZERORK_STATUS_UNKNOWN_ERROR;
  if ((size_t)n > h->cap_n) {
    if (!grow(&h->dT, n) || !grow(&h->dP, n) || !grow(&h->dsteps, n) || !grow(&h->dorder, n))
      return ZERORK_STATUS_UNKNOWN_ERROR;
    h->cap_n = n;
  }
  const size_t ny = (size_t)n * stride;
  if (ny > h->cap_y) {
    if (!grow(&h->dY, ny)) return ZERORK_STATUS_UNKNOWN_ERROR;
    h->cap_y = ny;
  }
  std::vector<double> T0(T, T + n);
  // the manager's own cost and temperature-change arrays when the caller
  // gives none: resized, keeping earlier values (SetInputVariables,
  // zerork_reactor_manager.cpp:503-518)
  if (!h->cost) h->rc_own.resize(n, 1.0);
  if (!h->tdelta) h->tdelta_own.resize(n, 0.0);
  double* rc = h->cost ? h->cost : h->rc_own.data();
  double* tdelta = h->tdelta ? h->tdelta : h->tdelta_own.data();
  // work order: descending previous cost (the reference's sort_reactors,
  // :522-524, as longest-processing-time scheduling)
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  const bool sorted = h->iopt["sort_reactors"] != 0;
  if (sorted)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rc[a] > rc[b]; });
  // reactor dumps (zerork_reactor_manager.cpp:837, 904-906, 1042-1045, 1124-1128):
  // ids are positions in the solve order
  h->n_calls++;
  const bool dump = h->iopt["dump_reactors"] != 0;
  if (dump)
    for (int j = 0; j < n; ++j) {
      const int k = order[j];
      dump_reactor(h, "pre", j, T[k], P[k], rc[k], h->gpu ? h->gpu[k] : 0.0,
                   mf + (size_t)k * stride, nsp);
    }
  hipStream_t st = h->stream;
  std::vector<double> P0(P, P + n);
  bool ok = hipMemcpyAsync(h->dT, T, sizeof(double) * n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(h->dP, P, sizeof(double) * n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(h->dY, mf, sizeof(double) * ny, hipMemcpyHostToDevice, st) == hipSuccess;
  if (sorted)
    ok = ok && hipMemcpyAsync(h->dorder, order.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                              st) == hipSuccess;
  // solve_temperature (zerork_reactor_manager.cpp:1109-1113): with
  // always_solve_temperature = 0 a reactor is integrated at fixed temperature
  // unless its temperature change of the previous solve is above 0
  const int always_T = h->iopt["always_solve_temperature"] != 0;
  if (!always_T) {
    if ((size_t)n > h->cap_td) {
      if (!grow(&h->dtdelta, n)) return ZERORK_STATUS_UNKNOWN_ERROR;
      h->cap_td = n;
    }
    ok = ok && hipMemcpyAsync(h->dtdelta, tdelta, sizeof(double) * n, hipMemcpyHostToDevice,
                              st) == hipSuccess;
  }
  int n_notemp = 0;
  if (!always_T)
    for (int k = 0; k < n; ++k) n_notemp += !(tdelta[k] > 0.0);
  // source terms from the DPDT, E_SRC and Y_SRC aux fields (per reactor; Y_SRC
  // in the layout of mf), as the manager passes them to InitializeState
  // (zerork_reactor_manager.cpp:909-1053)
  const bool srcs = h->dpdt || h->e_src || h->y_src;
  // IGNITION_TIME: the reactor's root time (zerork_reactor_manager.cpp:1131)
  // when delta_temperature_ignition > 0; copied in first so that a failed
  // reactor keeps the value it had, as the manager leaves it
  const bool ign = h->dopt["delta_temperature_ignition"] > 0.0 && h->ign != nullptr;
  const double *ddpdt = nullptr, *desrc = nullptr, *dysrc = nullptr;
  double* dign = nullptr;
  if (srcs || ign) {
    if ((size_t)(3 * n) > h->cap_src_n) {
      if (!grow(&h->dsrc_n, 3 * (size_t)n)) return ZERORK_STATUS_UNKNOWN_ERROR;
      h->cap_src_n = 3 * (size_t)n;
    }
    if (ign) {
      dign = h->dsrc_n + 2 * (size_t)n;
      ok = ok && hipMemcpyAsync(dign, h->ign, sizeof(double) * n, hipMemcpyHostToDevice, st) ==
                     hipSuccess;
    }
    if (h->dpdt) {
      ok = ok && hipMemcpyAsync(h->dsrc_n, h->dpdt, sizeof(double) * n, hipMemcpyHostToDevice,
                                st) == hipSuccess;
      ddpdt = h->dsrc_n;
    }
    if (h->e_src) {
      ok = ok && hipMemcpyAsync(h->dsrc_n + n, h->e_src, sizeof(double) * n,
                                hipMemcpyHostToDevice, st) == hipSuccess;
      desrc = h->dsrc_n + n;
    }
    if (h->y_src) {
      if (ny > h->cap_src_y) {
        if (!grow(&h->dysrc, ny)) return ZERORK_STATUS_UNKNOWN_ERROR;
        h->cap_src_y = ny;
      }
      ok = ok && hipMemcpyAsync(h->dysrc, h->y_src, sizeof(double) * ny, hipMemcpyHostToDevice,
                                st) == hipSuccess;
      dysrc = h->dysrc;
    }
  }
  if (!ok) return ZERORK_STATUS_UNKNOWN_ERROR;
  zrk_bdf_options o = solve_options(h, nsp);
  // the step callback: one reactor, load_balance 0 (zerork_reactor_manager.cpp:1082-1084)
  const bool use_cb = h->cb != nullptr && h->iopt["load_balance"] == 0 && n == 1;
  const int rec_len = 2 * nsp + 5;
  zrk::StepTrace trace{nullptr, 0, nullptr};
  std::vector<double> mf0;
  if (use_cb) {
    // records for max_steps calls per try (plus root returns) of up to
    // cvode_num_retries tries, within 256 MiB; a solve that makes more is
    // run again from its inputs with room for all of them
    const size_t tries = std::max(1, o.cvode_num_retries);
    const size_t want = std::min<size_t>(((size_t)std::max(o.max_steps, 1) + 136) * tries,
                                         std::max<size_t>(1, (256u << 20) / (8u * rec_len)));
    if (!h->dtrace_n && !grow(&h->dtrace_n, 1)) return ZERORK_STATUS_UNKNOWN_ERROR;
    if (want > h->cap_trace) {
      if (!grow(&h->dtrace, want * rec_len)) return ZERORK_STATUS_UNKNOWN_ERROR;
      h->cap_trace = want;
    }
    mf0.assign(mf, mf + ny);
  }
  std::vector<int> steps(n);
  int n_resolved = 0;
  int rc_dev = 0;
  std::chrono::steady_clock::time_point t0;
  for (;;) {
    trace = zrk::StepTrace{h->dtrace, (int)std::min<size_t>(h->cap_trace, INT32_MAX), h->dtrace_n};
    t0 = std::chrono::steady_clock::now();
    rc_dev = zrk::integrate_device(m, o, n, dt, nullptr, h->dT, h->dP, h->dY, stride, h->dsteps,
                                   nullptr, sorted ? h->dorder : nullptr, st, ddpdt, desrc, dysrc,
                                   dign, always_T ? nullptr : h->dtdelta, always_T, nullptr,
                                   use_cb ? &trace : nullptr);
    if (rc_dev != 0) return ZERORK_STATUS_UNKNOWN_ERROR;
    ok = hipMemcpyAsync(steps.data(), h->dsteps, sizeof(int) * n, hipMemcpyDeviceToHost, st) ==
         hipSuccess;
    if (hipStreamSynchronize(st) != hipSuccess || !ok) return ZERORK_STATUS_UNKNOWN_ERROR;
    // a reactor whose batched integration failed is solved again on its own,
    // as the manager re-solves a failed GPU batch on the CPU
    // (zerork_reactor_manager.cpp:1021-1034, 1090-1147): here with the
    // reference's own arithmetic, the FP64 Newton matrix, when the batch used
    // the FP32 one (failed reactors kept their input state); the step records
    // are then the re-solve's
    std::vector<int> failed;
    for (int k = 0; k < n; ++k)
      if (steps[k] < 0) failed.push_back(k);
    n_resolved = 0;
    if (!failed.empty() && o.fp32_preconditioner && nsp <= 191) {
      zrk_bdf_options o64 = o;
      o64.fp32_preconditioner = 0;
      ok = hipMemcpyAsync(h->dorder, failed.data(), sizeof(int) * failed.size(),
                          hipMemcpyHostToDevice, st) == hipSuccess;
      if (!ok) return ZERORK_STATUS_UNKNOWN_ERROR;
      rc_dev = zrk::integrate_device(m, o64, (int)failed.size(), dt, nullptr, h->dT, h->dP,
                                     h->dY, stride, h->dsteps, nullptr, h->dorder, st, ddpdt,
                                     desrc, dysrc, dign, always_T ? nullptr : h->dtdelta,
                                     always_T, nullptr, use_cb ? &trace : nullptr);
      if (rc_dev != 0) return ZERORK_STATUS_UNKNOWN_ERROR;
      ok = hipMemcpyAsync(steps.data(), h->dsteps, sizeof(int) * n, hipMemcpyDeviceToHost,
                          st) == hipSuccess;
      if (hipStreamSynchronize(st) != hipSuccess || !ok) return ZERORK_STATUS_UNKNOWN_ERROR;
      n_resolved = (int)failed.size();
    }
    if (!use_cb) break;
    int nrec = 0;
    ok = hipMemcpy(&nrec, h->dtrace_n, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return ZERORK_STATUS_UNKNOWN_ERROR;
    if ((size_t)nrec <= h->cap_trace) break;
    // more records than room: the same solve again (it is deterministic)
    // from the inputs, with room for all of them
    if (!grow(&h->dtrace, (size_t)nrec * rec_len)) return ZERORK_STATUS_UNKNOWN_ERROR;
    h->cap_trace = (size_t)nrec;
    ok = hipMemcpyAsync(h->dT, T0.data(), sizeof(double) * n, hipMemcpyHostToDevice, st) ==
         hipSuccess;
    ok = ok && hipMemcpyAsync(h->dP, P0.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                              st) == hipSuccess;
    ok = ok && hipMemcpyAsync(h->dY, mf0.data(), sizeof(double) * ny, hipMemcpyHostToDevice,
                              st) == hipSuccess;
    if (sorted)
      ok = ok && hipMemcpyAsync(h->dorder, order.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                                st) == hipSuccess;
    if (ign)
      ok = ok && hipMemcpyAsync(dign, h->ign, sizeof(double) * n, hipMemcpyHostToDevice, st) ==
                     hipSuccess;
    if (!ok) return ZERORK_STATUS_UNKNOWN_ERROR;
  }
  ok = hipMemcpyAsync(T, h->dT, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(mf, h->dY, sizeof(double) * ny, hipMemcpyDeviceToHost, st) == hipSuccess;
  if (o.constant_volume || srcs)  // end pressure (GetState)
    ok = ok && hipMemcpyAsync(P, h->dP, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess;
  if (ign)
    ok = ok && hipMemcpyAsync(h->ign, dign, sizeof(double) * n, hipMemcpyDeviceToHost, st) ==
                   hipSuccess;
  if (hipStreamSynchronize(st) != hipSuccess || !ok) return ZERORK_STATUS_UNKNOWN_ERROR;
  if (use_cb) {
    if (!replay_callback(h, o, nsp, rec_len, dt, T0[0], P0[0], mf0.data(), always_T || tdelta[0] > 0.0,
                         srcs, T, P, mf, &steps[0]))
      return ZERORK_STATUS_UNKNOWN_ERROR;
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  zerork_status_t flag = ZERORK_STATUS_SUCCESS;
  const double thr = h->dopt["solve_temperature_threshold"];
  long long nsteps_total = 0;
  int nsolved = 0;
  for (int k = 0; k < n; ++k) {
    if (steps[k] < 0) flag = ZERORK_STATUS_FAILED_SOLVE;
    rc[k] = steps[k];
    if (h->gpu) h->gpu[k] = secs / n;
    if (steps[k] >= 0) {
      // TEMPERATURE_DELTA: the change if at least the threshold, else 0
      const double d = T[k] - T0[k];
      tdelta[k] = d < thr ? 0.0 : d;
      nsteps_total += steps[k];
      ++nsolved;
    }
  }
  const bool dump_failed = h->iopt["dump_failed_reactors"] != 0;
  if (dump || dump_failed)
    for (int j = 0; j < n; ++j) {
      const int k = order[j];
      const double* y = mf + (size_t)k * stride;
      if (dump_failed && steps[k] < 0)  // the kept input state
        dump_reactor(h, "failed_state", j, T[k], P0[k], steps[k], secs / n, y, nsp);
      if (dump) dump_reactor(h, "postg", j, T[k], P[k], steps[k], secs / n, y, nsp);
    }
  if (h->timing_open) {
    // one line per solve (zerork_reactor_manager.cpp:1340-1376); the whole
    // batch counts as one GPU group, the re-solved failures as CPU solves
    const int n_gpu = nsolved, n_cpu = n_resolved;
    const long long avg = nsolved > 0 ? nsteps_total / nsolved : 0;
    h->timing_log << std::setw(13) << n_cycle << std::setw(17) << nsolved
                  << std::setw(17) << n_cpu << std::setw(17) << 0 << std::setw(17) << n_gpu
                  << std::setw(17) << n_notemp << std::setw(17) << 1 << std::setw(17) << avg
                  << std::setw(17) << 0 << std::setw(17) << avg << std::setw(17) << 0
                  << std::setw(17) << secs << std::setw(17) << 0
                  << std::setw(17) << (nsteps_total > 0 ? secs / nsteps_total : 0.0)
                  << std::setw(17) << secs << std::setw(17) << secs << std::endl;
  }
  return flag;
}

}  // namespace

extern "C" {

zerork_handle zerork_reactor_init() {
  try {
    return new zerork_handle_impl();
  } catch (...) {
    return nullptr;
  }
}

zerork_status_t zerork_reactor_read_options_file(const char* options_filename, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!options_filename) return ZERORK_STATUS_INVALID_FILENAME;
  try {
    return read_options(h, options_filename);
  } catch (...) {
    return ZERORK_STATUS_FAILED_OPTIONS_PARSE;
  }
}

zerork_status_t zerork_reactor_set_mechanism_files(const char* mech_file, const char* therm_file,
                                                   zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!mech_file || !therm_file) return ZERORK_STATUS_INVALID_FILENAME;
  h->sopt["mech_filename"] = mech_file;
  h->sopt["therm_filename"] = therm_file;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_load_mechanism(zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  try {
    return load_mechanism(h);
  } catch (...) {
    return ZERORK_STATUS_UNKNOWN_ERROR;
  }
}

zerork_status_t zerork_reactor_solve(const int n_cycle, const double time, const double dt,
                                     const int n_reactors, double* T, double* P, double* mf,
                                     zerork_handle h) {
  (void)time;
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  try {
    zerork_status_t f = finish_init(h);
    if (f != ZERORK_STATUS_SUCCESS) return f;
    return solve(h, n_cycle, dt, n_reactors, T, P, mf);
  } catch (...) {
    return ZERORK_STATUS_UNKNOWN_ERROR;
  }
}

zerork_status_t zerork_reactor_set_aux_field_pointer(zerork_field_t ft, double* p,
                                                     zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!p) return ZERORK_STATUS_INVALID_POINTER;
  switch (ft) {
    case ZERORK_FIELD_DPDT: h->dpdt = p; break;
    case ZERORK_FIELD_COST: h->cost = p; break;
    case ZERORK_FIELD_GPU: h->gpu = p; break;
    case ZERORK_FIELD_Y_SRC: h->y_src = p; break;
    case ZERORK_FIELD_E_SRC: h->e_src = p; break;
    case ZERORK_FIELD_IGNITION_TIME: h->ign = p; break;
    case ZERORK_FIELD_TEMPERATURE_DELTA: h->tdelta = p; break;
    default: return ZERORK_STATUS_INVALID_FIELD_NAME;
  }
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_set_int_option(const char* name, int v, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!name) return ZERORK_STATUS_INVALID_POINTER;
  h->iopt[name] = v;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_set_double_option(const char* name, double v, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!name) return ZERORK_STATUS_INVALID_POINTER;
  h->dopt[name] = v;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_set_string_option(const char* name, const char* v,
                                                 zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!name || !v) return ZERORK_STATUS_INVALID_POINTER;
  h->sopt[name] = v;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_get_int_option(const char* name, int* v, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!name || !v) return ZERORK_STATUS_INVALID_POINTER;
  auto it = h->iopt.find(name);
  if (it == h->iopt.end()) return ZERORK_STATUS_OPTION_NOT_DEFINED;
  *v = it->second;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_get_double_option(const char* name, double* v, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!name || !v) return ZERORK_STATUS_INVALID_POINTER;
  auto it = h->dopt.find(name);
  if (it == h->dopt.end()) return ZERORK_STATUS_OPTION_NOT_DEFINED;
  *v = it->second;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_get_string_option(const char* name, char** v, int len,
                                                 zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!name || !v || !*v) return ZERORK_STATUS_INVALID_POINTER;
  auto it = h->sopt.find(name);
  if (it == h->sopt.end()) return ZERORK_STATUS_OPTION_NOT_DEFINED;
  std::strncpy(*v, it->second.c_str(), len);
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_set_callback_fn(zerork_callback_fn fn, void* data, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!fn) return ZERORK_STATUS_INVALID_POINTER;
  h->cb = fn;
  h->cb_data = data;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_set_reactor_ids(int* ids, zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  if (!ids) return ZERORK_STATUS_INVALID_POINTER;
  h->ids = ids;
  return ZERORK_STATUS_SUCCESS;
}

zerork_status_t zerork_reactor_free(zerork_handle h) {
  if (!h) return ZERORK_STATUS_INVALID_HANDLE;
  delete h;
  return ZERORK_STATUS_SUCCESS;
}

}  // extern "C"
