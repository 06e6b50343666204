// Device-level C ABI (include/zrk_device.h): mechanism upload and launches.
#include <hip/hip_runtime_api.h>

#include <cstdio>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zrk_device.h"
#include "ck_mech.h"
#include "dev_mech.h"
#include "launch.h"
#include "zrk_internal.h"

namespace zrk {
namespace {

struct BlobBuilder {
  std::vector<char> bytes;
  template <class T>
  size_t add(const std::vector<T>& v) {
    size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + std::max<size_t>(v.size() * sizeof(T), 8), 0);
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

}  // namespace

int build_device_mech(MechHandle* h, std::string* err) {
  if (h->dev_blob) return 0;
  const Mechanism& M = h->host;
  const int nsp = M.nsp, nstep = M.nstep, nrxn = M.nrxn;
  std::vector<double> inv_mw(nsp), ru_mw(nsp);
  for (int j = 0; j < nsp; ++j) {
    inv_mw[j] = 1.0 / M.mw[j];
    ru_mw[j] = 8.314462618e3 / M.mw[j];
  }
  // reaction records, grouped by kind (plain, third-body, falloff)
  std::vector<int> order;
  for (int pass = 0; pass < 3; ++pass)
    for (int j = 0; j < nrxn; ++j) {
      const auto& r = M.rxns[j];
      const int kind = r.falloff ? 2 : (r.third_body ? 1 : 0);
      if (kind == pass) order.push_back(j);
    }
  const int nrec = (int)order.size();
  std::vector<int> fstep(nrec), rstep(nrec), flags(nrec), reac(nrec * kMaxPart, -1),
      prod(nrec * kMaxPart, -1), eptr(nrec + 1, 0), fspc(nrec, -1), esp;
  std::vector<double> fA(nrec), fB(nrec), fT(nrec), rA(nrec, 0.0), rB(nrec, 0.0), rT(nrec, 0.0),
      fo(nrec * 7, 0.0), eeff;
  for (int q = 0; q < nrec; ++q) {
    const int j = order[q];
    const auto& r = M.rxns[j];
    const int s = M.fwd_step[j];
    fstep[q] = s;
    rstep[q] = M.rev_step[j];
    fA[q] = M.step_lnA[s];
    fB[q] = M.step_beta[s];
    fT[q] = M.step_Ta[s];
    int fl = 0;
    if (M.rev_type[j] == kRevArrhenius) {
      fl |= kRecRevArr;
      const int sr = M.rev_step[j];
      rA[q] = M.step_lnA[sr];
      rB[q] = M.step_beta[sr];
      rT[q] = M.step_Ta[sr];
    } else if (M.rev_type[j] == kRevKeq) {
      fl |= kRecRevKeq;
      rA[q] = M.keq_ndelta[j];
    }
    if (r.third_body) fl |= kRecThirdBody;
    if (r.falloff) {
      fl |= kRecFalloff;
      if (M.falloff_type[j] == kTroe3 || M.falloff_type[j] == kTroe4) fl |= kRecTroe;
      if (M.falloff_type[j] == kTroe4) fl |= kRecTroe4;
      for (int k = 0; k < 7; ++k) fo[q * 7 + k] = M.falloff_param[(size_t)j * 7 + k];
      fspc[q] = M.falloff_spc[j];
    }
    flags[q] = fl;
    const int nr = M.reac_ptr[s + 1] - M.reac_ptr[s];
    const int np = M.prod_ptr[s + 1] - M.prod_ptr[s];
    if (nr > kMaxPart || np > kMaxPart) {
      *err = "reaction " + std::to_string(j + 1) + " has more than 4 expanded participants";
      return -1;
    }
    for (int k = 0; k < nr; ++k) reac[q * kMaxPart + k] = M.reac_sp[M.reac_ptr[s] + k];
    for (int k = 0; k < np; ++k) prod[q * kMaxPart + k] = M.prod_sp[M.prod_ptr[s] + k];
    for (int k = M.tb_sp_ptr[j]; k < M.tb_sp_ptr[j + 1]; ++k) {
      esp.push_back(M.tb_sp[k]);
      eeff.push_back(M.tb_eff[k]);
    }
    eptr[q + 1] = (int)esp.size();
  }
  // per-species gather lists in step order (reference summation order)
  std::vector<std::vector<int>> dl(nsp), cl(nsp);
  for (int s = 0; s < nstep; ++s) {
    for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) dl[M.reac_sp[k]].push_back(s);
    for (int k = M.prod_ptr[s]; k < M.prod_ptr[s + 1]; ++k) cl[M.prod_sp[k]].push_back(s);
  }
  std::vector<int> dptr(nsp + 1, 0), cptr(nsp + 1, 0), dstep, cstep;
  int maxd = 0, maxc = 0;
  for (int j = 0; j < nsp; ++j) {
    dstep.insert(dstep.end(), dl[j].begin(), dl[j].end());
    cstep.insert(cstep.end(), cl[j].begin(), cl[j].end());
    dptr[j + 1] = (int)dstep.size();
    cptr[j + 1] = (int)cstep.size();
    maxd = std::max(maxd, (int)dl[j].size());
    maxc = std::max(maxc, (int)cl[j].size());
  }
  // Jacobian species-block terms, entry (row i, col c): destruction terms
  // first then creation terms, each in step order (the reference's order of
  // accumulation, reactor_nvector_serial.cpp:149-176 and 655-669)
  std::vector<std::vector<int>> jd((size_t)nsp * nsp), jc((size_t)nsp * nsp);
  for (int s = 0; s < nstep; ++s)
    for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) {
      const int c = M.reac_sp[k];
      for (int q = M.reac_ptr[s]; q < M.reac_ptr[s + 1]; ++q)
        jd[(size_t)c * nsp + M.reac_sp[q]].push_back(-(s + 1));
      for (int q = M.prod_ptr[s]; q < M.prod_ptr[s + 1]; ++q)
        jc[(size_t)c * nsp + M.prod_sp[q]].push_back(s + 1);
    }
  std::vector<int> jptr((size_t)nsp * nsp + 1, 0), jterm;
  for (size_t e = 0; e < (size_t)nsp * nsp; ++e) {
    jterm.insert(jterm.end(), jd[e].begin(), jd[e].end());
    jterm.insert(jterm.end(), jc[e].begin(), jc[e].end());
    jptr[e + 1] = (int)jterm.size();
  }

  BlobBuilder b;
  const size_t o_mw = b.add(M.mw), o_imw = b.add(inv_mw), o_rmw = b.add(ru_mw),
               o_th = b.add(M.thermo), o_fs = b.add(fstep), o_rs = b.add(rstep),
               o_fl = b.add(flags), o_fA = b.add(fA), o_fB = b.add(fB), o_fT = b.add(fT),
               o_rA = b.add(rA), o_rB = b.add(rB), o_rT = b.add(rT), o_re = b.add(reac),
               o_pr = b.add(prod), o_ep = b.add(eptr), o_es = b.add(esp), o_ee = b.add(eeff),
               o_fsp = b.add(fspc), o_fo = b.add(fo), o_dp = b.add(dptr), o_ds = b.add(dstep),
               o_cp = b.add(cptr), o_cs = b.add(cstep), o_jp = b.add(jptr), o_jt = b.add(jterm);
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, b.bytes.size());
  if (e != hipSuccess) {
    *err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return (int)e;
  }
  e = hipMemcpy(base, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    *err = std::string("hipMemcpy: ") + hipGetErrorString(e);
    return (int)e;
  }
  char* p = (char*)base;
  DevMech& d = h->dev;
  d.nsp = nsp;
  d.nstep = nstep;
  d.nrxn = nrxn;
  d.nrec = nrec;
  d.max_dlist = maxd;
  d.max_clist = maxc;
  d.mw = (const double*)(p + o_mw);
  d.inv_mw = (const double*)(p + o_imw);
  d.ru_mw = (const double*)(p + o_rmw);
  d.thermo = (const double*)(p + o_th);
  d.r_fstep = (const int*)(p + o_fs);
  d.r_rstep = (const int*)(p + o_rs);
  d.r_flags = (const int*)(p + o_fl);
  d.r_fA = (const double*)(p + o_fA);
  d.r_fB = (const double*)(p + o_fB);
  d.r_fT = (const double*)(p + o_fT);
  d.r_rA = (const double*)(p + o_rA);
  d.r_rB = (const double*)(p + o_rB);
  d.r_rT = (const double*)(p + o_rT);
  d.r_reac = (const int*)(p + o_re);
  d.r_prod = (const int*)(p + o_pr);
  d.r_eptr = (const int*)(p + o_ep);
  d.e_sp = (const int*)(p + o_es);
  d.e_eff = (const double*)(p + o_ee);
  d.r_fspc = (const int*)(p + o_fsp);
  d.r_fo = (const double*)(p + o_fo);
  d.d_ptr = (const int*)(p + o_dp);
  d.d_step = (const int*)(p + o_ds);
  d.c_ptr = (const int*)(p + o_cp);
  d.c_step = (const int*)(p + o_cs);
  d.jt_ptr = (const int*)(p + o_jp);
  d.jt_term = (const int*)(p + o_jt);
  h->dev_blob = base;
  h->dev_bytes = b.bytes.size();
  if (hipGetDevice(&h->device) != hipSuccess) h->device = 0;
  return 0;
}

// Integrator workspace (per handle): saved Jacobians per resident wave slot
// and the work counter.  Grown on demand, released with the handle.
int ensure_workspace(MechHandle* h, int nslots) {
  if (h->ws_slots >= nslots && h->ws_jac) return 0;
  if (h->ws_jac) (void)hipFree(h->ws_jac);
  if (h->ws_counter) (void)hipFree(h->ws_counter);
  h->ws_jac = nullptr;
  h->ws_counter = nullptr;
  h->ws_slots = 0;
  hipError_t e = hipMalloc((void**)&h->ws_jac, sizeof(double) * (size_t)nslots * kNP * kNP);
  if (e != hipSuccess) return (int)e;
  e = hipMalloc((void**)&h->ws_counter, 256);
  if (e != hipSuccess) return (int)e;
  h->ws_slots = nslots;
  return 0;
}

BdfOpts to_bdf_opts(const zrk_bdf_options& o) {
  BdfOpts b;
  b.rtol = o.rel_tol;
  b.atol = o.abs_tol;
  b.hmax_inv = o.max_dt > 0.0 ? 1.0 / o.max_dt : 0.0;
  b.nlscoef = o.nonlinear_convergence_coeff;
  b.step_limiter = o.step_limiter;
  b.Tref = o.reference_temperature;
  b.min_mass_fraction = o.min_mass_fraction;
  b.retry_atol_adj = o.cvode_retry_absolute_tolerance_adjustment;
  b.retry_rtol_adj = o.cvode_retry_relative_tolerance_adjustment;
  b.max_steps = o.max_steps;
  b.num_retries = o.cvode_num_retries;
  return b;
}

MechHandle::~MechHandle() {
  if (ws_jac) (void)hipFree(ws_jac);
  if (ws_counter) (void)hipFree(ws_counter);
  if (dev_blob) (void)hipFree(dev_blob);
}

int ensure_device(MechHandle* h) {
  std::string msg;
  const int rc = build_device_mech(h, &msg);
  if (rc) std::fprintf(stderr, "zerork-hip: device mechanism upload failed: %s\n", msg.c_str());
  return rc;
}

int integrate_device(MechHandle* h, const zrk_bdf_options& o, int n, double dt, double* T,
                     const double* P, double* Y, int stride, int* nsteps, int* stats,
                     const int* order, hipStream_t st) {
  if (n < 0 || !T || !P || !Y || stride < h->host.nsp || !(dt > 0.0)) return -1;
  if (h->host.nsp > kNP - 1) return -2;
  if (n == 0) return 0;
  if (int rc = ensure_device(h)) return rc;
  if (h->resident_waves <= 0) h->resident_waves = bdf_max_resident_waves(h->dev);
  const int nw = std::max(1, std::min(h->resident_waves, n));
  int rc = ensure_workspace(h, nw);
  if (rc) return rc;
  hipError_t e = launch_bdf(h->dev, to_bdf_opts(o), n, dt, T, P, Y, stride, nsteps, stats,
                            order, h->ws_counter, h->ws_jac, nw, st);
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace zrk

using zrk::MechHandle;

extern "C" {

void zrk_bdf_default_options(zrk_bdf_options* o) {
  // zerork_reactor_manager.cpp:50-101
  o->rel_tol = 1.0e-8;
  o->abs_tol = 1.0e-20;
  o->max_dt = 0.05;
  o->nonlinear_convergence_coeff = 0.05;
  o->step_limiter = 1.0e22;
  o->reference_temperature = 1.0;
  o->min_mass_fraction = 1.0e-30;
  o->cvode_retry_absolute_tolerance_adjustment = 0.1;
  o->cvode_retry_relative_tolerance_adjustment = 1.0;
  o->max_steps = 5000;
  o->cvode_num_retries = 5;
}

int zrk_mech_load(const char* mech_file, const char* therm_file, zrk_mech* out, char* err,
                  int errlen) {
  if (!mech_file || !out) return -1;
  auto* h = new MechHandle();
  std::string msg;
  if (!zrk::load_mechanism(mech_file, therm_file ? therm_file : "", &h->host, &msg)) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", msg.c_str());
    delete h;
    return -3;
  }
  // the device image is built on first device use, so a mechanism can be
  // parsed and inspected without a GPU
  *out = reinterpret_cast<zrk_mech>(h);
  return 0;
}

void zrk_mech_free(zrk_mech m) { delete reinterpret_cast<MechHandle*>(m); }

int zrk_mech_info(zrk_mech m, int* nsp, int* nrxn, int* nstep) {
  if (!m) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (nsp) *nsp = h->host.nsp;
  if (nrxn) *nrxn = h->host.nrxn;
  if (nstep) *nstep = h->host.nstep;
  return 0;
}

int zrk_mech_species(zrk_mech m, int j, char* name, int len) {
  if (!m || !name || len <= 0) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (j < 0 || j >= h->host.nsp) return -1;
  std::snprintf(name, len, "%s", h->host.species[j].name.c_str());
  return 0;
}

int zrk_mech_molwt(zrk_mech m, double* mw) {
  if (!m || !mw) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  std::copy(h->host.mw.begin(), h->host.mw.end(), mw);
  return 0;
}

int zrk_mech_steps(zrk_mech m, int* step_rxn, int* step_dir, int* reac_ptr, int* reac_sp,
                   int* prod_ptr, int* prod_sp) {
  if (!m) return -1;
  const auto& H = reinterpret_cast<MechHandle*>(m)->host;
  if (step_rxn) std::copy(H.step_rxn.begin(), H.step_rxn.end(), step_rxn);
  if (step_dir) std::copy(H.step_dir.begin(), H.step_dir.end(), step_dir);
  if (reac_ptr) std::copy(H.reac_ptr.begin(), H.reac_ptr.end(), reac_ptr);
  if (reac_sp) std::copy(H.reac_sp.begin(), H.reac_sp.end(), reac_sp);
  if (prod_ptr) std::copy(H.prod_ptr.begin(), H.prod_ptr.end(), prod_ptr);
  if (prod_sp) std::copy(H.prod_sp.begin(), H.prod_sp.end(), prod_sp);
  return 0;
}

int zrk_mech_reactions(zrk_mech m, int* fwd_step, int* rev_step, int* rev_type, int* n_enhanced,
                       int* falloff_type) {
  if (!m) return -1;
  const auto& H = reinterpret_cast<MechHandle*>(m)->host;
  for (int j = 0; j < H.nrxn; ++j) {
    const auto& r = H.rxns[j];
    if (fwd_step) fwd_step[j] = H.fwd_step[j];
    if (rev_step) rev_step[j] = H.rev_step[j];
    if (rev_type) rev_type[j] = H.rev_type[j];
    if (n_enhanced)
      n_enhanced[j] = (r.third_body || r.falloff) ? H.tb_sp_ptr[j + 1] - H.tb_sp_ptr[j] : -1;
    if (falloff_type) falloff_type[j] = H.falloff_type[j];
  }
  return 0;
}

int zrk_rates_device(zrk_mech m, int n, const double* T, const double* P, const double* Y,
                     int stride, double step_limiter, double* ydot, double* rop, double* create,
                     double* destroy, double* net, void* stream) {
  if (!m || n < 0 || !T || !P || !Y) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (stride < h->host.nsp) return -1;
  if (int rc = zrk::ensure_device(h)) return rc;
  hipError_t e = zrk::launch_rates(h->dev, n, T, P, Y, stride, step_limiter, ydot, rop, create,
                                   destroy, net, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (int)e;
}

int zrk_jacobian_device(zrk_mech m, const zrk_bdf_options* o, int n, const double* T,
                        const double* P, const double* Y, int stride, double* J, void* stream) {
  if (!m || !o || n < 0 || !T || !P || !Y || !J) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (stride < h->host.nsp) return -1;
  if (h->host.nsp > zrk::kNP - 1) return -2;
  if (int rc = zrk::ensure_device(h)) return rc;
  const int nb = std::max(1, std::min(n, 4096));
  int rc = zrk::ensure_workspace(h, nb);
  if (rc) return rc;
  hipError_t e = zrk::launch_jacobian(h->dev, zrk::to_bdf_opts(*o), n, T, P, Y, stride, J,
                                      h->ws_jac, nb, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (int)e;
}

int zrk_integrate_device(zrk_mech m, const zrk_bdf_options* o, int n, double dt, double* T,
                         const double* P, double* Y, int stride, int* nsteps, int* stats,
                         void* stream) {
  if (!m || !o) return -1;
  return zrk::integrate_device(reinterpret_cast<MechHandle*>(m), *o, n, dt, T, P, Y, stride,
                               nsteps, stats, nullptr, (hipStream_t)stream);
}

int zrk_integrator_resident_waves(zrk_mech m) {
  if (!m) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (int rc = zrk::ensure_device(h)) return -rc;
  if (h->resident_waves <= 0) h->resident_waves = zrk::bdf_max_resident_waves(h->dev);
  return h->resident_waves;
}

}  // extern "C"
