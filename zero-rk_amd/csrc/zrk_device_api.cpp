// Device-level C ABI (include/zrk_device.h): mechanism upload and launches.
#include <hip/hip_runtime_api.h>

#include <cstdio>

#include <algorithm>
#include <cstdint>
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
  // reaction records, grouped by kind (falloff, third-body, plain), then the
  // records that need the general evaluation (PLOG, non-integer, SRI)
  auto is_ext = [&](int j) {
    return M.rxn_plog[j] >= 0 || M.step_ni[M.fwd_step[j]] >= 0 || M.falloff_type[j] == kSri;
  };
  std::vector<int> order;
  for (int pass = 2; pass >= 0; --pass)
    for (int j = 0; j < nrxn; ++j) {
      const auto& r = M.rxns[j];
      const int kind = is_ext(j) ? -1 : (r.falloff ? 2 : (r.third_body ? 1 : 0));
      if (kind == pass) order.push_back(j);
    }
  for (int j = 0; j < nrxn; ++j)
    if (is_ext(j)) order.push_back(j);
  const int nrec = (int)order.size();
  if (nrec == 0) {
    *err = "mechanism has no reactions";
    return -1;
  }
  if (nstep >= 0xFFFF || nsp >= 0x7FFF) {
    *err = "mechanism too large for 16-bit step/species indices";
    return -1;
  }
  std::vector<int> meta((size_t)nrec * 4), spk((size_t)nrec * 4), eptr(nrec + 1, 0), esp;
  std::vector<double> coef((size_t)nrec * 6, 0.0), fo((size_t)nrec * kFoParams, 0.0), eeff;
  // PLOG tables and non-integer exponent lists of the extended records
  std::vector<int> ext, ni_sp;
  std::vector<double> plog_t, ni_pow;
  auto add_rop_list = [&](int step, int* off, int* cnt) {
    const int q = M.step_ni[step];
    *off = (int)ni_sp.size();
    *cnt = (int)M.ni[q].rop_sp.size();
    ni_sp.insert(ni_sp.end(), M.ni[q].rop_sp.begin(), M.ni[q].rop_sp.end());
    ni_pow.insert(ni_pow.end(), M.ni[q].rop_pow.begin(), M.ni[q].rop_pow.end());
  };
  auto pack16 = [](int a, int b) { return (int)((unsigned)(a & 0xFFFF) | ((unsigned)(b & 0xFFFF) << 16)); };
  for (int q = 0; q < nrec; ++q) {
    const int j = order[q];
    const auto& r = M.rxns[j];
    const int s = M.fwd_step[j];
    double* c = &coef[(size_t)q * 6];
    c[0] = M.step_lnA[s];
    c[1] = M.step_beta[s];
    c[2] = M.step_Ta[s];
    int fl = 0, fsp = -1;
    if (M.rev_type[j] == kRevArrhenius) {
      fl |= kRecRevArr;
      const int sr = M.rev_step[j];
      c[3] = M.step_lnA[sr];
      c[4] = M.step_beta[sr];
      c[5] = M.step_Ta[sr];
    } else if (M.rev_type[j] == kRevKeq) {
      fl |= kRecRevKeq;
      c[3] = M.keq_ndelta[j];
    }
    if (r.third_body) fl |= kRecThirdBody;
    if (r.falloff) {
      fl |= kRecFalloff;
      if (M.falloff_type[j] == kTroe3 || M.falloff_type[j] == kTroe4) fl |= kRecTroe;
      if (M.falloff_type[j] == kTroe4) fl |= kRecTroe4;
      if (M.falloff_type[j] == kSri) fl |= kRecSri;
      for (int k = 0; k < kFoParams; ++k)
        fo[(size_t)q * kFoParams + k] = M.falloff_param[(size_t)j * kFoParams + k];
      fsp = M.falloff_spc[j];
    }
    const bool is_plog = M.rxn_plog[j] >= 0, is_ni = M.step_ni[s] >= 0;
    if (is_ext(j)) {
      int e4[4] = {-1, 0, -1, 0};
      if (is_plog) {
        fl |= kRecPlog;
        const PlogStep& ps = M.plog[M.rxn_plog[j]];
        const int np = (int)ps.p.size();
        e4[0] = (int)plog_t.size();
        plog_t.push_back((double)np);
        plog_t.insert(plog_t.end(), ps.p.begin(), ps.p.end());
        plog_t.insert(plog_t.end(), ps.lnp.begin(), ps.lnp.end());
        for (int k = 0; k <= np; ++k) plog_t.push_back((double)ps.line_ptr[k]);
        for (size_t k = 0; k < ps.lnA.size(); ++k) {
          plog_t.push_back(ps.lnA[k]);
          plog_t.push_back(ps.sign[k]);
          plog_t.push_back(ps.beta[k]);
          plog_t.push_back(ps.Ta[k]);
        }
      }
      if (is_ni) {
        fl |= kRecNonInt;
        int nf = 0, nr = 0;
        add_rop_list(s, &e4[1], &nf);
        if (M.rev_step[j] >= 0) add_rop_list(M.rev_step[j], &e4[2], &nr);
        if (nf > 0xFFFF || nr > 0x7FFF) {
          *err = "reaction " + std::to_string(j + 1) + " has too many reaction orders";
          return -1;
        }
        e4[3] = nf | (nr << 16);
      }
      fl |= (int)(ext.size() / 4) << kRecExtShift;
      ext.insert(ext.end(), e4, e4 + 4);
    }
    meta[(size_t)q * 4 + 0] = s;
    meta[(size_t)q * 4 + 1] = M.rev_step[j];
    meta[(size_t)q * 4 + 2] = fl;
    meta[(size_t)q * 4 + 3] = fsp;
    const int nr = M.reac_ptr[s + 1] - M.reac_ptr[s];
    const int np = M.prod_ptr[s + 1] - M.prod_ptr[s];
    if (nr > kMaxPart || np > kMaxPart) {
      *err = "reaction " + std::to_string(j + 1) + " has more than 4 expanded participants";
      return -1;
    }
    // absent participants: index nsp (the rate kernels read 1 and 0 there)
    int re[kMaxPart] = {nsp, nsp, nsp, nsp}, pr[kMaxPart] = {nsp, nsp, nsp, nsp};
    for (int k = 0; k < nr; ++k) re[k] = M.reac_sp[M.reac_ptr[s] + k];
    for (int k = 0; k < np; ++k) pr[k] = M.prod_sp[M.prod_ptr[s] + k];
    spk[(size_t)q * 4 + 0] = pack16(re[0], re[1]);
    spk[(size_t)q * 4 + 1] = pack16(re[2], re[3]);
    spk[(size_t)q * 4 + 2] = pack16(pr[0], pr[1]);
    spk[(size_t)q * 4 + 3] = pack16(pr[2], pr[3]);
    for (int k = M.tb_sp_ptr[j]; k < M.tb_sp_ptr[j + 1]; ++k) {
      esp.push_back(M.tb_sp[k]);
      eeff.push_back(M.tb_eff[k]);
    }
    while (esp.size() % 4) {  // groups of 4: species 0 with efficiency 0 adds +-0
      esp.push_back(0);
      eeff.push_back(0.0);
    }
    eptr[q + 1] = (int)esp.size();
  }
  if (esp.empty()) {  // one padding group: the rate loop always reads group 0
    esp.assign(4, 0);
    eeff.assign(4, 0.0);
  }
  // per-species gather lists in step order (reference summation order),
  // padded with the zero-rate slot nstep to a common length per 64-species block
  std::vector<std::vector<int>> dl(nsp), cl(nsp);
  for (int s = 0; s < nstep; ++s) {
    for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) dl[M.reac_sp[k]].push_back(s);
    for (int k = M.prod_ptr[s]; k < M.prod_ptr[s + 1]; ++k) cl[M.prod_sp[k]].push_back(s);
  }
  size_t maxd = 0, maxc = 0;
  for (int j = 0; j < nsp; ++j) {
    maxd = std::max(maxd, dl[j].size());
    maxc = std::max(maxc, cl[j].size());
  }
  const int nsb = (nsp + 63) / 64;
  const int gdb = (int)((maxd + 7) / 8), gcb = (int)((maxc + 7) / 8);
  auto pad_gather = [&](const std::vector<std::vector<int>>& lists, int nb) {
    std::vector<uint16_t> g((size_t)nsb * nb * 64 * 8, (uint16_t)nstep);
    for (int j = 0; j < nsp; ++j)
      for (size_t k = 0; k < lists[j].size(); ++k) {
        const size_t blk = (size_t)(j / 64) * nb + k / 8;
        g[(blk * 64 + (j % 64)) * 8 + k % 8] = (uint16_t)lists[j][k];
      }
    return g;
  };
  const std::vector<uint16_t> gd = pad_gather(dl, gdb), gc = pad_gather(cl, gcb);
  // the blocks each 64-species group actually uses ({destroy, create}): a
  // wavefront's gather loop stops there (the padding beyond adds the zero
  // rate, so stopping early changes no sum); the common length is set by the
  // most connected species of all, 15 to 60 times the typical one
  std::vector<int32_t> gnb(2 * (size_t)nsb, 0);
  for (int j = 0; j < nsp; ++j) {
    int32_t& bd = gnb[2 * (size_t)(j / 64)];
    int32_t& bc = gnb[2 * (size_t)(j / 64) + 1];
    bd = std::max(bd, (int32_t)((dl[j].size() + 7) / 8));
    bc = std::max(bc, (int32_t)((cl[j].size() + 7) / 8));
  }
  // non-integer destruction/creation terms by species, in step order and
  // within a step in reaction-text order (non_integer_reaction_network.cpp:
  // 295-333; perf_net.cpp:250-253 adds them before the integer terms)
  std::vector<std::vector<std::pair<double, int>>> nd(nsp), nc(nsp);
  for (const auto& st : M.ni) {
    for (size_t k = 0; k < st.re_sp.size(); ++k) nd[st.re_sp[k]].push_back({st.re_nu[k], st.step});
    for (size_t k = 0; k < st.pr_sp.size(); ++k) nc[st.pr_sp[k]].push_back({st.pr_nu[k], st.step});
  }
  std::vector<int> nig_dptr(nsp + 1, 0), nig_cptr(nsp + 1, 0);
  std::vector<double> nig;  // pairs {nu, step}
  for (int j = 0; j < nsp; ++j) {
    for (auto& e : nd[j]) {
      nig.push_back(e.first);
      nig.push_back((double)e.second);
    }
    nig_dptr[j + 1] = (int)(nig.size() / 2);
  }
  nig_cptr[0] = nig_dptr[nsp];  // creation entries follow the destruction entries
  for (int j = 0; j < nsp; ++j) {
    for (auto& e : nc[j]) {
      nig.push_back(e.first);
      nig.push_back((double)e.second);
    }
    nig_cptr[j + 1] = (int)(nig.size() / 2);
  }
  // Jacobian species-block terms by row (integrator teams of up to 3
  // wavefronts, nsp <= 191; row stride = team size):
  // entry (row i, col c) = destruction terms then creation terms, each in step
  // order (the reference's order of accumulation, reactor_nvector_serial.cpp:149-176
  // and 655-669); a row's entries are sorted by column
  std::vector<uint32_t> jrt;
  std::vector<double> ni_jt;  // pairs {multiplier, step}
  int jrb = 0;
  const int team = 64 * ((nsp + 1 + 63) / 64);
  if (nsp <= 191) {
    std::vector<std::vector<int>> jd((size_t)nsp * nsp), jc((size_t)nsp * nsp);
    for (int s = 0; s < nstep; ++s)
      for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) {
        const int c = M.reac_sp[k];
        for (int q = M.reac_ptr[s]; q < M.reac_ptr[s + 1]; ++q)
          jd[(size_t)M.reac_sp[q] * nsp + c].push_back(s);
        for (int q = M.prod_ptr[s]; q < M.prod_ptr[s + 1]; ++q)
          jc[(size_t)M.prod_sp[q] * nsp + c].push_back(s);
      }
    // non-integer terms (non_integer_reaction_network.cpp:774-905): for each
    // step, each exponent species c with a non-zero exponent p, the rows of
    // its reactants (-nu p) and products (+nu p); value (mult ROP) / C_c
    std::vector<std::vector<int>> jn((size_t)nsp * nsp);
    for (const auto& st : M.ni)
      for (size_t k = 0; k < st.rop_sp.size(); ++k) {
        if (st.rop_pow[k] == 0.0) continue;
        const int c = st.rop_sp[k];
        for (size_t q = 0; q < st.re_sp.size(); ++q) {
          jn[(size_t)st.re_sp[q] * nsp + c].push_back((int)(ni_jt.size() / 2));
          ni_jt.push_back(-st.re_nu[q] * st.rop_pow[k]);
          ni_jt.push_back((double)st.step);
        }
        for (size_t q = 0; q < st.pr_sp.size(); ++q) {
          jn[(size_t)st.pr_sp[q] * nsp + c].push_back((int)(ni_jt.size() / 2));
          ni_jt.push_back(st.pr_nu[q] * st.rop_pow[k]);
          ni_jt.push_back((double)st.step);
        }
      }
    if (ni_jt.size() / 2 > 0xFFFF) {
      *err = "too many non-integer Jacobian terms";
      return -1;
    }
    std::vector<std::vector<uint32_t>> rows(nsp);
    size_t maxr = 0;
    for (int i = 0; i < nsp; ++i) {
      for (int c = 0; c < nsp; ++c) {
        const size_t e = (size_t)i * nsp + c;
        for (int st : jd[e]) rows[i].push_back((uint32_t)st | ((uint32_t)c << 16) | 0x80000000u);
        for (int st : jc[e]) rows[i].push_back((uint32_t)st | ((uint32_t)c << 16));
        for (int t : jn[e]) rows[i].push_back((uint32_t)t | ((uint32_t)c << 16) | kJtNonInt);
      }
      maxr = std::max(maxr, rows[i].size());
    }
    jrb = (int)((maxr + 3) / 4);
    const uint32_t pad = (uint32_t)nstep | (0xFFu << 16);
    jrt.assign((size_t)jrb * team * 4, pad);
    for (int i = 0; i < nsp; ++i)
      for (size_t k = 0; k < rows[i].size(); ++k)
        jrt[((k / 4) * team + i) * 4 + k % 4] = rows[i][k];
  }
  if (jrt.empty()) jrt.assign(4, 0u);
  // the Jacobian's pattern by row, 3 x 64 bits (reactor_nvector_serial.cpp:122-200)
  std::vector<unsigned long long> jpat;
  if (nsp <= 191) {
    const int N = nsp + 1;
    jpat.assign((size_t)N * 3, 0ull);
    auto mark = [&](int row, int col) { jpat[(size_t)row * 3 + col / 64] |= 1ull << (col % 64); };
    for (int k = 0; k < N; ++k) {
      mark(k, k);
      mark(nsp, k);
      mark(k, nsp);
    }
    for (int s = 0; s < nstep; ++s)
      for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) {
        const int c = M.reac_sp[k];
        for (int q = M.reac_ptr[s]; q < M.reac_ptr[s + 1]; ++q) mark(M.reac_sp[q], c);
        for (int q = M.prod_ptr[s]; q < M.prod_ptr[s + 1]; ++q) mark(M.prod_sp[q], c);
      }
    for (const auto& st : M.ni)
      for (size_t k = 0; k < st.rop_sp.size(); ++k) {
        const int c = st.rop_sp[k];
        for (int r : st.re_sp) mark(r, c);
        for (int r : st.pr_sp) mark(r, c);
      }
  }
  if (jpat.empty()) jpat.assign(3, 0ull);
  const int n_ext = (int)(ext.size() / 4);
  const bool has_nig = !nig.empty();
  // keep every table non-empty so that each device pointer is valid
  if (ext.empty()) ext.assign(4, -1);
  if (plog_t.empty()) plog_t.assign(1, 0.0);
  if (ni_sp.empty()) {
    ni_sp.assign(1, 0);
    ni_pow.assign(1, 0.0);
  }
  if (nig.empty()) nig.assign(2, 0.0);
  if (ni_jt.empty()) ni_jt.assign(2, 0.0);

  BlobBuilder b;
  const size_t o_mw = b.add(M.mw), o_imw = b.add(inv_mw), o_rmw = b.add(ru_mw),
               o_th = b.add(M.thermo), o_meta = b.add(meta), o_sp = b.add(spk),
               o_coef = b.add(coef), o_ep = b.add(eptr), o_es = b.add(esp), o_ee = b.add(eeff),
               o_fo = b.add(fo), o_gd = b.add(gd), o_gc = b.add(gc), o_jr = b.add(jrt),
               o_ext = b.add(ext), o_pl = b.add(plog_t), o_nsp = b.add(ni_sp),
               o_npw = b.add(ni_pow), o_ndp = b.add(nig_dptr), o_ncp = b.add(nig_cptr),
               o_nig = b.add(nig), o_njt = b.add(ni_jt), o_jpat = b.add(jpat),
               o_gnb = b.add(gnb);
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
  d.gd_blocks = gdb;
  d.gc_blocks = gcb;
  d.jr_blocks = jrb;
  d.jr_stride = team;
  d.mw = (const double*)(p + o_mw);
  d.inv_mw = (const double*)(p + o_imw);
  d.ru_mw = (const double*)(p + o_rmw);
  d.thermo = (const double*)(p + o_th);
  d.r_meta = (const int4*)(p + o_meta);
  d.r_sp = (const int4*)(p + o_sp);
  d.r_coef = (const double2*)(p + o_coef);
  d.r_eptr = (const int*)(p + o_ep);
  d.e_sp = (const int*)(p + o_es);
  d.e_eff = (const double*)(p + o_ee);
  d.r_fo = (const double*)(p + o_fo);
  d.g_d = (const uint4*)(p + o_gd);
  d.g_c = (const uint4*)(p + o_gc);
  d.g_nb = (const int2*)(p + o_gnb);
  d.jr = (const uint4*)(p + o_jr);
  d.n_ext = n_ext;
  d.r_ext = (const int4*)(p + o_ext);
  d.plog = (const double*)(p + o_pl);
  d.ni_sp = (const int*)(p + o_nsp);
  d.ni_pow = (const double*)(p + o_npw);
  d.has_nig = has_nig ? 1 : 0;
  d.nig_dptr = (const int*)(p + o_ndp);
  d.nig_cptr = (const int*)(p + o_ncp);
  d.nig = (const double2*)(p + o_nig);
  d.ni_jt = (const double2*)(p + o_njt);
  d.jpat = (const unsigned long long*)(p + o_jpat);
  e = hipMalloc((void**)&h->dev_self, sizeof(DevMech));
  if (e == hipSuccess) {
    d.self = h->dev_self;
    e = hipMemcpy(h->dev_self, &d, sizeof(DevMech), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    (void)hipFree(base);
    if (h->dev_self) (void)hipFree(h->dev_self);
    h->dev_self = nullptr;
    *err = std::string("device mechanism copy: ") + hipGetErrorString(e);
    return (int)e;
  }
  h->dev_blob = base;
  h->dev_bytes = b.bytes.size();
  if (hipGetDevice(&h->device) != hipSuccess) h->device = 0;
  return 0;
}

// Integrator workspace (per handle): saved Jacobians per resident wave slot
// and the work counter.  Grown on demand, released with the handle.
int ensure_workspace(MechHandle* h, size_t doubles) {
  if (h->ws_doubles >= doubles && h->ws_jac) return 0;
  if (h->ws_jac) (void)hipFree(h->ws_jac);
  if (h->ws_counter) (void)hipFree(h->ws_counter);
  h->ws_jac = nullptr;
  h->ws_counter = nullptr;
  h->ws_doubles = 0;
  hipError_t e = hipMalloc((void**)&h->ws_jac, sizeof(double) * doubles);
  if (e != hipSuccess) return (int)e;
  e = hipMalloc((void**)&h->ws_counter, 256);
  if (e != hipSuccess) return (int)e;
  h->ws_doubles = doubles;
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
  b.constant_volume = o.constant_volume != 0;
  b.fp32_preconditioner = o.fp32_preconditioner != 0;
  b.ign_dT = o.delta_temperature_ignition > 0.0 ? o.delta_temperature_ignition : 0.0;
  b.stop_after_ignition = o.stop_after_ignition != 0;
  b.abstol_dens = o.abstol_dens != 0;
  b.krylov = !(o.dense == 1 && o.iterative == 0);
  b.jac = o.analytic ? kJacAnalytic : (b.krylov ? kJacReactorDQ : kJacCvodeDQ);
  // the preconditioner of the Krylov solver: the Jacobian's pattern with the
  // threshold (0 without iterative), or every entry for dense divided
  // differences (JacobianFactor, reactor_nvector_serial.cpp:373-397)
  b.pthr = (o.iterative ? o.preconditioner_threshold : 0.0);
  b.pfull = !b.krylov || (o.dense == 1 && !o.analytic);
  b.eps_lin = o.eps_lin;
  return b;
}

MechHandle::~MechHandle() {
  delete plan;
  if (sp_blob) (void)hipFree(sp_blob);
  if (ws_jac) (void)hipFree(ws_jac);
  if (ws_counter) (void)hipFree(ws_counter);
  if (dev_blob) (void)hipFree(dev_blob);
  if (dev_self) (void)hipFree(dev_self);
}

int ensure_device(MechHandle* h) {
  std::string msg;
  const int rc = build_device_mech(h, &msg);
  if (rc) std::fprintf(stderr, "zerork-hip: device mechanism upload failed: %s\n", msg.c_str());
  return rc;
}

int ensure_sparse(MechHandle* h) {
  if (h->sp_blob) return 0;
  if (int rc = ensure_device(h)) return rc;
  std::string err;
  SparsePlan* P = new SparsePlan();
  if (!build_sparse_plan(h->host, P, &err)) {
    std::fprintf(stderr, "zerork-hip: sparse plan failed: %s\n", err.c_str());
    delete P;
    return -5;
  }
  // one device allocation: int arrays, then the term list
  std::vector<const std::vector<int>*> parts = {&P->perm, &P->rp, &P->ci, &P->dpos, &P->lf_ptr,
                                                &P->lf_rows, &P->lb_ptr, &P->lb_rows,
                                                &P->jt_ptr, &P->pos_row, &P->pos_col,
                                                &P->cp_ptr, &P->cp_pos, &P->trow_pos,
                                                &P->tcol_pos, &P->ll_ptr, &P->ll_rows,
                                                &P->lbl_ptr, &P->lbl_rows};
  std::vector<size_t> off;
  size_t n = 0;
  for (auto* v : parts) {
    off.push_back(n);
    n += (v->size() + 63) & ~(size_t)63;
  }
  const size_t jt_off = n;
  n += P->jt.size();
  std::vector<int> blob(n, 0);
  for (size_t k = 0; k < parts.size(); ++k)
    std::copy(parts[k]->begin(), parts[k]->end(), blob.begin() + off[k]);
  std::memcpy(blob.data() + jt_off, P->jt.data(), sizeof(uint32_t) * P->jt.size());
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(int) * n);
  if (e != hipSuccess) {
    delete P;
    return (int)e;
  }
  e = hipMemcpy(d, blob.data(), sizeof(int) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    delete P;
    return (int)e;
  }
  const int* b = static_cast<const int*>(d);
  DevSparse& S = h->sp_dev;
  S.N = P->N;
  S.nnz = P->nnz;
  S.nlev_f = P->nlev_f;
  S.nlev_b = P->nlev_b;
  S.max_row = P->max_row;
  S.perm = b + off[0];
  S.rp = b + off[1];
  S.ci = b + off[2];
  S.dpos = b + off[3];
  S.lf_ptr = b + off[4];
  S.lf_rows = b + off[5];
  S.lb_ptr = b + off[6];
  S.lb_rows = b + off[7];
  S.jt_ptr = b + off[8];
  S.pos_row = b + off[9];
  S.pos_col = b + off[10];
  S.cp_ptr = b + off[11];
  S.cp_pos = b + off[12];
  S.trow_pos = b + off[13];
  S.tcol_pos = b + off[14];
  S.ll_ptr = b + off[15];
  S.ll_rows = b + off[16];
  S.lbl_ptr = b + off[17];
  S.lbl_rows = b + off[18];
  S.nlev_lb = P->nlev_lb;
  S.split = P->split;
  S.nlev_l = P->nlev_l;
  S.jt = reinterpret_cast<const uint32_t*>(b + jt_off);
  h->sp_blob = d;
  h->plan = P;
  h->sp_resident = sparse_max_resident(h->dev);
  return h->sp_resident > 0 ? 0 : -2;
}

namespace {

// the sparse path: one reactor per workgroup, FP32 sparse LU (see bdf_sparse.hip)
int integrate_sparse(MechHandle* h, const zrk_bdf_options& o, int n, double dt,
                     const double* dts, double* T, double* P, double* Y, int stride,
                     int* nsteps, int* stats, const int* order, hipStream_t st,
                     const double* dpdt, const double* esrc, const double* ysrc,
                     const StepTrace* trace) {
  if (!o.fp32_preconditioner) return -3;
  // source terms, the ignition root and PLOG / non-integer / SRI records are
  // not on this path
  if (dpdt || esrc || ysrc || o.delta_temperature_ignition > 0.0 || h->dev.n_ext > 0) return -4;
  if (n == 0) return 0;
  if (int rc = ensure_sparse(h)) return rc;
  // the dense trailing blocks of the reactors in flight are kept within half
  // of the 256 MB last-level cache: the block's right-looking factorisation
  // re-reads it O(nd) times, so spilling them to HBM costs more than the
  // concurrency gains
  const double nd = (double)(h->sp_dev.N - h->sp_dev.split);
  const int cap = nd > 0.0 ? std::max(1, (int)(128.0 * 1048576.0 / (4.0 * nd * nd))) : n;
  const int nb = std::max(1, std::min(std::min(h->sp_resident, cap), n));
  const size_t per = sparse_slot_bytes(h->dev, h->sp_dev) / sizeof(double);
  if (int rc = ensure_workspace(h, (size_t)nb * per)) return rc;
  BdfOpts bo = to_bdf_opts(o);
  if (trace) {
    bo.trace = trace->buf;
    bo.trace_cap = trace->cap;
    bo.trace_n = trace->count;
  }
  // bounded launches: each one integrates at most four rounds of the resident
  // workgroups, so a large batch of a large mechanism is a few launches of
  // seconds rather than one of minutes (reactors are independent; with an
  // order array the chunks are consecutive ranges of it)
  const int chunk = 4 * nb;
  for (int off = 0; off < n; off += chunk) {
    const int k = std::min(chunk, n - off);
    const hipError_t e =
        order ? launch_bdf_sparse(h->dev, h->sp_dev, bo, k, dt, dts, T, P, Y, stride, nsteps,
                                  stats, order + off, h->ws_counter, h->ws_jac, nb, st)
              : launch_bdf_sparse(h->dev, h->sp_dev, bo, k, dt, dts ? dts + off : nullptr,
                                  T + off, P + off, Y + (size_t)off * stride, stride,
                                  nsteps ? nsteps + off : nullptr,
                                  stats ? stats + (size_t)off * 8 : nullptr, nullptr,
                                  h->ws_counter, h->ws_jac, nb, st);
    if (e != hipSuccess) return (int)e;
  }
  return 0;
}

}  // namespace

int team_width_for(MechHandle* h, bool f32, bool iter) {
  int W = h->team_pref[f32] ? h->team_pref[f32] : bdf_team_width(h->host.nsp);
  if (!bdf_team_ok(h->host.nsp, f32, W)) return 0;
  int& res = h->resident[f32][iter][W];
  if (res <= 0) res = bdf_max_resident_teams(h->dev, f32, W, iter);
  return res > 0 ? W : 0;
}

int integrate_device(MechHandle* h, const zrk_bdf_options& o, int n, double dt,
                     const double* dts, double* T, double* P, double* Y, int stride, int* nsteps, int* stats,
                     const int* order, hipStream_t st, const double* dpdt, const double* esrc,
                     const double* ysrc, double* ign_time, const double* tdelta, int always_T,
                     const ExtWorkspace* ws, const StepTrace* trace) {
  if (n < 0 || !T || !P || !Y || stride < h->host.nsp || (!dts && !(dt > 0.0))) return -1;
  if (trace && (n != 1 || !trace->buf || !trace->count || trace->cap < 0)) return -1;
  BdfOpts bo = to_bdf_opts(o);
  const int W0 = bdf_team_width(h->host.nsp);
  if (ws && !h->dev_blob) return -6;  // no upload here: nothing may be allocated
  if (W0 == 0 || h->force_sparse) {  // above 191 species: the sparse path
    if (ws) return -4;
    // the sparse path has the direct analytic solver with solved temperature
    // and plain absolute tolerances only
    if (bo.krylov || bo.jac != kJacAnalytic || bo.abstol_dens || (!always_T && tdelta))
      return -4;
    if (n > 0)
      if (int rc = ensure_device(h)) return rc;
    return integrate_sparse(h, o, n, dt, dts, T, P, Y, stride, nsteps, stats, order, st, dpdt,
                            esrc, ysrc, trace);
  }
  if (n == 0) return 0;
  if (int rc = ensure_device(h)) return rc;
  const int f32 = o.fp32_preconditioner != 0;
  const int it = bdf_alt_modes(bo) ? 1 : 0;
  const int W = team_width_for(h, f32, it);
  if (W == 0) return -2;                                // does not fit in LDS
  const size_t slot = bdf_slot_doubles(h->dev, W, f32);
  int nw = std::max(1, std::min(h->resident[f32][it][W], n));
  double* jac = nullptr;
  unsigned int* counter = nullptr;
  if (ws) {  // the caller's scratch: no allocation
    if (!ws->jac || !ws->counter || ws->doubles < slot) return -7;
    nw = std::min(nw, (int)std::min<size_t>(ws->doubles / slot, (size_t)INT32_MAX));
    jac = ws->jac;
    counter = ws->counter;
  } else {
    if (int rc = ensure_workspace(h, (size_t)nw * slot)) return rc;
    jac = h->ws_jac;
    counter = h->ws_counter;
  }
  bo.dpdt = dpdt;
  bo.esrc = esrc;
  bo.ysrc = ysrc;
  bo.ign_time = bo.ign_dT > 0.0 ? ign_time : nullptr;
  bo.always_T = always_T ? 1 : 0;
  bo.tdelta = always_T ? nullptr : tdelta;
  if (trace) {
    bo.trace = trace->buf;
    bo.trace_cap = trace->cap;
    bo.trace_n = trace->count;
  }
  hipError_t e = launch_bdf(h->dev, bo, n, dt, dts, T, P, Y, stride, nsteps, stats, order,
                            counter, jac, nw, W, st);
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
  // the plugin defaults to constant volume (constant_volume=1); this
  // device-level default is the constant-pressure reactor
  o->constant_volume = 0;
  o->fp32_preconditioner = 0;
  o->delta_temperature_ignition = 0.0;
  o->stop_after_ignition = 0;
  o->abstol_dens = 0;
  o->dense = 1;
  o->analytic = 1;
  o->iterative = 0;
  o->eps_lin = 1.0e-3;
  o->preconditioner_threshold = 1.0e-3;
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
  double* scratch = nullptr;
  int grid = 0;
  if (n > 0 && zrk::rates_need_block(h->dev)) {  // large mechanisms: one reactor per workgroup
    grid = zrk::rates_block_grid(h->dev, n);
    if (grid <= 0) return -2;
    if (int rc = zrk::ensure_workspace(h, (size_t)grid * (h->host.nstep + 1))) return rc;
    scratch = h->ws_jac;
  }
  hipError_t e = zrk::launch_rates(h->dev, n, T, P, Y, stride, step_limiter, ydot, rop, create,
                                   destroy, net, (hipStream_t)stream, scratch, grid);
  return e == hipSuccess ? 0 : (int)e;
}

int zrk_jacobian_device(zrk_mech m, const zrk_bdf_options* o, int n, const double* T,
                        const double* P, const double* Y, int stride, double* J, void* stream) {
  if (!m || !o || n < 0 || !T || !P || !Y || !J) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (stride < h->host.nsp) return -1;
  if (zrk::bdf_team_width(h->host.nsp) == 0) return -2;
  if (int rc = zrk::ensure_device(h)) return rc;
  const size_t N = (size_t)h->host.nsp + 1;
  const int nb = std::max(1, std::min(n, N > (size_t)zrk::kNP ? 512 : 4096));
  const size_t per = std::max((size_t)zrk::kNP * zrk::kNP, 2 * N * N);
  int rc = zrk::ensure_workspace(h, (size_t)nb * per);
  if (rc) return rc;
  hipError_t e = zrk::launch_jacobian(h->dev, zrk::to_bdf_opts(*o), n, T, P, Y, stride, J,
                                      h->ws_jac, nb, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (int)e;
}

int zrk_integrate_device(zrk_mech m, const zrk_bdf_options* o, int n, double dt, double* T,
                         double* P, double* Y, int stride, int* nsteps, int* stats,
                         void* stream) {
  if (!m || !o) return -1;
  return zrk::integrate_device(reinterpret_cast<MechHandle*>(m), *o, n, dt, nullptr, T, P, Y,
                               stride, nsteps, stats, nullptr, (hipStream_t)stream);
}

int zrk_integrate_device_aux(zrk_mech m, const zrk_bdf_options* o, int n, double dt, double* T,
                             double* P, double* Y, int stride, const double* dpdt,
                             const double* esrc, const double* ysrc, double* ignition_time,
                             int* nsteps, int* stats, void* stream) {
  if (!m || !o) return -1;
  return zrk::integrate_device(reinterpret_cast<MechHandle*>(m), *o, n, dt, nullptr, T, P, Y,
                               stride, nsteps, stats, nullptr, (hipStream_t)stream, dpdt, esrc,
                               ysrc, ignition_time);
}

int zrk_autotune_integrator(zrk_mech m, const zrk_bdf_options* o, int n, double dt,
                            const double* T, const double* P, const double* Y, int stride,
                            int* choice, float* ms, void* stream) {
  if (!m || !o || n <= 0 || !T || !P || !Y || !(dt > 0.0)) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  const int nsp = h->host.nsp;
  if (stride < nsp) return -1;
  if (int rc = zrk::ensure_device(h)) return rc;
  const bool f32 = o->fp32_preconditioner != 0;
  hipStream_t st = (hipStream_t)stream;
  double *dT = nullptr, *dP = nullptr, *dY = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const size_t ny = (size_t)n * stride;
  int rc = 0, best = 0;
  float best_ms = 0.0f;
  auto hrc = [&](hipError_t e) { if (e != hipSuccess && rc == 0) rc = (int)e; return e == hipSuccess; };
  if (hrc(hipMalloc((void**)&dT, sizeof(double) * n)) && hrc(hipMalloc((void**)&dP, sizeof(double) * n)) &&
      hrc(hipMalloc((void**)&dY, sizeof(double) * ny)) && hrc(hipEventCreate(&e0)) &&
      hrc(hipEventCreate(&e1))) {
    const int saved = h->team_pref[f32];
    for (int W = 1; W <= 3 && rc == 0; ++W) {
      if (ms) ms[W] = -1.0f;
      if (!zrk::bdf_team_ok(nsp, f32, W)) continue;
      h->team_pref[f32] = W;
      if (zrk::team_width_for(h, f32) != W) continue;  // does not fit
      float t = 0.0f;
      for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // first run warms up
        hrc(hipMemcpyAsync(dT, T, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        hrc(hipMemcpyAsync(dP, P, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        hrc(hipMemcpyAsync(dY, Y, sizeof(double) * ny, hipMemcpyDeviceToDevice, st));
        hrc(hipEventRecord(e0, st));
        if (int r = zrk::integrate_device(h, *o, n, dt, nullptr, dT, dP, dY, stride, nullptr,
                                          nullptr, nullptr, st))
          if (rc == 0) rc = r;
        hrc(hipEventRecord(e1, st));
        hrc(hipEventSynchronize(e1));
        hrc(hipEventElapsedTime(&t, e0, e1));
      }
      if (rc) break;
      if (ms) ms[W] = t;
      if (best == 0 || t < best_ms) {
        best = W;
        best_ms = t;
      }
    }
    h->team_pref[f32] = rc ? saved : best;
  }
  if (dT) (void)hipFree(dT);
  if (dP) (void)hipFree(dP);
  if (dY) (void)hipFree(dY);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (choice) *choice = best;
  return rc;
}

int zrk_integrator_set_sparse(zrk_mech m, int on) {
  if (!m) return -1;
  reinterpret_cast<MechHandle*>(m)->force_sparse = on != 0;
  return 0;
}

int zrk_sparse_plan_info(zrk_mech m, int* nnz, int* levels, long long* factor_flops) {
  if (!m) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (int rc = zrk::ensure_sparse(h)) return rc;
  if (nnz) *nnz = h->plan->nnz;
  if (levels) *levels = h->plan->nlev_f;
  if (factor_flops) *factor_flops = h->plan->factor_flops;
  return 0;
}

int zrk_integrator_set_team_width(zrk_mech m, int fp32_preconditioner, int W) {
  if (!m) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  const bool f32 = fp32_preconditioner != 0;
  if (W != 0 && !zrk::bdf_team_ok(h->host.nsp, f32, W)) return -1;
  h->team_pref[f32] = W;
  return 0;
}

int zrk_integrator_resident_waves(zrk_mech m, const zrk_bdf_options* o) {
  if (!m) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (int rc = zrk::ensure_device(h)) return -rc;
  const int f32 = (o && o->fp32_preconditioner) ? 1 : 0;
  if (zrk::bdf_team_width(h->host.nsp) == 0 || h->force_sparse) {
    if (!f32 || zrk::ensure_sparse(h)) return 0;
    return h->sp_resident * zrk::kSparseWaves;
  }
  const int it = (o && zrk::bdf_alt_modes(zrk::to_bdf_opts(*o))) ? 1 : 0;
  const int W = zrk::team_width_for(h, f32, it);
  return W ? h->resident[f32][it][W] * W : 0;
}

}  // extern "C"
