// C ABI of the learned-surrogate path (include/zrk_surrogate.h).
#include <algorithm>
#include <cmath>

#include "../../include/zrk_surrogate.h"
#include "launch.h"
#include "zrk_internal.h"

using zrk::MechHandle;

namespace {
inline int rc_of(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }
}  // namespace

extern "C" {

int zrk_integrate_device_select(zrk_mech m, const zrk_bdf_options* o, int n_sel, const int* sel,
                                const double* dt, double* T, double* P, double* Y, int stride,
                                int* nsteps, int* stats, void* stream) {
  if (!m || !o || !sel || !dt || n_sel < 0) return -1;
  return zrk::integrate_device(reinterpret_cast<MechHandle*>(m), *o, n_sel, 0.0, dt, T, P, Y,
                               stride, nsteps, stats, sel, (hipStream_t)stream);
}

int zrk_mlp_features(zrk_mech m, int n, const double* T, const double* P, const double* Y,
                     int stride, const double* dt, double dt_all, const float* mu,
                     const float* sig, void* X, int ldx, void* stream) {
  if (!m || n < 0 || !T || !P || !Y || !mu || !sig || !X) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  const int nsp = h->host.nsp, nfeat = nsp + 3;
  if (stride < nsp || ldx < nfeat || (!dt && !(dt_all > 0.0))) return -1;
  return rc_of(zrk::launch_mlp_features(n, nsp, T, P, Y, stride, dt, dt_all, mu, sig,
                                        static_cast<unsigned short*>(X), ldx, nfeat,
                                        (hipStream_t)stream));
}

int zrk_gemm_bf16(int M, int N, int K, const void* A, long sAm, long sAk, const void* B, long sBk,
                  long sBn, float* C, long sCm, long sCn, int accumulate, void* stream) {
  if (M < 0 || N < 0 || K < 0 || !A || !B || !C) return -1;
  return rc_of(zrk::launch_gemm_bf16(M, N, K, static_cast<const unsigned short*>(A), sAm, sAk,
                                     static_cast<const unsigned short*>(B), sBk, sBn, C, sCm,
                                     sCn, accumulate, (hipStream_t)stream));
}

int zrk_mlp_bias_act(int M, int N, const float* C, const float* b, float* Z, void* Yb, float* Yf,
                     int act, void* stream) {
  if (M < 0 || N < 0 || !C || !b) return -1;
  return rc_of(zrk::launch_bias_act(M, N, C, b, Z, static_cast<unsigned short*>(Yb), Yf, act,
                                    (hipStream_t)stream));
}

int zrk_mlp_act_backward(int M, int N, const float* dY, const float* Z, float* dZ, void* dZb,
                         int act, void* stream) {
  if (M < 0 || N < 0 || !dY || (act && !Z) || !dZ || !dZb) return -1;
  return rc_of(zrk::launch_act_backward(M, N, dY, Z, dZ, static_cast<unsigned short*>(dZb), act,
                                        (hipStream_t)stream));
}

int zrk_mlp_col_sum(int M, int N, const float* dZ, float* db, void* stream) {
  if (M < 0 || N < 0 || !dZ || !db) return -1;
  return rc_of(zrk::launch_col_sum(M, N, dZ, db, (hipStream_t)stream));
}

int zrk_mlp_adam(long n, float* w, const float* g, float* m, float* v, void* wb, float lr,
                 float beta1, float beta2, float eps, int step, void* stream) {
  if (n < 0 || !w || !g || !m || !v || step < 1) return -1;
  const float c1 = 1.0f - std::pow(beta1, (float)step);
  const float c2 = 1.0f - std::pow(beta2, (float)step);
  return rc_of(zrk::launch_adam(n, w, g, m, v, static_cast<unsigned short*>(wb), lr, beta1, beta2,
                                eps, c1, c2, (hipStream_t)stream));
}

int zrk_to_bf16(long n, const float* x, void* y, void* stream) {
  if (n < 0 || !x || !y) return -1;
  return rc_of(zrk::launch_to_bf16(n, x, static_cast<unsigned short*>(y), (hipStream_t)stream));
}

int zrk_surrogate_accept(zrk_mech m, const zrk_bdf_options* o, int n, const double* dt,
                         double dt_all, const double* T0, const double* P, const double* Y0,
                         int nodes, const double* tau, double* Tp, double* Yp, int stride,
                         int max_newton, int* accept, double* res, void* stream) {
  if (!m || !o || n < 0 || !T0 || !P || !Y0 || !tau || !Tp || !Yp || !accept) return -1;
  if (!dt && !(dt_all > 0.0)) return -1;
  if (nodes < 3 || nodes > 8 || max_newton < 1) return -1;
  auto* h = reinterpret_cast<MechHandle*>(m);
  if (stride < h->host.nsp) return -1;
  if (h->host.nsp > zrk::kNP - 1) return -2;
  // the test integrates the constant-pressure reactor without source terms
  if (o->constant_volume) return -4;
  if (int rc = zrk::ensure_device(h)) return rc;
  const int nb = std::max(1, std::min(n, 2048));
  if (int rc = zrk::ensure_workspace(h, (size_t)nb * zrk::kNP * zrk::kNP)) return rc;
  return rc_of(zrk::launch_accept(h->dev, zrk::to_bdf_opts(*o), n, dt, dt_all, T0, P, Y0, nodes,
                                  tau, Tp, Yp, stride, max_newton, accept, res, h->ws_jac, nb,
                                  (hipStream_t)stream));
}

}  // extern "C"
