/* zrk_surrogate.h - C ABI of the learned-surrogate path on MI355X.
 *
 * None of these entry points has a counterpart in the reference (the
 * reference integrates every reactor with CVODE).  They implement a small
 * MLP that proposes a reactor's end state, trained and evaluated with
 * BF16-input / FP32-accumulate MFMA kernels, an FP64 acceptance test of each
 * proposed trajectory, and the plain BDF integration of the rejected
 * reactors (each with its own dt).  Device pointers and a HIP stream, as in zrk_device.h.
 * bf16 arrays are passed as void* (16-bit elements).  Return value: 0, a
 * negative code on invalid arguments, or a positive hipError_t.
 */
#ifndef ZRK_SURROGATE_H
#define ZRK_SURROGATE_H

#include "zrk_device.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Integrate the n_sel reactors listed in sel[] (indices into T/P/Y), reactor
 * r over [0, dt[r]]; otherwise as zrk_integrate_device. */
ZRK_EXPORT int zrk_integrate_device_select(zrk_mech m, const zrk_bdf_options* o, int n_sel,
                                           const int* sel, const double* dt, double* T,
                                           double* P, double* Y, int stride, int* nsteps,
                                           int* stats, void* stream);

/* X[r*ldx + f] (bf16) = (x_f - mu[f]) / sig[f] for the features
 * x = [ln T, ln P, ln max(Y_j, 1e-20) (nsp), ln dt] (nfeat = nsp + 3), zero
 * for f >= nfeat.  dt: per-reactor array or NULL for dt_all. */
ZRK_EXPORT int zrk_mlp_features(zrk_mech m, int n, const double* T, const double* P,
                                const double* Y, int stride, const double* dt, double dt_all,
                                const float* mu, const float* sig, void* X, int ldx,
                                void* stream);

/* C[M x N] (f32) = A . B (accumulate = 0) or C += A . B; A[i,k] at
 * A[i*sAm + k*sAk], B[k,j] at B[k*sBk + j*sBn], C[i,j] at C[i*sCm + j*sCn]. */
ZRK_EXPORT int zrk_gemm_bf16(int M, int N, int K, const void* A, long sAm, long sAk,
                             const void* B, long sBk, long sBn, float* C, long sCm, long sCn,
                             int accumulate, void* stream);

/* Z = C + b (row-broadcast bias), Y = gelu(Z) (act = 1) or Z (act = 0);
 * Z, Yb (bf16) and Yf (f32) each optional.  Row-major M x N. */
ZRK_EXPORT int zrk_mlp_bias_act(int M, int N, const float* C, const float* b, float* Z, void* Yb,
                                float* Yf, int act, void* stream);

/* dZ = dY * gelu'(Z) (act = 1) or dY (act = 0), as f32 and bf16. */
ZRK_EXPORT int zrk_mlp_act_backward(int M, int N, const float* dY, const float* Z, float* dZ,
                                    void* dZb, int act, void* stream);

/* db[j] = sum_i dZ[i*N + j]. */
ZRK_EXPORT int zrk_mlp_col_sum(int M, int N, const float* dZ, float* db, void* stream);

/* One Adam step (bias correction for step number `step` >= 1) on FP32 master
 * weights w with gradient g and moments m, v; wb (optional) receives bf16(w). */
ZRK_EXPORT int zrk_mlp_adam(long n, float* w, const float* g, float* m, float* v, void* wb,
                            float lr, float beta1, float beta2, float eps, int step,
                            void* stream);

/* y (bf16) = x (f32), round to nearest even. */
ZRK_EXPORT int zrk_to_bf16(long n, const float* x, void* y, void* stream);

/* FP64 acceptance test of proposed trajectories of the constant-pressure
 * step [0, dt] from (T0, Y0) at pressure P (-4 with constant_volume set).
 * The proposal gives the state at the nodes t_k = tau[k-1] dt, k = 1..nodes
 * (3 <= nodes <= 8, tau a device array, increasing, tau[nodes-1] = 1):
 * Tp[(k-1)*n + r], Yp[((k-1)*n + r)*stride + j].  Interval k (from t_{k-1},
 * t_0 = 0 with the initial state) is the backward Euler equation (k = 1) or
 * the trapezoidal one (k > 1), solved by at most max_newton Newton
 * iterations with the iteration matrix I - c J (c = h or h/2, J the reference
 * Jacobian at the start point); it converges when the correction's weighted
 * RMS norm is at most 0.1.  The start point is the proposal at t_k, or the
 * converged state at t_{k-1} when its residual -h f(y_{k-1}) has the smaller
 * weighted norm.  The converged nodes' local errors are estimated from
 * divided differences of f over the nodes after t_0 ((h^2/2) f[t_1, t_2] for
 * the first interval, (h^3/12) 2 f[t_a, t_b, t_c] for the others), each
 * weighted with the weights 1 / (rel_tol |y| + abs_tol) of the node that
 * completes it, and summed componentwise.  accept[r] = 1 when every interval
 * converged and that sum has weighted RMS norm <= 1 (the test stops at the
 * first node where it exceeds 1); the converged end state then replaces the
 * last node's proposal.
 * res[4r..4r+3] (optional): most Newton iterations of an interval, largest
 * first correction, the error norm where the test stopped (infinite after a
 * failed interval), first interval whose Newton iterations failed (0: none,
 * -1: the RHS failed).  Uses the handle's workspace: calls on one handle must
 * not overlap.  Requires nsp <= 63. */
ZRK_EXPORT int zrk_surrogate_accept(zrk_mech m, const zrk_bdf_options* o, int n,
                                    const double* dt, double dt_all, const double* T0,
                                    const double* P, const double* Y0, int nodes,
                                    const double* tau, double* Tp, double* Yp, int stride,
                                    int max_newton, int* accept, double* res, void* stream);

#ifdef __cplusplus
}
#endif

#endif
