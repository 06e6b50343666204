// Learned surrogate of one reactor step: a small MLP trained and evaluated
// with BF16-input / FP32-accumulate MFMA kernels (v_mfma_f32_16x16x16_bf16),
// and the FP64 acceptance test of its proposals.
//
//   features     [log T, log P, log max(Y_j, 1e-20) (nsp), log dt] -> normalised BF16
//   gemm         C(f32) [+]= A(bf16) . B(bf16), any strides; one wavefront per
//                16x16 tile of C, K walked 16 at a time through the MFMA
//   bias_act     Z = C + b (f32), Y = gelu(Z) or Z (bf16 and f32)
//   act_backward dZ = dY * gelu'(Z) (f32 and bf16)
//   col_sum      db = sum over rows of dZ
//   adam         FP32 master weights, BF16 copies
//   accept       FP64 acceptance test of a proposed trajectory (one wavefront
//                per reactor, the integrator's RHS)
#include <hip/hip_runtime.h>

#include "bdf_device.h"
#include "launch.h"

namespace zrk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
  // round to nearest even (finite inputs; NaN stays NaN)
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __uint_as_float(((unsigned int)h) << 16);
}

// ---------------------------------------------------------------------------
__global__ void mlp_features_kernel(int n, int nsp, const double* __restrict__ T,
                                    const double* __restrict__ P, const double* __restrict__ Y,
                                    int stride, const double* __restrict__ dt, double dt_all,
                                    const float* __restrict__ mu, const float* __restrict__ sig,
                                    unsigned short* __restrict__ X, int ldx, int nfeat) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * ldx) return;
  const int r = (int)(i / ldx), f = (int)(i % ldx);
  float v = 0.0f;
  if (f < nfeat) {
    double x;
    if (f == 0) x = log(T[r]);
    else if (f == 1) x = log(P[r]);
    else if (f < 2 + nsp) x = log(fmax(Y[(size_t)r * stride + (f - 2)], 1.0e-20));
    else x = log(dt ? dt[r] : dt_all);
    v = (float)((x - (double)mu[f]) / (double)sig[f]);
  }
  X[i] = f32_to_bf16(v);
}

// C[M x N] (f32) = or += A[M x K] . B[K x N]; A, B bf16 with element strides.
// MFMA 16x16x16 bf16 operand layout: lane l supplies A[row l%16][k 4(l/16)+j]
// and B[k 4(l/16)+j][col l%16], j = 0..3, and receives D[row 4(l/16)+j][col l%16].
__global__ void __launch_bounds__(256)
gemm_bf16_kernel(int M, int N, int K, const unsigned short* __restrict__ A, long sAm, long sAk,
                 const unsigned short* __restrict__ B, long sBk, long sBn, float* __restrict__ C,
                 long sCm, long sCn, int accumulate) {
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  const int tilesN = (N + 15) / 16;
  const long tm = wave / tilesN;
  const int tn = (int)(wave % tilesN);
  if (tm * 16 >= M) return;
  const long ra = tm * 16 + (lane & 15);
  const int cb = tn * 16 + (lane & 15);
  const int kq = (lane >> 4) * 4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + kq + j;
      a[j] = (ra < M && k < K) ? (short)A[ra * sAm + (long)k * sAk] : (short)0;
      b[j] = (cb < N && k < K) ? (short)B[(long)k * sBk + (long)cb * sBn] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int col = tn * 16 + (lane & 15);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long row = tm * 16 + (lane >> 4) * 4 + j;
    if (row < M && col < N) {
      float* c = C + row * sCm + (long)col * sCn;
      *c = accumulate ? *c + acc[j] : acc[j];
    }
  }
}

// LDS-tiled form of gemm_bf16_kernel for the network's shapes: a workgroup of
// 4 wavefronts computes a 64 x 64 block of C, each wavefront a 32 x 32 quarter
// (2 x 2 MFMA tiles).  K is walked 32 at a time: the A and B tiles are read
// from HBM with 16-byte loads along whichever of their dimensions is
// contiguous (kAk: A[m, k] has k contiguous, else m; kBk: B[k, n] has k
// contiguous, else n) and stored k-contiguous in LDS (As[m][k], Bs[n][k], rows
// padded to 40 elements), so every MFMA operand is one 8-byte LDS read.
// Requires N % 64 == 0, the contiguous dimension of each operand a multiple
// of 8 elements, the other stride a multiple of 8 and 16-byte aligned bases
// (launch_gemm_bf16 checks this and otherwise uses gemm_bf16_kernel).  Rows
// of A past M and k past K read as zero.
constexpr int kGemmBM = 64, kGemmBN = 64, kGemmBK = 32, kGemmLd = kGemmBK + 8;

template <bool kAk, bool kBk>
__global__ void __launch_bounds__(256)
gemm_bf16_tiled_kernel(int M, int N, int K, const unsigned short* __restrict__ A, long sAm,
                       long sAk, const unsigned short* __restrict__ B, long sBk, long sBn,
                       float* __restrict__ C, long sCm, long sCn, int accumulate) {
  __shared__ __align__(16) unsigned short As[kGemmBM * kGemmLd];
  __shared__ __align__(16) unsigned short Bs[kGemmBN * kGemmLd];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const long m0 = (long)blockIdx.x * kGemmBM;
  const int n0 = blockIdx.y * kGemmBN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += kGemmBK) {
    // A tile (64 m x 32 k): 256 chunks of 8 elements, one per thread
    {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (kAk) {
        const int mm = t >> 2, c = (t & 3) * 8;
        const long gm = m0 + mm;
        const int gk = k0 + c;
        if (gm < M && gk < K) v = *reinterpret_cast<const uint4*>(A + gm * sAm + gk);
        *reinterpret_cast<uint4*>(As + mm * kGemmLd + c) = v;
      } else {
        const int kk = t >> 3, c = (t & 7) * 8;
        const long gm = m0 + c;
        const int gk = k0 + kk;
        if (gm < M && gk < K) v = *reinterpret_cast<const uint4*>(A + (long)gk * sAk + gm);
        const unsigned short* e = reinterpret_cast<const unsigned short*>(&v);
#pragma unroll
        for (int j = 0; j < 8; ++j) As[(c + j) * kGemmLd + kk] = e[j];
      }
    }
    // B tile (32 k x 64 n)
    {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (kBk) {
        const int nn = t >> 2, c = (t & 3) * 8;
        const int gk = k0 + c;
        if (gk < K) v = *reinterpret_cast<const uint4*>(B + (long)(n0 + nn) * sBn + gk);
        *reinterpret_cast<uint4*>(Bs + nn * kGemmLd + c) = v;
      } else {
        const int kk = t >> 3, c = (t & 7) * 8;
        const int gk = k0 + kk;
        if (gk < K) v = *reinterpret_cast<const uint4*>(B + (long)gk * sBk + n0 + c);
        const unsigned short* e = reinterpret_cast<const unsigned short*>(&v);
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[(c + j) * kGemmLd + kk] = e[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGemmBK; ks += 16) {
      const int kq = ks + (lane >> 4) * 4;
      bf16x4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *reinterpret_cast<const bf16x4*>(As + (wm + 16 * i + (lane & 15)) * kGemmLd + kq);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf16x4*>(Bs + (wn + 16 * j + (lane & 15)) * kGemmLd + kq);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long row = m0 + wm + 16 * i + (lane >> 4) * 4 + v;
        if (row < M) {
          float* c = C + row * sCm + (long)col * sCn;
          *c = accumulate ? *c + acc[i][j][v] : acc[i][j][v];
        }
      }
    }
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// Z = C + b; Y = act(Z) as bf16 (and f32 if Yf).  act: 1 gelu, 0 identity.
__global__ void bias_act_kernel(int M, int N, const float* __restrict__ C, const float* __restrict__ b,
                                float* __restrict__ Z, unsigned short* __restrict__ Yb,
                                float* __restrict__ Yf, int act) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const float z = C[i] + b[i % N];
  if (Z) Z[i] = z;
  const float y = act ? gelu(z) : z;
  if (Yb) Yb[i] = f32_to_bf16(y);
  if (Yf) Yf[i] = y;
}

// dZ = dY * act'(Z) (f32 and bf16 copies)
__global__ void act_backward_kernel(int M, int N, const float* __restrict__ dY,
                                    const float* __restrict__ Z, float* __restrict__ dZ,
                                    unsigned short* __restrict__ dZb, int act) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const float g = act ? dY[i] * gelu_grad(Z[i]) : dY[i];
  dZ[i] = g;
  dZb[i] = f32_to_bf16(g);
}

// db[j] = sum_i dZ[i*N + j]: one thread per column, rows in order
__global__ void col_sum_kernel(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.0f;
  for (int i = 0; i < M; ++i) s += dZ[(long)i * N + j];
  db[j] = s;
}

// Adam (bias-corrected) on FP32 master weights; refreshes the BF16 copy
__global__ void adam_kernel(long n, float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v,
                            unsigned short* __restrict__ wb, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f32_to_bf16(wi);
}

__global__ void to_bf16_kernel(long n, const float* __restrict__ x, unsigned short* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f32_to_bf16(x[i]);
}

// ---------------------------------------------------------------------------
// Acceptance of a proposed trajectory of one reactor step [0, dt] (one
// wavefront per reactor, constant pressure, FP64).  The proposal gives the
// state at the nodes t_k = tau_k dt (k = 1..m, tau increasing, tau_m = 1):
// Tp[(k-1)*n + r] and Yp[((k-1)*n + r)*stride + j].  Interval k, from t_{k-1}
// (t_0 = 0, the initial state) to t_k, is the implicit equation
//   G(y) = y - y_{k-1} - h f(y)                      (k = 1, backward Euler)
//   G(y) = y - y_{k-1} - h/2 (f(y_{k-1}) + f(y))     (k > 1, trapezoidal rule)
// solved by Newton iterations with the iteration matrix I - c J (c = h or
// h/2, J the reference Jacobian at the start point): at most max_newton
// iterations, converged when the correction's weighted RMS norm is at most
// kAcceptConv.  The start point is the proposal at t_k, or the converged state
// at t_{k-1} when its residual (-h f(y_{k-1})) has the smaller weighted norm
// (weights at y_{k-1}); a start too far from the solution for that many
// iterations fails the test.  Local errors come from divided differences of f
// over the converged nodes after t_0 (the initial state may lie off the slow
// manifold that the first interval's backward Euler step damps away):
//   backward Euler   (h^2 / 2) f[t_1, t_2]
//   trapezoidal      (h^3 / 12) 2 f[t_a, t_b, t_c], three consecutive nodes
// Each interval's |local error| is weighted with the weights
// 1 / (rtol |y| + atol) of the node that completes its divided difference and
// summed componentwise; the step is accepted when every interval converged
// and that sum has weighted RMS norm at most 1.  The sum only grows, so the
// test stops at the first node where its norm exceeds 1.  An accepted
// reactor's converged end state is written to the last node's arrays.
// res[4r..4r+3] (optional): most Newton iterations of an interval, largest
// first correction, the error norm where the test stopped (infinite if an
// interval failed first), first interval whose Newton iterations failed
// (0: none, -1: the RHS failed).
// LDS: Ml[kNP*kNP], rvec[kNP], t2l[kNP], then C, G, rop.
// ---------------------------------------------------------------------------
constexpr int kAcceptMaxNodes = 8;
constexpr double kAcceptConv = 0.1;

template <bool kExt>
__global__ void __launch_bounds__(kWave)
accept_kernel(DevMech m, BdfOpts o, int n, const double* __restrict__ dts, double dt_all,
              const double* __restrict__ T0, const double* __restrict__ P,
              const double* __restrict__ Y0, int nodes, const double* __restrict__ tau,
              double* __restrict__ Tp, double* __restrict__ Yp, int stride, int max_newton,
              int* __restrict__ accept, double* __restrict__ res,
              double* __restrict__ jscratch) {
  extern __shared__ double lds[];
  const int lane = lane_id();
  const int nsp = m.nsp;
  const int N = nsp + 1;
  double* Ml = lds;
  double* rvec = Ml + kNP * kNP;
  double* t2l = rvec + kNP;
  double* rl = t2l + kNP;
  RhsLds s = rhs_lds_at(rl, nsp);
  double* saved = jscratch + (size_t)blockIdx.x * kNP * kNP;
  const bool act = lane <= nsp;
  const double inv = 1.0 / (double)N;
  const int rot = gj_rotation(N);  // register c of the inverse holds column (c + rot) mod 64
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    const double dt = dts ? dts[r] : dt_all;
    const double Pr = P[r];
    double y0 = 0.0;
    if (lane < nsp) y0 = Y0[(size_t)r * stride + lane];
    if (lane == nsp) y0 = T0[r] / o.Tref;
    RhsAux aux;
    double yprev = y0, fprev = 0.0;
    int bad = rhs64<kExt>(m, s, Pr, 0.0, o.Tref, o.step_limiter, y0, fprev, aux);
    int failed = 0, most_it = 0;
    double first_max = 0.0;
    double esum = 0.0;  // this lane's weighted |local errors| summed so far
    double en = 0.0;
    bool stopped = false;
    // f and t at the two converged nodes before node k (sliding window)
    double fm2 = 0.0, fm1 = 0.0, tm2 = 0.0, tm1 = 0.0;
    for (int k = 0; k < nodes && !bad && !failed && !stopped; ++k) {
      const size_t row = (size_t)k * n + r;
      double y = 0.0;
      if (lane < nsp) y = Yp[row * stride + lane];
      if (lane == nsp) y = Tp[row] / o.Tref;
      const double tk = tau[k] * dt;
      const double h = tk - (k > 0 ? tm1 : 0.0);
      const double c = k == 0 ? h : 0.5 * h;
      // the start point: the proposal, or y_{k-1} if its residual is smaller
      double fy;
      const int badp = rhs64<kExt>(m, s, Pr, 0.0, o.Tref, o.step_limiter, y, fy, aux);
      const double wp = act ? 1.0 / (o.rtol * fabs(yprev) + o.atol) : 0.0;
      const double gp = (!act || badp) ? 0.0
                        : (k == 0 ? y - yprev - h * fy : y - yprev - 0.5 * h * (fprev + fy));
      const double rp = badp ? INFINITY : wave_sum(gp * gp * wp * wp);
      const double rq = wave_sum(act ? h * h * fprev * fprev * wp * wp : 0.0);
      if (rq < rp) {
        y = yprev;
        bad = rhs64<kExt>(m, s, Pr, 0.0, o.Tref, o.step_limiter, y, fy, aux);
        if (bad) continue;
      }
      // Newton matrix I - c J at the start point, inverted in registers
      jacobian_lds<double, kExt>(m, s, Pr, 0.0, o, y, fy, aux.inv_rho, Ml, kNP, saved, t2l);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double R[kNP];
#pragma unroll
      for (int col = 0; col < kNP; ++col) {
        double v = (col <= nsp && lane <= nsp) ? -c * Ml[col * kNP + lane] : 0.0;
        if (col == lane) v += 1.0;
        R[col] = v;
      }
      const bool ok = gj_invert<double>(R, N);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int col = 0; col < kNP; ++col) Ml[col * kNP + lane] = R[col];
      if (!ok) {
        failed = k + 1;
        continue;
      }
      const double w = act ? 1.0 / (o.rtol * fabs(y) + o.atol) : 0.0;
      bool conv = false;
      int it = 0;
      for (; it < max_newton && !conv && !bad; ++it) {
        if (it > 0) bad = rhs64<kExt>(m, s, Pr, 0.0, o.Tref, o.step_limiter, y, fy, aux);
        if (bad) continue;
        const double G = !act ? 0.0
                         : (k == 0 ? y - yprev - h * fy : y - yprev - 0.5 * h * (fprev + fy));
        rvec[lane] = G;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double d = 0.0;
#pragma unroll 8
        for (int col = 0; col < kNP; ++col)
          d = fma(Ml[col * kNP + lane], rvec[(col + rot) & (kNP - 1)], d);
        __builtin_amdgcn_wave_barrier();
        if (act) y -= d;
        const double dn = sqrt(wave_sum(act ? d * d * w * w : 0.0) * inv);
        if (it == 0) first_max = fmax(first_max, dn);
        conv = dn <= kAcceptConv;
      }
      if (bad) continue;
      most_it = max(most_it, it);
      if (!conv) {
        failed = k + 1;
        continue;
      }
      bad = rhs64<kExt>(m, s, Pr, 0.0, o.Tref, o.step_limiter, y, fy, aux);
      if (bad) continue;
      yprev = y;
      fprev = fy;
      // the local errors that node k completes, in the weights at node k
      double e = 0.0;
      if (k == 1) {  // backward Euler on [0, t_1]: (h^2 / 2) y'' with y'' = f[t_1, t_2]
        e = fabs(0.5 * tm1 * tm1 * (fy - fm1) / (tk - tm1));
      } else if (k >= 2) {
        // trapezoidal intervals: (h^3 / 12) 2 f[t_a, t_b, t_c] over the three
        // consecutive nodes ending at k; the first three nodes also serve the
        // trapezoidal interval before, which has no node of its own before it
        const double dd = ((fy - fm1) / (tk - tm1) - (fm1 - fm2) / (tm1 - tm2)) / (tk - tm2);
        e = fabs(h * h * h / 12.0 * 2.0 * dd);
        if (k == 2) {
          const double h1 = tm1 - tm2;
          e += fabs(h1 * h1 * h1 / 12.0 * 2.0 * dd);
        }
      }
      fm2 = fm1;
      fm1 = fy;
      tm2 = tm1;
      tm1 = tk;
      esum += act ? e / (o.rtol * fabs(y) + o.atol) : 0.0;
      en = sqrt(wave_sum(esum * esum) * inv);
      if (en > 1.0) stopped = true;
    }
    if (bad || failed) en = INFINITY;
    const bool acc = !bad && !failed && en <= 1.0;
    if (acc) {
      const size_t row = (size_t)(nodes - 1) * n + r;
      if (lane < nsp) Yp[row * stride + lane] = yprev;
      if (lane == nsp) Tp[row] = yprev * o.Tref;
    }
    if (lane == 0) {
      accept[r] = acc ? 1 : 0;
      if (res) {
        res[4 * (size_t)r] = (double)most_it;
        res[4 * (size_t)r + 1] = first_max;
        res[4 * (size_t)r + 2] = en;
        res[4 * (size_t)r + 3] = bad ? -1.0 : (double)failed;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
inline unsigned blocks_for(long n, int b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

hipError_t launch_mlp_features(int n, int nsp, const double* T, const double* P, const double* Y,
                               int stride, const double* dt, double dt_all, const float* mu,
                               const float* sig, unsigned short* X, int ldx, int nfeat,
                               hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_features_kernel, dim3(blocks_for((long)n * ldx, 256)), dim3(256), 0, st,
                     n, nsp, T, P, Y, stride, dt, dt_all, mu, sig, X, ldx, nfeat);
  return hipGetLastError();
}

hipError_t launch_gemm_bf16(int M, int N, int K, const unsigned short* A, long sAm, long sAk,
                            const unsigned short* B, long sBk, long sBn, float* C, long sCm,
                            long sCn, int accumulate, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  // the tiled kernel where its layout conditions hold (16-byte loads along
  // the contiguous dimension of each operand), else one wavefront per tile
  const auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool aK = sAk == 1 && sAm % 8 == 0 && K % 8 == 0;
  const bool aM = sAm == 1 && sAk % 8 == 0 && M % 8 == 0;
  const bool bK = sBk == 1 && sBn % 8 == 0 && K % 8 == 0;
  const bool bN = sBn == 1 && sBk % 8 == 0;
  if (N % kGemmBN == 0 && K > 0 && (aK || aM) && (bK || bN) && al16(A) && al16(B)) {
    const dim3 grid((unsigned)((M + kGemmBM - 1) / kGemmBM), (unsigned)(N / kGemmBN));
    if (aK && bN)
      hipLaunchKernelGGL((gemm_bf16_tiled_kernel<true, false>), grid, dim3(256), 0, st, M, N, K, A,
                         sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate);
    else if (aK)
      hipLaunchKernelGGL((gemm_bf16_tiled_kernel<true, true>), grid, dim3(256), 0, st, M, N, K, A,
                         sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate);
    else if (bN)
      hipLaunchKernelGGL((gemm_bf16_tiled_kernel<false, false>), grid, dim3(256), 0, st, M, N, K,
                         A, sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate);
    else
      hipLaunchKernelGGL((gemm_bf16_tiled_kernel<false, true>), grid, dim3(256), 0, st, M, N, K, A,
                         sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate);
    return hipGetLastError();
  }
  const long tiles = (long)((M + 15) / 16) * ((N + 15) / 16);
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(blocks_for(tiles * kWave, 256)), dim3(256), 0, st, M,
                     N, K, A, sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate);
  return hipGetLastError();
}

hipError_t launch_bias_act(int M, int N, const float* C, const float* b, float* Z,
                           unsigned short* Yb, float* Yf, int act, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for((long)M * N, 256)), dim3(256), 0, st, M, N,
                     C, b, Z, Yb, Yf, act);
  return hipGetLastError();
}

hipError_t launch_act_backward(int M, int N, const float* dY, const float* Z, float* dZ,
                               unsigned short* dZb, int act, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(act_backward_kernel, dim3(blocks_for((long)M * N, 256)), dim3(256), 0, st, M,
                     N, dY, Z, dZ, dZb, act);
  return hipGetLastError();
}

hipError_t launch_col_sum(int M, int N, const float* dZ, float* db, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(col_sum_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, st, M, N, dZ, db);
  return hipGetLastError();
}

hipError_t launch_adam(long n, float* w, const float* g, float* m, float* v, unsigned short* wb,
                       float lr, float b1, float b2, float eps, float c1, float c2,
                       hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, w, g, m, v, wb,
                     lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_to_bf16(long n, const float* x, unsigned short* y, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

hipError_t launch_accept(const DevMech& m, const BdfOpts& o, int n, const double* dts,
                         double dt_all, const double* T0, const double* P, const double* Y0,
                         int nodes, const double* tau, double* Tp, double* Yp, int stride,
                         int max_newton, int* accept, double* res, double* jscratch,
                         int nblocks, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (m.nsp > kNP - 1 || nodes < 3 || nodes > kAcceptMaxNodes) return hipErrorInvalidValue;
  const int grid = n < nblocks ? n : nblocks;
  const size_t lds = sizeof(double) * (size_t)(kNP * kNP + 2 * kNP) + rhs_lds_bytes(m);
  if (m.n_ext > 0)
    hipLaunchKernelGGL((accept_kernel<true>), dim3(grid), dim3(kWave), lds, st, m, o, n, dts,
                       dt_all, T0, P, Y0, nodes, tau, Tp, Yp, stride, max_newton, accept, res,
                       jscratch);
  else
    hipLaunchKernelGGL((accept_kernel<false>), dim3(grid), dim3(kWave), lds, st, m, o, n, dts,
                       dt_all, T0, P, Y0, nodes, tau, Tp, Yp, stride, max_newton, accept, res,
                       jscratch);
  return hipGetLastError();
}

}  // namespace zrk
