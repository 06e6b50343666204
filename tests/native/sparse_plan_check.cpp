// CPU check of the sparse plan (zero-rk_amd/csrc/sparse_plan.cpp): index
// ranges and level orders, and the GPU kernel's factorisation (leading rows
// by levels, trailing rows against the leading columns, the dense block
// right-looking, copied back to the CSR positions) emulated in FP64 on a
// random matrix with the plan's pattern, solved over the plan's levels and
// compared with the matrix itself (residual).  Usage: sparse_plan_check
// mech therm [emulate: 0/1].
#include "sparse_plan.h"
#include <cstdio>
#include <cmath>
#include <random>
#include <cstdlib>
using namespace zrk;
#define CHECK(c) do { if (!(c)) { printf("FAIL %s line %d\n", #c, __LINE__); return 1; } } while (0)
int main(int argc, char** argv) {
  Mechanism M; std::string err;
  if (!load_mechanism(argv[1], argv[2], &M, &err)) { printf("parse: %s\n", err.c_str()); return 1; }
  SparsePlan P;
  if (!build_sparse_plan(M, &P, &err)) { printf("plan: %s\n", err.c_str()); return 1; }
  const int N = P.N, nsp = N - 1, nnz = P.nnz;
  CHECK((int)P.perm.size() == N && P.perm[N-1] == nsp);
  CHECK(P.rp[0] == 0 && P.rp[N] == nnz);
  for (int i = 0; i < N; ++i) { CHECK(P.rp[i] < P.rp[i+1]); CHECK(P.dpos[i] >= P.rp[i] && P.dpos[i] < P.rp[i+1] && P.ci[P.dpos[i]] == i);
    for (int p = P.rp[i]; p < P.rp[i+1]; ++p) { CHECK(P.ci[p] >= 0 && P.ci[p] < N); if (p > P.rp[i]) CHECK(P.ci[p] > P.ci[p-1]); } }
  std::vector<int> seen(N, 0);
  CHECK(P.lf_ptr[0] == 0 && P.lf_ptr[P.nlev_f] == N && P.lb_ptr[P.nlev_b] == N);
  std::vector<int> lev(N, -1);
  for (int l = 0; l < P.nlev_f; ++l) for (int x = P.lf_ptr[l]; x < P.lf_ptr[l+1]; ++x) { int i = P.lf_rows[x]; CHECK(i >= 0 && i < N); seen[i]++; lev[i] = l; }
  for (int i = 0; i < N; ++i) { CHECK(seen[i] == 1); for (int p = P.rp[i]; p < P.dpos[i]; ++p) CHECK(lev[P.ci[p]] < lev[i]); }
  std::fill(lev.begin(), lev.end(), -1);
  for (int l = 0; l < P.nlev_b; ++l) for (int x = P.lb_ptr[l]; x < P.lb_ptr[l+1]; ++x) lev[P.lb_rows[x]] = l;
  for (int i = 0; i < N; ++i) { CHECK(lev[i] >= 0); for (int p = P.dpos[i]+1; p < P.rp[i+1]; ++p) CHECK(lev[P.ci[p]] < lev[i]); }
  CHECK((int)P.jt_ptr.size() == nnz + 1 && P.jt_ptr[nnz] <= (int)P.jt.size());
  for (uint32_t t : P.jt) CHECK((int)(t & 0x7FFFFFFFu) < M.nstep);
  for (int p = 0; p < nnz; ++p) { CHECK(P.pos_row[p] >= 0 && P.pos_row[p] < N && P.pos_col[p] >= 0 && P.pos_col[p] < N); }
  CHECK((int)P.cp_ptr.size() == nsp + 1);
  for (int j = 0; j < nsp; ++j) for (int k = P.cp_ptr[j]; k < P.cp_ptr[j+1]; ++k) { int p = P.cp_pos[k]; CHECK(p >= 0 && p < nnz && P.pos_col[p] == j && P.pos_row[p] < nsp); if (k > P.cp_ptr[j]) CHECK(P.pos_row[p] > P.pos_row[P.cp_pos[k-1]]); }
  for (int j = 0; j < nsp; ++j) { CHECK(P.pos_row[P.trow_pos[j]] == nsp && P.pos_col[P.trow_pos[j]] == j); }
  for (int k = 0; k < N; ++k) { CHECK(P.tcol_pos[k] >= 0 && P.pos_row[P.tcol_pos[k]] == k && P.pos_col[P.tcol_pos[k]] == nsp); }
  printf("split %d (block %d) leading levels %d\n", P.split, N - P.split, P.nlev_l);
  for (int l = 0; l < P.nlev_l; ++l) for (int x = P.ll_ptr[l]; x < P.ll_ptr[l+1]; ++x) CHECK(P.ll_rows[x] < P.split);
  // numeric emulation on a small system only
  if (argc > 3 && atoi(argv[3]) != 0) {
    std::mt19937 g(5); std::uniform_real_distribution<double> u(-1, 1);
    std::vector<double> A((size_t)N*N, 0.0), lu(nnz, 0.0);
    for (int p = 0; p < nnz; ++p) { int r = P.pos_row[p], c = P.pos_col[p]; bool orig = (r == c) || r == nsp || c == nsp || P.jt_ptr[p+1] > P.jt_ptr[p]; double v = orig ? (r == c ? 4.0 + u(g) : 0.3*u(g)) : 0.0; A[(size_t)r*N+c] = v; lu[p] = v; }
    std::vector<double> w(N, 0.0);
    const int sp = P.split, nd = N - sp;
    // phase 1: leading rows by levels
    for (int l = 0; l < P.nlev_l; ++l) for (int x = P.ll_ptr[l]; x < P.ll_ptr[l+1]; ++x) { int i = P.ll_rows[x];
      for (int p = P.rp[i]; p < P.rp[i+1]; ++p) w[P.ci[p]] = lu[p];
      for (int p = P.rp[i]; p < P.dpos[i]; ++p) { int k = P.ci[p]; double lk = w[k] / lu[P.dpos[k]]; w[k] = lk; for (int q = P.dpos[k]+1; q < P.rp[k+1]; ++q) w[P.ci[q]] -= lk * lu[q]; }
      for (int p = P.rp[i]; p < P.rp[i+1]; ++p) { lu[p] = w[P.ci[p]]; w[P.ci[p]] = 0; } }
    // phase 2: trailing rows against the leading columns, into the dense block
    std::vector<double> D((size_t)nd * nd, 0.0);
    for (int i = sp; i < N; ++i) {
      for (int p = P.rp[i]; p < P.rp[i+1]; ++p) w[P.ci[p]] = lu[p];
      for (int p = P.rp[i]; p < P.rp[i+1] && P.ci[p] < sp; ++p) { int k = P.ci[p]; double lk = w[k] / lu[P.dpos[k]]; w[k] = lk; for (int q = P.dpos[k]+1; q < P.rp[k+1]; ++q) w[P.ci[q]] -= lk * lu[q]; }
      for (int p = P.rp[i]; p < P.rp[i+1]; ++p) { if (P.ci[p] < sp) lu[p] = w[P.ci[p]]; }
      for (int c = sp; c < N; ++c) { D[(size_t)(i-sp)*nd + (c-sp)] = w[c]; }
      for (int p = P.rp[i]; p < P.rp[i+1]; ++p) w[P.ci[p]] = 0; }
    // phase 3: dense right-looking LU of the block
    for (int k = 0; k < nd; ++k) { double piv = D[(size_t)k*nd+k];
      for (int i = k+1; i < nd; ++i) { double a = D[(size_t)i*nd+k]; if (a == 0.0) continue; double l = a / piv; D[(size_t)i*nd+k] = l; for (int j = k+1; j < nd; ++j) D[(size_t)i*nd+j] -= l * D[(size_t)k*nd+j]; } }
    // phase 4: block back into the CSR positions; every other block entry is zero
    for (int i = sp; i < N; ++i) { for (int p = P.rp[i]; p < P.rp[i+1]; ++p) if (P.ci[p] >= sp) { lu[p] = D[(size_t)(i-sp)*nd + P.ci[p]-sp]; D[(size_t)(i-sp)*nd + P.ci[p]-sp] = 0.0; } }
    double stray = 0; for (double v : D) stray = fmax(stray, fabs(v));
    printf("block entries outside the pattern after factorisation: max %.3e\n", stray);
    CHECK(stray == 0.0);
    std::vector<double> b(N), x(N), bb(N);
    for (int i = 0; i < N; ++i) b[i] = u(g);
    for (int i = 0; i < N; ++i) x[i] = b[P.perm[i]];
    for (int l = 0; l < P.nlev_f; ++l) for (int xx = P.lf_ptr[l]; xx < P.lf_ptr[l+1]; ++xx) { int i = P.lf_rows[xx]; double s = 0; for (int p = P.rp[i]; p < P.dpos[i]; ++p) s += lu[p]*x[P.ci[p]]; x[i] -= s; }
    for (int l = 0; l < P.nlev_b; ++l) for (int xx = P.lb_ptr[l]; xx < P.lb_ptr[l+1]; ++xx) { int i = P.lb_rows[xx]; double s = 0; for (int p = P.dpos[i]+1; p < P.rp[i+1]; ++p) s += lu[p]*x[P.ci[p]]; x[i] = (x[i]-s)/lu[P.dpos[i]]; }
    for (int i = 0; i < N; ++i) bb[P.perm[i]] = x[i];
    double res = 0, bn = 0;
    for (int r = 0; r < N; ++r) { double s = 0; for (int c = 0; c < N; ++c) s += A[(size_t)r*N+c]*bb[c]; res = fmax(res, fabs(s - b[r])); bn = fmax(bn, fabs(b[r])); }
    printf("residual %.3e (|b| %.3e)\n", res, bn);
    CHECK(res < 1e-10 * bn * N);
  }
  printf("N %d nnz %d levels %d/%d ok\n", N, nnz, P.nlev_f, P.nlev_b);
  return 0;
}
