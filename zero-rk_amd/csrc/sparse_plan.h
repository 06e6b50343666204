// Sparse Newton-matrix plan for large mechanisms (the integrator's sparse
// path): the analytic Jacobian's pattern (reactor_nvector_serial.cpp:122-281:
// diagonal, full temperature row and column, and for every step each reactant
// column against the rows of the step's reactants and products), a fill-
// reducing symmetric ordering with the temperature last, the filled LU
// pattern, level sets for the row-by-row factorisation and the two
// triangular solves, the trailing block that fill makes nearly dense (it is
// factorised as a dense matrix), and the Jacobian term lists by LU position.
//
// Diagonal pivots: the reference's SuperLU options ask for them
// (DiagPivotThresh = 0, applications/cfd_plugin/interfaces/superlu_manager/
// superlu_manager.cpp:13-16); its MC64 row permutation and equilibration are
// not restated (I - gamma J has a dominant unit diagonal part for the step
// sizes the integrator takes; a zero pivot is reported as a recoverable
// setup failure, as CVODE's linear solver does).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "ck_mech.h"

namespace zrk {

// Device view of a plan (all arrays in one device allocation).  Indices: old =
// the mechanism's species order with the temperature at nsp; new = the
// factorisation order.  LU rows are in new order, CSR with sorted columns.
struct DevSparse {
  int N, nnz;
  int nlev_f, nlev_b;     // level counts: factorisation / forward solve, backward solve
  int max_row;            // longest LU row
  // rows and columns >= split form the trailing block, factorised densely
  // (split = N: no dense block); the leading rows' own level sets
  int split, nlev_l, nlev_lb;
  const int* ll_ptr;      // [nlev_l + 1]
  const int* ll_rows;     // [split]
  const int* lbl_ptr;     // [nlev_lb + 1] backward levels of the leading rows
  const int* lbl_rows;    // [split]
  const int* perm;        // [N] new -> old
  const int* rp;          // [N + 1]
  const int* ci;          // [nnz] new column
  const int* dpos;        // [N] position of the diagonal
  const int* lf_ptr;      // [nlev_f + 1] rows of each forward level in lf_rows
  const int* lf_rows;     // [N]
  const int* lb_ptr;      // [nlev_b + 1]
  const int* lb_rows;     // [N]
  const int* jt_ptr;      // [nnz + 1] Jacobian terms of each position
  const uint32_t* jt;     // step | 1u << 31 for a destruction term
  const int* pos_row;     // [nnz] old row (nsp: temperature)
  const int* pos_col;     // [nnz] old column
  const int* cp_ptr;      // [nsp + 1] positions of species rows in old column j,
  const int* cp_pos;      //   ascending old row
  const int* trow_pos;    // [nsp] position of (temperature row, old column j)
  const int* tcol_pos;    // [N] position of (old row k, temperature column)
};

struct SparsePlan {
  int N = 0, nnz = 0, nlev_f = 0, nlev_b = 0, max_row = 0, split = 0, nlev_l = 0, nlev_lb = 0;
  std::vector<int> ll_ptr, ll_rows, lbl_ptr, lbl_rows, perm, rp, ci, dpos, lf_ptr, lf_rows, lb_ptr, lb_rows, jt_ptr, pos_row,
      pos_col, cp_ptr, cp_pos, trow_pos, tcol_pos;
  std::vector<uint32_t> jt;
  long long factor_flops = 0;  // multiply-adds of one numeric factorisation
};

// Builds the plan of mechanism M (false with *err set when it cannot).
bool build_sparse_plan(const Mechanism& M, SparsePlan* plan, std::string* err);

}  // namespace zrk
