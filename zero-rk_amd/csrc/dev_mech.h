// Device image of a mechanism: flat arrays in one HBM allocation.
//
// Reaction records are reordered on the device by kind (plain, third-body,
// falloff) so that the lanes of a wavefront working on consecutive records
// follow the same branch; each record keeps its forward/reverse step index so
// step-ordered outputs (rates of progress) are unchanged.
#pragma once

#include <cstdint>

namespace zrk {

constexpr int kMaxPart = 4;  // max expanded reactants/products per step on device
constexpr int kNP = 64;      // padded state length of the integrator (one wavefront)

// Integrator options in kernel form (see zrk_bdf_options).
struct BdfOpts {
  double rtol, atol, hmax_inv, nlscoef, step_limiter, Tref, min_mass_fraction;
  double retry_atol_adj, retry_rtol_adj;
  int max_steps, num_retries;
};

enum RecFlags : int {
  kRecRevArr = 1,      // explicit reverse Arrhenius (REV)
  kRecRevKeq = 2,      // reverse from equilibrium constant
  kRecThirdBody = 4,   // +M
  kRecFalloff = 8,     // (+M) / (+species)
  kRecTroe = 16,       // Troe broadening (3 or 4 parameters)
  kRecTroe4 = 32,      // 4-parameter Troe
};

struct DevMech {
  int nsp, nstep, nrxn, nrec;
  int max_dlist, max_clist;  // longest per-species destroy/create list
  // species [nsp]
  const double* mw;
  const double* inv_mw;   // 1/mw
  const double* ru_mw;    // Ru/mw
  const double* thermo;   // [nsp*16]
  // reaction records [nrec]
  const int* r_fstep;
  const int* r_rstep;     // -1 if irreversible
  const int* r_flags;
  const double* r_fA;     // ln A, beta, Ta of the forward step
  const double* r_fB;
  const double* r_fT;
  const double* r_rA;     // reverse: ln A, beta, Ta (REV) or ndelta in r_rA (Keq)
  const double* r_rB;
  const double* r_rT;
  const int* r_reac;      // [nrec*kMaxPart], -1 padded (forward-direction reactants)
  const int* r_prod;      // [nrec*kMaxPart]
  const int* r_eptr;      // [nrec+1] enhanced third-body list
  const int* e_sp;
  const double* e_eff;    // efficiency - 1
  const int* r_fspc;      // falloff single species or -1
  const double* r_fo;     // [nrec*7] falloff parameters
  // per-species gather lists (step indices, multiplicity repeated, step order)
  const int* d_ptr;       // [nsp+1]
  const int* d_step;
  const int* c_ptr;       // [nsp+1]
  const int* c_step;
  // Jacobian species-block terms: entry (row i, col c) -> terms
  // jt_ptr[c*nsp + i] .. jt_ptr[c*nsp+i+1]; term = -(s+1) destruction, +(s+1) creation
  const int* jt_ptr;
  const int* jt_term;
};

}  // namespace zrk
