// Device image of a mechanism: flat arrays in one HBM allocation.
//
// Reaction records are reordered on the device by kind (plain, third-body,
// falloff) so that the lanes of a wavefront working on consecutive records
// follow the same branch; each record keeps its forward/reverse step index so
// step-ordered outputs (rates of progress) are unchanged.
#pragma once

#include <cstdint>

#include <hip/hip_vector_types.h>

namespace zrk {

constexpr int kMaxPart = 4;  // max expanded reactants/products per step on device
constexpr int kNP = 64;      // padded state length of the integrator (one wavefront)

// Integrator options in kernel form (see zrk_bdf_options).
struct BdfOpts {
  double rtol, atol, hmax_inv, nlscoef, step_limiter, Tref, min_mass_fraction;
  double retry_atol_adj, retry_rtol_adj;
  int max_steps, num_retries;
  int constant_volume;
  int fp32_preconditioner;
  // optional per-reactor source terms (the plugin's DPDT, E_SRC, Y_SRC fields):
  // dpdt[n] Pa/s, esrc[n] J/kg/s, ysrc[n*stride] 1/s (null: none)
  const double* dpdt = nullptr;
  const double* esrc = nullptr;
  const double* ysrc = nullptr;
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
  const DevMech* self;  // this struct's own copy in device memory
  int nsp, nstep, nrxn, nrec;
  int gd_blocks, gc_blocks;  // 8-entry blocks of the padded destroy/create gathers
  int jr_blocks;             // 4-entry blocks of the per-row Jacobian term lists
  int jr_stride;             // rows per block of jr (64 * ceil((nsp + 1) / 64))
  // species [nsp]
  const double* mw;
  const double* inv_mw;   // 1/mw
  const double* ru_mw;    // Ru/mw
  const double* thermo;   // [nsp*16]
  // reaction records [nrec]
  const int4* r_meta;     // {forward step, reverse step or -1, flags, falloff species or -1}
  const int4* r_sp;       // 16-bit species indices (nsp = none): {re0|re1, re2|re3, pr0|pr1, pr2|pr3}
  const double2* r_coef;  // [nrec*3]: {fwd lnA, beta}, {fwd Ta, rev lnA}, {rev beta, rev Ta};
                          // for a Keq reverse the reverse lnA slot holds ndelta
  const int* r_eptr;      // [nrec+1] enhanced third-body list (multiples of 4)
  const int* e_sp;        // in groups of 4 (read as int4); padding: species 0
  const double* e_eff;    // efficiency - 1, in pairs (read as double2); padding: 0
  const double* r_fo;     // [nrec*7] falloff parameters
  // per-species gathers, padded to a common length so the 64 lanes of a
  // species block read one contiguous line per block of 8 entries:
  // g_d[(jb*gd_blocks + b)*64 + (j&63)] holds 8 x uint16 step indices of
  // species j = 64*jb + (j&63), in step order; padding index nstep (rate 0)
  const uint4* g_d;
  const uint4* g_c;
  // Jacobian species-block terms by row (nsp <= 191): jr[b*S + row] (S = 64,
  // 128 or 192: the integrator's team size) holds 4
  // terms, each bits 0-15 step, 16-23 column (0xFF = padding), bit 31 set for
  // a destruction term; sorted by column, then destruction before creation,
  // each in step order (the reference's accumulation order)
  const uint4* jr;
};

}  // namespace zrk
