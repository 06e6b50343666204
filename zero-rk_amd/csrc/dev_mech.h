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
  // ignition: root of T - (T0 + ign_dT) (off when ign_dT <= 0); ign_time[n]
  // receives the last root time (0 if none); stop_after_ignition ends the
  // integration at the root
  double ign_dT = 0.0;
  int stop_after_ignition = 0;
  double* ign_time = nullptr;
  // absolute tolerances scaled by the initial molar densities
  int abstol_dens = 0;
  // linear solver: 0 the direct solve of I - gamma J; 1 SPGMR preconditioned
  // by it (pfull: every entry kept, else the Jacobian's pattern and
  // |entry| > pthr off the diagonal), tolerance factor eps_lin
  int krylov = 0;
  int jac = 0;  // kJacAnalytic, kJacCvodeDQ or kJacReactorDQ
  int pfull = 1;
  double pthr = 0.0, eps_lin = 0.0;
  // fixed temperature: with always_T = 0 a reactor whose tdelta entry is not
  // above 0 keeps its temperature (not a state variable)
  int always_T = 1;
  const double* tdelta = nullptr;
  // per-step record of a one-reactor batch (the plugin's step callback,
  // solver_cvode.cpp:247-257): after each successful CVode call record k =
  // {CVode calls so far, t, t - t_prev, y[nsp+1], ydot[nsp+1]} (y and ydot
  // CVodeGetDky(t, 0 and 1), nvar entries used) at trace + k * (2 nsp + 5);
  // records past trace_cap are counted, not written; *trace_n = records
  double* trace = nullptr;
  int trace_cap = 0;
  int* trace_n = nullptr;
};

enum JacKind : int {
  kJacAnalytic = 0,   // reactor_nvector_serial.cpp:610-771
  kJacCvodeDQ = 1,    // CVODE's dense difference quotients (cvLsDenseDQJac)
  kJacReactorDQ = 2,  // the reactor's divided differences (:816-835)
};

// whether these options need the integrator's alternate-mode kernels (SPGMR
// or a difference-quotient Jacobian); the direct analytic solve does not
inline bool bdf_alt_modes(const BdfOpts& o) { return o.krylov != 0 || o.jac != kJacAnalytic; }

enum RecFlags : int {
  kRecRevArr = 1,      // explicit reverse Arrhenius (REV)
  kRecRevKeq = 2,      // reverse from equilibrium constant
  kRecThirdBody = 4,   // +M
  kRecFalloff = 8,     // (+M) / (+species)
  kRecTroe = 16,       // Troe broadening (3 or 4 parameters)
  kRecTroe4 = 32,      // 4-parameter Troe
  kRecSri = 64,        // SRI blending function
  kRecPlog = 128,      // forward rate from PLOG tables
  kRecNonInt = 256,    // non-integer reaction (real rate-of-progress exponents)
};
// Records with kRecPlog or kRecNonInt carry an index into r_ext in the bits of
// their flags from kRecExtShift up.
constexpr int kRecExtShift = 12;
// A Jacobian term with this bit set is a non-integer term: bits 0-15 index ni_jt.
constexpr unsigned kJtNonInt = 0x40000000u;

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
  const double* r_fo;     // [nrec*8] falloff parameters (lnAlow, beta, Ta, Troe or SRI)
  // PLOG and non-integer reactions (absent from most mechanisms): r_ext[e] =
  // {PLOG table offset or -1, forward exponent list offset, reverse exponent
  // list offset or -1, forward count | reverse count << 16}.  A PLOG table at
  // offset o in plog: npts, pressures [npts] (Pa), their logs [npts], first
  // line of each pressure [npts + 1], then lines of 4 doubles {ln|A|, sign of
  // A, beta, Ta}.  Exponent lists: species ni_sp[k], exponent ni_pow[k].
  int n_ext;
  const int4* r_ext;
  const double* plog;
  const int* ni_sp;
  const double* ni_pow;
  // non-integer destruction and creation terms by species (added before the
  // gathers below, as the reference does): entries nig[k] = {nu, step} for
  // k in [nig_dptr[j], nig_dptr[j+1]) and [nig_cptr[j], nig_cptr[j+1])
  int has_nig;
  const int* nig_dptr;
  const int* nig_cptr;
  const double2* nig;
  const double2* ni_jt;   // Jacobian non-integer terms {multiplier, step}
  // per-species gathers, padded to a common length so the 64 lanes of a
  // species block read one contiguous line per block of 8 entries:
  // g_d[(jb*gd_blocks + b)*64 + (j&63)] holds 8 x uint16 step indices of
  // species j = 64*jb + (j&63), in step order; padding index nstep (rate 0)
  const uint4* g_d;
  const uint4* g_c;
  // per 64-species group: {destroy, create} blocks in use (at most gd_blocks,
  // gc_blocks); the entries beyond are padding
  const int2* g_nb;
  // Jacobian species-block terms by row (nsp <= 191): jr[b*S + row] (S = 64,
  // 128 or 192: the integrator's team size) holds 4
  // terms, each bits 0-15 step, 16-23 column (0xFF = padding), bit 31 set for
  // a destruction term, or bit 30 (kJtNonInt) for a non-integer term; sorted
  // by column, then destruction before creation, each in step order, then the
  // non-integer terms (the reference's accumulation order)
  const uint4* jr;
  // the analytic Jacobian's pattern (nsp <= 191), for the Krylov solver's
  // preconditioner: bit c % 64 of jpat[row * 3 + c / 64] is set when column c
  // of that row is in the pattern (diagonal, temperature row and column,
  // reactant columns against reactant and product rows, non-integer terms)
  const unsigned long long* jpat;
};

}  // namespace zrk
