// Host-side mechanism: CHEMKIN-II reader and the step/rate tables the device
// kernels consume.
//
// Semantics follow the reference reader and mechanism build:
//   * text format: src/CKconverter/CKParser.cpp:861-1046 (thermo cards),
//     :1094-1110 (REACTIONS units), :1477-1668 (LOW/TROE/REV/DUP/efficiencies)
//   * molecular weights: src/zerork/species.cpp:38-45 with IUPAC-2007 masses
//   * step numbering and expanded participants: src/zerork/info_net.cpp:46-300
//   * rate-constant conversions, Arrhenius merging, Keq/third-body/falloff
//     lists: src/zerork/rate_const.cpp:50-134, 212-537, 1030-1128
//   * Jacobian term lists: applications/cfd_plugin/reactor_nvector_serial.cpp:122-281
//   * SRI falloff: rate_const.cpp:508-531 (parameters), :975-1004 (evaluation)
//   * PLOG: CKParser.cpp:1643-1669 and :312-345 (text, exclusions),
//     rate_const.cpp:1146-1186 (units), plog_reaction.cpp:10-135 (line sort,
//     duplicate pressures within 1e-8 relative)
//   * FORD/RORD and non-integer stoichiometry: CKParser.cpp:1615-1639,
//     info_net.cpp:411-470 (which reactions), non_integer_reaction_network.cpp:
//     53-262 (rate-of-progress exponents, step direction), rate_const.cpp:
//     334-386 (Keq change in moles from the exponents)
#pragma once

#include <array>
#include <map>
#include <string>
#include <vector>

namespace zrk {

enum FalloffType { kLindemann = 0, kTroe3 = 1, kTroe4 = 2, kSri = 3 };
constexpr int kFoParams = 8;  // per falloff reaction: lnAlow, beta, Ta, then Troe or SRI
enum RevType { kRevNone = 0, kRevArrhenius = 1, kRevKeq = 2 };

struct Species {
  std::string name;
  std::vector<std::pair<std::string, int>> comp;
  double tmid = 1000.0;
  double low[7] = {0};
  double high[7] = {0};
};

struct ReactionText {
  std::vector<std::pair<std::string, double>> reac, prod;
  bool reversible = true;
  double A = 0, n = 0, E = 0;  // E already in kelvin
  bool third_body = false;     // +M
  bool falloff = false;        // (+M) or (+species)
  std::string falloff_species; // empty for (+M)
  bool has_low = false;
  double low[3] = {0, 0, 0};
  std::vector<double> troe;
  bool has_rev = false;
  double rev[3] = {0, 0, 0};
  std::vector<double> sri;            // SRI a, b, c [, d [, e]]
  std::vector<std::array<double, 4>> plog;  // PLOG lines: p [atm], A, n, E [K]
  std::map<std::string, double> ford, rord;  // real reaction orders by species name
  bool real_order = false;            // FORD or RORD present
  std::map<std::string, double> eff;  // sorted by species name, like CKReader's map
};

// A step of a non-integer reaction (fractional stoichiometry or FORD/RORD):
// it has no expanded participant lists; its rate of progress is
// K * prod sign(C_s) |C_s|^p_s over rop_sp, and it destroys nu * ROP of each
// reactant and creates nu * ROP of each product (lists in reaction-text order,
// repeated names kept separate).  rop lists are in species-name order.
struct NonIntStep {
  int step = -1;
  std::vector<int> rop_sp;
  std::vector<double> rop_pow;
  std::vector<int> re_sp, pr_sp;
  std::vector<double> re_nu, pr_nu;
};

// PLOG forward rate of one step: distinct pressures (Pa) ascending, each with
// a range of Arrhenius lines (ln|A| with the unit conversion, sign of A, beta,
// Ta) that are summed.
struct PlogStep {
  int step = -1;
  std::vector<double> p, lnp;
  std::vector<int> line_ptr;  // [npts + 1]
  std::vector<double> lnA, sign, beta, Ta;
};

// Reference-layout tables (host copy).
struct Mechanism {
  std::vector<std::string> elements;
  std::vector<Species> species;
  std::vector<ReactionText> rxns;
  int nsp = 0, nrxn = 0, nstep = 0;
  std::vector<double> mw;            // [nsp]
  std::vector<double> thermo;        // [nsp*16] Tmid, low a1..a7, high a1..a7
  std::vector<int> step_rxn, step_dir;
  std::vector<int> reac_ptr, reac_sp, prod_ptr, prod_sp;
  std::vector<int> fwd_step, rev_step;  // per reaction (-1 if none)
  // per step Arrhenius (merged representative) for steps that have one
  std::vector<int> step_has_arr;
  std::vector<double> step_lnA, step_beta, step_Ta;
  std::vector<int> rev_type;           // per reaction
  std::vector<double> keq_ndelta;      // per reaction (Keq reverse)
  std::vector<int> tb_sp_ptr, tb_sp;   // per reaction enhanced lists (tb or falloff)
  std::vector<double> tb_eff;          // efficiency - 1
  std::vector<int> falloff_type;       // per reaction, -1 if not falloff
  std::vector<int> falloff_spc;        // per reaction, -1 for (+M)
  std::vector<double> falloff_param;   // per reaction [kFoParams]: lnAlow, beta, Ta,
                                       // Troe alpha, T***, T*, T** or SRI a, b, 1/c
                                       // (-1 if c <= 0), d, e
  std::vector<int> step_ni;            // per step: index into ni or -1
  std::vector<NonIntStep> ni;          // in step order
  std::vector<int> rxn_plog;           // per reaction: index into plog or -1
  std::vector<PlogStep> plog;
};

// Parse CHEMKIN mechanism (+ optional separate thermo file) and build tables.
// Returns false with a message on failure.
bool load_mechanism(const std::string& mech_file, const std::string& therm_file,
                    Mechanism* out, std::string* err);

}  // namespace zrk
