// Golden-vector probe over the reference Zero-RK rate core.
//
// Test infrastructure only: linked against the reference's own
// src/zerork + src/CKconverter sources (compiled by oracle/Makefile into
// oracle/_ref/), it evaluates the reference functions on given states and
// prints the results as text.  Nothing in the product links or runs it.
//
// Usage: zrk_ref_probe <mech> <therm> <states.txt> <out.txt>
//   states.txt: first line "n nsp", then n lines "T P Y_0 .. Y_{nsp-1}"
//               (T in K, P in Pa, Y mass fractions).
// Output blocks (all doubles printed with %.17g):
//   header: nsp nrxn nstep
//   mw[nsp]
//   per step: rxn_idx_with_dir  n_reac reac...  n_prod prod...
//   per state: T P, then cp_R[nsp] h_RT[nsp] g_RT[nsp],
//              conc[nsp], kfwd[nrxn] krev[nrxn],
//              rop_limited[nstep] create[nsp] destroy[nsp] net[nsp]
//              (limiter 1e22 as in reactor_nvector_serial.cpp:56),
//              rop_nolimit[nstep] net_nolimit[nsp]
//
// Reference calls: mechanism ctor (src/zerork/mechanism.cpp:17),
// getDensityFromTPY (:279), getCfromVY (:323), getCp_R/getH_RT/getG_RT via
// nasa_poly_group (nasa_poly.cpp:50-172), getKrxnFromTC (mechanism.cpp:549),
// getReactionRatesLimiter (mechanism.cpp:610), getReactionRates (:556).
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>

#include "zerork/mechanism.h"

static void pr(FILE* f, const std::vector<double>& v) {
  for (size_t i = 0; i < v.size(); ++i) fprintf(f, "%.17g ", v[i]);
  fprintf(f, "\n");
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s mech therm states out\n", argv[0]);
    return 2;
  }
  zerork::mechanism mech(argv[1], argv[2], "", 0);
  const int nsp = mech.getNumSpecies();
  const int nrxn = mech.getNumReactions();
  const int nstep = mech.getNumSteps();

  FILE* fs = fopen(argv[3], "r");
  if (!fs) { perror("states"); return 1; }
  int n = 0, nsp_in = 0;
  if (fscanf(fs, "%d %d", &n, &nsp_in) != 2 || nsp_in != nsp) {
    fprintf(stderr, "bad states header (nsp %d vs mech %d)\n", nsp_in, nsp);
    return 1;
  }
  FILE* fo = fopen(argv[4], "w");
  if (!fo) { perror("out"); return 1; }
  fprintf(fo, "%d %d %d\n", nsp, nrxn, nstep);
  std::vector<double> mw(nsp);
  mech.getMolWtSpc(mw.data());
  pr(fo, mw);
  for (int s = 0; s < nstep; ++s) {
    int nr = mech.getOrderOfStep(s), np = mech.getNumProductsOfStep(s);
    fprintf(fo, "%d %d", mech.getRxnIdxOfStep(s), nr);
    for (int k = 0; k < nr; ++k) fprintf(fo, " %d", mech.getSpecIdxOfStepReactant(s, k));
    fprintf(fo, " %d", np);
    for (int k = 0; k < np; ++k) fprintf(fo, " %d", mech.getSpecIdxOfStepProduct(s, k));
    fprintf(fo, "\n");
  }
  std::vector<double> y(nsp), c(nsp), cp(nsp), h(nsp), g(nsp), kf(nrxn), kr(nrxn);
  std::vector<double> rop(nstep), cr(nsp), de(nsp), net(nsp), lim(nstep, 1.0e22);
  std::vector<double> rop2(nstep), cr2(nsp), de2(nsp), net2(nsp);
  for (int i = 0; i < n; ++i) {
    double T, P;
    if (fscanf(fs, "%lf %lf", &T, &P) != 2) return 1;
    for (int j = 0; j < nsp; ++j)
      if (fscanf(fs, "%lf", &y[j]) != 1) return 1;
    const double rho = mech.getDensityFromTPY(T, P, y.data());
    mech.getCfromVY(1.0 / rho, y.data(), c.data());
    mech.getCp_R_Enthalpy_RT(T, cp.data(), h.data());
    mech.getNonDimGibbsFromT(T, g.data());
    mech.getKrxnFromTC(T, c.data(), kf.data(), kr.data());
    mech.getReactionRatesLimiter(T, c.data(), lim.data(), net.data(), cr.data(),
                                 de.data(), rop.data());
    mech.getReactionRates(T, c.data(), net2.data(), cr2.data(), de2.data(), rop2.data());
    fprintf(fo, "%.17g %.17g\n", T, P);
    pr(fo, cp); pr(fo, h); pr(fo, g); pr(fo, c); pr(fo, kf); pr(fo, kr);
    pr(fo, rop); pr(fo, cr); pr(fo, de); pr(fo, net);
    pr(fo, rop2); pr(fo, net2);
  }
  fclose(fo);
  fclose(fs);
  return 0;
}
