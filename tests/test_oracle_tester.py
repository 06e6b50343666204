"""Pins the oracle's whole reactor integration to the reference's own runs:
the cfd_plugin tester's outputs (examples/cfd_plugin_tester/outputs.ref, run
by examples/cfd_plugin_tester/run.sh.in; copies under tests/golden/tester).

Each run integrates 32 identical constant-volume reactors (1450 K, 2.0265e6
Pa, stoichiometric fuel/air from inputs/<name>_fracs.log, composition built as
zerork_cfd_plugin_tester.cpp:457-530 does) through 10 solves of dt = 2 x the
ignition delay / 10, rtol 1e-8, atol 1e-18, ignition root at +400 K, and
prints reactor 0's T, p and three mass fractions after each solve with 6
significant digits.  Mechanisms: H2 (10 species), DME (79 species), n-heptane
(159 species).

These outputs were made with the gas constant of the reference's earlier
releases, CODATA 2014's 8.3144598e3 J/kmol/K (also quoted in
data/mechanisms/ideal/non_integer_test.mech); with it, the oracle reproduces
every printed value, ignition included, to within the print rounding
(5e-6 relative) plus five times the oracle's own global-error estimate
(its rtol 1e-8 run against an rtol 1e-11 run), with an absolute floor of 1e-15
for burned-out species.  With the current constant (src/zerork/constants.h:10)
the fuel is consumed faster by ~1e-5 relative per solve, as a 3.4e-7 change in
the activation temperatures predicts: that case is checked to fail the same
bound, so the comparison can tell the two apart.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

TESTER = os.path.join(GOLDEN, "tester")
RU_2014 = 8.3144598e3
IDT = {"h2": 8.2565998e-07, "dme": 1.1e-05, "nc7h16-skel": 1.35e-05}
RUNS = {"h2": ["h2_dense"], "dme": ["dme_dense", "dme_sparse"],
        "nc7h16-skel": ["nc7h16-skel_sparse"]}


def initial_mass_fractions(o, name, mech):
    """zerork_cfd_plugin_tester.cpp:457-530 at phi = 1, no EGR: fuel and
    oxidizer mole fractions (each normalised), oxidizer moles per mole of fuel
    from the atomic-oxygen balance (mechanism.cpp:731-742), then X -> Y."""
    from oracle import ckparse
    sp = o.tables.species
    th = {}
    for key, lines, rest in ckparse._sections(open(mech[1]).read()):
        if key == "THER":
            th.update(ckparse._parse_thermo(lines, sp))
    for key, lines, rest in ckparse._sections(open(mech[0]).read()):
        if key == "THER":
            th.update(ckparse._parse_thermo(lines, sp))
    thl = {k.upper(): v for k, v in th.items()}
    comp = [dict(th[s][3] if s in th else thl[s.upper()][3]) for s in sp]
    idx = {s: j for j, s in enumerate(sp)}
    fuel, oxid = np.zeros(len(sp)), np.zeros(len(sp))
    for line in open(os.path.join(TESTER, f"{name}_fracs.txt")):
        if line.startswith("#") or not line.strip():
            continue
        k, v = line.split(":")
        k = k.strip()
        (oxid if k.lower() in ("o2", "n2") else fuel)[idx[k]] = float(v)
    fuel /= fuel.sum()
    oxid /= oxid.sum()
    bal = lambda x: sum(x[j] * (comp[j].get("O", 0) - 2.0 * comp[j].get("C", 0)
                                - 0.5 * comp[j].get("H", 0)) for j in range(len(sp)))
    X = 1.0 * fuel + abs(bal(fuel)) / bal(oxid) * oxid
    X /= X.sum()
    return X * o.tables.mw / np.sum(X * o.tables.mw)


def trajectory(oracle_mod, o, Y0, dt, rtol, atol):
    T, P, Y = np.array([1450.0]), np.array([2.02650e6]), Y0[None].copy()
    oo = oracle_mod.default_opts(rtol=rtol, atol=atol, constant_volume=1, max_steps=30000,
                                 delta_temperature_ignition=400.0)
    out = []
    for _ in range(10):
        T, Y, st = o.integrate(T, P, Y, dt, oo)
        P = st["P"]
        assert st["nsteps"][0] > 0
        out.append((T[0], P[0], Y[0].copy()))
    return out


def compare(o, name, run, traj, traj_fine):
    path = os.path.join(TESTER, f"{run}_000.hist")
    hdr = open(path).readline().split()[5:]
    ref = np.loadtxt(path)
    idx = {s: j for j, s in enumerate(o.tables.species)}
    worst = 0.0
    for s in range(10):
        pick = lambda tr: np.array([tr[s][0], tr[s][1]] + [tr[s][2][idx[k]] for k in hdr])
        v, vf = pick(traj), pick(traj_fine)
        bound = 5e-6 * np.abs(ref[s, 2:]) + 1e-15 + 5.0 * np.abs(v - vf)
        worst = max(worst, float(np.max(np.abs(v - ref[s, 2:]) / bound)))
    return worst


@pytest.mark.parametrize("name", ["h2", "dme", "nc7h16-skel"])
def test_oracle_reproduces_reference_tester(oracle_mod, name):
    mech = (os.path.join(TESTER, f"{name}.mech"), os.path.join(TESTER, f"{name}.therm"))
    o = oracle_mod.Oracle(*mech, ru=RU_2014)
    Y0 = initial_mass_fractions(o, name, mech)
    dt = 2.0 * IDT[name] / 10.0
    traj = trajectory(oracle_mod, o, Y0, dt, 1e-8, 1e-18)
    fine = trajectory(oracle_mod, o, Y0, dt, 1e-11, 1e-21)
    for run in RUNS[name]:
        w = compare(o, name, run, traj, fine)
        print(f"{run}: worst |oracle - reference| / bound = {w:.2f}")
        assert w <= 1.0
    # the current gas constant moves the ignition by more than the bound allows
    o2 = oracle_mod.Oracle(*mech)
    Y2 = initial_mass_fractions(o2, name, mech)
    traj2 = trajectory(oracle_mod, o2, Y2, dt, 1e-8, 1e-18)
    fine2 = trajectory(oracle_mod, o2, Y2, dt, 1e-11, 1e-21)
    w2 = compare(o2, name, RUNS[name][0], traj2, fine2)
    print(f"{name} with the current gas constant: {w2:.2f}")
    assert w2 > 1.0


# The tester's runs used the options file inputs/base_plugin.yml with, per run,
# dense/analytic/iterative set by run.sh.in: "dense" = CVODE's dense direct
# solver with its difference-quotient Jacobian, "sparse" = SPGMR (eps_lin 0.1)
# preconditioned by the thresholded analytic Newton matrix.  Their timing logs
# (reactor_timing_log) hold the CVode call count of every solve
# (n_steps_avg: all 32 reactors are identical).
SOLVER = {"dense": dict(dense=1, analytic=0, iterative=0),
          "sparse": dict(dense=0, analytic=1, iterative=1)}
THRESHOLD = {"h2": 3.2e-5, "dme": 6.4e-5, "nc7h16-skel": 1.28e-4}  # run.sh.in


def solve_counts(oracle_mod, name, solver):
    mech = (os.path.join(TESTER, f"{name}.mech"), os.path.join(TESTER, f"{name}.therm"))
    o = oracle_mod.Oracle(*mech, ru=RU_2014)
    Y0 = initial_mass_fractions(o, name, mech)
    dt = 2.0 * IDT[name] / 10.0
    oo = oracle_mod.default_opts(rtol=1e-8, atol=1e-18, constant_volume=1, max_steps=3000,
                                 num_retries=3, delta_temperature_ignition=400.0, eps_lin=0.1,
                                 preconditioner_threshold=THRESHOLD[name], **SOLVER[solver])
    T, P, Y = np.array([1450.0]), np.array([2.02650e6]), Y0[None].copy()
    out = []
    for _ in range(10):
        T, Y, st = o.integrate(T, P, Y, dt, oo)
        P = st["P"]
        out.append(int(st["nsteps"][0]))
    return np.array(out)


def reference_counts(run):
    rows = [l.split() for l in open(os.path.join(TESTER, f"{run}_log.txt")) if l[0] != "#"]
    return np.array([int(float(r[5])) for r in rows])


@pytest.mark.parametrize("run", ["h2_dense", "h2_sparse", "dme_dense", "dme_sparse",
                                 "nc7h16-skel_sparse"])
def test_oracle_solver_modes_take_the_reference_steps(oracle_mod, run):
    """The oracle's restatements of the reference's two solver configurations
    take the reference's own number of CVode calls per solve: every solve
    within max(4, 12 %) of the logged count and the ten-solve total within
    5 % (the SPGMR runs match 5 to 9 of the 10 solves exactly; the largest
    differences are in the solve that ignites, where step selection reacts to
    round-off, and the logs come from an unstated SUNDIALS release and
    LAPACK/SuperLU build).  The same scenario in the other configuration
    misses the log by far more, so the counts tell the modes apart."""
    name, solver = run.rsplit("_", 1)
    ref = reference_counts(run)
    got = solve_counts(oracle_mod, name, solver)
    print(run, "oracle", got.tolist(), "reference", ref.tolist())
    assert np.all(np.abs(got - ref) <= np.maximum(4, 0.12 * ref)), (got, ref)
    assert abs(got.sum() - ref.sum()) <= 0.05 * ref.sum()
    other = solve_counts(oracle_mod, name, "sparse" if solver == "dense" else "dense")
    assert abs(other.sum() - ref.sum()) > 0.1 * ref.sum() or \
        np.any(np.abs(other - ref) > np.maximum(4, 0.12 * ref)), other
