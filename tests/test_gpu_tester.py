"""The reference tester's scenario through the drop-in plugin on the GPU
(examples/cfd_plugin_tester/run.sh.in with inputs/base_tester.yml and
base_plugin.yml): 32 identical constant-volume reactors, 1450 K, 2.0265e6 Pa,
stoichiometric fuel/air, 10 solves of 2 x the ignition delay / 10, rtol 1e-8,
atol 1e-18, ignition root at +400 K, with the reference's H2 (10 species:
one-wavefront FP64 integrator), DME (79 species) and n-heptane (159 species:
two- and three-wavefront teams, FP32 Newton matrices).

Reactor 0's T, p and mass fractions after each solve are compared with the
oracle's trajectory (current gas constant; the oracle itself is pinned to the
reference tester's printed outputs in test_oracle_tester.py): within five
times the oracle's global-error estimate (its rtol 1e-8 run against rtol
1e-11) plus 1e-7 relative and an absolute 1e-15 floor.  All 32 reactors end
bitwise equal, every solve returns success, and the IGNITION_TIME field holds
the time of the solve in which T crossed T0 + 400 K.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_oracle_tester import IDT, TESTER, initial_mass_fractions, trajectory

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["h2", "dme", "nc7h16-skel"])
def test_plugin_reproduces_tester_trajectory(torch_gpu, oracle_mod, name):
    from zerork_hip import capi
    mech = (os.path.join(TESTER, f"{name}.mech"), os.path.join(TESTER, f"{name}.therm"))
    o = oracle_mod.Oracle(*mech)
    Y0 = initial_mass_fractions(o, name, mech)
    dt = 2.0 * IDT[name] / 10.0
    traj = trajectory(oracle_mod, o, Y0, dt, 1e-8, 1e-18)
    fine = trajectory(oracle_mod, o, Y0, dt, 1e-11, 1e-21)
    print("case tester", name, "nsp", o.nsp, flush=True)
    p = capi.Plugin()
    assert p.set_mechanism_files(*mech) == 0
    for k, v in (("rel_tol", 1e-8), ("abs_tol", 1e-18), ("max_steps", 3000),
                 ("delta_temperature_ignition", 400.0)):
        assert (p.set_int(k, v) if isinstance(v, int) else p.set_double(k, v)) == 0, k
    assert p.load_mechanism() == 0
    n = 32
    T = np.full(n, 1450.0)
    P = np.full(n, 2.02650e6)
    Y = np.tile(Y0, (n, 1))
    ign = np.zeros(n)
    assert p.set_aux("IGNITION_TIME", ign) == 0
    worst = 0.0
    ign_solves = []
    for s in range(10):
        T0 = T.copy()
        assert p.solve(s, s * dt, dt, T, P, Y) == 0
        assert np.all(T == T[0]) and np.all(Y == Y[0]) and np.all(P == P[0])
        v = np.concatenate([[T[0], P[0]], Y[0]])
        vo = np.concatenate([[traj[s][0], traj[s][1]], traj[s][2]])
        vf = np.concatenate([[fine[s][0], fine[s][1]], fine[s][2]])
        bound = 5.0 * np.abs(vo - vf) + 1e-7 * np.abs(vo) + 1e-15
        r = float(np.max(np.abs(v - vo) / bound))
        worst = max(worst, r)
        print(f"  solve {s}: T {T[0]:.6f} (oracle {traj[s][0]:.6f}); worst |gpu - oracle| / bound "
              f"{r:.2f}", flush=True)
        if T[0] - T0[0] > 400.0:
            ign_solves.append(s)
            assert 0.0 < ign[0] <= dt
    assert ign_solves, "the mixture ignites within the run"
    assert worst <= 1.0
