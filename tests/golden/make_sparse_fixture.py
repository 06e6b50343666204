"""Writes tests/golden/sparse_2000sp_fixture.npz: end states of the C oracle
(FP64 CVODE restatement, dense LU with partial pivoting) for a few reactors
of the seeded 2,000-species workload over its full dt = 1e-4 s, at rtol 1e-6,
atol 1e-12, constant pressure; and the same run with every initial
temperature moved up by one ulp, which measures the oracle's own round-off
sensitivity of those end states.  The oracle needs minutes per reactor at
this size, so the GPU test of the sparse path reads these vectors instead of
running it.  Run in the build container (about 7 minutes on 8 threads):

    python tests/golden/make_sparse_fixture.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "zero-rk_amd"))

from zerork_hip import mechgen  # noqa: E402

SEED, N, DT, RTOL, ATOL = 62, 4, 1.0e-4, 1.0e-6, 1.0e-12


def main():
    import oracle
    mp = os.path.join(HERE, "mech", "synth_2000sp.mech")
    tp = os.path.join(HERE, "mech", "synth_2000sp.therm")
    o = oracle.Oracle(mp, tp)
    st = mechgen.workload_states("2000sp", mechgen.workload_mechanism("2000sp"), N, seed=SEED)
    opts = oracle.default_opts(rtol=RTOL, atol=ATOL)
    T0 = np.concatenate([st.T, np.nextafter(st.T, np.inf)])
    P0 = np.concatenate([st.P, st.P])
    Y0 = np.concatenate([st.Y, st.Y])
    T, Y, s = o.integrate(T0, P0, Y0, DT, opts, nthreads=2 * N)
    np.savez_compressed(os.path.join(HERE, "sparse_2000sp_fixture.npz"),
                        seed=np.array(SEED), dt=np.array(DT), rtol=np.array(RTOL),
                        atol=np.array(ATOL), T0=st.T, P0=st.P, Y0=st.Y,
                        T=T[:N], Y=Y[:N], nsteps=s["nsteps"][:N],
                        T_ulp=T[N:], Y_ulp=Y[N:], nsteps_ulp=s["nsteps"][N:])
    print("steps", s["nsteps"], "T", T0[:N], "->", T[:N])


if __name__ == "__main__":
    main()
