import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "zero-rk_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
MECHS = os.path.join(GOLDEN, "mech")
GOLDEN_CASES = ["synth_10sp", "synth_53sp", "synth_160sp", "ref_h2", "ref_dme",
                "ref_h2_no_falloff", "ref_troe_test", "ref_sri", "ref_plog", "ref_non_integer",
                "ref_ford_propane", "ref_rord_propane", "ref_h2_features", "synth_2000sp",
                "ref_ic8"]
# the reference's CHEMKIN-feature test mechanisms (SRI, PLOG, fractional
# stoichiometry, FORD, RORD) and the H2 mechanism with all of them added
FEATURE_CASES = ["ref_sri", "ref_plog", "ref_non_integer", "ref_ford_propane",
                 "ref_rord_propane", "ref_h2_features"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def mech_paths(name):
    return os.path.join(MECHS, name + ".mech"), os.path.join(MECHS, name + ".therm")


def plugin_solver_options(p):
    """The linear-solver options a plugin handle integrates with (the
    reference's defaults unless set), as keyword arguments of
    capi.default_options, so that a device-path call reproduces the plugin."""
    kw = {}
    for k in ("dense", "analytic", "iterative"):
        rc, v = p.get_int(k)
        assert rc == 0, k
        kw[k] = v
    for k in ("eps_lin", "preconditioner_threshold"):
        rc, v = p.get_double(k)
        assert rc == 0, k
        kw[k] = v
    return kw


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def product_lib():
    from zerork_hip import capi
    return capi.lib()


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch
