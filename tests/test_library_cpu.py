"""CPU checks of the product library (no GPU calls):

* every symbol declared in include/*.h is exported by libzerork_cfd_plugin.so;
* the library's CHEMKIN reader produces the reference step structure and
  molecular weights (compared with the golden vectors of the reference core);
* plugin option and status semantics of zerork_cfd_plugin.cpp:21-216 /
  zerork_reactor_manager.cpp:50-101,135-191.
"""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT, load_golden, mech_paths


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(zerork_reactor_\w+|zrk_\w+)\s*\(", txt)))


@pytest.mark.parametrize("header", sorted(h for h in os.listdir(os.path.join(ROOT, "include"))
                                    if h.endswith(".h")))
def test_exports(product_lib, header):
    names = _declared(header)
    assert names
    for n in names:
        assert hasattr(product_lib, n), n


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_parser_matches_reference(product_lib, name):
    from zerork_hip import capi
    g = load_golden(name)
    m = capi.Mech(*mech_paths(name))
    assert m.nsp == len(g["mw"]) and m.nstep == len(g["step_rxn"])
    np.testing.assert_array_equal(m.molwt(), g["mw"])
    s = m.steps()
    np.testing.assert_array_equal(s["step_rxn"], g["step_rxn"])
    for k in ("reac_ptr", "reac_sp", "prod_ptr", "prod_sp"):
        np.testing.assert_array_equal(s[k], g[k])


def test_parser_errors(product_lib, tmp_path):
    from zerork_hip import capi
    with pytest.raises(RuntimeError):
        capi.Mech(str(tmp_path / "missing.mech"), str(tmp_path / "missing.therm"))
    bad = tmp_path / "bad.mech"
    bad.write_text("ELEMENTS\nH\nEND\nSPECIES\nH2 XX\nEND\nREACTIONS\nH2 = XX 1 0 0\nEND\n")
    therm = mech_paths("ref_h2")[1]
    with pytest.raises(RuntimeError, match="no thermo data"):
        capi.Mech(str(bad), therm)


def test_plugin_options_and_status(product_lib, tmp_path):
    from zerork_hip import capi
    p = capi.Plugin()
    # defaults of the reference manager
    assert p.get_int("max_steps") == (0, 5000)
    assert p.get_double("rel_tol") == (0, 1.0e-8)
    assert p.get_double("abs_tol") == (0, 1.0e-20)
    assert p.get_int("constant_volume") == (0, 1)
    assert p.get_int("no_such_option")[0] == 8  # OPTION_NOT_DEFINED
    assert p.set_int("max_steps", 77) == 0 and p.get_int("max_steps") == (0, 77)
    assert p.set_double("rel_tol", 1e-6) == 0 and p.get_double("rel_tol") == (0, 1e-6)
    assert p.set_string("mech_filename", "x.mech") == 0
    assert p.get_string("mech_filename") == (0, "x.mech")
    # options file (flat YAML, reference key names)
    f = tmp_path / "opts.yml"
    f.write_text("relative_tolerance: 1.0e-7\nabsolute_tolerance: 1e-14\nmax_steps: 321\n"
                 "constant_volume: 0\npreconditioner_threshold: 1e-4\n")
    assert p.read_options_file(str(f)) == 0
    assert p.get_double("rel_tol") == (0, 1e-7) and p.get_int("max_steps") == (0, 321)
    g = tmp_path / "bad.yml"
    g.write_text("max_steps: many\npreconditioner_threshold: 1e-4\n")
    assert p.read_options_file(str(g)) == 6  # FAILED_OPTIONS_PARSE
    # mechanism errors are status codes, never exceptions across the ABI
    assert p.set_mechanism_files(str(tmp_path / "nope.mech"), str(tmp_path / "nope.therm")) == 0
    assert p.load_mechanism() == 2  # INVALID_FILENAME
    bad = tmp_path / "bad.mech"
    bad.write_text("ELEMENTS\nH\nEND\nSPECIES\nH2\nEND\nREACTIONS\nH2 => H2 + ZZ 1 0 0\nEND\n")
    assert p.set_mechanism_files(str(bad), mech_paths("ref_h2")[1]) == 0
    assert p.load_mechanism() == 5  # FAILED_MECHANISM_PARSE
    m, t = mech_paths("synth_10sp")
    assert p.set_mechanism_files(m, t) == 0
    assert p.load_mechanism() == 0
    assert p.free() == 0
    L = capi.lib()
    assert L.zerork_reactor_solve(0, 0.0, 1e-4, 0, None, None, None, None) == 1  # INVALID_HANDLE
    assert L.zerork_reactor_free(None) == 1


def _therm_card(name, comp, coef_lines):
    """one 4-line NASA-7 card: name cols 1-18, element/count pairs cols
    25-44, phase col 45, T ranges cols 46-73, card number col 80"""
    els = "".join(f"{e:<2s}{n:>3d}" for e, n in comp).ljust(20)
    head = f"{name:<18s}{'000000':6s}{els}G{300.0:10.3f}{5000.0:10.3f}{1000.0:8.2f}"
    return [head.ljust(79) + "1"] + coef_lines


def test_element_table_beyond_light_elements(product_lib, oracle_mod, tmp_path):
    """Molecular weights of species made of heavy elements (Fe, Xe, U) are the
    element-count sums of the IUPAC 2007 masses that src/zerork/atomicMassDB.h
    tabulates (Fe 55.845, Xe 131.293, U 238.02891, O 15.9994), in both the
    product's reader and the oracle's; lower-case symbols match too."""
    from zerork_hip import capi
    h2 = open(mech_paths("ref_h2")[1]).read().splitlines()
    k = next(i for i, l in enumerate(h2) if l.lower().startswith("h2 "))
    coef = h2[k + 1:k + 4]
    cards = []
    for name, comp in (("FE", [("FE", 1)]), ("FEO", [("Fe", 1), ("O", 1)]), ("O", [("O", 1)]),
                       ("O2", [("O", 2)]), ("XE", [("XE", 1)]), ("UO2", [("u", 1), ("o", 2)])):
        cards += _therm_card(name, comp, coef)
    therm = tmp_path / "heavy.therm"
    therm.write_text("THERMO\n   300.000  1000.000  5000.000\n" + "\n".join(cards) + "\nEND\n")
    mech = tmp_path / "heavy.mech"
    mech.write_text("ELEMENTS\nFE O XE U\nEND\nSPECIES\nFE FEO O O2 XE UO2\nEND\n"
                    "REACTIONS\nFE + O2 <=> FEO + O   1.0E13 0.0 1000.0\n"
                    "O + O + XE <=> O2 + XE   1.0E15 0.0 0.0\nEND\n")
    expect = np.array([55.845, 55.845 + 15.9994, 15.9994, 2 * 15.9994, 131.293,
                       238.02891 + 2 * 15.9994])
    m = capi.Mech(str(mech), str(therm))
    np.testing.assert_allclose(m.molwt(), expect, rtol=1e-15)
    o = oracle_mod.Oracle(str(mech), str(therm))
    np.testing.assert_allclose(o.tables.mw, expect, rtol=1e-15)


def test_options_file_schema(product_lib, tmp_path):
    """ZeroRKReactorManager::ReadOptionsFile (zerork_reactor_manager.cpp:135-191)
    with the ZeroRKCFDPluginIFP schema (generateZeroRKCFDPluginIFP.py): the
    reference tester's own options file sets the dense DQ-Jacobian direct
    solver; keys it leaves out take the schema's defaults, not the manager
    constructor's; preconditioner_threshold is required; values outside a
    key's bounds or allowed set fail the parse and change nothing."""
    from zerork_hip import capi
    from conftest import GOLDEN
    p = capi.Plugin()
    assert p.get_int("dense") == (0, 0) and p.get_double("eps_lin") == (0, 1e-3)
    assert p.get_int("cvode_num_retries") == (0, 5)
    assert p.read_options_file(os.path.join(GOLDEN, "tester", "base_plugin.yml")) == 0
    for k, v in (("dense", 1), ("analytic", 0), ("iterative", 0), ("max_steps", 3000),
                 ("integrator", 0), ("gpu", 0), ("cvode_num_retries", 3),
                 ("n_reactors_max", 2048), ("n_reactors_min", 256), ("constant_volume", 1),
                 ("always_solve_temperature", 1), ("abstol_dens", 0)):
        assert p.get_int(k) == (0, v), k
    for k, v in (("rel_tol", 1e-8), ("abs_tol", 1e-18), ("preconditioner_threshold", 3.2e-5),
                 ("eps_lin", 0.1), ("nonlinear_convergence_coeff", 0.05), ("max_dt", 0.05),
                 ("solve_temperature_threshold", 2.0), ("step_limiter", 1e22)):
        assert p.get_double(k) == (0, v), k
    assert p.get_string("reactor_timing_log_filename") == (0, "reactor.log")
    assert p.get_string("mechanism_parsing_log_filename") == (0, "mech.cklog")
    for body in ("relative_tolerance: 1e-6\n",                                  # required key
                 "preconditioner_threshold: 1e-4\ndense: 2\n",                  # not 0/1
                 "preconditioner_threshold: 1e-4\ncvode_num_retries: 21\n",     # above 20
                 "preconditioner_threshold: 1e-4\ndelta_temperature_ignition: -1\n"):
        f = tmp_path / "x.yml"
        f.write_text(body)
        assert p.read_options_file(str(f)) == 6, body
    assert p.get_int("dense") == (0, 1) and p.get_double("rel_tol") == (0, 1e-8)
    p.free()
