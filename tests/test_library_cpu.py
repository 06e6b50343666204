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
                 "constant_volume: 0\n")
    assert p.read_options_file(str(f)) == 0
    assert p.get_double("rel_tol") == (0, 1e-7) and p.get_int("max_steps") == (0, 321)
    g = tmp_path / "bad.yml"
    g.write_text("max_steps: many\n")
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
