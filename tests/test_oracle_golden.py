"""Pins the C oracle against golden vectors of the reference rate core.

The vectors in tests/golden/*.npz were produced by the reference's own
src/zerork + src/CKconverter (tests/golden/make_golden.py): NASA-7 thermo,
concentrations, rate constants, limited and unlimited rates of progress and
species production, on seeded states of synthetic mechanisms and of the
reference's H2, DME, hydrogen_no_falloff and troe_test mechanisms.
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, load_golden, mech_paths

REL = 1.0e-12


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_tables_match_reference(name, oracle_mod):
    g = load_golden(name)
    o = oracle_mod.Oracle(*mech_paths(name))
    t = o.tables
    assert np.array_equal(t.mw, g["mw"])
    assert np.array_equal(t.step_rxn, g["step_rxn"])
    assert np.array_equal(t.reac_ptr, g["reac_ptr"]) and np.array_equal(t.reac_sp, g["reac_sp"])
    assert np.array_equal(t.prod_ptr, g["prod_ptr"]) and np.array_equal(t.prod_sp, g["prod_sp"])


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_thermo_and_rates_match_reference(name, oracle_mod):
    g = load_golden(name)
    o = oracle_mod.Oracle(*mech_paths(name))
    for i in range(len(g["T"])):
        T, P, Y = g["T"][i], g["P"][i], g["Y"][i]
        cp, h, gg = o.thermo(T)
        np.testing.assert_allclose(cp, g["cp_R"][i], rtol=1e-15, atol=0)
        np.testing.assert_allclose(h, g["h_RT"][i], rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(gg, g["g_RT"][i], rtol=1e-14, atol=1e-14)
        C = o.conc(T, P, Y)
        np.testing.assert_allclose(C, g["conc"][i], rtol=1e-14, atol=0)
        # rates from the reference's own concentrations: isolates rate parity
        Cg = g["conc"][i]
        rop, cr, de, net = o.rates(T, Cg, 1.0e22)
        np.testing.assert_allclose(rop, g["rop"][i], rtol=REL, atol=0)
        np.testing.assert_allclose(cr, g["create"][i], rtol=REL, atol=0)
        np.testing.assert_allclose(de, g["destroy"][i], rtol=REL, atol=0)
        scale = np.maximum(np.abs(g["create"][i]), np.abs(g["destroy"][i]))
        assert np.all(np.abs(net - g["net"][i]) <= REL * scale)
        rop2, _, _, net2 = o.rates(T, Cg, 0.0)  # limiter off
        np.testing.assert_allclose(rop2, g["rop_nolimit"][i], rtol=REL, atol=0)
        # forward/reverse rate constants per reaction (getKrxnFromTC)
        K = o.rate_constants(T, Cg)
        kf = np.zeros(o.tables.nrxn)
        kr = np.zeros(o.tables.nrxn)
        for s, (rx, d) in enumerate(zip(o.tables.step_rxn, o.tables.step_dir)):
            (kf if d == 1 else kr)[rx] = K[s]
        np.testing.assert_allclose(kf, g["kfwd"][i], rtol=REL, atol=0)
        np.testing.assert_allclose(kr, g["krev"][i], rtol=REL, atol=0)


def test_troe_closed_form(oracle_mod):
    """Troe fall-off of tests/reactions/troe_gtest.cpp:114-175 restated in closed
    form (3- and 4-parameter) for the reference's troe_test mechanism."""
    o = oracle_mod.Oracle(*mech_paths("ref_troe_test"))
    g = load_golden("ref_troe_test")
    for i in range(len(g["T"])):
        T = g["T"][i]
        C = g["conc"][i]
        K = o.rate_constants(T, C)
        kinf = 2.0e12 * 1000.0 ** (1 - 1)
        klow = 2.0e9 * 1000.0 ** (-1)
        M = C.sum()
        Pr = klow * M / kinf
        for s, four in ((0, False), (1, True)):
            a, t3, t1 = 0.68, 78.0, 1995.0
            fc = (1 - a) * np.exp(-T / t3) + a * np.exp(-T / t1)
            if four:
                fc += np.exp(-5590.0 / T)
            c = -0.4 - 0.67 * np.log10(fc)
            n = 0.75 - 1.27 * np.log10(fc)
            f1 = (np.log10(Pr) + c) / (n - 0.14 * (np.log10(Pr) + c))
            F = 10.0 ** (np.log10(fc) / (1 + f1 * f1))
            expect = kinf * Pr / (1 + Pr) * F
            assert abs(K[s] - expect) <= 1e-13 * expect
