"""Pin the oracle's mechanism preprocessing against closed forms and tables.

The reference (preprocess/load.py:8-135) reads these through Cantera, which
is absent; the checks below fix the unit conversions, the coefficient layout
(flip/roll of the NASA-7 arrays) and the molar masses independently.
"""
import numpy as np
import pytest

from oracle.mech import load_mechanism

RU = 8.314463


@pytest.fixture(scope="module")
def mech(mech_path):
    return load_mechanism(mech_path)


def test_species_order_and_inert(mech):
    rp, th, nd = mech
    assert rp["species"] == ["H2", "O2", "OH", "O", "H", "H2O", "HO2", "H2O2", "N2"]
    assert rp["num_of_species"] == 9 and rp["num_of_inert_species"] == 1
    assert rp["vf"].shape == (19, 8)


def test_molecular_weights(mech):
    _, th, _ = mech
    # IUPAC standard atomic weights H 1.008, O 15.999, N 14.007 (g/mol)
    expect = [2.016, 31.998, 17.007, 15.999, 1.008, 18.015, 33.006, 34.014, 28.014]
    np.testing.assert_allclose(th["molecular_weight"], expect, rtol=0, atol=1e-12)
    np.testing.assert_allclose(th["species_M"] * 1000 * RU, expect, rtol=1e-15)


def test_rate_constants_closed_form(mech):
    rp, _, _ = mech
    # A_janc = A_cgs * (1e-6)^(order-1) * M0^(1 - sum(nu'))   (load.py:47,56)
    A_cgs = {0: (1.0e14, 2, 2), 2: (1.8e10, 2, 2), 5: (2.2e22, 3, 2), 6: (6.4e17, 3, 2),
             17: (1.2e17, 2, 1), 18: (6.0e13, 3, 2)}
    for i, (a, order, vfs) in A_cgs.items():
        expect = a * 1e-6 ** (order - 1) * RU ** (1 - vfs)
        assert rp["A"][i] == pytest.approx(expect, rel=1e-14), i
    # Ea/Ru = Ea_cal * 4.184 / Ru
    for i, ea in ((0, 5.6e4), (1, 1.68e4), (8, -1000.0), (18, -1800.0), (5, 0.0)):
        assert rp["Ea/Ru"][i] == pytest.approx(ea * 4.184 / RU, rel=1e-14, abs=0)
    assert list(rp["B"][[2, 5, 6, 7]]) == [1.0, -2.0, -1.0, -0.6]
    # three-body flags and efficiencies (x rho0/M0 = 1/Ru)
    assert list(np.nonzero(rp["is_third_body"])[0]) == [5, 6, 7, 8, 17, 18]
    eff = rp["third_body_coeffs"] * RU
    assert eff[5, 5] == pytest.approx(6.0) and eff[5, 0] == pytest.approx(1.0)
    assert eff[6, 0] == pytest.approx(2.0) and eff[17, 5] == pytest.approx(15.0)
    assert np.all(eff[:5] == 0)
    np.testing.assert_array_equal(rp["vsum"][[0, 5, 17]], [0.0, -1.0, 1.0])


def _molar(th, key, sp, T):
    k = ["H2", "O2", "OH", "O", "H", "H2O", "HO2", "H2O2", "N2"].index(sp)
    r = "low" if T < th["Tcr"][k] else "high"
    return np.polyval(th[key + "_" + r][k], T), k, r


@pytest.mark.parametrize("sp,cp", [("N2", 29.12), ("O2", 29.38), ("H2", 28.84),
                                   ("H2O", 33.58), ("H", 20.786)])
def test_nasa7_cp_298(mech, sp, cp):
    _, th, _ = mech
    val, k, _ = _molar(th, "cp", sp, 298.15)
    assert val * th["molecular_weight"][k] / 1000 == pytest.approx(cp, rel=3e-3)


@pytest.mark.parametrize("sp,hf", [("H2O", -241.826), ("H", 217.998), ("O", 249.18),
                                   ("H2", 0.0), ("O2", 0.0), ("N2", 0.0)])
def test_nasa7_formation_enthalpy(mech, sp, hf):
    _, th, _ = mech
    val, k, r = _molar(th, "h", sp, 298.15)
    h_molar = val * th["molecular_weight"][k] / 1000 / 1000       # kJ/mol
    assert h_molar == pytest.approx(hf, abs=0.3)
    # the chem (per mole / R) form is the same polynomial
    hc = np.polyval(th["h_chem_" + r][k], 298.15) * RU / 1000
    assert hc == pytest.approx(h_molar, rel=1e-12, abs=1e-9)


@pytest.mark.parametrize("sp,s0", [("H2O", 188.83), ("O2", 205.15), ("H2", 130.68),
                                   ("N2", 191.61), ("H", 114.72)])
def test_nasa7_entropy(mech, sp, s0):
    _, th, _ = mech
    k = ["H2", "O2", "OH", "O", "H", "H2O", "HO2", "H2O2", "N2"].index(sp)
    T = 298.15
    s_R = np.polyval(th["s_low"][k], T) + th["logcof_low"][k] * np.log(T)
    assert s_R * RU == pytest.approx(s0, abs=0.15)


def test_dcp_is_derivative(mech):
    _, th, _ = mech
    for r in ("low", "high"):
        for k in range(9):
            d = np.polyder(th["cp_" + r][k])
            np.testing.assert_allclose(th["dcp_" + r][k], d, rtol=1e-15)
            dh = np.polyder(th["h_" + r][k])
            np.testing.assert_allclose(dh, th["cp_" + r][k], rtol=1e-14)
