"""Pin the oracle's NASA-7 thermo and Halley solve (thermo_model.py:80-190)."""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle.mech import load_mechanism
from oracle.thermo import Thermo

RU = 8.314463


@pytest.fixture(scope="module")
def th(mech_path):
    _, t, _ = load_mechanism(mech_path)
    return Thermo(t), t


def _state(n=40, seed=0):
    rng = np.random.default_rng(seed)
    T = rng.uniform(250.0, 3300.0, (1, n))
    Y = rng.dirichlet(np.ones(9), n).T                     # (9, n)
    return T, Y[:8]


def _e_closed_form(raw_lo, raw_hi, M, T, Y9):
    """e = sum_k Y_k (h_k - R_k T) straight from the NASA-7 a1..a6 form."""
    out = 0.0
    for k in range(9):
        a = raw_lo[k] if T < 1000.0 else raw_hi[k]
        h_RT = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
        Rk = RU * 1000 / M[k]
        out += Y9[k] * (h_RT - 1.0) * Rk * T
    return out


def test_energy_matches_closed_form(th):
    thermo, t = th
    T, Y = _state()
    cp, gamma, h, R, dcp = thermo.get_thermo(T, Y)
    Y9 = thermo.fill_Y(Y)
    for j in range(T.shape[1]):
        e_ref = _e_closed_form(t["raw_low"], t["raw_high"], t["molecular_weight"], T[0, j], Y9[:, j])
        assert h[0, j] - R[0, j] * T[0, j] == pytest.approx(e_ref, rel=1e-12, abs=1e-6)
    assert np.all(gamma > 1.0) and np.all(gamma < 1.7)


def test_halley_recovers_T_and_matches_brentq(th):
    thermo, t = th
    T, Y = _state(seed=1)
    cp, gamma, h, R, _ = thermo.get_thermo(T, Y)
    e = h - R * T
    T0 = T * (1.0 + 0.15 * np.sin(np.arange(T.shape[1])))
    aux = thermo.get_T(e, Y, T0)
    np.testing.assert_allclose(aux[1], T[0], rtol=1e-12)
    np.testing.assert_allclose(aux[0], gamma[0], rtol=1e-12)
    Y9 = thermo.fill_Y(Y)
    for j in range(0, T.shape[1], 7):
        f = lambda x: _e_closed_form(t["raw_low"], t["raw_high"], t["molecular_weight"], x,
                                     Y9[:, j]) - e[0, j]
        root = brentq(f, 150.0, 4000.0, xtol=1e-12)
        assert aux[1, j] == pytest.approx(root, rel=1e-10)


def test_global_and_per_cell_stop_agree(th):
    thermo, _ = th
    T, Y = _state(n=200, seed=2)
    cp, gamma, h, R, _ = thermo.get_thermo(T, Y)
    e = h - R * T
    T0 = T * 1.05
    a = thermo.get_T(e, Y, T0)
    b = thermo.get_T(e, Y, T0, per_cell=True)
    np.testing.assert_allclose(a, b, rtol=1e-14)


def test_get_T_bwd_matches_finite_difference(th):
    thermo, _ = th
    T, Y = _state(n=12, seed=3)
    cp, gamma, h, R, _ = thermo.get_thermo(T, Y)
    e = h - R * T
    aux = thermo.get_T(e, Y, T)
    g = np.array([[0.7] * 12, [1.3] * 12])
    dL_de, dL_dY, dL_dT0 = thermo.get_T_bwd(e, Y, aux, g)
    assert np.all(dL_dT0 == 0)
    eps = 10.0
    ap = thermo.get_T(e + eps, Y, T)
    am = thermo.get_T(e - eps, Y, T)
    fd = np.sum(g * (ap - am), axis=0) / (2 * eps)
    np.testing.assert_allclose(dL_de[0], fd, rtol=1e-6)
    for k in (0, 5):
        Yp, Ym = Y.copy(), Y.copy()
        Yp[k] += 1e-5
        Ym[k] -= 1e-5
        fd = np.sum(g * (thermo.get_T(e, Yp, T) - thermo.get_T(e, Ym, T)), axis=0) / 2e-5
        np.testing.assert_allclose(dL_dY[k], fd, rtol=1e-5)
