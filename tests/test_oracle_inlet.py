"""Pins the oracle's injection inlet (oracle/inlet.py) against the
injector model's own conservation laws, an independent root solve, its
limits and the side/mirror symmetry."""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle.inlet import (A1, A2, A3, GAMMA_INJ, inj_model, make_injection_inlet,
                          stoichiometric_injection_Y)
from oracle.mech import load_mechanism
from oracle.thermo import Thermo

P0, T0 = 2.0e5, 300.0


@pytest.fixture(scope="module")
def thermo(mech_path):
    _, th, _ = load_mechanism(mech_path)
    return Thermo(th)


def _Y(thermo):
    return stoichiometric_injection_Y(len(thermo.Mex))


def _col(x):
    return np.asarray(x, dtype=np.float64).reshape(1, -1, 1)


def test_stoichiometric_injection_mixture(thermo):
    """Example 2's mixture: H2 : O2 = 1 : 8 by mass, O2 : N2 of air by mass
    (0.232 : 0.768), trace 1e-20 for the other reacting species."""
    Y = _Y(thermo)
    Y9 = np.concatenate([Y, [1 - Y.sum()]])
    assert Y9[0] / Y9[1] == pytest.approx(1.0 / 8.0, rel=1e-14)
    assert Y9[1] / (Y9[1] + Y9[-1]) == pytest.approx(0.232, rel=1e-12)
    assert np.all(Y[2:] == 1e-20) and Y9.sum() == pytest.approx(1.0, rel=1e-15)


def _model_pieces(M, p, R):
    g = GAMMA_INJ
    f = 1 + (g - 1) / 2 * M ** 2
    P1 = P0 * f ** (-g / (g - 1))
    V1 = M * f ** (-0.5) * np.sqrt(g * R * T0)
    MFC = A1 * P0 / np.sqrt(T0) * np.sqrt(g / R) * M * f ** (-(g + 1) / 2 / (g - 1))
    B = g / (g - 1) * p * A3 / MFC
    C = -g / (g - 1) * R * T0
    V3 = (-B + np.sqrt(B ** 2 - 2 * C)) / 1.0
    P2 = (MFC * (V3 - V1) - P1 * A1 + p * A3) / A2
    return P1, P2, V3, MFC


@pytest.mark.parametrize("ratio", [0.2, 0.5, 0.8, 0.95])
def test_injector_model_conservation_and_root(thermo, ratio):
    Y = _Y(thermo)
    p = _col([ratio * P0])
    Yc = Y.reshape(-1, 1, 1)
    V3, T, h, gamma = inj_model(p, Yc, P0, T0, thermo)
    R = float(thermo.get_R(Yc).ravel()[0])
    g = GAMMA_INJ
    rho = float(p.ravel()[0]) / (R * float(T.ravel()[0]))
    V = float(V3.ravel()[0])
    # energy: V^2/2 + g/(g-1) p/rho = g/(g-1) R T0 (the quadratic for V3)
    lhs = 0.5 * V ** 2 + g / (g - 1) * ratio * P0 / rho
    assert lhs == pytest.approx(g / (g - 1) * R * T0, rel=1e-12)
    # independent root of the momentum balance P2(M) = P1(M) (brentq)
    fM = lambda M: _model_pieces(M, ratio * P0, R)[1] - _model_pieces(M, ratio * P0, R)[0]  # noqa
    if fM(1.0 - 1e-12) * fM(1e-9) < 0:
        Mr = brentq(fM, 1e-9, 1.0 - 1e-12, xtol=1e-15)
        _, _, V3r, MFCr = _model_pieces(Mr, ratio * P0, R)
        assert V == pytest.approx(V3r, rel=2e-5)
        # mass: rho V3 A3 = MFC
        assert rho * V * A3 == pytest.approx(MFCr, rel=5e-5)
    else:   # no interior root: bisection pinned to the choked end M -> 1
        _, _, V3r, _ = _model_pieces(1.0 - 2.0 ** -20, ratio * P0, R)
        assert V == pytest.approx(V3r, rel=1e-12)


def test_injection_weakens_toward_manifold_pressure(thermo):
    Y = _Y(thermo).reshape(-1, 1, 1)
    V = [float(inj_model(_col([r * P0]), Y, P0, T0, thermo)[0].ravel()[0])
         for r in (0.3, 0.6, 0.9, 0.99)]
    assert all(a > b for a, b in zip(V, V[1:])) and V[-1] > 0


def _state(thermo, nx, p_wall, T=300.0, v=12.0, Y=None):
    Y = _Y(thermo) * 0.9 if Y is None else Y
    Yc = np.broadcast_to(Y.reshape(-1, 1, 1), (len(Y), nx, 3)).copy()
    Tf = np.full((1, nx, 3), T)
    R = thermo.get_R(Yc)
    p = np.broadcast_to(np.asarray(p_wall, dtype=np.float64).reshape(1, -1, 1), (1, nx, 3))
    rho = p / (R * Tf)
    cp, gamma, h, _, _ = thermo.get_thermo(Tf, Yc)
    u = np.full_like(rho, 3.0)
    vv = np.full_like(rho, v)
    U = np.concatenate([rho, rho * u, rho * vv, rho * h - p + 0.5 * rho * (u ** 2 + vv ** 2),
                        rho * Yc])
    return U, np.concatenate([gamma, Tf])


def test_blocked_and_open_cells(thermo):
    bc = make_injection_inlet(thermo, "bottom", P0, T0, _Y(thermo))
    U, aux = _state(thermo, 4, [0.5 * P0, 0.9 * P0, P0, 3 * P0])
    Ug, ag = bc(U, aux)
    assert Ug.shape == (U.shape[0], 4, 3) and ag.shape == (2, 4, 3)
    assert np.array_equal(Ug[:, :, 0], Ug[:, :, 1]) and np.array_equal(Ug[:, :, 0], Ug[:, :, 2])
    # open cells: no tangential velocity, inflow (+v), injected composition
    for i in (0, 1):
        assert Ug[1, i, 0] == 0.0 and Ug[2, i, 0] > 0.0
        np.testing.assert_allclose(Ug[4:, i, 0] / Ug[0, i, 0], _Y(thermo), rtol=1e-14)
    # blocked cells (p >= P0): wall cell with normal velocity negated
    for i in (2, 3):
        np.testing.assert_allclose(Ug[[0, 1, 3], i, 0], U[[0, 1, 3], i, 0], rtol=1e-14)
        assert Ug[2, i, 0] == pytest.approx(-U[2, i, 0], rel=1e-14)
        np.testing.assert_allclose(ag[:, i, 0], aux[:, i, 0], rtol=1e-14)
    # the ghost pressure is the wall cell's in both branches
    R = thermo.get_R(Ug[4:, :, 0:1] / Ug[0:1, :, 0:1])
    np.testing.assert_allclose(Ug[0:1, :, 0:1] * R * ag[1:2, :, 0:1],
                               [[[0.5 * P0], [0.9 * P0], [P0], [3 * P0]]], rtol=1e-12)


def test_side_symmetry(thermo):
    """The top inlet of the y-mirrored state is the mirror of the bottom one."""
    U, aux = _state(thermo, 5, [0.3 * P0, 0.6 * P0, 0.8 * P0, 1.2 * P0, 0.5 * P0], v=-7.0)
    gb = make_injection_inlet(thermo, "bottom", P0, T0, _Y(thermo))
    gt = make_injection_inlet(thermo, "top", P0, T0, _Y(thermo))
    Ub, ab = gb(U, aux)
    Um = U[:, :, ::-1].copy()
    Um[2] *= -1
    Ut, at = gt(Um, aux[:, :, ::-1].copy())
    Ut = Ut[:, :, ::-1].copy()
    Ut[2] *= -1
    np.testing.assert_allclose(Ut, Ub, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(at[:, :, ::-1], ab, rtol=1e-14)
    # left inlet of the transposed state = transpose of the bottom inlet
    gl = make_injection_inlet(thermo, "left", P0, T0, _Y(thermo))
    UT = U.transpose(0, 2, 1)[[0, 2, 1] + list(range(3, U.shape[0]))].copy()
    Ul, al = gl(UT, aux.transpose(0, 2, 1).copy())
    Ul = Ul.transpose(0, 2, 1)[[0, 2, 1] + list(range(3, U.shape[0]))]
    np.testing.assert_allclose(Ul, Ub, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(al.transpose(0, 2, 1), ab, rtol=1e-14)
