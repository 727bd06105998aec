"""Pin the oracle's point-implicit chemistry (reaction_model.py:78-193)."""
import copy

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle.mech import load_mechanism
from oracle.reaction import Reaction
from oracle.thermo import Thermo

RU = 8.314463
SPECIES = ["H2", "O2", "OH", "O", "H", "H2O", "HO2", "H2O2", "N2"]
ATOMS = {"H2": (2, 0, 0), "O2": (0, 2, 0), "OH": (1, 1, 0), "O": (0, 1, 0), "H": (1, 0, 0),
         "H2O": (2, 1, 0), "HO2": (1, 2, 0), "H2O2": (2, 2, 0), "N2": (0, 0, 2)}


@pytest.fixture(scope="module")
def chem(mech_path):
    rp, th, nd = load_mechanism(mech_path)
    return Reaction(rp, th, nd), Thermo(th), rp, th, nd


def _state(n=16, seed=0, Tlo=900.0, Thi=3000.0):
    rng = np.random.default_rng(seed)
    T = rng.uniform(Tlo, Thi, (1, n))
    rho = rng.uniform(0.2, 3.0, (1, n))
    Y = rng.dirichlet(np.ones(9) * 0.7, n).T
    return T, rho, Y


def test_element_conservation(chem):
    r, thermo, rp, th, _ = chem
    T, rho, Y = _state()
    X = rho * Y / th["species_M"][:, None]
    w, _ = r.rates_and_jacobian(T, X)
    M = th["molecular_weight"][:8]
    for e in range(2):
        a = np.array([ATOMS[s][e] for s in SPECIES[:8]])
        net = np.sum((a / M)[:, None] * w, axis=0)
        scale = np.sum(np.abs((a / M)[:, None] * w), axis=0)
        assert np.all(np.abs(net) <= 1e-12 * scale + 1e-300)
    # mass: sum_k w_k = 0
    assert np.all(np.abs(w.sum(axis=0)) <= 1e-12 * np.abs(w).sum(axis=0))


def _gibbs_RT(raw_lo, raw_hi, k, T):
    a = raw_lo[k] if T < 1000 else raw_hi[k]
    h_RT = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    s_R = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return h_RT - s_R


@pytest.mark.parametrize("i", range(19))
def test_detailed_balance(chem, i):
    """With only reaction i active, a state whose quotient equals
    Kc = exp(-dG/RT) (P_atm/(Ru T))^dnu has zero net rate."""
    r, thermo, rp, th, nd = chem
    rp1 = copy.deepcopy(rp)
    keep = np.zeros(19)
    keep[i] = 1.0
    rp1["A"] = rp["A"] * keep
    r1 = Reaction(rp1, th, nd)
    T = np.array([[1700.0]])
    dG = sum((rp["vb"][i, k] - rp["vf"][i, k]) * _gibbs_RT(th["raw_low"], th["raw_high"], k, 1700.0)
             for k in range(8))
    dnu = rp["vsum"][i]
    Kc = np.exp(-dG) * (101325.0 / (RU * 1700.0)) ** dnu      # mol/m^3 units
    C = np.full(9, 5.0)                                       # mol/m^3
    prods = [k for k in range(8) if rp["vb"][i, k] > 0]
    # adjust one product concentration so that Q = Kc
    Q = np.prod(C[:8] ** rp["vb"][i]) / np.prod(C[:8] ** rp["vf"][i])
    kp = prods[0]
    C[kp] *= (Kc / Q) ** (1.0 / rp["vb"][i, kp])
    X = (C * RU)[:, None]                                     # X = rho Y / M = Ru * C
    w, _ = r1.rates_and_jacobian(T, X)
    # compare with the size of the forward rate alone
    rp2 = copy.deepcopy(rp1)
    r2 = Reaction(rp2, th, nd)
    kf_only = np.abs(w).max()
    X2 = X.copy()
    X2[kp] *= 10.0
    w2, _ = r2.rates_and_jacobian(T, X2)
    assert np.abs(w).max() <= 1e-10 * np.abs(w2).max()


def test_jacobian_vs_central_differences(chem):
    r, thermo, rp, th, _ = chem
    T, rho, Y = _state(n=6, seed=4, Tlo=1200.0)
    Y = np.maximum(Y, 1e-4)
    X = rho * Y / th["species_M"][:, None]
    w, J = r.rates_and_jacobian(T, X)
    M = th["species_M"]
    for n in range(8):
        h = 1e-6 * X[n]
        Xp, Xm = X.copy(), X.copy()
        Xp[n] += h
        Xm[n] -= h
        wp, _ = r.rates_and_jacobian(T, Xp)
        wm, _ = r.rates_and_jacobian(T, Xm)
        fd = (wp - wm) / (2 * h) / M[n]                     # d w / d(rho Y_n)
        scale = np.abs(J).max(axis=(0, 1))
        assert np.all(np.abs(fd - J[:, n]) <= 2e-6 * scale + 1e-6 * np.abs(fd)), n


def test_small_dt_limit_is_explicit_rate(chem):
    r, thermo, rp, th, _ = chem
    T, rho, Y = _state(n=8, seed=5)
    U = np.concatenate([rho, 0 * rho, 0 * rho, 0 * rho, rho * Y[:8]])
    aux = np.concatenate([1.3 + 0 * T, T])
    X = rho * Y / th["species_M"][:, None]
    w, J = r.rates_and_jacobian(T, X)
    # dt such that dt*|J| ~ 1e-7: (I/dt - J)^-1 w = dt w (1 + O(dt |J|))
    dt = 1e-7 / np.abs(J).max()
    S = r.detailed_reaction(U, aux, dt, thermo.fill_Y)
    np.testing.assert_allclose(S[4:], dt * w, rtol=1e-5, atol=1e-30)
    assert np.all(S[:4] == 0)


def test_point_implicit_vs_bdf(chem):
    """Linearised backward Euler is first order: for a dt well below the
    chemical time scale it tracks a tight BDF integration of the same rates
    (T held fixed, as in the reference's split step)."""
    r, thermo, rp, th, _ = chem
    T = np.array([[1500.0]])
    rho = np.array([[0.5]])
    Y9 = np.array([0.028, 0.226, 1e-3, 1e-4, 1e-4, 1e-2, 1e-5, 1e-6, 0.0])
    Y9[-1] = 1 - Y9[:-1].sum()
    rY0 = (rho[0, 0] * Y9[:8]).copy()
    M = th["species_M"]

    def rhs(t, rY):
        Y = np.concatenate([rY / rho[0, 0], [1 - rY.sum() / rho[0, 0]]])
        X = (rho[0, 0] * Y / M)[:, None]
        w, _ = r.rates_and_jacobian(T, X)
        return w[:, 0]

    U = np.concatenate([rho, [[0]], [[0]], [[0]], rY0[:, None]])
    aux = np.array([[1.3], [1500.0]])
    errs = []
    for dt in (1e-10, 1e-9):
        sol = solve_ivp(rhs, (0, dt), rY0, method="BDF", rtol=1e-11, atol=1e-18)
        S = r.detailed_reaction(U, aux, dt, thermo.fill_Y)
        d_bdf = sol.y[:, -1] - rY0
        errs.append(np.max(np.abs(S[4:, 0] - d_bdf)) / np.max(np.abs(d_bdf)))
    # the fastest eigenvalue of J here is ~3e7 1/s: dt*lambda = 3e-3 at 1e-10
    assert errs[0] < 5e-3
    assert 5.0 < errs[1] / errs[0] < 20.0          # first order in dt


def test_clip_bounds_increment(chem):
    r, thermo, rp, th, _ = chem
    T, rho, Y = _state(n=32, seed=6, Tlo=2500, Thi=3200)
    U = np.concatenate([rho, 0 * rho, 0 * rho, 0 * rho, rho * Y[:8]])
    aux = np.concatenate([1.3 + 0 * T, T])
    S = r.detailed_reaction(U, aux, 1e-3, thermo.fill_Y)    # huge dt: clip active
    Ynew = (U[4:] + S[4:]) / rho
    assert np.all(Ynew >= -1e-15) and np.all(Ynew <= 1 + 1e-15)
