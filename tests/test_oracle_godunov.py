"""Pin the oracle's Godunov path: MUSCL reconstruction and the HLLC flux.

The checks come from what the mathematics fixes, not from the formulas
themselves: consistency F(q, q) = F(q) against the flux-splitting path's
physical flux, the HLL integral condition that fixes the contact speed and
the star states, exact resolution of an isolated contact, upwinding of
supersonic states, mirror symmetry, MUSCL's exactness on linear data and its
minmod limiting, and telescoping / uniform-flow preservation of the full
divergence.
"""
import numpy as np
import pytest

from oracle import flux as F
from oracle.step import Oracle
from janc_amd import cases
from janc_amd.mechanism import Mechanism

PER = {s + "_boundary": "periodic" for s in ("left", "right", "bottom", "top")}
GOD = {"solver_type": "godunov", "interface_reconstruction": "WENO5_JS",
       "riemann_solver": "HLLC"}


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


@pytest.fixture(scope="module")
def thermo(mech_path):
    return Oracle(mech_path, PER, 1.0, 1.0).thermo


def _states(mech, rng, n, T=(300.0, 2500.0), p=(5e4, 3e6), vel=300.0):
    """Random primitive states q = [rho, u, v, p, Y(8)] of shape (12, n, 1)."""
    Tv = rng.uniform(*T, size=(n, 1))
    pv = rng.uniform(*p, size=(n, 1))
    phi = rng.uniform(0.5, 1.5, size=(n, 1))
    Y9 = cases.premixed_Y(mech, phi)
    # move some mass into radicals and products so every species is present
    Y9[2:8] += rng.uniform(0.0, 0.01, size=(6, n, 1))
    Y9 /= Y9.sum(axis=0, keepdims=True)
    R = (Y9 / mech.species_M[:, None, None]).sum(axis=0)
    rho = pv / (R * Tv)
    u = rng.uniform(-vel, vel, size=(n, 1))
    v = rng.uniform(-vel, vel, size=(n, 1))
    return np.concatenate([rho[None], u[None], v[None], pv[None], Y9[:8]], axis=0)


def _conserved(q, thermo):
    rho, u, v, p, Y = q[0:1], q[1:2], q[2:3], q[3:4], q[4:]
    T = p / (thermo.get_R(Y) * rho)
    _, gamma, h, _, _ = thermo.get_thermo(T, Y)
    E = rho * h - p + 0.5 * rho * (u ** 2 + v ** 2)
    return np.concatenate([rho, rho * u, rho * v, E, rho * Y], axis=0), gamma, T


# ------------------------------------------------------------------ HLLC
@pytest.mark.parametrize("ixy", [1, 2])
def test_hllc_consistency_with_physical_flux(mech, thermo, ixy):
    """F_HLLC(q, q) equals the physical flux of the flux-splitting path."""
    q = _states(mech, np.random.default_rng(1), 64)
    U, gamma, T = _conserved(q, thermo)
    aux = np.concatenate([gamma, T], axis=0)
    Hp, Hm = F.split_flux_lf(ixy, U, aux, thermo)
    Fphys = Hp + Hm
    Fh = F.hllc_dir(q, q, thermo, ixy)
    scale = np.abs(Fphys).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(Fh - Fphys) <= 1e-12 * scale)


@pytest.mark.parametrize("ixy", [1, 2])
def test_hllc_integral_condition(mech, thermo, ixy):
    """(s_M-s_L) U*_L + (s_R-s_M) U*_R = s_R U_R - s_L U_L - (F_R - F_L).

    The HLL integral form of the conservation law over the Riemann fan; with
    it, F*_R - F*_L = s_M (U*_R - U*_L) (Rankine-Hugoniot across the contact).
    A wrong contact speed or star-state entry breaks it.
    """
    rng = np.random.default_rng(2)
    qL = _states(mech, rng, 200)
    qR = _states(mech, rng, 200)
    h = F.hllc_parts(qL, qR, thermo, ixy)
    sL, sM, sR = h["s_L"], h["s_M"], h["s_R"]
    lhs = (sM - sL) * h["UM_L"] + (sR - sM) * h["UM_R"]
    rhs = sR * h["U_R"] - sL * h["U_L"] - (h["F_R"] - h["F_L"])
    scale = (np.abs(sR * h["U_R"]) + np.abs(sL * h["U_L"]) + np.abs(h["F_R"])
             + np.abs(h["F_L"]))
    assert np.all(np.abs(lhs - rhs) <= 1e-10 * scale)
    jump = h["FM_R"] - h["FM_L"] - sM * (h["UM_R"] - h["UM_L"])
    assert np.all(np.abs(jump) <= 1e-9 * scale)
    # pressure is continuous across the contact: p* from either side agrees
    n = 1 if ixy == 1 else 2
    pL = qL[3:4] + qL[0:1] * (sL - qL[n:n + 1]) * (sM - qL[n:n + 1])
    pR = qR[3:4] + qR[0:1] * (sR - qR[n:n + 1]) * (sM - qR[n:n + 1])
    np.testing.assert_allclose(pL, pR, rtol=1e-9)
    assert np.all(sL <= sM) and np.all(sM <= sR)


@pytest.mark.parametrize("ixy", [1, 2])
def test_hllc_resolves_isolated_contact(mech, thermo, ixy):
    """Equal normal velocity and pressure: s_M = u_n and the flux is the
    upwind side's physical flux (HLLC keeps contacts sharp)."""
    rng = np.random.default_rng(3)
    qL = _states(mech, rng, 50, vel=100.0)
    qR = _states(mech, rng, 50, vel=100.0)
    n = 1 if ixy == 1 else 2
    qR[n] = qL[n]
    qR[3] = qL[3]
    h = F.hllc_parts(qL, qR, thermo, ixy)
    np.testing.assert_allclose(h["s_M"], qL[n:n + 1], rtol=1e-9, atol=1e-9)
    Fh = F.hllc_dir(qL, qR, thermo, ixy)
    up = qL[n:n + 1] > 0
    expect = np.where(up, h["F_L"], h["F_R"])
    # the star state equals the upwind state up to rounding, amplified by s
    scale = np.abs(expect) + np.maximum(np.abs(h["s_L"]), np.abs(h["s_R"])) * (
        np.abs(h["U_L"]) + np.abs(h["U_R"]))
    assert np.all(np.abs(Fh - expect) <= 1e-12 * scale)


def test_hllc_supersonic_upwinding(mech, thermo):
    rng = np.random.default_rng(4)
    qL = _states(mech, rng, 30, T=(300.0, 400.0), vel=1.0)
    qR = _states(mech, rng, 30, T=(300.0, 400.0), vel=1.0)
    for sign in (+1, -1):
        qa, qb = qL.copy(), qR.copy()
        qa[1] = qb[1] = sign * 2500.0                # Mach ~7 along x
        h = F.hllc_parts(qa, qb, thermo, 1)
        Fh = F.hllc_dir(qa, qb, thermo, 1)
        side = "F_L" if sign > 0 else "F_R"
        assert np.all(h["s_L"] > 0) if sign > 0 else np.all(h["s_R"] < 0)
        np.testing.assert_array_equal(Fh, h[side])


def test_hllc_mirror_symmetry(mech, thermo):
    """Reflecting x -> -x swaps the sides and negates u and the odd fluxes."""
    rng = np.random.default_rng(5)
    qL = _states(mech, rng, 100)
    qR = _states(mech, rng, 100)
    Fh = F.hllc_dir(qL, qR, thermo, 1)
    mL, mR = qR.copy(), qL.copy()
    mL[1] *= -1
    mR[1] *= -1
    Fm = F.hllc_dir(mL, mR, thermo, 1)
    sgn = -np.ones((12, 1, 1))
    sgn[1] = 1                                       # x-momentum flux is even
    scale = np.abs(Fh).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(Fm - sgn * Fh) <= 1e-10 * scale)
    # x and y faces agree after swapping u and v
    sw = [0, 2, 1] + list(range(3, 12))
    Gy = F.hllc_dir(qL[sw], qR[sw], thermo, 2)
    assert np.all(np.abs(Gy[sw] - Fh) <= 1e-14 * scale)


# ------------------------------------------------------------------ MUSCL
def test_muscl_exact_on_linear_data_and_index_windows():
    """Linear data: minmod keeps the slope, both sides equal the face value.
    Padded row i holds value i, so face f (between padded rows f+2 and f+3)
    must read f + 2.5 on both sides."""
    nx, ny = 9, 5
    q = np.zeros((2, nx + 6, ny + 6))
    q[0] = np.arange(nx + 6)[:, None] * 1.0
    q[1] = 3.0 - 0.5 * np.arange(ny + 6)[None, :]
    L, R = F.muscl_L_x(q), F.muscl_R_x(q)
    assert L.shape == (2, nx + 1, ny) and R.shape == (2, nx + 1, ny)
    np.testing.assert_allclose(L[0], (np.arange(nx + 1) + 2.5)[:, None] * np.ones(ny))
    np.testing.assert_allclose(R[0], L[0])
    Ly, Ry = F.muscl_L_y(q), F.muscl_R_y(q)
    assert Ly.shape == (2, nx, ny + 1)
    np.testing.assert_allclose(Ly[1], (3.0 - 0.5 * (np.arange(ny + 1) + 2.5))[None, :]
                               * np.ones((nx, 1)))
    np.testing.assert_allclose(Ry[1], Ly[1])


def _minmod(a, b):
    return 0.0 if a * b <= 0 else (a if abs(a) < abs(b) else b)


def test_muscl_minmod_limiting():
    """beta = 1: both limited slopes are minmod(d-, d+), so the face values
    are the classical minmod-limited linear reconstruction."""
    mm = _minmod
    rng = np.random.default_rng(6)
    q = rng.normal(size=(3, 14, 7))
    L, R = F.muscl_L_x(q), F.muscl_R_x(q)
    for f in range(L.shape[1]):
        for j in range(L.shape[2]):
            for v in range(3):
                col = q[v, :, j + 3]
                iL, iR = f + 2, f + 3
                sL = mm(col[iL] - col[iL - 1], col[iL + 1] - col[iL])
                sR = mm(col[iR] - col[iR - 1], col[iR + 1] - col[iR])
                assert L[v, f, j] == pytest.approx(col[iL] + 0.5 * sL, rel=1e-13, abs=1e-14)
                assert R[v, f, j] == pytest.approx(col[iR] - 0.5 * sR, rel=1e-13, abs=1e-14)


def test_muscl_is_bounded_by_neighbours():
    rng = np.random.default_rng(7)
    q = rng.normal(size=(1, 30, 7))
    L = F.muscl_L_x(q)[0]
    lo = np.minimum(q[0, 2:-3, 3:-3], q[0, 3:-2, 3:-3])
    hi = np.maximum(q[0, 2:-3, 3:-3], q[0, 3:-2, 3:-3])
    assert np.all(L >= lo - 1e-14) and np.all(L <= hi + 1e-14)


# ------------------------------------------------------------------ divergence
@pytest.mark.parametrize("recon", ["WENO5_JS", "MUSCL"])
def test_godunov_periodic_telescoping(mech_path, mech, recon):
    U, aux, geo = cases.periodic_box(mech, nx=20, ny=14)
    o = Oracle(mech_path, PER, geo["dx"], geo["dy"],
               flux_config=dict(GOD, interface_reconstruction=recon))
    rhs = o.flux_func(U, aux, 1.0)
    assert np.all(np.isfinite(rhs))
    tot = rhs[:, :-1, :-1].sum(axis=(1, 2))
    scale = np.abs(rhs).sum(axis=(1, 2))
    assert np.all(np.abs(tot) <= 1e-11 * scale)


@pytest.mark.parametrize("solver,recon", [("godunov", "WENO5_JS"), ("godunov", "MUSCL"),
                                          ("flux_splitting", "MUSCL")])
def test_uniform_flow_is_preserved(mech_path, mech, solver, recon):
    T = np.full((10, 9), 900.0)
    p = np.full((10, 9), 2e5)
    Y9 = cases.premixed_Y(mech, np.ones((10, 9)))
    U, aux = cases.build_state(mech, T, p, Y9, np.full((10, 9), 120.0), np.full((10, 9), -40.0))
    o = Oracle(mech_path, PER, 1e-3, 1e-3,
               flux_config={"solver_type": solver, "interface_reconstruction": recon,
                            "riemann_solver": "HLLC", "split_method": "LF"})
    rhs = o.flux_func(U, aux, 1e-7)
    assert np.abs(rhs).max() <= 1e-12 * np.abs(U).max()


def _smooth_rhs(mech_path, mech, nx, flux_config):
    ny = 8
    x = np.arange(nx) / (nx - 1)          # the padding's period is nx - 1 cells
    T = 900.0 + 30.0 * np.sin(2 * np.pi * x)[:, None] * np.ones((1, ny))
    p = np.full((nx, ny), 2e5)
    Y9 = cases.premixed_Y(mech, np.ones((nx, ny)))
    U, aux = cases.build_state(mech, T, p, Y9, np.full((nx, ny), 50.0), np.zeros((nx, ny)))
    o = Oracle(mech_path, PER, 1e-3, 1e-3, flux_config=flux_config)
    return o.flux_func(U, aux, 1.0)[0]


def test_godunov_and_splitting_converge_together(mech_path, mech):
    """Both schemes discretise the same divergence: on a smooth field the
    difference of their mass right-hand sides is WENO's truncation error and
    shrinks at fifth order under refinement (a wrong HLLC flux entry would
    leave an O(1) difference)."""
    d = []
    for nx in (24, 48):
        g = _smooth_rhs(mech_path, mech, nx, GOD)
        s = _smooth_rhs(mech_path, mech, nx, dict(GOD, solver_type="flux_splitting"))
        d.append(np.abs(g - s).max() / np.abs(s).max())
    assert d[1] < 1e-3 and d[0] / d[1] > 16.0, d


# ------------------------------------------------------------------ Steger-Warming
@pytest.mark.parametrize("ixy", [1, 2])
def test_sw_split_sums_to_the_flux_of_a_perfect_gas(mech, thermo, ixy):
    """F+ + F- = lambda1 H1 + lambda2 H2 + lambda4 H4 = F exactly when
    rho E = p/(gamma-1) + rho|u|^2/2 (the Steger-Warming identity): pins every
    entry of H1, H2, H4 and the eigenvalues."""
    rng = np.random.default_rng(8)
    q = _states(mech, rng, 80)
    rho, u, v, p, Y = q[0:1], q[1:2], q[2:3], q[3:4], q[4:]
    gamma = rng.uniform(1.2, 1.6, size=rho.shape)
    E = p / (gamma - 1) + 0.5 * rho * (u ** 2 + v ** 2)
    U = np.concatenate([rho, rho * u, rho * v, E, rho * Y], axis=0)
    T = p / (thermo.get_R(Y) * rho)
    aux = np.concatenate([gamma, T], axis=0)
    Hp, Hm = F.split_flux_sw(ixy, U, aux, thermo)
    Hl, Hr = F.split_flux_lf(ixy, U, aux, thermo)
    Fphys = Hl + Hr
    scale = np.abs(Fphys).max(axis=(1, 2), keepdims=True) + np.abs(U).max(axis=(1, 2),
                                                                       keepdims=True) * 1e3
    assert np.all(np.abs(Hp + Hm - Fphys) <= 1e-12 * scale)
    # the eigenvalue split: H+ carries the non-negative, H- the non-positive
    # characteristic speeds; a supersonic state has H- ~ 0
    qs = q.copy()
    qs[ixy] = 3000.0
    rho, u, v, p, Y = qs[0:1], qs[1:2], qs[2:3], qs[3:4], qs[4:]
    E = p / (gamma - 1) + 0.5 * rho * (u ** 2 + v ** 2)
    U = np.concatenate([rho, rho * u, rho * v, E, rho * Y], axis=0)
    Hp, Hm = F.split_flux_sw(ixy, U, aux, thermo)
    assert np.all(np.abs(Hm) <= 1e-14 * np.abs(Hp).max(axis=(1, 2), keepdims=True))


def test_sw_split_mirror_symmetry(mech, thermo):
    rng = np.random.default_rng(9)
    q = _states(mech, rng, 50)
    U, gamma, T = _conserved(q, thermo)
    aux = np.concatenate([gamma, T], axis=0)
    Hp, Hm = F.split_flux_sw(1, U, aux, thermo)
    Um = U.copy()
    Um[1] *= -1
    Mp, Mm = F.split_flux_sw(1, Um, aux, thermo)
    sgn = -np.ones((12, 1, 1))
    sgn[1] = 1
    scale = np.abs(Hp).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(Mp - sgn * Hm) <= 1e-12 * scale)
    assert np.all(np.abs(Mm - sgn * Hp) <= 1e-12 * scale)


def test_sw_divergence_telescopes_and_preserves_uniform_flow(mech_path, mech):
    U, aux, geo = cases.periodic_box(mech, nx=20, ny=14)
    o = Oracle(mech_path, PER, geo["dx"], geo["dy"],
               flux_config={"solver_type": "flux_splitting", "split_method": "SW",
                            "interface_reconstruction": "WENO5_JS"})
    rhs = o.flux_func(U, aux, 1.0)
    tot = rhs[:, :-1, :-1].sum(axis=(1, 2))
    assert np.all(np.abs(tot) <= 1e-11 * np.abs(rhs).sum(axis=(1, 2)))
    T = np.full((10, 9), 900.0)
    Y9 = cases.premixed_Y(mech, np.ones((10, 9)))
    U, aux = cases.build_state(mech, T, np.full((10, 9), 2e5), Y9, np.full((10, 9), 120.0),
                               np.full((10, 9), -40.0))
    assert np.abs(o.flux_func(U, aux, 1e-7)).max() <= 1e-12 * np.abs(U).max()
