"""The per-cell device functions (``__host__ __device__`` in janc_amd/csrc)
compiled for the host with g++ and checked on the CPU: chemistry increment
against the oracle, its VJP against autograd of the torch twin, the
temperature-solve VJP against the oracle's get_T_bwd, WENO5 (FP64, FP32)
against the oracle, the WENO5 gradient against autograd, and the Godunov
path's HLLC flux and MUSCL / WENO5 face values against the oracle."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism

HERE = os.path.dirname(os.path.abspath(__file__))
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("host") / "shim.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", os.path.join(HERE, "host", "shim.cpp"), "-o", out],
                   check=True)
    lib = ctypes.CDLL(out)
    lib.host_weno.restype = ctypes.c_double
    lib.host_weno_f32.restype = ctypes.c_float
    lib.host_weno_grad.restype = ctypes.c_double
    return lib


P = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return np.ascontiguousarray(a).ctypes.data_as(P)


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def _cells(mech):
    """Unburned / burned / partly reacted cells over 1-30 atm and 900-3000 K
    (with and without radicals, so pivoting and the clip both occur)."""
    rng = np.random.default_rng(3)
    n = 64
    T = rng.uniform(900.0, 3000.0, (n, 1))
    p = rng.uniform(1.0, 30.0, (n, 1)) * 101325.0
    Y9 = cases.premixed_Y(mech, rng.uniform(0.6, 1.4, (n, 1)))
    rad = np.where(rng.uniform(size=(n, 1)) < 0.3, 0.0, 10 ** rng.uniform(-8, -2, (n, 1)))
    Y9[2:8] = rad * rng.uniform(0.1, 1.0, (6, n, 1))
    Y9 /= Y9.sum(0, keepdims=True)
    return cases.build_state(mech, T, p, Y9)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("dt", [2e-9, 1e-7])
def test_chemistry_increment_and_vjp(shim, mech, mech_path, dt):
    from oracle.step import Oracle
    from oracle.torch_step import TorchOracle
    U, aux = _cells(mech)
    n = U.shape[1] * U.shape[2]
    mc = mech.const_array()
    d = np.zeros((8, n))
    shim.host_chem(_p(mc), n, _p(U), _p(aux[1]), ctypes.c_double(dt), _p(d))
    bc = {s + "_boundary": "periodic" for s in ("left", "right", "bottom", "top")}
    S = Oracle(mech_path, bc, 1e-4, 1e-4).reaction_source_terms(U, aux, dt)
    assert _rel(d, S[4:].reshape(8, -1)) < 1e-12
    to = TorchOracle(mech_path, bc, 1e-4, 1e-4)
    Sb = np.random.default_rng(1).standard_normal((8,) + U.shape[1:])
    Ut = torch.tensor(U, requires_grad=True)
    at = torch.tensor(aux, requires_grad=True)
    (to.reaction(Ut, at, dt)[4:] * torch.tensor(Sb)).sum().backward()
    rb, yb, tb = np.zeros(n), np.zeros((8, n)), np.zeros(n)
    shim.host_chem_vjp(_p(mc), n, _p(U), _p(aux[1]), ctypes.c_double(dt), _p(Sb), _p(rb),
                       _p(yb), _p(tb))
    g = Ut.grad.numpy().reshape(12, -1)
    assert _rel(rb, g[0]) < 1e-11
    assert _rel(tb, at.grad.numpy()[1].ravel()) < 1e-11
    # species with a non-negligible share: the gradient with respect to a
    # ~1e-9 radical cancels terms of ~1/X in both forms and is not compared
    for k in range(8):
        big = (U[4 + k].ravel() / U[0].ravel()) > 1e-6
        assert _rel(yb[k][big], g[4 + k][big]) < 1e-9, k


def test_temperature_vjp(shim, mech, mech_path):
    from oracle.mech import load_mechanism
    from oracle.thermo import Thermo
    U, aux = _cells(mech)
    n = U.shape[1] * U.shape[2]
    Y = (U[4:] / U[0]).reshape(8, -1)
    rng = np.random.default_rng(2)
    gb, Tbar = rng.standard_normal(n), rng.standard_normal(n)
    eb, Yb = np.zeros(n), np.zeros((8, n))
    shim.host_getT_vjp(_p(mech.const_array()), n, _p(Y), _p(aux[1].ravel()), _p(gb), _p(Tbar),
                       _p(eb), _p(Yb))
    _, th, _ = load_mechanism(mech_path)
    de, dY, _ = Thermo(th).get_T_bwd(None, Y, aux.reshape(2, -1), np.stack([gb, Tbar]))
    assert _rel(eb, de.ravel()) < 1e-12
    assert _rel(Yb, dY) < 1e-12


def test_temperature_solve(shim, mech, mech_path):
    """halley_T (mixture polynomials from the cp tables, per-cell stop)
    against the oracle's get_T with the per-cell stop, from initial guesses
    10 % off, on both sides of the 1000 K range switch."""
    from oracle.mech import load_mechanism
    from oracle.thermo import Thermo
    U, aux = _cells(mech)
    n = U.shape[1] * U.shape[2]
    rho = U[0].ravel()
    Y = (U[4:] / U[0]).reshape(8, -1)
    u, v = U[1].ravel() / rho, U[2].ravel() / rho
    e = U[3].ravel() / rho - 0.5 * (u * u + v * v)
    T0 = aux[1].ravel() * np.random.default_rng(4).uniform(0.9, 1.1, n)
    g, T = np.zeros(n), np.zeros(n)
    shim.host_halley(_p(mech.const_array()), n, _p(Y), _p(e), _p(T0), _p(g), _p(T))
    _, th, _ = load_mechanism(mech_path)
    ref = Thermo(th).get_T(e[None], Y, T0[None], per_cell=True)
    assert _rel(T, ref[1]) < 1e-12
    assert _rel(g, ref[0]) < 1e-12
    assert np.abs(T - aux[1].ravel()).max() < 1e-6 * aux[1].max()


def test_weno(shim):
    from oracle.flux import _weno_core
    rng = np.random.default_rng(0)
    for _ in range(500):
        v = rng.standard_normal(5) * 10 ** rng.uniform(-3, 10)
        ref = _weno_core(*v)
        assert abs(shim.host_weno(_p(v)) - ref) <= 1e-13 * max(np.abs(v).max(), 1e-300)
        v32 = v.astype(np.float32)
        got32 = shim.host_weno_f32(np.ascontiguousarray(v32).ctypes.data_as(
            ctypes.POINTER(ctypes.c_float)))
        assert abs(got32 - _weno_core(*v32.astype(np.float64))) <= 2e-6 * np.abs(v32).max()


def test_weno_stencil_kinds(shim):
    """The FP64 form (differences around f0, candidates as f0 plus a
    correction) and the FP32 form against the oracle's WENO5_JS formulas on
    smooth, rough, discontinuous, constant and sign-alternating stencils over
    20 decades of scale, in both orientations."""
    from oracle.flux import _weno_core
    rng = np.random.default_rng(5)
    x = np.arange(5.0)
    stencils = [np.zeros(5)]
    for _ in range(2000):
        kind = rng.integers(5)
        if kind == 0:
            f = np.sin(0.3 * x + rng.uniform(0, 6)) + rng.uniform(-2, 2)
        elif kind == 1:
            f = rng.standard_normal(5)
        elif kind == 2:
            f = np.where(x < rng.integers(1, 5), rng.standard_normal(), rng.standard_normal())
        elif kind == 3:
            f = np.full(5, rng.standard_normal())
        else:
            f = (-1.0) ** x * rng.uniform(0.5, 1.5, 5)
        stencils.append(f * 10 ** rng.uniform(-6, 14))
    F32 = ctypes.POINTER(ctypes.c_float)
    for f in stencils:
        tol = 2e-15 * max(np.abs(f).max(), 1e-300)
        for g in (f, f[::-1]):
            g = np.ascontiguousarray(g)
            assert abs(shim.host_weno(_p(g)) - _weno_core(*g)) <= tol
            # FP32 form (mixed precision): normalised weights, no overflow up
            # to detonation-front fluxes; checked against the FP64 formula on
            # the FP32-rounded stencil
            if np.abs(g).max() < 1e13:
                g32 = g.astype(np.float32)
                got = shim.host_weno_f32(g32.ctypes.data_as(F32))
                ref = _weno_core(*g32.astype(np.float64))
                assert abs(got - ref) <= 1e-6 * max(np.abs(g32).max(), 1e-30)


def test_weno_run_f32_equals_per_face(shim):
    """The FP32 run form (four neighbouring faces sharing differences) gives
    the per-face FP32 weno5 values bit for bit, both orientations."""
    F32 = ctypes.POINTER(ctypes.c_float)
    shim.host_weno_f32.restype = ctypes.c_float
    rng = np.random.default_rng(8)
    for it in range(2000):
        f = (rng.standard_normal(8) * 10 ** rng.uniform(-6, 12)).astype(np.float32)
        if it % 4 == 0:
            f = np.where(np.arange(8) < rng.integers(1, 8), f[0], f[7]).astype(np.float32)
        L, R = np.zeros(4, np.float32), np.zeros(4, np.float32)
        shim.host_weno_run_f32(f.ctypes.data_as(F32), L.ctypes.data_as(F32),
                               R.ctypes.data_as(F32))
        for i in range(4):
            w = np.ascontiguousarray(f[i:i + 5])
            assert L[i] == shim.host_weno_f32(w.ctypes.data_as(F32))
            w = np.ascontiguousarray(f[i:i + 5][::-1])
            assert R[i] == shim.host_weno_f32(w.ctypes.data_as(F32))


def test_weno_gradient(shim):
    rng = np.random.default_rng(1)
    for it in range(300):
        v = rng.standard_normal(5) * 10 ** rng.uniform(-4, 4)
        if it % 3 == 0:   # a jump inside the stencil
            v = np.where(np.arange(5) < rng.integers(1, 5), v[0], v[4])
        t = torch.tensor(v, requires_grad=True)
        f = t
        IS1 = 0.25 * (f[0] - 4 * f[1] + 3 * f[2]) ** 2 + 13 / 12 * (f[0] - 2 * f[1] + f[2]) ** 2
        IS2 = 0.25 * (f[1] - f[3]) ** 2 + 13 / 12 * (f[1] - 2 * f[2] + f[3]) ** 2
        IS3 = 0.25 * (3 * f[2] - 4 * f[3] + f[4]) ** 2 + 13 / 12 * (f[2] - 2 * f[3] + f[4]) ** 2
        a = torch.stack([0.1 / (1e-6 + IS1) ** 2, 0.6 / (1e-6 + IS2) ** 2,
                         0.3 / (1e-6 + IS3) ** 2])
        w = a / a.sum()
        q = torch.stack([f[0] / 3 - 7 * f[1] / 6 + 11 * f[2] / 6,
                         -f[1] / 6 + 5 * f[2] / 6 + f[3] / 3,
                         f[2] / 3 + 5 * f[3] / 6 - f[4] / 6])
        val = (w * q).sum()
        val.backward()
        g = np.zeros(5)
        got = shim.host_weno_grad(_p(v), _p(g))
        np.testing.assert_allclose(g, t.grad.numpy(), rtol=1e-9, atol=1e-12)
        assert abs(got - float(val.detach())) <= 1e-13 * np.abs(v).max()


def _face_states(mech, rng, n, vel):
    T = rng.uniform(300.0, 3000.0, (n, 1))
    p = rng.uniform(0.3, 30.0, (n, 1)) * 101325.0
    Y9 = cases.premixed_Y(mech, rng.uniform(0.6, 1.4, (n, 1)))
    Y9[2:8] += rng.uniform(0.0, 0.02, (6, n, 1))
    Y9 /= Y9.sum(0, keepdims=True)
    R = (Y9 / mech.species_M[:, None, None]).sum(0)
    rho = p / (R * T)
    u = rng.uniform(-vel, vel, (n, 1))
    v = rng.uniform(-vel, vel, (n, 1))
    return np.concatenate([rho[None], u[None], v[None], p[None], Y9[:8]], axis=0)


@pytest.mark.parametrize("dir_", [0, 1])
@pytest.mark.parametrize("vel", [200.0, 3000.0])
def test_hllc_flux(shim, mech, mech_path, dir_, vel):
    """Subsonic (all four branches) and supersonic face states."""
    from oracle.flux import hllc_dir
    from oracle.step import Oracle
    th = Oracle(mech_path, {s + "_boundary": "periodic" for s in
                            ("left", "right", "bottom", "top")}, 1.0, 1.0).thermo
    rng = np.random.default_rng(10 + dir_)
    n = 400
    qL, qR = _face_states(mech, rng, n, vel), _face_states(mech, rng, n, vel)
    # a few contacts and sonic points: equal normal velocity and pressure
    qR[1 + dir_, :20] = qL[1 + dir_, :20]
    qR[3, :20] = qL[3, :20]
    ref = hllc_dir(qL, qR, th, dir_ + 1).reshape(12, n)
    F = np.zeros((12, n))
    shim.host_hllc(_p(mech.const_array()), n, _p(qL.reshape(12, n)), _p(qR.reshape(12, n)),
                   dir_, _p(F))
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(F - ref) / scale
    assert err.max() < 1e-12, err.max(axis=1)


@pytest.mark.parametrize("recon", [0, 1])
def test_face_reconstructions(shim, recon):
    from oracle import flux as OF
    rng = np.random.default_rng(20 + recon)
    # one padded column of 6 cells gives one face; stack many along y
    m = 300
    q = rng.standard_normal((1, 6, m + 6)) * 10 ** rng.uniform(-3, 6, (1, 1, m + 6))
    q[0, :, :50] = np.round(q[0, :, :50])               # ties and flat stretches
    Lx = OF.RECONSTRUCTION["MUSCL" if recon else "WENO5_JS"][0](q)[0, 0]
    Rx = OF.RECONSTRUCTION["MUSCL" if recon else "WENO5_JS"][1](q)[0, 0]
    out = np.zeros(2)
    for j in range(m):
        v = np.ascontiguousarray(q[0, :, j + 3])
        shim.host_recon(recon, _p(v), _p(out))
        tol = 1e-13 * max(np.abs(v).max(), 1e-300)
        assert abs(out[0] - Lx[j]) <= tol and abs(out[1] - Rx[j]) <= tol, j


def _halley_inputs(mech_si, nondim=None, seed=6, spread=0.1):
    """Cells of _cells (built with the SI mechanism) in the given unit
    system, with initial temperature guesses `spread` off (relative)."""
    from janc_amd.mechanism import NondimScales
    U, aux = _cells(mech_si)
    n = U.shape[1] * U.shape[2]
    rho = U[0].ravel()
    Y = (U[4:] / U[0]).reshape(8, -1)
    u, v = U[1].ravel() / rho, U[2].ravel() / rho
    e = U[3].ravel() / rho - 0.5 * (u * u + v * v)
    T = aux[1].ravel()
    if nondim is not None:
        s = NondimScales(nondim)
        e = e / s.e0
        T = T / s.T0
    T0 = T * np.random.default_rng(seed).uniform(1.0 - spread, 1.0 + spread, n)
    return n, Y, e, T0


@pytest.mark.parametrize("nondim", [None, {"P0": 1.0e5, "T0": 300.0, "R0": 287.0, "x0": 1e-3}])
def test_grid_stop_reproduction(shim, mech_path, nondim):
    """The grid-wide stop (thermo_model.py:140-142) as the kernels reproduce
    it: N* = max of the per-cell update counts, a cell with n < N* takes its
    state after one more update.  Against the oracle's grid-wide get_T at
    1e-13, in SI and nondimensional units; with nondimensional scales the
    per-cell stop alone is visibly off (~1e-10 relative), in SI it is not."""
    from oracle.mech import load_mechanism
    from oracle.thermo import Thermo
    mech = Mechanism(mech_path, nondim)
    n, Y, e, T0 = _halley_inputs(Mechanism(mech_path), nondim)
    g, T, g1, T1 = (np.zeros(n) for _ in range(4))
    nit, conv, ok1 = (np.zeros(n, dtype=np.int32) for _ in range(3))
    I = ctypes.POINTER(ctypes.c_int)
    shim.host_halley_more(_p(mech.const_array()), n, _p(Y), _p(e), _p(T0), _p(g), _p(T),
                          nit.ctypes.data_as(I), _p(g1), _p(T1), conv.ctypes.data_as(I),
                          ok1.ctypes.data_as(I))
    assert conv.all() and ok1.all()
    ns = nit.max()
    assert (nit < ns).any()                      # some cells stop early
    Tg = np.where(nit == ns, T, T1)
    gg = np.where(nit == ns, g, g1)
    _, th, _ = load_mechanism(mech_path, nondim)
    ref = Thermo(th).get_T(e[None], Y, T0[None])
    assert _rel(Tg, ref[1]) < 1e-13
    assert _rel(gg, ref[0]) < 1e-13
    per_cell = _rel(T, ref[1])
    if nondim is None:
        assert per_cell < 1e-14
    else:
        assert per_cell > 1e-12


def test_lu_solve_pivot_check(shim):
    """lu_solve against numpy.linalg.solve on pivoting systems; its pivot
    check flags a singular system (a zero column) and a non-finite one."""
    rng = np.random.default_rng(8)
    for _ in range(20):
        A = rng.standard_normal((8, 8)) * 10.0 ** rng.uniform(-3, 3, (8, 1))
        b = rng.standard_normal(8)
        x = b.copy()
        assert shim.host_lu_solve8(_p(A), _p(x)) == 1
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
    A = rng.standard_normal((8, 8))
    A[:, 3] = 0.0
    assert shim.host_lu_solve8(_p(A), _p(np.ones(8))) == 0
    A = rng.standard_normal((8, 8))
    A[5, 2] = np.nan
    assert shim.host_lu_solve8(_p(A), _p(np.ones(8))) == 0
