"""GPU parity of the hand-written adjoint (jh_step_vjp, janc_amd.adjoint).

Oracle: reverse mode of the differentiable torch twin of the oracle step
(oracle/torch_step.py, itself pinned against finite differences of the numpy
oracle).  Bars: per-variable max-norm relative error <= 1e-9 against the
twin; directional derivative within 1e-6 of a central finite difference of
the GPU step (step 1e-7 relative per component); VJP linear in the cotangent
to 1e-12; checkpointed rollout identical to the chained single-step VJPs.
"""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

from oracle.step import Oracle  # noqa: E402
from oracle.torch_step import TorchOracle  # noqa: E402
from test_oracle_adjoint import tie_free_state  # noqa: E402

TWIN_TOL = 1e-9
DT = 2e-9


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def rel_err(got, ref):
    got = got.detach().cpu().numpy() if hasattr(got, "detach") else np.asarray(got)
    ref = ref.detach().cpu().numpy() if hasattr(ref, "detach") else np.asarray(ref)
    d = np.abs(got - ref).reshape(ref.shape[0], -1).max(axis=1)
    s = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    return d / np.maximum(s, 1e-300)


def _setup(mech_path, mech, name):
    from janc_amd.solver import Simulator
    if name == "box_ties":
        # exact LF-maximum ties: a fast cell on row x = 0 and one on column
        # y = 0, duplicated onto x = N-1 / y = N-1 (the period-(N-1) copies
        # the reference's periodic padding expects); no other near-ties
        U, aux, g = tie_free_state(mech_path, mech, "box", 23, 19)
        rho = U[0]
        for (ix, iy, comp) in ((0, 5, 1), (7, 0, 2)):
            old = U[comp, ix, iy]
            U[comp, ix, iy] = rho[ix, iy] * 400.0
            U[3, ix, iy] += 0.5 * (U[comp, ix, iy] ** 2 - old ** 2) / rho[ix, iy]
        U[:, -1, :] = U[:, 0, :]
        U[:, :, -1] = U[:, :, 0]
        aux[:, -1, :] = aux[:, 0, :]
        aux[:, :, -1] = aux[:, :, 0]
        aux = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"]).update_func(U, aux)
    elif name == "tube":
        U, aux, g = tie_free_state(mech_path, mech, "tube", 26, 13)
    elif name == "wide":
        # 76 padded columns: two 56-column segments of the flux VJP's y pass,
        # the second ragged
        U, aux, g = tie_free_state(mech_path, mech, "tube", 21, 70)
    elif name == "tall":
        # 81 padded rows: three 32-row segments of the flux VJP's x pass (the
        # last ragged), with wall ghost rows (momentum flips) at both ends
        U, aux, g = tie_free_state(mech_path, mech, "tube", 75, 20)
        g["boundary_config"] = {"left_boundary": "no_slip_wall",
                                "right_boundary": "slip_wall",
                                "bottom_boundary": "slip_wall",
                                "top_boundary": "zero_gradient"}
    elif name == "box_small":
        # a 6 x 5 grid: one partial wave in each pass (5 of 64 columns in the
        # x pass, 11 padded columns in the y pass)
        U, aux, g = tie_free_state(mech_path, mech, "box", 6, 5)
    elif name == "box_tall":
        # periodic x ghost rows in the first and last of three x segments
        U, aux, g = tie_free_state(mech_path, mech, "box", 70, 16)
    elif name == "walls":
        U, aux, g = tie_free_state(mech_path, mech, "tube", 21, 17)
        g["boundary_config"] = {"left_boundary": "no_slip_wall",
                                "right_boundary": "slip_wall",
                                "bottom_boundary": "zero_gradient",
                                "top_boundary": "slip_wall"}
    else:
        U, aux, g = tie_free_state(mech_path, mech, "box", 22, 15)
    cfg = cases.simulation_config(g, mech_path, cfl=0.3, dt=DT)
    sim = Simulator(cfg)
    print("case", name, U.shape, flush=True)
    return sim, g, U, aux


def _cotangents(U, aux, seed=0):
    rng = np.random.default_rng(seed)
    sU = np.abs(U).reshape(U.shape[0], -1).max(1)[:, None, None]
    sU = np.where(sU > 0, sU, 1.0)         # species absent from the whole field
    sa = np.abs(aux).reshape(2, -1).max(1)[:, None, None]
    return rng.standard_normal(U.shape) / sU, rng.standard_normal(aux.shape) / sa


@pytest.mark.parametrize("name", ["box", "box_ties", "tube", "walls", "wide", "tall", "box_tall",
                                  "box_small"])
def test_step_vjp_matches_twin(mech_path, mech, name):
    sim, g, U, aux = _setup(mech_path, mech, name)
    W, Wa = _cotangents(U, aux)
    to = TorchOracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    Uv = torch.tensor(U, requires_grad=True)
    av = torch.tensor(aux, requires_grad=True)
    U4, a4 = to.advance_one_step(Uv, av, DT)
    ((U4 * torch.tensor(W)).sum() + (a4 * torch.tensor(Wa)).sum()).backward()
    gU, ga, Uo, ao = sim.step_vjp(U, aux, DT, W, Wa, with_output=True)
    eU, ea = rel_err(gU, Uv.grad), rel_err(ga, av.grad)
    print("U_bar rel err", eU.max(), "aux_bar rel err", ea.max(), flush=True)
    assert eU.max() <= TWIN_TOL and ea.max() <= TWIN_TOL
    assert rel_err(Uo, U4).max() <= 1e-12 and rel_err(ao, a4).max() <= 1e-12


@pytest.mark.parametrize("name,nx,ny", [("tube", 26, 13), ("box", 256, 160)])
def test_step_vjp_directional_fd(mech_path, mech, name, nx, ny):
    from janc_amd.solver import Simulator
    U, aux, g = tie_free_state(mech_path, mech, name, nx, ny)
    sim = Simulator(cases.simulation_config(g, mech_path, cfl=0.3, dt=DT))
    print("case", name, U.shape, flush=True)
    W, Wa = _cotangents(U, aux, 1)
    Wt, Wat = torch.tensor(W, device="cuda"), torch.tensor(Wa, device="cuda")
    gU, ga = sim.step_vjp(U, aux, DT, Wt, Wat)
    rng = np.random.default_rng(2)
    Ut, at = torch.tensor(U, device="cuda"), torch.tensor(aux, device="cuda")
    dU = torch.tensor(rng.standard_normal(U.shape) * np.abs(U) * 1e-7, device="cuda")

    def L(Ux):
        A, B = sim.advance_one_step(Ux, at, sim.dx, sim.dy, DT)
        return float((A * Wt).sum() + (B * Wat).sum())

    fd = (L(Ut + dU) - L(Ut - dU)) / 2
    ad = float((gU * dU).sum())
    print("fd", fd, "ad", ad, flush=True)
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def test_step_vjp_linear_in_cotangent(mech_path, mech):
    sim, g, U, aux = _setup(mech_path, mech, "walls")
    W1, Wa1 = _cotangents(U, aux, 3)
    W2, Wa2 = _cotangents(U, aux, 4)
    g1, a1 = sim.step_vjp(U, aux, DT, W1, Wa1)
    g2, a2 = sim.step_vjp(U, aux, DT, W2, Wa2)
    g3, a3 = sim.step_vjp(U, aux, DT, 2.0 * W1 - W2, 2.0 * Wa1 - Wa2)
    assert rel_err(g3, 2.0 * g1 - g2).max() <= 1e-12
    assert rel_err(a3, 2.0 * a1 - a2).max() <= 1e-12


@pytest.mark.parametrize("name", ["walls", "box_ties"])
def test_step_vjp_is_deterministic(mech_path, mech, name):
    """Two VJPs of the same inputs agree bit for bit: the cotangents reach
    their cells through fixed-order gathers (ghost folds, block sums), not
    through atomics."""
    sim, g, U, aux = _setup(mech_path, mech, name)
    W, Wa = _cotangents(U, aux, 9)
    g1, a1 = sim.step_vjp(U, aux, DT, W, Wa)
    g2, a2 = sim.step_vjp(U, aux, DT, W, Wa)
    assert torch.equal(g1, g2) and torch.equal(a1, a2)


def test_vjp_with_known_output_is_identical(mech_path, mech):
    """jh_step_vjp_known (the step's output supplied, no chemistry forward
    pass) gives the same cotangents bit for bit as jh_step_vjp."""
    sim, g, U, aux = _setup(mech_path, mech, "tube")
    W, Wa = _cotangents(U, aux, 7)
    gU, ga, Uo, ao = sim.step_vjp(U, aux, DT, W, Wa, with_output=True)
    kU, ka = sim.step_vjp(U, aux, DT, W, Wa, known_output=(Uo, ao))
    assert torch.equal(gU, kU) and torch.equal(ga, ka)


@pytest.mark.parametrize("name", ["tube", "walls"])
def test_vjp_from_residuals_is_identical(mech_path, mech, name):
    """jh_step_keep's output is the step's (jh_step_vjp's U_out) and
    jh_step_vjp_res from its residuals gives jh_step_vjp_known's cotangents,
    bit for bit; the residuals are not modified by the VJP."""
    sim, g, U, aux = _setup(mech_path, mech, name)
    W, Wa = _cotangents(U, aux, 11)
    gU, ga, Uo, ao = sim.step_vjp(U, aux, DT, W, Wa, with_output=True)
    Uk, ak, res = sim.step_keep(U, aux, DT)
    assert torch.equal(Uk, Uo) and torch.equal(ak, ao)
    r0 = res.clone()
    rU, ra = sim.step_vjp(U, aux, DT, W, Wa, known_output=(Uk, ak), residuals=res)
    assert torch.equal(rU, gU) and torch.equal(ra, ga)
    assert torch.equal(res, r0)
    with pytest.raises(ValueError):
        sim.step_vjp(U, aux, DT, W, Wa, known_output=(Uk, ak), residuals=res[:-1].contiguous())


def test_rollout_checkpointing_matches_chained_steps(mech_path, mech):
    from janc_amd import adjoint
    sim, g, U, aux = _setup(mech_path, mech, "tube")
    W, Wa = _cotangents(U, aux, 5)
    Wt, Wat = torch.tensor(W, device="cuda"), torch.tensor(Wa, device="cuda")
    n = 5
    U0 = torch.tensor(U, device="cuda", requires_grad=True)
    a0 = torch.tensor(aux, device="cuda", requires_grad=True)
    Ux, ax = U0, a0
    for _ in range(n):
        Ux, ax = adjoint.step(sim, Ux, ax, DT)
    ((Ux * Wt).sum() + (ax * Wat).sum()).backward()
    U1 = torch.tensor(U, device="cuda", requires_grad=True)
    a1 = torch.tensor(aux, device="cuda", requires_grad=True)
    Uy, ay = adjoint.rollout(sim, U1, a1, DT, n, checkpoint_every=2)
    ((Uy * Wt).sum() + (ay * Wat).sum()).backward()
    assert torch.equal(Ux.detach(), Uy.detach())
    assert torch.equal(U0.grad, U1.grad) and torch.equal(a0.grad, a1.grad)


def test_step_vjp_rejects_unsupported(mech_path, mech):
    from janc_amd import _lib
    from janc_amd.solver import Simulator
    U, aux, g = cases.detonation_tube(mech, nx=20, ny=12, driver=5)
    W, Wa = _cotangents(U, aux)
    g["boundary_config"]["top_boundary"] = "pressure_outlet"
    sim = Simulator(cases.simulation_config(g, mech_path, cfl=0.3, dt=DT,
                                            theta={"Pb": 101325.0}))
    with pytest.raises(_lib.JancError):
        sim.step_vjp(U, aux, DT, W, Wa)
    g["boundary_config"]["top_boundary"] = "slip_wall"
    sim = Simulator(cases.simulation_config(g, mech_path, cfl=0.3, dt=DT, precision="mixed"))
    with pytest.raises(NotImplementedError):
        sim.step_vjp(U.astype(np.float32), aux.astype(np.float32), DT, W, Wa)


# ------------------------------------------------- gradients w.r.t. T and Y
def _tY_case(mech, nx, ny, seed=9):
    """Tube with 0.5 % temperature noise (so the LF maxima have single
    maximisers) and a few m/s of velocity noise (so no |u| or |v| sits at its
    kink at 0) as (T, p, Y, u, v) fields."""
    U, aux, g = cases.detonation_tube(mech, nx=nx, ny=ny, driver=max(4, nx // 8), seed=seed)
    rng = np.random.default_rng(seed)
    T = aux[1] * (1.0 + 0.005 * rng.uniform(-1, 1, aux[1].shape))
    rho = U[0]
    Y = U[4:] / rho
    Y9 = np.concatenate([Y, (1.0 - Y.sum(0))[None]])
    R, _, _, _ = cases.mixture(mech, aux[1], Y9)
    p = rho * R * aux[1]
    u = 3.0 * rng.standard_normal(T.shape)
    v = 3.0 * rng.standard_normal(T.shape)
    return T, p, Y, u, v, g


def _loss(mech, U0, UN, aN, dx, dy, p_target, q_scale, p_ref):
    from janc_amd import adjoint
    q = adjoint.heat_release(mech, U0, UN, dx, dy) / q_scale
    pN = adjoint.pressure(mech, UN, aN)
    return q + (((pN - p_target) / p_ref) ** 2).mean()


def test_initial_field_gradient_matches_twin(mech_path, mech):
    """d loss / d (T, Y) through the GPU rollout vs autograd of the torch twin."""
    from janc_amd import adjoint
    from janc_amd.solver import Simulator
    T, p, Y, u, v, g = _tY_case(mech, 40, 16)
    steps, dt = 3, 2e-9
    sim = Simulator(cases.simulation_config(g, mech_path, dt=dt))
    to = TorchOracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    rng = np.random.default_rng(1)
    p_target = p * (1.0 + 0.01 * rng.standard_normal(p.shape))
    grads = []
    for dev in ("cuda", "cpu"):
        Tt = torch.tensor(T, device=dev, requires_grad=True)
        Yt = torch.tensor(Y, device=dev, requires_grad=True)
        U0, a0 = adjoint.initial_state(mech, Tt, torch.tensor(p, device=dev), Yt,
                                       torch.tensor(u, device=dev), torch.tensor(v, device=dev))
        if dev == "cuda":
            UN, aN = adjoint.rollout(sim, U0, a0, dt, steps, checkpoint_every=2)
        else:
            UN, aN = U0, a0
            for _ in range(steps):
                UN, aN = to.advance_one_step(UN, aN, dt)
        L = _loss(mech, U0, UN, aN, g["dx"], g["dy"], torch.tensor(p_target, device=dev),
                  1.0, 101325.0)
        L.backward()
        grads.append((Tt.grad.cpu(), Yt.grad.cpu()))
    eT = rel_err(grads[0][0][None], grads[1][0][None]).max()
    eY = rel_err(grads[0][1], grads[1][1]).max()
    print("dL/dT rel err", eT, "dL/dY rel err", eY, flush=True)
    assert eT <= TWIN_TOL and eY <= TWIN_TOL


def test_initial_field_gradient_matches_fd(mech_path, mech):
    """Central differences (eps 1e-6) on random directions in (T, Y) space
    through a checkpointed multi-step rollout."""
    from janc_amd import adjoint
    from janc_amd.solver import Simulator
    T, p, Y, u, v, g = _tY_case(mech, 96, 24)
    probe = Simulator(cases.simulation_config(g, mech_path, cfl=0.3))
    steps = 12
    Tt = torch.tensor(T, device="cuda")
    Yt = torch.tensor(Y, device="cuda")
    pt = torch.tensor(p, device="cuda")
    ut, vt = torch.tensor(u, device="cuda"), torch.tensor(v, device="cuda")
    U0, a0 = adjoint.initial_state(mech, Tt, pt, Yt, ut, vt)
    dt = 0.8 * float(probe.CFL(U0, a0))
    sim = Simulator(cases.simulation_config(g, mech_path, dt=dt))
    rng = np.random.default_rng(2)
    p_target = pt * (1.0 + 0.01 * torch.tensor(rng.standard_normal(p.shape), device="cuda"))
    UN, aN = sim.advance(U0.clone(), a0.clone(), steps)[:2]
    q_scale = float(adjoint.heat_release(mech, U0, UN, g["dx"], g["dy"]).abs()) + 1e-30

    def L(Tx, Yx, grad=False):
        U0, a0 = adjoint.initial_state(mech, Tx, pt, Yx, ut, vt)
        if grad:
            UN, aN = adjoint.rollout(sim, U0, a0, dt, steps, checkpoint_every=5)
        else:
            UN, aN = sim.advance(U0.clone(), a0.clone(), steps)[:2]
        return _loss(mech, U0, UN, aN, g["dx"], g["dy"], p_target, q_scale, 101325.0)

    Tg = Tt.clone().requires_grad_(True)
    Yg = Yt.clone().requires_grad_(True)
    L(Tg, Yg, grad=True).backward()
    eps = 1e-6
    for d in range(4):
        dT = Tt * torch.tensor(rng.standard_normal(T.shape), device="cuda")
        dY = Yt * torch.tensor(rng.standard_normal(Y.shape), device="cuda")
        fd = (float(L(Tt + eps * dT, Yt + eps * dY)) -
              float(L(Tt - eps * dT, Yt - eps * dY))) / (2 * eps)
        ad = float((Tg.grad * dT).sum() + (Yg.grad * dY).sum())
        print("direction", d, "fd", fd, "ad", ad, flush=True)
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad))
