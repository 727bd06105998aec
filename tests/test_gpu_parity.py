"""GPU parity: every C-ABI entry point against the FP64 oracle.

Bar: FP64 form <= 1e-12 relative error per variable (max-norm of the
difference over max-norm of the oracle's field); mixed form: the stated
L2/Linf bounds below.  Grid sizes are chosen so the oracle finishes in
seconds and the tiles are ragged (nx not a multiple of 8, ny not of 64).
"""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism
from oracle.step import Oracle

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP64_TOL = 1e-12


def rel_err(got, ref):
    got = got.detach().cpu().numpy() if hasattr(got, "detach") else np.asarray(got)
    ref = np.asarray(ref)
    d = np.abs(got - ref).reshape(ref.shape[0], -1).max(axis=1)
    s = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    return float(np.max(d / np.maximum(s, 1e-300)))


def _case(name, mech):
    if name == "box":
        U, aux, g = cases.periodic_box(mech, nx=37, ny=70)
        theta = None
    elif name == "tube":
        U, aux, g = cases.detonation_tube(mech, nx=45, ny=22, driver=8)
        theta = None
    elif name == "outlet":   # walls of every kind + a subsonic pressure outlet
        U, aux, g = cases.detonation_tube(mech, nx=29, ny=67, driver=6, seed=5)
        g["boundary_config"] = {"left_boundary": "no_slip_wall",
                                "right_boundary": "pressure_outlet",
                                "bottom_boundary": "zero_gradient",
                                "top_boundary": "pressure_outlet"}
        theta = {"Pb": 1.5 * 101325.0}
    elif name == "walls":   # copy kinds only, both wall kinds
        U, aux, g = cases.detonation_tube(mech, nx=33, ny=41, driver=6, seed=3)
        g["boundary_config"] = {"left_boundary": "no_slip_wall",
                                "right_boundary": "zero_gradient",
                                "bottom_boundary": "slip_wall",
                                "top_boundary": "no_slip_wall"}
        theta = None
    else:   # annulus crop: periodic x, injection inlet below (open and blocked
        # cells), pressure outlet above
        U, aux, g, theta = cases.rdc_annulus(mech, nx=61, ny=37, seed_cols=12)
    return U, aux, g, theta


CASES = ["box", "tube", "outlet", "inlet", "walls"]


def make_oracle(mech_path, g, theta, flux_config=None):
    """Oracle for a geometry; injection-inlet sides get the restated example
    boundary function (a user callable in the reference)."""
    bcfg = dict(g["boundary_config"])
    inlets = [s for s in ("left", "right", "bottom", "top")
              if bcfg[s + "_boundary"] == "injection_inlet"]
    for s in inlets:
        bcfg[s + "_boundary"] = "slip_wall"
    o = Oracle(mech_path, bcfg, g["dx"], g["dy"], flux_config=flux_config)
    if inlets:
        from oracle.inlet import attach_inlet, stoichiometric_injection_Y
        Y = theta.get("Yinj", stoichiometric_injection_Y(len(o.thermo.Mex)))
        for s in inlets:
            attach_inlet(o, s, theta["P0_inj"], theta.get("T0_inj", 300.0), Y)
    return o


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def _pair(mech_path, mech, name, precision="fp64", dt=None):
    from janc_amd.solver import Simulator
    U, aux, g, theta = _case(name, mech)
    cfg = cases.simulation_config(g, mech_path, cfl=0.3, dt=dt, precision=precision,
                                  theta=theta)
    sim = Simulator(cfg)
    o = make_oracle(mech_path, g, theta)
    print("case", name, precision, U.shape, flush=True)
    return sim, o, U, aux, theta


@pytest.mark.parametrize("name", CASES)
def test_flux_func(mech_path, mech, name):
    sim, o, U, aux, theta = _pair(mech_path, mech, name)
    dt = 3e-9
    got = sim.flux_func(U, aux, sim.dx, sim.dy, dt, theta)
    ref = o.flux_func(U, aux, dt, theta)
    assert rel_err(got, ref) <= FP64_TOL


@pytest.mark.parametrize("name", CASES)
def test_lf_alpha_and_cfl(mech_path, mech, name):
    sim, o, U, aux, theta = _pair(mech_path, mech, name)
    a = sim.lf_alpha(U, aux, theta).cpu().numpy()
    ax, ay = o.lf_alpha(U, aux, theta)
    np.testing.assert_allclose(a, [ax, ay], rtol=1e-14)
    dt = float(sim.CFL(U, aux, 0.3))
    assert dt == pytest.approx(o.cfl_dt(U, aux, 0.3), rel=1e-14)


@pytest.mark.parametrize("name", CASES)
def test_update_aux(mech_path, mech, name):
    sim, o, U, aux, theta = _pair(mech_path, mech, name)
    guess = aux.copy()
    guess[1] *= 1.0 + 0.03 * np.cos(np.arange(guess[1].size)).reshape(guess[1].shape)
    got = sim.update_func(U, guess)
    ref = o.update_func(U, guess)
    assert rel_err(got, ref) <= FP64_TOL


@pytest.mark.parametrize("name", ["box", "tube"])
def test_reaction_source(mech_path, mech, name):
    sim, o, U, aux, theta = _pair(mech_path, mech, name)
    for dt in (1e-9, 1e-7):
        got = sim.reaction_source_terms(U, aux, dt)
        ref = o.reaction_source_terms(U, aux, dt)
        assert np.all(got[:4].cpu().numpy() == 0)
        assert rel_err(got[4:], ref[4:]) <= FP64_TOL


@pytest.mark.parametrize("name", CASES)
def test_advance_func_one_step(mech_path, mech, name):
    sim, o, U, aux, theta = _pair(mech_path, mech, name)
    Ug, ag, t, dt = sim.advance_func(U, aux, 0.0, theta)
    Ur, ar, tr, dtr = o.advance_func(U, aux, 0.0, 0.3, theta)
    assert float(dt) == pytest.approx(dtr, rel=1e-14)
    assert rel_err(Ug, Ur) <= FP64_TOL
    assert rel_err(ag, ar) <= FP64_TOL


@pytest.mark.parametrize("name", ["tube", "outlet", "inlet"])
def test_fixed_dt_five_steps(mech_path, mech, name):
    sim, o, U, aux, theta = _pair(mech_path, mech, name, dt=4e-9)
    Ug, ag = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    Ur, ar = U, aux
    for _ in range(5):
        Ug, ag = sim.advance_one_step(Ug, ag, sim.dx, sim.dy, 4e-9, theta)
        Ur, ar = o.advance_one_step(Ur, ar, 4e-9, theta)
    assert rel_err(Ug, Ur) <= FP64_TOL
    assert rel_err(ag, ar) <= FP64_TOL


def test_fused_steps_equal_single_steps(mech_path, mech):
    sim, o, U, aux, theta = _pair(mech_path, mech, "tube")
    Ua, aa = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    for _ in range(3):
        Ua, aa, _, _ = sim.advance_func(Ua, aa, 0.0)
    Ub, ab = torch.as_tensor(U).cuda().clone(), torch.as_tensor(aux).cuda().clone()
    _, _, dts = sim.advance(Ub, ab, 3)
    assert torch.equal(Ua, Ub) and torch.equal(aa, ab)
    assert dts.shape == (3,) and bool(torch.all(dts > 0))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["box", "tube", "walls"])
def test_halo_from_source_cells_equals_strips(mech_path, mech, name, precision):
    """With every side a copy kind the fused step's stage kernels read their
    halo from the source cells and no bc_kernel runs; tracing keeps the
    ghost-strip path.  Both give the same state bit for bit."""
    sim, o, U, aux, theta = _pair(mech_path, mech, name, precision)
    dt = torch.float64 if precision == "fp64" else torch.float32
    U0 = torch.as_tensor(U).cuda().to(dt)
    a0 = torch.as_tensor(aux).cuda().to(dt)
    Ua, aa = U0.clone(), a0.clone()
    sim.timing(Ua, enable=1)
    sim.advance(Ua, aa, 3)
    kt = sim.timing(Ua)
    sim.timing(Ua, enable=False)
    assert kt["bc"][1] == 0 and kt["stage"][1] == 9
    Ub, ab = U0.clone(), a0.clone()
    sim.trace(Ub, torch.zeros((3, 22), dtype=torch.int64, device="cuda"))
    sim.timing(Ub, enable=1)
    sim.advance(Ub, ab, 3)
    kt = sim.timing(Ub)
    sim.timing(Ub, enable=False)
    sim.trace(Ub, None)
    assert kt["bc"][1] == 9
    assert torch.equal(Ua, Ub) and torch.equal(aa, ab)


MIXED_TOL = {"flux_func": 1e-6, "update_aux": 1e-6, "reaction": 1e-6, "step": 1e-5}


def _f32(x):
    return torch.as_tensor(x, dtype=torch.float32).cuda()


def state_err(got, ref, U):
    """Max-norm error per variable relative to the size of the state it is
    added to: max|U_v| for rho and rho*E, max(|rho u|, rho_max sqrt(E_max))
    for the momenta (a fluid at rest has none) and max(rho) for the species (a
    trace species is judged against the mixture density)."""
    got = got.detach().double().cpu().numpy() if hasattr(got, "detach") else np.asarray(got)
    d = np.abs(got - ref).reshape(ref.shape[0], -1).max(axis=1)
    sU = np.abs(U).reshape(U.shape[0], -1).max(axis=1)
    off = U.shape[0] - ref.shape[0]
    mom = np.sqrt(sU[0] * sU[3])          # rho * sqrt(E): a momentum scale that is never 0
    base = [sU[0], max(sU[1], mom), max(sU[2], mom), sU[3]]
    scale = np.array([base[v + off] if v + off < 4 else sU[0] for v in range(ref.shape[0])])
    return float(np.max(d / scale))


@pytest.mark.parametrize("name", CASES)
def test_mixed_pieces(mech_path, mech, name):
    """Mixed form of each entry point against the FP64 oracle on FP32-rounded
    inputs; errors relative to the state they update (state_err) within
    MIXED_TOL, temperature relative within 1e-6."""
    sim, o, U, aux, theta = _pair(mech_path, mech, name, precision="mixed")
    U = U.astype(np.float32).astype(np.float64)
    aux = aux.astype(np.float32).astype(np.float64)
    a = sim.lf_alpha(_f32(U), _f32(aux), theta).cpu().numpy()
    np.testing.assert_allclose(a, o.lf_alpha(U, aux, theta), rtol=1e-6)
    got = sim.flux_func(_f32(U), _f32(aux), sim.dx, sim.dy, 3e-9, theta)
    assert bool(torch.isfinite(got).all())
    assert state_err(got, o.flux_func(U, aux, 3e-9, theta), U) <= MIXED_TOL["flux_func"]
    guess = aux.copy()
    guess[1] *= 1.02
    got = sim.update_func(_f32(U), _f32(guess))
    ref = o.update_func(U, guess.astype(np.float32).astype(np.float64))
    assert rel_err(got.double(), ref) <= MIXED_TOL["update_aux"]
    got = sim.reaction_source_terms(_f32(U), _f32(aux), 1e-8)
    ref = o.reaction_source_terms(U, aux, 1e-8)
    assert state_err(got[4:], ref[4:], U) <= MIXED_TOL["reaction"]
    Ug, ag, _, dt = sim.advance_func(_f32(U), _f32(aux), 0.0, theta)
    Ur, ar, _, dtr = o.advance_func(U, aux, 0.0, 0.3, theta)
    assert float(dt) == pytest.approx(dtr, rel=1e-6)
    assert state_err(Ug, Ur, Ur) <= MIXED_TOL["step"]
    assert rel_err(ag.double(), ar) <= MIXED_TOL["step"]


def test_mixed_precision_against_fp64_oracle(mech_path, mech):
    """Mixed form (FP32 state/WENO/flux, FP64 divergence, RK combine,
    chemistry and T solve) after 10 CFL steps of the detonation tube:
    relative L2 error of rho, p, T and every Y_k vs the FP64 oracle <= 1e-5
    and Linf (relative to the field max) <= 1e-4 (measured: <= 5.3e-7 and
    <= 2.4e-6)."""
    sim, o, U, aux, theta = _pair(mech_path, mech, "tube", precision="mixed")
    Ug = torch.as_tensor(U, dtype=torch.float32).cuda()
    ag = torch.as_tensor(aux, dtype=torch.float32).cuda()
    Ur, ar = U, aux
    for _ in range(10):
        Ug, ag, _, _ = sim.advance_func(Ug, ag, 0.0)
        Ur, ar, _, _ = o.advance_func(Ur, ar, 0.0, 0.3)
    Ug = Ug.double().cpu().numpy()
    ag = ag.double().cpu().numpy()
    fields = {"rho": (Ug[0], Ur[0]), "T": (ag[1], ar[1])}
    for k in range(8):
        fields["Y%d" % k] = (Ug[4 + k] / Ug[0], Ur[4 + k] / Ur[0])
    for name, (g, r) in fields.items():
        l2 = np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300)
        linf = np.abs(g - r).max() / max(np.abs(r).max(), 1e-300)
        print(name, "max", np.abs(r).max(), "L2", l2, "Linf", linf)
        assert l2 <= 1e-5 and linf <= 1e-4, name


# ---------------------------------------------------------- callable sides
def _blowing(xp_copy):
    """Bottom-side boundary function: mirrored layers with the wall-normal
    momentum reflected about a blowing velocity theta['vb']."""
    def fn(U_bd, aux_bd, theta):
        Ug = xp_copy(U_bd[:, :, [2, 1, 0]])
        ag = xp_copy(aux_bd[:, :, [2, 1, 0]])
        Ug[2] = -Ug[2] + 2.0 * Ug[0] * theta["vb"]
        return Ug, ag
    return fn


def _mirror_x(xp_copy):
    """Left-side boundary function: zero-gradient mirror written by hand."""
    def fn(U_bd, aux_bd, theta):
        return xp_copy(U_bd[:, [2, 1, 0], :]), xp_copy(aux_bd[:, [2, 1, 0], :])
    return fn


def _user_pair(mech_path, mech, precision="fp64", dt=None):
    from janc_amd.solver import Simulator
    U, aux, g = cases.detonation_tube(mech, nx=33, ny=26, driver=6, seed=8)
    theta = {"vb": 15.0}
    bc_np = dict(g["boundary_config"], bottom_boundary=_blowing(np.copy),
                 left_boundary=_mirror_x(np.copy))
    bc_t = dict(g["boundary_config"], bottom_boundary=_blowing(torch.clone),
                left_boundary=_mirror_x(torch.clone))
    cfg = cases.simulation_config(dict(g, boundary_config=bc_t), mech_path, cfl=0.3, dt=dt,
                                  precision=precision, theta=theta)
    sim = Simulator(cfg)
    o = Oracle(mech_path, bc_np, g["dx"], g["dy"])
    print("case user", precision, U.shape, flush=True)
    return sim, o, U, aux, theta


def test_callable_boundaries(mech_path, mech):
    sim, o, U, aux, theta = _user_pair(mech_path, mech)
    got = sim.flux_func(U, aux, sim.dx, sim.dy, 3e-9)
    assert rel_err(got, o.flux_func(U, aux, 3e-9, theta)) <= FP64_TOL
    a = sim.lf_alpha(U, aux).cpu().numpy()
    np.testing.assert_allclose(a, o.lf_alpha(U, aux, theta), rtol=1e-14)
    Ug, ag, _, dt = sim.advance_func(U, aux, 0.0)
    Ur, ar, _, dtr = o.advance_func(U, aux, 0.0, 0.3, theta)
    assert float(dt) == pytest.approx(dtr, rel=1e-14)
    assert rel_err(Ug, Ur) <= FP64_TOL and rel_err(ag, ar) <= FP64_TOL
    Ug, ag = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    _, _, dts = sim.advance(Ug, ag, 3)
    Ur, ar, t = U, aux, 0.0
    for k in range(3):
        Ur, ar, t, dtr = o.advance_func(Ur, ar, t, 0.3, theta)
        assert float(dts[k]) == pytest.approx(dtr, rel=1e-13)
    assert rel_err(Ug, Ur) <= FP64_TOL and rel_err(ag, ar) <= FP64_TOL


def test_callable_boundaries_need_fresh_strips(mech_path, mech):
    """A stage whose user strip was not written since the last use fails."""
    from janc_amd import _lib
    from janc_amd.solver import _ptr
    sim, o, U, aux, theta = _user_pair(mech_path, mech, dt=2e-9)
    Ut, at = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    ctx = sim._context(Ut, None)
    out = torch.empty_like(Ut)
    with pytest.raises(_lib.JancError):
        ctx.call("jh_rhs", _ptr(Ut), _ptr(at), 2e-9, _ptr(out))
