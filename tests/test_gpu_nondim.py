"""Nondimensional configurations (nondim_config, preprocess/nondim.py:14-25)
on the GPU against the FP64 oracle with the same scales: the state, grid,
dt and back pressure in nondimensional units, the mechanism constants
rescaled (A, Ea/R, molar masses, NASA-7 T0 in the entropy's log term).

The temperature solve's tolerance, 5e-9 (thermo_model.py:18), is absolute in
the configuration's energy unit.  The reference iterates every cell until
the grid-wide max|res| is within it (:140-142); with these scales (e ~ 10) a
cell that converged before the slowest one is still ~1e-10 relative from its
root, one update short of where the reference leaves it.  A nondimensional
configuration therefore runs the grid-wide stop (the Simulator's 'auto'
choice, JH_HALLEY_GRID), held at the 1e-12 bar against the oracle's
reference stop; the per-cell stop is checked against the oracle's per-cell
form at 1e-12 and shown to be ~1e-9 away from the reference stop (so the
grid-wide tests exercise cells that stop early)."""
import numpy as np
import pytest

from janc_amd import _lib, cases
from janc_amd.mechanism import Mechanism, NondimScales

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

from test_gpu_parity import FP64_TOL, rel_err  # noqa: E402
from oracle.step import Oracle  # noqa: E402

ND = {"P0": 1.0e5, "T0": 300.0, "R0": 287.0, "x0": 1.0e-3}
# the rotating-detonation example's scales (examples/janc_basic_example2.ipynb:
# manifold pressure and temperature, H2-air gas constant, combustor length)
ND_RDC = {"P0": 10 * 101325.0, "T0": 300.0, "R0": 369.0, "x0": 0.08}


def _nondim_state(U, aux, s):
    Un = U.copy()
    Un[0] /= s.rho0
    Un[1:3] /= s.rho0 * s.u0
    Un[3] /= s.rho0 * s.e0
    Un[4:] /= s.rho0
    an = aux.copy()
    an[1] /= s.T0
    return Un, an


def _case(mech, name):
    """(state, geometry, theta, scales) of a nondimensional case."""
    if name == "annulus":
        U, aux, g, th = cases.rdc_annulus(mech, nx=61, ny=37, seed_cols=12)
        nd = ND_RDC
        s = NondimScales(nd)
        theta = {"Pb": th["Pb"] / nd["P0"], "P0_inj": th["P0_inj"] / nd["P0"],
                 "T0_inj": th["T0_inj"] / nd["T0"], "Yinj": th["Yinj"]}
    else:
        U, aux, g = cases.detonation_tube(mech, nx=37, ny=29, driver=6, seed=11)
        nd = ND
        s = NondimScales(nd)
        theta = None
        if name == "outlet":
            g["boundary_config"]["right_boundary"] = "pressure_outlet"
            theta = {"Pb": 1.5 * 101325.0 / nd["P0"]}
    Un, an = _nondim_state(U, aux, s)
    gn = dict(g, dx=g["dx"] / s.x0, dy=g["dy"] / s.x0)
    return Un, an, gn, theta, nd, s


def _oracle(mech_path, g, theta, nd, per_cell=False):
    """Nondimensional oracle for a geometry; an injection-inlet side gets the
    restated example boundary function (a user callable in the reference)."""
    bcfg = dict(g["boundary_config"])
    inlets = [s for s in ("left", "right", "bottom", "top")
              if bcfg[s + "_boundary"] == "injection_inlet"]
    for s in inlets:
        bcfg[s + "_boundary"] = "slip_wall"
    o = Oracle(mech_path, bcfg, g["dx"], g["dy"], nondim_config=nd, per_cell_halley=per_cell)
    if inlets:
        from oracle.inlet import attach_inlet
        for s in inlets:
            attach_inlet(o, s, theta["P0_inj"], theta["T0_inj"], theta["Yinj"])
    return o


def _sim(mech_path, g, nd, theta, stop=None, **kw):
    from janc_amd.solver import Simulator
    cfg = cases.simulation_config(g, mech_path, theta=theta, **kw)
    cfg["nondim_config"] = nd
    if stop is not None:
        cfg["computation_config"]["halley_stop"] = stop
    return Simulator(cfg)


@pytest.mark.parametrize("name", ["tube", "outlet", "annulus"])
def test_nondim_step_grid_stop(mech_path, name):
    mech = Mechanism(mech_path)
    Un, an, gn, theta, nd, s = _case(mech, name)
    dt = 3e-9 / s.t0
    sim = _sim(mech_path, gn, nd, theta, dt=dt)
    assert sim.halley_stop == "grid"            # 'auto' on a nondimensional configuration
    o = _oracle(mech_path, gn, theta, nd)
    print("case nondim grid stop", name, Un.shape, flush=True)
    got = sim.flux_func(Un, an, sim.dx, sim.dy, dt, theta)
    assert rel_err(got, o.flux_func(Un, an, dt, theta)) <= FP64_TOL
    Ug, ag = sim.advance_one_step(Un, an, sim.dx, sim.dy, dt, theta)
    Ur, ar = o.advance_one_step(Un, an, dt, theta)
    assert rel_err(Ug, Ur) <= FP64_TOL and rel_err(ag, ar) <= FP64_TOL
    # five fused steps
    U5 = torch.as_tensor(Un).cuda()
    a5 = torch.as_tensor(an).cuda()
    sim.advance(U5, a5, 5, theta)
    Ur, ar = Un, an
    for _ in range(5):
        Ur, ar = o.advance_one_step(Ur, ar, dt, theta)
    assert rel_err(U5, Ur) <= FP64_TOL and rel_err(a5, ar) <= FP64_TOL
    assert sim.health(U5) == 0, _lib.health_names(sim.health(U5))
    # the temperature field is the dimensional one over T0
    assert 1.0 < float(a5[1].max()) < 20.0


@pytest.mark.parametrize("name", ["tube", "annulus"])
def test_nondim_cfl_steps_grid_stop(mech_path, name):
    """CFL steps: dt and the next step's LF maxima come from the candidate
    maxima of the N* the fix-up applies."""
    mech = Mechanism(mech_path)
    Un, an, gn, theta, nd, s = _case(mech, name)
    sim = _sim(mech_path, gn, nd, theta, cfl=0.3)
    o = _oracle(mech_path, gn, theta, nd)
    Ug, ag = torch.as_tensor(Un).cuda(), torch.as_tensor(an).cuda()
    dth = torch.empty(4, dtype=torch.float64, device="cuda")
    sim.advance(Ug, ag, 4, theta, dt_hist=dth)
    Ur, ar, t = Un, an, 0.0
    dts = []
    for _ in range(4):
        Ur, ar, t, dt = o.advance_func(Ur, ar, t, 0.3, theta)
        dts.append(dt)
    np.testing.assert_allclose(dth.cpu().numpy(), dts, rtol=1e-13)
    assert rel_err(Ug, Ur) <= FP64_TOL and rel_err(ag, ar) <= FP64_TOL


def test_nondim_per_cell_stop_differs_from_reference(mech_path):
    mech = Mechanism(mech_path)
    Un, an, gn, theta, nd, s = _case(mech, "tube")
    dt = 3e-9 / s.t0
    sim = _sim(mech_path, gn, nd, theta, stop="per_cell", dt=dt)
    op = _oracle(mech_path, gn, theta, nd, per_cell=True)
    o = _oracle(mech_path, gn, theta, nd)
    Ug, ag = sim.advance_one_step(Un, an, sim.dx, sim.dy, dt, theta)
    Up, ap = op.advance_one_step(Un, an, dt, theta)
    assert rel_err(Ug, Up) <= FP64_TOL and rel_err(ag, ap) <= FP64_TOL
    Ur, ar = o.advance_one_step(Un, an, dt, theta)
    d = max(rel_err(Ug, Ur), rel_err(ag, ar))
    assert 1e-11 < d < 1e-8, d


def test_nondim_update_aux_grid_stop(mech_path):
    mech = Mechanism(mech_path)
    Un, an, gn, theta, nd, s = _case(mech, "tube")
    sim = _sim(mech_path, gn, nd, theta, dt=1e-3)
    o = _oracle(mech_path, gn, theta, nd)
    # a state whose temperature guess is off by a few per cent
    a0 = an.copy()
    a0[1] *= 1.0 + 0.03 * np.sin(np.arange(a0[1].size).reshape(a0[1].shape))
    got = sim.update_func(Un, a0)
    ref = o.update_func(Un, a0)
    assert rel_err(got, ref) <= FP64_TOL
