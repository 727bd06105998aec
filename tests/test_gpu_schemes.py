"""GPU parity of the other flux schemes of set_flux_solver (flux.py:19-40):
the Godunov path (primitive reconstruction + HLLC, flux.py:43-54) with
WENO5_JS and MUSCL, LF flux splitting with MUSCL, and Steger-Warming flux
splitting (flux_splitting/SW.py) with both, against the FP64 oracle.

Bar as for the default scheme: FP64 <= 1e-12 relative error per variable;
mixed form within the mixed-precision bounds of test_gpu_parity.
"""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

from test_gpu_parity import (FP64_TOL, MIXED_TOL, _case, _f32, make_oracle,  # noqa: E402
                             rel_err, state_err)

SCHEMES = {
    "hllc_weno": {"solver_type": "godunov", "interface_reconstruction": "WENO5_JS",
                  "riemann_solver": "HLLC"},
    "hllc_muscl": {"solver_type": "godunov", "interface_reconstruction": "MUSCL",
                   "riemann_solver": "HLLC"},
    "lf_muscl": {"solver_type": "flux_splitting", "interface_reconstruction": "MUSCL",
                 "split_method": "LF"},
    "sw_weno": {"solver_type": "flux_splitting", "interface_reconstruction": "WENO5_JS",
                "split_method": "SW"},
    "sw_muscl": {"solver_type": "flux_splitting", "interface_reconstruction": "MUSCL",
                 "split_method": "SW"},
}
CASES = ["box", "tube", "outlet", "inlet"]


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def _pair(mech_path, mech, name, scheme, precision="fp64", dt=None):
    from janc_amd.solver import Simulator
    U, aux, g, theta = _case(name, mech)
    fc = SCHEMES[scheme]
    cfg = cases.simulation_config(g, mech_path, cfl=0.3, dt=dt, precision=precision,
                                  theta=theta, flux=fc)
    sim = Simulator(cfg)
    o = make_oracle(mech_path, g, theta, flux_config=fc)
    print("case", name, scheme, precision, U.shape, flush=True)
    return sim, o, U, aux, theta


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("name", CASES)
def test_flux_func(mech_path, mech, name, scheme):
    sim, o, U, aux, theta = _pair(mech_path, mech, name, scheme)
    got = sim.flux_func(U, aux, sim.dx, sim.dy, 3e-9, theta)
    ref = o.flux_func(U, aux, 3e-9, theta)
    assert rel_err(got, ref) <= FP64_TOL


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("name", CASES)
def test_advance_func_one_step(mech_path, mech, name, scheme):
    sim, o, U, aux, theta = _pair(mech_path, mech, name, scheme)
    Ug, ag, t, dt = sim.advance_func(U, aux, 0.0, theta)
    Ur, ar, tr, dtr = o.advance_func(U, aux, 0.0, 0.3, theta)
    assert float(dt) == pytest.approx(dtr, rel=1e-14)
    assert rel_err(Ug, Ur) <= FP64_TOL
    assert rel_err(ag, ar) <= FP64_TOL


@pytest.mark.parametrize("scheme", ["hllc_weno", "hllc_muscl", "sw_weno"])
@pytest.mark.parametrize("name", ["tube", "inlet"])
def test_fixed_dt_five_steps(mech_path, mech, name, scheme):
    sim, o, U, aux, theta = _pair(mech_path, mech, name, scheme, dt=4e-9)
    Ug, ag = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    Ur, ar = U, aux
    for _ in range(5):
        Ug, ag = sim.advance_one_step(Ug, ag, sim.dx, sim.dy, 4e-9, theta)
        Ur, ar = o.advance_one_step(Ur, ar, 4e-9, theta)
    assert rel_err(Ug, Ur) <= FP64_TOL
    assert rel_err(ag, ar) <= FP64_TOL


@pytest.mark.parametrize("scheme", ["hllc_weno", "lf_muscl", "sw_weno"])
@pytest.mark.parametrize("name", ["tube", "inlet"])
def test_mixed(mech_path, mech, name, scheme):
    sim, o, U, aux, theta = _pair(mech_path, mech, name, scheme, precision="mixed")
    U = U.astype(np.float32).astype(np.float64)
    aux = aux.astype(np.float32).astype(np.float64)
    got = sim.flux_func(_f32(U), _f32(aux), sim.dx, sim.dy, 3e-9, theta)
    assert bool(torch.isfinite(got).all())
    assert state_err(got, o.flux_func(U, aux, 3e-9, theta), U) <= MIXED_TOL["flux_func"]
    Ug, ag, _, dt = sim.advance_func(_f32(U), _f32(aux), 0.0, theta)
    Ur, ar, _, dtr = o.advance_func(U, aux, 0.0, 0.3, theta)
    assert float(dt) == pytest.approx(dtr, rel=1e-6)
    assert state_err(Ug, Ur, Ur) <= MIXED_TOL["step"]
    assert rel_err(ag.double(), ar) <= MIXED_TOL["step"]


def test_adjoint_and_batch_refuse_other_schemes(mech_path, mech):
    sim, o, U, aux, theta = _pair(mech_path, mech, "box", "hllc_weno", dt=2e-9)
    with pytest.raises(NotImplementedError):
        sim.step_vjp(U, aux, 2e-9, np.zeros_like(U), np.zeros_like(aux))
