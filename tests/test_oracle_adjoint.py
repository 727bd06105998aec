"""Pins the differentiable torch twin of the oracle step (oracle/torch_step.py).

Forward: the twin agrees with the numpy oracle's fixed-dt step.
Reverse: autograd's directional derivative agrees with a central finite
difference of the numpy oracle (step 1e-7 relative per component, bound
1e-6 relative; measured ~1e-8) on tie-free states, for slip, no-slip and
zero-gradient walls and for periodic sides.
"""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism
from oracle.step import Oracle

torch = pytest.importorskip("torch")
from oracle.torch_step import TorchOracle  # noqa: E402


def tie_free_state(mech_path, mech, name, nx, ny):
    """A state whose LF maxima have single maximisers and no |u| or |v| at
    zero: velocity noise of a few m/s and temperature noise of ~0.1 %, with
    aux recomputed from U.  (At exact ties the max is not differentiable:
    reverse mode splits the cotangent, a finite difference picks a side.)"""
    if name == "tube":
        U, aux, g = cases.detonation_tube(mech, nx=nx, ny=ny, driver=5, seed=3)
        g["boundary_config"]["top_boundary"] = "no_slip_wall"
        g["boundary_config"]["right_boundary"] = "zero_gradient"
    else:
        U, aux, g = cases.periodic_box(mech, nx=nx, ny=ny)
    rng = np.random.default_rng(11)
    rho = U[0]
    e_kin = 0.5 * (U[1] ** 2 + U[2] ** 2) / rho
    U = U.copy()
    U[1] += rho * 3.0 * rng.standard_normal(rho.shape)
    U[2] += rho * 3.0 * rng.standard_normal(rho.shape)
    e_int = U[3] - e_kin
    U[3] = e_int * (1.0 + 1e-3 * rng.standard_normal(rho.shape)) \
        + 0.5 * (U[1] ** 2 + U[2] ** 2) / rho
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    aux = o.update_func(U, aux)
    return U, aux, g


def _state(mech_path, mech, name):
    return tie_free_state(mech_path, mech, name, 18, 11) if name == "tube" else \
        tie_free_state(mech_path, mech, name, 14, 12)


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


@pytest.mark.parametrize("name", ["tube", "box"])
def test_twin_forward_matches_oracle(mech_path, mech, name):
    U, aux, g = _state(mech_path, mech, name)
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    to = TorchOracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    dt = 2e-9
    Ur, ar = o.advance_one_step(U, aux, dt)
    Ut, at = to.advance_one_step(torch.tensor(U), torch.tensor(aux), dt)
    d = np.abs(Ut.numpy() - Ur).reshape(12, -1).max(1) / np.abs(Ur).reshape(12, -1).max(1)
    assert d.max() < 1e-12
    assert np.abs(at.numpy() - ar).max() / np.abs(ar).max() < 1e-12


@pytest.mark.parametrize("name", ["tube", "box"])
def test_twin_gradient_matches_finite_difference(mech_path, mech, name):
    U, aux, g = _state(mech_path, mech, name)
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    to = TorchOracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    dt = 2e-9
    rng = np.random.default_rng(5)
    W = rng.standard_normal(U.shape)
    Wa = rng.standard_normal(aux.shape)
    Uv = torch.tensor(U, requires_grad=True)
    av = torch.tensor(aux, requires_grad=True)
    U4, a4 = to.advance_one_step(Uv, av, dt)
    ((U4 * torch.tensor(W)).sum() + (a4 * torch.tensor(Wa)).sum()).backward()

    def L(Ux, ax):
        A, B = o.advance_one_step(Ux, ax, dt)
        return float((A * W).sum() + (B * Wa).sum())

    # direction in U (relative size 1e-7 per cell) and in aux's gamma / T
    dU = rng.standard_normal(U.shape) * np.abs(U) * 1e-7
    fd = (L(U + dU, aux) - L(U - dU, aux)) / 2
    ad = float((Uv.grad.numpy() * dU).sum())
    assert abs(fd - ad) <= 1e-6 * abs(fd)
    da = np.zeros_like(aux)
    da[0] = 1e-7 * aux[0] * rng.standard_normal(aux.shape[1:])
    da[1] = 1e-7 * aux[1] * rng.standard_normal(aux.shape[1:])
    fd = (L(U, aux + da) - L(U, aux - da)) / 2
    ad = float((av.grad.numpy() * da).sum())
    assert abs(fd - ad) <= 1e-6 * abs(fd)
