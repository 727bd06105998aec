"""Simulator.run's HDF5 snapshots (set_simulator.py:457-494 with H5Saver,
:126-318) on the GPU: the saved states equal the states of the same steps
taken one by one, a restart from a snapshot continues bit-identically, and
H5Saver.load's primitives match the oracle's update_aux from the reference's
initial guess (gamma 1.4, T 500 K)."""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

DT = 2.0 ** -28          # sums of steps are exact, so save times are too


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def _sim(mech_path, g, path, t_end, save_dt):
    from janc_amd.solver import Simulator
    cfg = cases.simulation_config(g, mech_path, dt=DT, t_end=t_end)
    cfg["computation_config"]["output_settings"] = {"save_dt": save_dt, "results_path": path}
    return Simulator(cfg)


def test_run_snapshots_and_restart(mech_path, mech, tmp_path):
    from janc_amd.h5 import H5File, H5Saver
    U0, a0, g = cases.detonation_tube(mech, nx=40, ny=24, driver=6)
    path = str(tmp_path / "run.h5")
    sim = _sim(mech_path, g, path, t_end=7.5 * DT, save_dt=2 * DT)
    print("run 8 steps, snapshots every 2", flush=True)
    U, aux, t = sim.run(U0, a0, save_aux=True)
    # the same steps one by one
    Ur, ar = torch.as_tensor(U0).cuda(), torch.as_tensor(a0).cuda()
    states = {0: (Ur.clone(), ar.clone())}
    for k in range(1, 9):
        Ur, ar, _, _ = sim.advance_func(Ur, ar, 0.0)
        states[k] = (Ur.clone(), ar.clone())
    assert torch.equal(U, Ur) and torch.equal(aux, ar)
    f = H5File(path, "r")
    keys = sorted(f.keys(), key=lambda s: int(s[9:]))
    # saves at t = 0 and at the first step with t >= 2, 4, 6 dt, then t_end
    assert keys == ["save_step%d" % k for k in range(5)]
    steps = []
    for key in keys:
        data, attrs = f.read_group(key)
        k = int(attrs["step"])
        steps.append(k)
        assert attrs["time"] == pytest.approx(k * DT, rel=1e-12)
        np.testing.assert_array_equal(data["u"], states[k][0].cpu().numpy())
        np.testing.assert_array_equal(data["aux"], states[k][1].cpu().numpy())
    assert steps == [0, 2, 4, 6, 8]
    f.close()
    # restart from the step-4 snapshot, 4 more steps
    s = H5Saver(path, "2D", simulator=sim)
    t4, k4, U4, a4 = s.load_state("save_step2")
    s.close()
    assert k4 == 4
    Uc, ac = torch.as_tensor(U4).cuda(), torch.as_tensor(a4).cuda()
    sim.advance(Uc, ac, 4)
    assert torch.equal(Uc, states[8][0]) and torch.equal(ac, states[8][1])


def test_load_primitives(mech_path, mech, tmp_path):
    from janc_amd.h5 import H5Saver
    from test_gpu_parity import make_oracle
    U0, a0, g = cases.detonation_tube(mech, nx=30, ny=20, driver=5)
    path = str(tmp_path / "p.h5")
    sim = _sim(mech_path, g, path, t_end=DT, save_dt=DT)
    s = H5Saver(path, "2D", simulator=sim)
    s.save(0, 0.0, 0, u=torch.as_tensor(U0).cuda())
    t, prim = s.load("save_step0")
    times, allp = s.load_all()
    s.close()
    o = make_oracle(mech_path, g, None)
    guess = np.stack([np.full(U0.shape[1:], 1.4), np.full(U0.shape[1:], 500.0)])
    T = o.update_func(U0, guess)[1]
    R = o.thermo.get_R(U0[4:] / U0[0])[0]
    assert t == 0.0 and list(times) == [0.0]
    np.testing.assert_array_equal(prim["density"].cpu().numpy(), U0[0])
    np.testing.assert_allclose(prim["tempreature"].cpu().numpy(), T, rtol=1e-12)
    np.testing.assert_allclose(prim["pressure"].cpu().numpy(), U0[0] * R * T, rtol=1e-12)
    np.testing.assert_allclose(prim["x-velocity"].cpu().numpy(), U0[1] / U0[0], rtol=1e-14)
    assert allp["species"].shape == (1, 8) + U0.shape[1:]
