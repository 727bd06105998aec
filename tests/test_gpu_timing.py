"""Per-launch HIP-event timing of the fused step (jh_timing_enable): every
launch, or with a sampling period k the launches of every k-th step; the
sampled kernel durations agree with the unsampled ones, and the state the
step produces does not depend on the timing."""
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism
from janc_amd.solver import Simulator

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]


def _sim():
    mech = Mechanism()
    U, aux, g, theta = cases.rdc_annulus(mech, nx=256, ny=128, seed_cols=20)
    sim = Simulator(cases.simulation_config(g, cfl=0.3, precision="fp64", theta=theta))
    return sim, torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()


def test_sampled_launch_counts_and_durations():
    sim, U, aux = _sim()
    Ua, aa = U.clone(), aux.clone()
    sim.advance(Ua, aa, 2)
    for every, steps, sampled in ((1, 9, 9), (3, 9, 3), (4, 9, 3), (0, 9, 0)):
        sim.timing(Ua, enable=every)
        sim.advance(Ua, aa, steps)
        kt = sim.timing(Ua)
        sim.timing(Ua, enable=False)
        print("every", every, kt, flush=True)
        assert kt["stage"][1] == 3 * sampled and kt["bc"][1] == 3 * sampled
        assert kt["chem"][1] == sampled
        assert kt["reduce"][1] == (1 if every else 0)
        if sampled:
            assert kt["stage"][0] > 0.0 and kt["chem"][0] > 0.0
    # sampled and unsampled averages of the same kernels agree
    avg = {}
    for every in (1, 4):
        sim.timing(Ua, enable=every)
        sim.advance(Ua, aa, 16)
        kt = sim.timing(Ua)
        sim.timing(Ua, enable=False)
        avg[every] = kt["stage"][0] / kt["stage"][1]
    assert abs(avg[4] - avg[1]) <= 0.25 * avg[1]


def test_timing_does_not_change_the_state():
    sim, U, aux = _sim()
    Ua, aa, Ub, ab = U.clone(), aux.clone(), U.clone(), aux.clone()
    sim.advance(Ua, aa, 6)
    sim.timing(Ub, enable=2)
    sim.advance(Ub, ab, 6)
    sim.timing(Ub)
    sim.timing(Ub, enable=False)
    assert torch.equal(Ua, Ub) and torch.equal(aa, ab)
