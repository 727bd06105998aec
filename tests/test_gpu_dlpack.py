"""DLPack bridge with a non-torch producer and consumer: state arrays in
hipMalloc memory exported by tests/dlpack_ctypes.py (the role a JAX ROCm
array plays) go into the solver zero-copy and are advanced in place; the
solver's output tensor is read back through its own DLPack capsule by the
ctypes consumer (the role jax.dlpack.from_dlpack plays)."""
import gc

import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]


@pytest.fixture(scope="module")
def tube(mech_path):
    from janc_amd.solver import Simulator
    mech = Mechanism(mech_path)
    U, aux, g = cases.detonation_tube(mech, nx=40, ny=24, driver=6, seed=21)
    sim = Simulator(cases.simulation_config(g, mech_path, dt=2e-9))
    return sim, U, aux


def test_foreign_producer_step_and_export(tube):
    from dlpack_ctypes import Producer, consume, kDLROCM
    sim, U, aux = tube
    pu, pa = Producer(U, device=0), Producer(aux, device=0)
    try:
        tu = torch.from_dlpack(pu)
        assert tu.is_cuda and tu.data_ptr() == pu.ptr          # zero copy
        del tu
        print("dlpack: one step from hipMalloc'd state", flush=True)
        U1, a1 = sim.advance_one_step(pu, pa, sim.dx, sim.dy, 2e-9)
        R1, r1 = sim.advance_one_step(torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda(),
                                      sim.dx, sim.dy, 2e-9)
        assert torch.equal(U1, R1) and torch.equal(a1, r1)
        # in place on the producer's memory: two fused steps
        print("dlpack: advance in place", flush=True)
        sim.advance(pu, pa, 2)
        Rt, rt = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
        sim.advance(Rt, rt, 2)
        torch.cuda.synchronize()
        assert np.array_equal(pu.read(), Rt.cpu().numpy())
        assert np.array_equal(pa.read(), rt.cpu().numpy())
        # export: the solver's output through its DLPack capsule
        out, (dtype, did, ptr) = consume(torch.utils.dlpack.to_dlpack(U1))
        assert dtype == kDLROCM and did == 0 and ptr == U1.data_ptr()
        assert np.array_equal(out, U1.cpu().numpy())
    finally:
        gc.collect()
        torch.cuda.synchronize()
        pu.free()
        pa.free()
