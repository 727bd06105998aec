"""GPU tests of the batched step (jh_batch_*, janc_amd.batch).

Bar: bit-identical to advancing each case alone through jh_advance (the
single-case path is itself held to the oracle in test_gpu_parity.py), for
uneven grid sizes, every boundary kind, CFL and fixed dt, FP64 and mixed;
HIP-graph replays identical to eager launches, with eager torch work and
synchronisations between replays.
"""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def _cases(mech, mech_path, precision, dt=None, copy_only=False):
    out = []
    sizes = [(37, 20), (64, 40), (29, 67), (128, 32), (20, 70), (45, 22)]
    for k, (nx, ny) in enumerate(sizes):
        if k == 1:
            U, aux, g = cases.periodic_box(mech, nx=nx, ny=ny)
            theta = None
        else:
            U, aux, g = cases.detonation_tube(mech, nx=nx, ny=ny, driver=5, seed=k,
                                              phi=0.8 + 0.1 * k, T_drv=2500.0 + 100 * k)
            theta = None
        if k == 2 and copy_only:   # walls of both kinds, no computed ghosts
            g["boundary_config"] = {"left_boundary": "no_slip_wall",
                                    "right_boundary": "zero_gradient",
                                    "bottom_boundary": "zero_gradient",
                                    "top_boundary": "slip_wall"}
        elif k == 2:
            g["boundary_config"] = {"left_boundary": "no_slip_wall",
                                    "right_boundary": "pressure_outlet",
                                    "bottom_boundary": "zero_gradient",
                                    "top_boundary": "pressure_outlet"}
            theta = {"Pb": 1.5 * 101325.0}
        cfg = cases.simulation_config(g, mech_path, cfl=0.3, dt=dt, precision=precision,
                                      theta=theta)
        if theta is not None:
            cfg["solver_parameters"] = theta
        out.append((cfg, U, aux))
    return out


def _dev(x, dtype):
    return torch.tensor(x, dtype=dtype, device="cuda")


@pytest.mark.parametrize("precision,dt", [("fp64", None), ("mixed", None), ("fp64", 2e-9)])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_cases(mech_path, mech, precision, dt, graph):
    from janc_amd.batch import BatchSimulator
    from janc_amd.solver import Simulator
    cs = _cases(mech, mech_path, precision, dt)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    Us = [_dev(U, dtype) for _, U, _ in cs]
    As = [_dev(a, dtype) for _, _, a in cs]
    print("batch", precision, dt, "graph", graph, [tuple(U.shape[1:]) for U in Us], flush=True)
    b = BatchSimulator([c for c, _, _ in cs], Us, As)
    nsteps = 4
    hist = b.advance(nsteps, graph=graph, dt_hist=True)
    torch.cuda.synchronize()
    for k, (cfg, U, aux) in enumerate(cs):
        sim = Simulator(cfg)
        Ur, Ar = _dev(U, dtype), _dev(aux, dtype)
        _, _, h1 = sim.advance(Ur, Ar, nsteps)
        assert torch.equal(Us[k], Ur), "case %d U" % k
        assert torch.equal(As[k], Ar), "case %d aux" % k
        assert torch.equal(hist[k], h1), "case %d dt" % k
    assert b.health() == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_of_copy_kinds_matches_single_cases(mech_path, mech, precision):
    """Every side of every case a copy kind: the batched stage kernel reads
    the halo from the source cells and no ghost-strip kernel runs; still
    bit-identical to the single cases."""
    from janc_amd.batch import BatchSimulator
    from janc_amd.solver import Simulator
    cs = _cases(mech, mech_path, precision, copy_only=True)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    Us = [_dev(U, dtype) for _, U, _ in cs]
    As = [_dev(a, dtype) for _, _, a in cs]
    print("batch copy kinds", precision, flush=True)
    b = BatchSimulator([c for c, _, _ in cs], Us, As)
    hist = b.advance(4, graph=True, dt_hist=True)
    torch.cuda.synchronize()
    for k, (cfg, U, aux) in enumerate(cs):
        Ur, Ar = _dev(U, dtype), _dev(aux, dtype)
        _, _, h1 = Simulator(cfg).advance(Ur, Ar, 4)
        assert torch.equal(Us[k], Ur), "case %d U" % k
        assert torch.equal(As[k], Ar), "case %d aux" % k
        assert torch.equal(hist[k], h1), "case %d dt" % k
    assert b.health() == 0


def test_graph_replays_match_eager(mech_path, mech):
    from janc_amd.batch import BatchSimulator
    cs = _cases(mech, mech_path, "mixed")
    f32 = torch.float32
    Ug = [_dev(U, f32) for _, U, _ in cs]
    Ag = [_dev(a, f32) for _, _, a in cs]
    Ue = [U.clone() for U in Ug]
    Ae = [a.clone() for a in Ag]
    bg = BatchSimulator([c for c, _, _ in cs], Ug, Ag)
    be = BatchSimulator([c for c, _, _ in cs], Ue, Ae)
    for rep in range(4):
        print("replay", rep, flush=True)
        bg.advance(3, graph=True)
        be.advance(3, graph=False)
        junk = torch.randn(1 << 16, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for k in range(len(cs)):
            assert torch.equal(Ug[k], Ue[k]) and torch.equal(Ag[k], Ae[k]), (rep, k)
    # rebinding invalidates the captured graph
    Un = [_dev(U, f32) for _, U, _ in cs]
    An = [_dev(a, f32) for _, _, a in cs]
    Ur = [U.clone() for U in Un]
    Ar = [a.clone() for a in An]
    bg.bind(Un, An)
    bg.advance(3, graph=True)
    be.bind(Ur, Ar)
    be.advance(3, graph=False)
    torch.cuda.synchronize()
    for k in range(len(cs)):
        assert torch.equal(Un[k], Ur[k]) and torch.equal(An[k], Ar[k])


def test_batch_tile_variants_agree(mech_path, mech):
    from janc_amd.autotune import variants
    from janc_amd.batch import BatchSimulator
    cs = _cases(mech, mech_path, "mixed")
    f32 = torch.float32
    ref = None
    for v in variants():
        Us = [_dev(U, f32) for _, U, _ in cs]
        As = [_dev(a, f32) for _, _, a in cs]
        b = BatchSimulator([c for c, _, _ in cs], Us, As)
        b.set_tile(v["variant"])
        b.advance(2, graph=True)
        torch.cuda.synchronize()
        if ref is None:
            ref = (Us, As)
        else:
            for k in range(len(cs)):
                assert torch.equal(Us[k], ref[0][k]) and torch.equal(As[k], ref[1][k])


def test_batch_errors(mech_path, mech):
    import ctypes
    from janc_amd import _lib
    from janc_amd.batch import BatchSimulator
    cs = _cases(mech, mech_path, "fp64")[:2]
    Us = [_dev(U, torch.float64) for _, U, _ in cs]
    As = [_dev(a, torch.float64) for _, _, a in cs]
    b = BatchSimulator([c for c, _, _ in cs], Us, As)
    with pytest.raises(_lib.JancError):
        b.advance(0)
    with pytest.raises(ValueError):
        b.bind(Us[::-1], As[::-1])          # shapes do not match the cases
    with pytest.raises(_lib.JancError):
        b.set_tile(99)
    lib = _lib.load()
    assert lib.jh_batch_advance(None, 1e-9, 0.0, 1, None, 0) == _lib.JH_ERR_ARG
    h = ctypes.c_void_p()
    one = (ctypes.c_int * 1)(2)
    d = (ctypes.c_double * 1)(1e-4)
    k = (ctypes.c_int * 4)(0, 0, 0, 0)
    c = mech.const_array()
    rc = lib.jh_batch_create(1, one, one, d, d, k, None, 0,
                             c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), c.size, None,
                             ctypes.byref(h))
    assert rc == _lib.JH_ERR_ARG and b"nx and ny" in lib.jh_batch_last_error(None)
