"""AMR blocks (jh_blocks_*, janc_amd.amr) against the oracle's restatement
of the reference's AMR branch (oracle/amr.py: set_simulator.py:37-42,
:107-111, time_step/TVD_RK3.py:15-37) at the FP64 bar (1e-12 relative per
variable).

The AMR hierarchy (jaxamr's get_ghost_block_data / update_external_boundary)
is a third-party library that is absent here; the tests stand in a uniform
block grid that tiles one domain: a block's ghost layers are its neighbours'
cells, and the domain's boundary padding at the domain's edges.  Blocks are
ragged against the kernel's tile (21 x 37 cells; 8-row, 64-column tiles)."""
import numpy as np
import pytest

from janc_amd import _lib, cases
from janc_amd.mechanism import Mechanism
from oracle import amr as oamr
from oracle.step import Oracle

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

TOL = 1e-12
BNX, BNY, NBX, NBY = 21, 37, 3, 2


def rel_err(got, ref):
    got = got.detach().cpu().numpy() if hasattr(got, "detach") else np.asarray(got)
    ref = np.asarray(ref)
    nv = ref.shape[1]
    g = np.moveaxis(got, 1, 0).reshape(nv, -1)
    r = np.moveaxis(ref, 1, 0).reshape(nv, -1)
    d = np.abs(g - r).max(1)
    s = np.abs(r).max(1)
    return float(np.max(d / np.maximum(s, 1e-300)))


class BlockGrid:
    """NBX x NBY blocks of BNX x BNY cells tiling one domain (block k = (i, j),
    k = i * NBY + j): the test's stand-in for the AMR hierarchy's ghost
    refill (neighbours' cells; the domain's boundary padding at its edges)."""

    def __init__(self, o):
        self.o = o

    def split(self, full):
        out = np.empty((NBX * NBY, full.shape[0], BNX, BNY))
        for i in range(NBX):
            for j in range(NBY):
                out[i * NBY + j] = full[:, i * BNX:(i + 1) * BNX, j * BNY:(j + 1) * BNY]
        return out

    def assemble(self, blk):
        full = np.empty((blk.shape[1], NBX * BNX, NBY * BNY))
        for i in range(NBX):
            for j in range(NBY):
                full[:, i * BNX:(i + 1) * BNX, j * BNY:(j + 1) * BNY] = blk[i * NBY + j]
        return full

    def ghost(self, blk):
        full = self.assemble(blk)
        Ug, ag = self.o.boundary(full[:-2], full[-2:])
        pad = np.concatenate([Ug, ag])
        out = np.empty((blk.shape[0], blk.shape[1], BNX + 6, BNY + 6))
        for i in range(NBX):
            for j in range(NBY):
                out[i * NBY + j] = pad[:, i * BNX:i * BNX + BNX + 6, j * BNY:j * BNY + BNY + 6]
        return out

    # the two functions advance_func_amr takes from the AMR hierarchy
    def get_ghost_block_data(self, blk, info):
        dt = blk.dtype
        return torch.as_tensor(self.ghost(blk.double().cpu().numpy()), dtype=dt).cuda()

    def update_external_boundary(self, level, parent, interiors, info):
        return interiors.contiguous()


@pytest.fixture(scope="module")
def setup(mech_path):
    mech = Mechanism(mech_path)
    U, aux, g = cases.detonation_tube(mech, nx=NBX * BNX, ny=NBY * BNY, driver=10, seed=13)
    # velocity noise so every block has its own LF maxima
    rng = np.random.default_rng(14)
    U[1] += U[0] * 40.0 * rng.standard_normal(U[0].shape)
    U[2] += U[0] * 40.0 * rng.standard_normal(U[0].shape)
    U[3] += 0.5 * (U[1] ** 2 + U[2] ** 2) / U[0]
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    grid = BlockGrid(o)
    blk = grid.split(np.concatenate([U, aux]))
    return mech, g, o, grid, blk


def _solver(mech_path, g, precision="fp64"):
    from janc_amd.amr import BlockSolver
    return BlockSolver(cases.simulation_config(g, mech_path, dt=2e-9, precision=precision))


def _dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


def test_blocks_flux(mech_path, setup):
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g)
    g0 = grid.ghost(blk)
    print("blocks flux", g0.shape, flush=True)
    got = bs.flux_func(_dev(g0), g["dx"], g["dy"], 2e-9)
    ref = oamr.flux_func(o, g0[:, :-2], g0[:, -2:], 2e-9)
    assert rel_err(got, ref) <= TOL
    assert float(got[:, :, :3].abs().max()) == 0.0 and float(got[:, :, :, -3:].abs().max()) == 0.0


@pytest.mark.parametrize("stage", [0, 1, 2])
def test_blocks_stage(mech_path, setup, stage):
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g)
    g0 = grid.ghost(blk)
    # a stage input that differs from the step input (one stage further)
    gs = grid.ghost(oamr.stage(o, 0, g0, g0, 2e-9)[..., 3:-3, 3:-3]) if stage else g0
    print("blocks stage", stage, flush=True)
    got = bs.stage(stage, _dev(gs), _dev(g0), g["dx"], g["dy"], 2e-9)
    ref = oamr.stage(o, stage, gs, g0, 2e-9)
    assert rel_err(got, ref) <= TOL
    assert bs.health() == 0


def test_blocks_chemistry(mech_path, setup):
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g)
    b = _dev(blk)
    bs.chemistry(b, g["dx"], g["dy"], 2e-9)
    ref = oamr.chemistry(o, blk, 2e-9)
    assert rel_err(b, ref) <= TOL


def test_blocks_step(mech_path, setup):
    """Two steps of advance_func_amr with the block grid's ghost refill."""
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g)
    b = _dev(blk)
    ref = blk
    for _ in range(2):
        b = bs.advance_one_step(1, None, g["dx"], g["dy"], 2e-9, b, None, amr=grid)
        ref = oamr.advance_one_step(o, ref, 2e-9, grid.ghost)
    assert rel_err(b, ref) <= TOL
    assert bs.health() == 0


def test_blocks_have_their_own_maxima(mech_path, setup):
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g)
    g0 = grid.ghost(blk)
    g1 = g0.copy()
    g1[4, 1, 10, 10] += g1[4, 0, 10, 10] * 900.0      # a fast cell in block 4
    a = bs.flux_func(_dev(g0), g["dx"], g["dy"], 2e-9)
    b = bs.flux_func(_dev(g1), g["dx"], g["dy"], 2e-9)
    assert torch.equal(a[:4], b[:4]) and torch.equal(a[5:], b[5:])
    assert not torch.equal(a[4], b[4])


def test_blocks_mixed_precision(mech_path, setup):
    """The mixed form of the block step (FP32 blocks) within the mixed bounds
    of the single-block tests against the FP64 oracle after one step."""
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g, precision="mixed")
    b = bs.advance_one_step(1, None, g["dx"], g["dy"], 2e-9, _dev(blk, torch.float32), None,
                            amr=grid)
    ref = oamr.advance_one_step(o, blk, 2e-9, grid.ghost)
    got = b.double().cpu().numpy()
    for k in (0, 3, 13):                              # rho, rho E, T
        d = np.linalg.norm(got[:, k] - ref[:, k]) / np.linalg.norm(ref[:, k])
        assert d < 1e-6, (k, d)


def test_blocks_health_flag(mech_path, setup):
    mech, g, o, grid, blk = setup
    bs = _solver(mech_path, g)
    b = blk.copy()
    b[2, 0, 5, 5] = -b[2, 0, 5, 5]                    # rho < 0 in one block's cell
    bt = _dev(b)
    bs.chemistry(bt, g["dx"], g["dy"], 2e-9)
    f = bs.health()
    assert f & _lib.HEALTH_FLAGS["negative_rho"], _lib.health_names(f)
    assert bs.health() == 0                           # cleared by the read


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gpu_ghost_refill_equals_torch_gather(dtype):
    """UniformBlockGrid's GPU refill (jh_blocks_fill_ghosts) equals its torch
    gather on the CPU bit for bit, from contiguous interior blocks and from
    the interior view of padded blocks, for every copy kind."""
    from janc_amd.amr import UniformBlockGrid
    bc = {"left_boundary": "no_slip_wall", "right_boundary": "zero_gradient",
          "bottom_boundary": "periodic", "top_boundary": "periodic"}
    nbx, nby, bnx, bny = 3, 4, 7, 5
    g_gpu = UniformBlockGrid(bc, nbx, nby, bnx, bny, device="cuda")
    g_cpu = UniformBlockGrid(bc, nbx, nby, bnx, bny, device="cpu")
    gen = torch.Generator().manual_seed(2)
    blk = torch.randn((nbx * nby, 14, bnx, bny), generator=gen, dtype=torch.float64).to(dtype)
    ref = g_cpu.get_ghost_block_data(blk)
    got = g_gpu.get_ghost_block_data(blk.cuda())
    assert torch.equal(got.cpu(), ref)
    pad = torch.randn((nbx * nby, 14, bnx + 6, bny + 6), generator=gen,
                      dtype=torch.float64).to(dtype)
    pad[..., 3:-3, 3:-3] = blk
    view = g_gpu.update_external_boundary(1, None, pad.cuda()[..., 3:-3, 3:-3])
    assert torch.equal(g_gpu.get_ghost_block_data(view).cpu(), ref)
