"""janc_amd.amr.UniformBlockGrid (the ghost refill of a uniform single-level
block grid, torch on the CPU here) against the oracle's boundary padding
(oracle/boundary.py: boundary.py:126-132, boundary_padding.py:20-25 and the
copied BC kinds): every block's padded window equals the padded domain's
window bit for bit, for every copied BC kind on every side, including the
periodic corners; split/assemble are inverse."""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism
from oracle.step import Oracle

torch = pytest.importorskip("torch")

SIDES = ("left", "right", "bottom", "top")
COMBOS = [("slip_wall", "zero_gradient", "slip_wall", "slip_wall"),
          ("periodic", "periodic", "no_slip_wall", "zero_gradient"),
          ("zero_gradient", "no_slip_wall", "periodic", "periodic"),
          ("no_slip_wall", "slip_wall", "zero_gradient", "no_slip_wall")]


@pytest.mark.parametrize("kinds", COMBOS)
def test_block_ghosts_are_the_domain_padding(mech_path, kinds):
    from janc_amd.amr import UniformBlockGrid
    mech = Mechanism(mech_path)
    bc = {s + "_boundary": k for s, k in zip(SIDES, kinds)}
    nbx, nby, bnx, bny = 3, 4, 5, 4
    U, aux, g = cases.detonation_tube(mech, nx=nbx * bnx, ny=nby * bny, driver=3)
    rng = np.random.default_rng(1)
    U[1] += U[0] * 10.0 * rng.standard_normal(U[0].shape)
    U[2] += U[0] * 7.0 * rng.standard_normal(U[0].shape)
    o = Oracle(mech_path, bc, 1e-4, 1e-4)
    Ug, ag = o.boundary(U, aux)
    pad = np.concatenate([Ug, ag])
    grid = UniformBlockGrid(bc, nbx, nby, bnx, bny, device="cpu")
    blk = grid.split(torch.tensor(U), torch.tensor(aux))
    gb = grid.get_ghost_block_data(blk).numpy()
    for i in range(nbx):
        for j in range(nby):
            w = pad[:, i * bnx:i * bnx + bnx + 6, j * bny:j * bny + bny + 6]
            assert np.array_equal(gb[i * nby + j], w), (i, j)
    Ua, aa = grid.assemble(blk)
    assert np.array_equal(Ua.numpy(), U) and np.array_equal(aa.numpy(), aux)


def test_computed_kinds_are_refused():
    from janc_amd.amr import UniformBlockGrid
    bc = {s + "_boundary": "slip_wall" for s in SIDES}
    bc["top_boundary"] = "pressure_outlet"
    with pytest.raises(NotImplementedError):
        UniformBlockGrid(bc, 2, 2, 4, 4, device="cpu")
