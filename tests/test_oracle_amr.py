"""The oracle's restatement of the AMR-block branch (oracle/amr.py) pinned
against the single-block oracle: a block whose ghost layers are the
single-block boundary padding (boundary/boundary_padding.py) gives the
single-block flux_func exactly, and the whole AMR step with that padding as
the ghost refill gives the single-block step; a block's result does not
depend on the other blocks of the batch (vmap: per-block Lax-Friedrichs
maxima), while a global maximum would change it."""
import numpy as np
import pytest

from janc_amd import cases
from janc_amd.mechanism import Mechanism
from oracle import amr
from oracle import flux as oflux
from oracle.step import Oracle


@pytest.fixture(scope="module")
def tube(mech_path):
    mech = Mechanism(mech_path)
    U, aux, g = cases.detonation_tube(mech, nx=30, ny=14, driver=6, seed=9)
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    return o, U, aux


def _padded(o, U, aux):
    Ug, ag = o.boundary(U, aux)
    return np.concatenate([Ug, ag])[None]


def test_one_block_flux_is_the_single_block_flux(tube):
    o, U, aux = tube
    blk = _padded(o, U, aux)
    r = amr.flux_func(o, blk[:, :-2], blk[:, -2:], 2e-9)
    assert r.shape == (1, 12, 36, 20)
    assert np.array_equal(r[0, :, 3:-3, 3:-3], o.flux_func(U, aux, 2e-9))
    assert not r[0, :, :3].any() and not r[0, :, :, -3:].any()


def test_one_block_step_is_the_single_block_step(tube):
    o, U, aux = tube

    def ghost_fn(b):
        return np.concatenate([_padded(o, b[k, :-2], b[k, -2:]) for k in range(b.shape[0])])

    dt = 2e-9
    out = amr.advance_one_step(o, np.concatenate([U, aux])[None], dt, ghost_fn)
    Ur, ar = o.advance_one_step(U, aux, dt)
    # SI units: the padded solves' stop (ghost cells in max|res|) moves T by
    # rounding only
    for got, ref in ((out[0, :-2], Ur), (out[0, -2:], ar)):
        d = np.abs(got - ref).reshape(ref.shape[0], -1).max(1)
        s = np.abs(ref).reshape(ref.shape[0], -1).max(1)
        assert np.max(d / s) < 1e-13


def test_blocks_have_their_own_lf_maxima(tube):
    """Block 2 carries a faster cell (a larger LF maximum); block 1's result
    in the pair equals block 1 alone, and differs from block 1 under block
    2's maximum (what one global maximum over the batch would give)."""
    o, U, aux = tube
    b1 = _padded(o, U, aux)
    U2 = U.copy()
    U2[1, 3, 7] += U2[0, 3, 7] * 600.0               # +600 m/s in one driver cell
    b2 = _padded(o, U2, aux)
    a1 = o.lf_alpha(U, aux)
    a2 = o.lf_alpha(U2, aux)
    assert a2[0] > a1[0] * 1.2
    pair = np.concatenate([b1, b2])
    r = amr.flux_func(o, pair[:, :-2], pair[:, -2:], 2e-9)
    alone = amr.flux_func(o, b1[:, :-2], b1[:, -2:], 2e-9)
    assert np.array_equal(r[0], alone[0])
    # block 1 under block 2's x maximum
    Ub, ab = b1[0, :-2], b1[0, -2:]
    Fp, Fm = oflux.split_flux_lf(1, Ub, ab, o.thermo)
    F = Fp + Fm
    Fx = oflux.weno5_L_x(0.5 * (F + a2[0] * Ub)) + oflux.weno5_R_x(0.5 * (F - a2[0] * Ub))
    Gp, Gm = oflux.split_flux_lf(2, Ub, ab, o.thermo)
    Gy = oflux.weno5_L_y(Gp) + oflux.weno5_R_y(Gm)
    glob = -((Fx[:, 1:] - Fx[:, :-1]) / o.dx + (Gy[:, :, 1:] - Gy[:, :, :-1]) / o.dy) * 2e-9
    assert np.abs(glob - alone[0, :, 3:-3, 3:-3]).max() > 1e-6 * np.abs(glob).max()
