"""Oracle restatement of the AMR-block branch of the 2-D step (test
infrastructure only).

Follows ``src/janc_v2/simulation/set_simulator.py``: ``set_rhs`` with
``is_amr`` (:37-42) -- ``flux_func`` and ``update_func`` are ``vmap``-ed over
a leading block axis, ``flux_func`` takes blocks that already carry their 3
ghost layers (no boundary conditions are applied) and returns
``pad(dt * total_flux(U, aux, dx, dy), 3)``, so every block has its own
Lax-Friedrichs maxima (the ``nanmax`` of LF.py:10-11 over that block's padded
field) and, under ``vmap``, its own temperature-solve stop (the ``while_loop``
runs each block until that block's max|res| is within the tolerance);
``advance_one_step``'s detailed-chemistry AMR branch (:107-111) --
``vmap(reaction_source_terms)`` on the blocks' interiors, then
``update_func``; and ``time_step/TVD_RK3.py:15-37`` (``advance_one_step_amr``).

``advance_one_step_amr`` refills the ghost layers between RK stages with
``jaxamr.amr.get_ghost_block_data`` and keeps the interiors with
``amr.update_external_boundary``; jaxamr (``git+https://github.com/JA4S/
JAX-AMR.git``, no revision pinned, ``setup.py:11``) is absent here, so those two
are arguments: ``ghost_fn(interior_blocks) -> padded blocks`` and the
interior slice.  Blocks are ``(n_blk, 14, bnx [+6], bny [+6])`` = U (12) and
aux (gamma, T) concatenated, as ``blk_data``.
"""
import numpy as np

from . import flux as _flux


def flux_func(o, U, aux, dt):
    """set_simulator.py:37-40: U (n_blk, 12, PX, PY), aux (n_blk, 2, PX, PY)
    with caller-filled ghosts -> (n_blk, 12, PX, PY), zero on the ghosts."""
    out = np.zeros_like(U)
    for b in range(U.shape[0]):
        out[b, :, 3:-3, 3:-3] = dt * _flux.scheme_total_flux(
            U[b], aux[b], o.dx, o.dy, o.thermo, o.solver_type, o.reconstruction, o.riemann,
            o.split)
    return out


def update_func(o, U, aux):
    """set_simulator.py:41: vmap(update_aux) -- each block its own stop."""
    out = np.empty_like(aux)
    for b in range(U.shape[0]):
        out[b] = _flux.update_aux(U[b], aux[b], o.thermo, per_cell=o.per_cell)
    return out


def stage(o, s, blk_s, blk_n, dt):
    """One stage of advance_one_step_amr (TVD_RK3.py:21-35) on padded
    blocks: blk_s the stage's input, blk_n the step's input (both with their
    ghost layers); returns the padded [U_s+1, aux_s+1]."""
    U, aux = blk_s[:, :-2], blk_s[:, -2:]
    Un = blk_n[:, :-2]
    r = flux_func(o, U, aux, dt)
    if s == 0:
        Uo = U + r
    elif s == 1:
        Uo = 3 / 4 * Un + 1 / 4 * (U + r)
    else:
        Uo = 1 / 3 * Un + 2 / 3 * (U + r)
    ao = update_func(o, Uo, aux)
    return np.concatenate([Uo, ao], axis=1)


def chemistry(o, blk, dt):
    """set_simulator.py:107-111 on interior blocks (n_blk, 14, bnx, bny)."""
    U, aux = blk[:, :-2], blk[:, -2:]
    dU = np.stack([o.reaction_source_terms(U[b], aux[b], dt) for b in range(U.shape[0])])
    U = U + dU
    aux = update_func(o, U, aux)
    return np.concatenate([U, aux], axis=1)


def advance_one_step(o, blk, dt, ghost_fn):
    """advance_one_step_amr (TVD_RK3.py:15-37) then the chemistry branch
    (set_simulator.py:107-111).  blk: interior blocks; ghost_fn(interior
    blocks) -> the same blocks with their ghost layers (jaxamr's
    get_ghost_block_data for the level); update_external_boundary is the
    interior slice."""
    g0 = ghost_fn(blk)
    b1 = stage(o, 0, g0, g0, dt)[..., 3:-3, 3:-3]
    b2 = stage(o, 1, ghost_fn(b1), g0, dt)[..., 3:-3, 3:-3]
    b3 = stage(o, 2, ghost_fn(b2), g0, dt)[..., 3:-3, 3:-3]
    return chemistry(o, b3, dt)
