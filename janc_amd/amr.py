"""The AMR branch of the reference's step: one refinement level's same-size
blocks advanced together on the GPU.

Mirrors ``AMR_Simulator`` (src/janc_v2/simulation/set_simulator.py:496-532):
``set_rhs(..., is_amr=True)`` vmaps ``flux_func`` and ``update_func`` over a
leading block axis (:37-42) -- blocks arrive with their 3 ghost layers
already filled, no boundary conditions are applied, every block has its own
Lax-Friedrichs maxima --, ``advance_one_step_amr`` (time_step/TVD_RK3.py:
15-37) runs the RK stages with the ghost layers refilled between them, and
the detailed-chemistry branch of ``advance_one_step`` (:107-111) reacts the
blocks' interiors.  The block arithmetic runs in libjanc_amd.so
(``jh_blocks_*``); the AMR hierarchy itself -- ``get_ghost_block_data`` and
``update_external_boundary`` of jaxamr (``amr`` in the reference), a
third-party library -- is the caller's, passed in as an object with those two
functions.

Arrays are torch tensors on the GPU in the reference's ``blk_data`` layout:
(n_blk, 14, bnx, bny) = U (12) and aux (gamma, T); with ghost layers
(n_blk, 14, bnx + 6, bny + 6).
"""
import ctypes

import torch

from . import _lib
from .solver import Simulator, _ptr


class _Blocks:
    """Owns one jh_blocks (block shape, level spacing, precision)."""

    def __init__(self, bnx, bny, dx, dy, precision, mech):
        lib = _lib.load()
        consts = mech.const_array()
        carr = consts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        stream = torch.cuda.current_stream().cuda_stream
        rc = lib.jh_blocks_create(bnx, bny, dx, dy, precision, carr, consts.size,
                                  ctypes.c_void_p(stream), ctypes.byref(h))
        if rc != _lib.JH_OK:
            raise _lib.JancError("jh_blocks_create: " + lib.jh_blocks_last_error(None).decode())
        self.h, self.lib = h, lib

    def call(self, name, *args):
        self.lib.jh_blocks_set_stream(self.h,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != _lib.JH_OK:
            raise _lib.JancError("%s: %s" % (name, self.lib.jh_blocks_last_error(self.h).decode()))

    def close(self):
        if self.h:
            self.lib.jh_blocks_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BlockSolver:
    """The vmapped block functions of one configuration (the reference's
    simulation_config dict, read as ``Simulator`` reads it; LF flux splitting
    with WENO5_JS, detailed chemistry).  ``dx``/``dy`` are per call, as in
    ``advance_func_amr`` (each refinement level has its own spacing)."""

    def __init__(self, simulation_config):
        sim = Simulator(simulation_config)
        if sim.flux != (0, 0, 0, 0):
            raise NotImplementedError("AMR blocks run LF flux splitting with WENO5_JS")
        if sim.halley_stop == "grid":
            raise NotImplementedError(
                "AMR blocks stop the temperature solve per cell; set "
                "computation_config['halley_stop'] = 'per_cell'")
        self.sim = sim
        self.mech, self.precision, self.dtype = sim.mech, sim.precision, sim.dtype
        self._ctx = {}

    def _blocks(self, bnx, bny, dx, dy):
        key = (bnx, bny, float(dx), float(dy))
        b = self._ctx.get(key)
        if b is None:
            b = _Blocks(bnx, bny, float(dx), float(dy), self.precision, self.mech)
            self._ctx[key] = b
        return b

    def _check(self, blk, padded):
        if not isinstance(blk, torch.Tensor) and hasattr(blk, "__dlpack__"):
            blk = torch.from_dlpack(blk)
        if blk.device.type != "cuda" or blk.dtype != self.dtype or not blk.is_contiguous():
            raise TypeError("blocks must be contiguous %s CUDA tensors" % self.dtype)
        if blk.dim() != 4 or blk.shape[1] != 4 + self.mech.ns - 1 + 2:
            raise ValueError("blocks must be (n_blk, %d, nx, ny)" % (4 + self.mech.ns + 1))
        g = 6 if padded else 0
        return blk, int(blk.shape[0]), int(blk.shape[2]) - g, int(blk.shape[3]) - g

    def flux_func(self, blk, dx, dy, dt):
        """set_rhs(is_amr=True).flux_func (set_simulator.py:37-40) on padded
        blocks (U and aux concatenated): (n_blk, 12, bnx+6, bny+6)."""
        blk, nb, bnx, bny = self._check(blk, True)
        out = torch.empty((nb, blk.shape[1] - 2) + tuple(blk.shape[2:]), dtype=blk.dtype,
                          device=blk.device)
        self._blocks(bnx, bny, dx, dy).call("jh_blocks_flux", nb, _ptr(blk), float(dt),
                                            _ptr(out))
        return out

    def stage(self, s, blk_s, blk_n, dx, dy, dt):
        """One stage of advance_one_step_amr (TVD_RK3.py:21-35) on padded
        blocks; returns the padded stage output."""
        blk_s, nb, bnx, bny = self._check(blk_s, True)
        blk_n, nb2, _, _ = self._check(blk_n, True)
        if blk_n.shape != blk_s.shape:
            raise ValueError("blk_s and blk_n must have the same shape")
        out = torch.empty_like(blk_s)
        self._blocks(bnx, bny, dx, dy).call("jh_blocks_stage", nb, int(s), _ptr(blk_s),
                                            _ptr(blk_n), float(dt), _ptr(out))
        return out

    def chemistry(self, blk, dx, dy, dt):
        """The detailed-chemistry AMR branch (set_simulator.py:107-111) on
        interior blocks, in place."""
        blk, nb, bnx, bny = self._check(blk, False)
        self._blocks(bnx, bny, dx, dy).call("jh_blocks_chem", nb, _ptr(blk), float(dt))
        return blk

    def advance_one_step(self, level, blk_data, dx, dy, dt, ref_blk_data, ref_blk_info,
                         theta=None, amr=None):
        """advance_func_amr (set_simulator.py:100-111 with TVD_RK3.py:15-37):
        blk_data = the parent level's blocks, ref_blk_data / ref_blk_info = this
        level's blocks and their block information; ``amr`` provides
        get_ghost_block_data(blocks, info) and
        update_external_boundary(level, parent_blocks, interiors, info)."""
        if amr is None:
            raise ValueError("advance_one_step needs the AMR hierarchy's ghost functions "
                             "(jaxamr's amr module or an object with the same two functions)")
        g0 = amr.get_ghost_block_data(ref_blk_data, ref_blk_info)
        b = None
        gs = g0
        for s in range(3):
            out = self.stage(s, gs, g0, dx, dy, dt)
            b = amr.update_external_boundary(level, blk_data, out[..., 3:-3, 3:-3], ref_blk_info)
            if s < 2:
                gs = amr.get_ghost_block_data(b, ref_blk_info)
        b = b.contiguous()
        return self.chemistry(b, dx, dy, dt)

    def health(self):
        """OR of every block context's health flags since the last query."""
        f = 0
        for b in self._ctx.values():
            v = ctypes.c_uint()
            b.call("jh_blocks_health", ctypes.byref(v))
            f |= int(v.value)
        return f


def AMR_Simulator(simulation_config):
    """set_simulator.py:496-532: (advance_func_amr, advance_func_base).  The
    first takes the reference's arguments plus ``amr`` (the hierarchy's ghost
    functions); the second advances the base level's single block
    (blk_data[0] = (14, nx, ny)) with the boundary conditions of the
    configuration, as the reference's advance_func_base."""
    solver = BlockSolver(simulation_config)
    sim = solver.sim

    def advance_func_base(blk_data, dx, dy, dt, theta=None):
        n = 4 + sim.mech.ns - 1
        U, aux = blk_data[0, :n], blk_data[0, n:]
        U, aux = sim.advance_one_step(U.contiguous(), aux.contiguous(), dx, dy, dt, theta)
        return torch.cat([U, aux], 0)[None]

    return solver.advance_one_step, advance_func_base


class UniformBlockGrid:
    """One refinement level tiled by nbx x nby same-size blocks of bnx x bny
    cells (block k = (i, j), k = i * nby + j), with the configuration's
    boundary kinds at the domain's edges: the ghost refill advance_func_amr
    takes from the AMR hierarchy (get_ghost_block_data,
    update_external_boundary), for a uniform single-level grid.

    A block's ghost layers are its neighbours' cells; at the domain's edges
    they are the reference's boundary padding (boundary.py:126-132): the
    off-by-one periodic pre-pad (boundary_padding.py:20-25; left ghosts =
    cells N-4..N-2, right = 1..3), then each non-periodic side's strip over
    the other axis' interior range -- zero_gradient mirrors (zero_gradient.py),
    slip_wall copies the wall cell with the normal momentum negated
    (slip_wall.py), no_slip_wall mirrors with both momenta negated
    (no_slip_wall.py); corners keep the periodic pre-pad.  Pressure outlets,
    injection inlets and callables compute their ghosts and are not offered.
    One gather per refill: on the GPU the library's blocks_fill_kernel
    (jh_blocks_fill_ghosts, reading the interiors through their strides, so
    update_external_boundary hands on the stage output's interior view
    without a copy); elsewhere torch's index_select."""

    KINDS = ("periodic", "zero_gradient", "slip_wall", "no_slip_wall")

    def __init__(self, boundary_config, nbx, nby, bnx, bny, device="cuda"):
        kinds = [boundary_config[s + "_boundary"] for s in ("left", "right", "bottom", "top")]
        for k in kinds:
            if callable(k) or k not in self.KINDS:
                raise NotImplementedError("UniformBlockGrid fills copied ghost kinds only (%s), "
                                          "not %r" % (", ".join(self.KINDS), k))
        self.nbx, self.nby, self.bnx, self.bny = nbx, nby, bnx, bny
        NX, NY = nbx * bnx, nby * bny
        gx = torch.arange(-3, NX + 3)
        gy = torch.arange(-3, NY + 3)

        def per(g, N):                     # boundary_padding.py:20-25
            return torch.where(g < 0, N - 1 + g, torch.where(g >= N, g - N + 1, g))

        def side(g, N, lo, hi):
            """source index and (normal, tangential) momentum signs along one axis."""
            src = per(g, N)
            sn = torch.ones_like(g, dtype=torch.float64)
            st = torch.ones_like(g, dtype=torch.float64)
            for kind, mask, mirror, wall in ((lo, g < 0, -g - 1, torch.zeros_like(g)),
                                             (hi, g >= N, 2 * N - 1 - g,
                                              torch.full_like(g, N - 1))):
                if kind == "zero_gradient":
                    src = torch.where(mask, mirror, src)
                elif kind == "slip_wall":
                    src = torch.where(mask, wall, src)
                    sn = torch.where(mask, -1.0, sn)
                elif kind == "no_slip_wall":
                    src = torch.where(mask, mirror, src)
                    sn = torch.where(mask, -1.0, sn)
                    st = torch.where(mask, -1.0, st)
            return src, sn, st

        sx, snx, stx = side(gx, NX, kinds[0], kinds[1])
        sy, sny, sty = side(gy, NY, kinds[2], kinds[3])
        inx = (gx >= 0) & (gx < NX)
        iny = (gy >= 0) & (gy < NY)
        # corners (both axes outside) keep the periodic pre-pad; a strip uses
        # its own axis' kind over the other axis' interior
        px = per(gx, NX)
        py = per(gy, NY)
        X = torch.where(iny[None, :], sx[:, None], px[:, None])
        Y = torch.where(inx[:, None], sy[None, :], py[None, :])
        one = torch.ones((), dtype=torch.float64)
        s_u = torch.where(iny[None, :], snx[:, None], one) * \
            torch.where(inx[:, None], sty[None, :], one)
        s_v = torch.where(inx[:, None], sny[None, :], one) * \
            torch.where(iny[None, :], stx[:, None], one)
        # column of domain cell (x, y) in the block-major interior flattening
        col = ((X // bnx) * nby + Y // bny) * (bnx * bny) + (X % bnx) * bny + Y % bny
        PX, PY = bnx + 6, bny + 6
        bi = torch.arange(nbx)[:, None, None, None] * bnx + torch.arange(PX)[None, None, :, None]
        bj = torch.arange(nby)[None, :, None, None] * bny + torch.arange(PY)[None, None, None, :]
        idx = col[bi, bj].reshape(-1)
        # every entry indexes the level's interiors (the gather kernel reads
        # through them unchecked)
        assert int(idx.min()) >= 0 and int(idx.max()) < nbx * nby * bnx * bny
        self.idx = idx.to(device)                                  # (n_blk * PX * PY,)
        self._idx32 = idx.to(torch.int32).to(device)
        self._signs = {}
        self.s_u = s_u[bi, bj].reshape(nbx * nby, PX, PY).to(device)
        self.s_v = s_v[bi, bj].reshape(nbx * nby, PX, PY).to(device)
        self.n_blk = nbx * nby

    def split(self, U, aux):
        """Domain state (U (12, NX, NY), aux (2, NX, NY)) -> interior blocks."""
        full = torch.cat([U, aux], 0)
        nv = full.shape[0]
        b = full.reshape(nv, self.nbx, self.bnx, self.nby, self.bny)
        return b.permute(1, 3, 0, 2, 4).reshape(self.n_blk, nv, self.bnx, self.bny).contiguous()

    def assemble(self, blk):
        """Interior blocks -> the domain's (U, aux)."""
        nv = blk.shape[1]
        b = blk.reshape(self.nbx, self.nby, nv, self.bnx, self.bny)
        full = b.permute(2, 0, 3, 1, 4).reshape(nv, self.nbx * self.bnx, self.nby * self.bny)
        return full[:-2], full[-2:]

    def _fill_gpu(self, blk):
        from .solver import _ptr
        if blk.dtype not in (torch.float64, torch.float32) or blk.dim() != 4 \
                or tuple(blk.shape) != (self.n_blk, 14, self.bnx, self.bny) \
                or blk.stride(3) != 1 or blk.device != self._idx32.device:
            raise ValueError("blocks must be (%d, 14, %d, %d) float CUDA tensors with "
                             "contiguous columns on the grid's device"
                             % (self.n_blk, self.bnx, self.bny))
        if blk.dtype not in self._signs:
            self._signs[blk.dtype] = (self.s_u.to(blk.dtype).contiguous(),
                                      self.s_v.to(blk.dtype).contiguous())
        su, sv = self._signs[blk.dtype]
        out = torch.empty((self.n_blk, 14, self.bnx + 6, self.bny + 6), dtype=blk.dtype,
                          device=blk.device)
        lib = _lib.load()
        prec = _lib.PREC_FP64 if blk.dtype == torch.float64 else _lib.PREC_MIXED
        rc = lib.jh_blocks_fill_ghosts(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                                       prec, self.n_blk, self.bnx, self.bny, _ptr(blk),
                                       blk.stride(0), blk.stride(1), blk.stride(2),
                                       _ptr(self._idx32), _ptr(su), _ptr(sv), _ptr(out))
        if rc != _lib.JH_OK:
            raise _lib.JancError("jh_blocks_fill_ghosts: "
                                 + lib.jh_blocks_last_error(None).decode())
        return out

    def get_ghost_block_data(self, blk, info=None):
        """Interior blocks (n_blk, 14, bnx, bny) -> padded blocks (n_blk, 14,
        bnx + 6, bny + 6)."""
        if blk.is_cuda:
            return self._fill_gpu(blk)
        nb, nv = blk.shape[0], blk.shape[1]
        flat = blk.reshape(nb, nv, -1).permute(1, 0, 2).reshape(nv, -1)
        g = flat.index_select(1, self.idx).reshape(nv, nb, self.bnx + 6, self.bny + 6)
        g[1] *= self.s_u.to(g.dtype)
        g[2] *= self.s_v.to(g.dtype)
        return g.permute(1, 0, 2, 3).contiguous()

    def update_external_boundary(self, level, parent, interiors, info=None):
        """A single level: the interiors are the blocks (on the GPU the view
        itself; get_ghost_block_data reads it through its strides)."""
        return interiors if interiors.is_cuda else interiors.contiguous()
