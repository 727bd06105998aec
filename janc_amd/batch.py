"""Many independent cases of uneven grid sizes advanced together.

``BatchSimulator`` takes one reference-style simulation_config per case (the
same dict ``Simulator`` takes), binds one device state per case and advances
all of them with the batched kernels of ``jh_batch_*``: one launch per kernel
kind for every case, the stage kernel persistent over the tile list of all
cases, and optionally the whole multi-step launch sequence captured in a HIP
graph and replayed.  Every case evolves bit-identically to advancing it alone
with ``Simulator.advance``.

Against the reference this replaces a Python loop of ``advance_func`` calls
(simulation/set_simulator.py:444-447), one per case; its own batching is
jax.vmap over same-size blocks (set_simulator.py:37-42, 66-70).
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from .solver import Simulator, _ptr


class BatchSimulator:
    """One simulation_config and one initial (U, aux) per case; grid shapes
    are taken from the states (the reference's grid_config holds dx, dy
    only).  The states are bound at once and advanced in place."""

    def __init__(self, simulation_configs, Us, auxs):
        if not simulation_configs:
            raise ValueError("need at least one case")
        if len(Us) != len(simulation_configs) or len(auxs) != len(simulation_configs):
            raise ValueError("need one state per case")
        sims = [Simulator(cfg) for cfg in simulation_configs]
        s0 = sims[0]
        for s in sims[1:]:
            if s.precision != s0.precision:
                raise ValueError("all cases must share the precision form")
            if s.mech.structure_digest() != s0.mech.structure_digest() or \
                    not np.array_equal(s.mech.const_array(), s0.mech.const_array()):
                raise ValueError("all cases must share the mechanism")
            if (s.fixed_dt, s.cfl) != (s0.fixed_dt, s0.cfl):
                raise ValueError("all cases must share the time control (dt or CFL)")
        if any(s.flux != (0, 0, 0, 0) for s in sims):
            raise NotImplementedError("batches run LF flux splitting with WENO5_JS")
        self.sims = sims
        self.ncase = len(sims)
        self.precision, self.dtype, self.device = s0.precision, s0.dtype, s0.device
        self.fixed_dt, self.cfl = s0.fixed_dt, s0.cfl
        nc = self.ncase
        self.shapes = []
        self._states = None
        lib = _lib.load()
        self.lib = lib
        nx = (ctypes.c_int * nc)()
        ny = (ctypes.c_int * nc)()
        dx = (ctypes.c_double * nc)()
        dy = (ctypes.c_double * nc)()
        kinds = (ctypes.c_int * (4 * nc))()
        params = (ctypes.c_double * (4 * nc))()
        for k, s in enumerate(sims):
            shape = tuple(int(v) for v in Us[k].shape[1:])
            if len(shape) != 2:
                raise ValueError("U must be (%d, nx, ny)" % (4 + s.mech.ns - 1))
            nx[k], ny[k] = shape
            self.shapes.append(shape)
            dx[k], dy[k] = s.dx, s.dy
            for j, (kind, par) in enumerate(zip(s.bc_kinds, s._bc_params(None))):
                kinds[4 * k + j] = kind
                params[4 * k + j] = par
        consts = s0.mech.const_array()
        carr = consts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        rc = lib.jh_batch_create(nc, nx, ny, dx, dy, kinds, params, s0.precision, carr,
                                 consts.size,
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                                 ctypes.byref(h))
        if rc != _lib.JH_OK:
            raise _lib.JancError("jh_batch_create: " + lib.jh_batch_last_error(None).decode())
        self.h = h
        for k, s in enumerate(sims):
            if _lib.BC_KINDS["injection_inlet"] in s.bc_kinds:
                T0, Y = s._inlet(None)
                arr = (ctypes.c_double * len(Y))(*Y)
                self._check(lib.jh_batch_set_inlet(h, k, T0, arr, len(Y)))
        self.bind(Us, auxs)

    # ------------------------------------------------------------ plumbing
    def _check(self, rc):
        if rc != _lib.JH_OK:
            raise _lib.JancError("janc_amd batch error %d: %s"
                                 % (rc, self.lib.jh_batch_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.jh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ API
    def bind(self, Us, auxs):
        """Bind one (U, aux) device state per case; they are advanced in place."""
        if len(Us) != self.ncase or len(auxs) != self.ncase:
            raise ValueError("need one state per case")
        Us2, auxs2 = [], []
        for k, (U, aux) in enumerate(zip(Us, auxs)):
            U, aux = self.sims[k]._state(U, aux)
            if tuple(U.shape[1:]) != self.shapes[k]:
                raise ValueError("case %d: state shape %s, grid %s"
                                 % (k, tuple(U.shape[1:]), self.shapes[k]))
            Us2.append(U)
            auxs2.append(aux)
        pu = (ctypes.c_void_p * self.ncase)(*[U.data_ptr() for U in Us2])
        pa = (ctypes.c_void_p * self.ncase)(*[a.data_ptr() for a in auxs2])
        self._check(self.lib.jh_batch_bind(self.h, pu, pa))
        self._states = (Us2, auxs2)      # keep the storage alive
        return Us2, auxs2

    @property
    def states(self):
        return self._states

    def advance(self, nsteps, graph=True, dt_hist=None):
        """nsteps steps of every bound case; returns the (ncase, nsteps) dt
        history when ``dt_hist`` is True or a device tensor to fill."""
        if self._states is None:
            raise RuntimeError("bind() the states first")
        if dt_hist is True:
            dt_hist = torch.empty((self.ncase, nsteps), dtype=torch.float64, device=self.device)
        self._check(self.lib.jh_batch_set_stream(
            self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        dt = self.fixed_dt if self.fixed_dt is not None else 0.0
        cfl = self.cfl if self.fixed_dt is None else 0.0
        self._check(self.lib.jh_batch_advance(self.h, float(dt), float(cfl), int(nsteps),
                                              _ptr(dt_hist) if dt_hist is not None else None,
                                              int(bool(graph))))
        return dt_hist

    def set_tile(self, variant):
        self._check(self.lib.jh_batch_set_tile(self.h, int(variant)))

    def get_tile(self):
        return int(self.lib.jh_batch_get_tile(self.h))

    def health(self):
        f = ctypes.c_uint()
        self._check(self.lib.jh_batch_health(self.h, ctypes.byref(f)))
        return int(f.value)

    def cells(self):
        return sum(nx * ny for nx, ny in self.shapes)

    def autotune(self, steps=5, graph=True):
        """Time every stage-kernel variant on copies of the bound states (one
        warm-up advance, then one timed advance of ``steps``) in two
        interleaved passes and keep the fastest, each variant at its best
        pass.  Returns {"variant": best, "ms_per_step": {variant: ms}}."""
        from .autotune import variants
        if self._states is None:
            raise RuntimeError("bind() the states first")
        Us, auxs = self._states
        times = {}
        for r in range(2):
            for v in variants():
                self.set_tile(v["variant"])
                self.bind([U.clone() for U in Us], [a.clone() for a in auxs])
                self.advance(steps, graph=graph)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self.advance(steps, graph=graph)
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) / steps * 1e3
                times[v["variant"]] = min(ms, times.get(v["variant"], ms))
        best = min(times, key=times.get)
        self.set_tile(best)
        self.bind(Us, auxs)
        return {"variant": best, "ms_per_step": times}
