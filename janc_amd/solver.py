"""Host-side mirror of the reference's 2-D Simulator for the hot path.

``Simulator(simulation_config)`` accepts the reference's configuration dict
(``src/janc_v2/simulation/set_simulator.py:321-455``) and exposes the same
callables -- ``advance_func(U, aux, t, theta)``, ``advance_one_step(U, aux,
dx, dy, dt, theta)``, ``run(U_init, aux_init)`` -- plus the unit pieces
``flux_func`` (set_simulator.py:44-47), ``update_func`` (aux_func.py:21-29),
``reaction_source_terms`` (reaction_model.py:202-203) and ``CFL``
(set_simulator.py:438-443).  Arrays are torch tensors on the GPU laid out like
the reference's JAX arrays: U (12, nx, ny), aux (2, nx, ny); anything that
speaks DLPack (a JAX GPU array, a torch tensor) is accepted, numpy arrays are
copied to the device.  Every computation runs in libjanc_amd.so; there is no
CPU path.

Configuration errors follow the reference: unsupported keys raise
``KeyError``/``AssertionError``/``FileNotFoundError`` at construction.
Paths the library does not implement (AUSMPW+, viscous fluxes, 1-D/3-D,
constant-gamma thermo, user-defined reaction sources, multi-device
decomposition) raise ``NotImplementedError``; the AMR branch's block step is
``janc_amd.amr``.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .mechanism import Mechanism

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the runtime image
    torch = None


def _as_device_tensor(x, dtype, device):
    if isinstance(x, np.ndarray):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device)
    if not isinstance(x, torch.Tensor) and hasattr(x, "__dlpack__"):
        x = torch.from_dlpack(x)
    if x.device.type != "cuda":
        x = x.to(device)
    if x.dtype != dtype:
        raise TypeError("expected %s state, got %s" % (dtype, x.dtype))
    return x.contiguous()


class _Context:
    """Owns one jh_ctx (one grid shape, precision and boundary set)."""

    def __init__(self, nx, ny, dx, dy, bc_kinds, bc_params, precision, mech):
        lib = _lib.load()
        consts = mech.const_array()
        if consts.size != lib.jh_mech_const_count():
            raise _lib.JancError("mechanism constant block has %d values, library expects %d"
                                 % (consts.size, lib.jh_mech_const_count()))
        digest = lib.jh_mech_digest().decode()
        if digest != mech.structure_digest():
            raise _lib.JancError(
                "libjanc_amd.so was built for mechanism %s, configuration needs %s; "
                "rebuild with that mechanism" % (digest, mech.structure_digest()))
        kinds = (ctypes.c_int * 4)(*bc_kinds)
        params = (ctypes.c_double * 4)(*bc_params)
        carr = consts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        stream = torch.cuda.current_stream().cuda_stream
        rc = lib.jh_ctx_create(nx, ny, dx, dy, kinds, params, precision, carr, consts.size,
                               ctypes.c_void_p(stream), ctypes.byref(h))
        if rc != _lib.JH_OK:
            raise _lib.JancError("jh_ctx_create: " + lib.jh_last_error(None).decode())
        self.h = h
        self.lib = lib
        self.nx, self.ny = nx, ny

    def call(self, name, *args):
        self.lib.jh_ctx_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        rc = getattr(self.lib, name)(self.h, *args)
        _lib.check(rc, self.h)

    def close(self):
        if self.h:
            self.lib.jh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Simulator:
    """``Simulator`` of simulation/set_simulator.py:320-494, 2-D detailed chemistry."""

    def __init__(self, simulation_config):
        cfg = simulation_config
        dim = cfg["dimension"]
        if dim != "2D":
            raise NotImplementedError("janc_amd implements the 2-D solver only")
        grid = cfg["grid_config"]
        thermo_config = cfg["thermo_config"]
        reaction_config = cfg["reaction_config"]
        flux_config = cfg["flux_config"]
        boundary_config = cfg["boundary_config"]
        time_control = cfg["time_config"]
        self.t_end = time_control["t_end"]
        self.theta = cfg.get("solver_parameters")
        nondim_config = cfg.get("nondim_config")
        comp = cfg.get("computation_config", {}) or {}
        if comp.get("is_parallel", False):
            raise NotImplementedError("multi-device x-slab decomposition is not built")
        # set_simulator.py:370-382: snapshot interval and file of run()
        out = comp.get("output_settings")
        if out is not None:
            self.save_dt = out["save_dt"]
            self.results_path = out["results_path"]
        else:
            self.save_dt = self.t_end
            self.results_path = "results.h5"
        # stop rule of the Halley temperature solve (get_T_nasa7,
        # thermo_model.py:135-153): 'grid' reproduces the reference's
        # grid-wide max|res| test, 'per_cell' stops each cell on its own
        # residual (equal to rounding in SI units); 'auto' (default) takes the
        # grid-wide rule for nondimensional configurations, where the absolute
        # tolerance is a larger fraction of the energy
        stop = comp.get("halley_stop", "auto")
        if stop not in ("auto", "grid", "per_cell"):
            raise KeyError("halley_stop must be 'auto', 'grid' or 'per_cell'")
        if stop == "auto":
            stop = "grid" if nondim_config is not None else "per_cell"
        self.halley_stop = stop
        precision = comp.get("precision", "fp64")
        if precision not in ("fp64", "mixed"):
            raise KeyError("precision must be 'fp64' or 'mixed'")
        self.precision = _lib.PREC_FP64 if precision == "fp64" else _lib.PREC_MIXED
        self.dtype = torch.float64 if precision == "fp64" else torch.float32

        # thermo_model.py:42-62
        if thermo_config["thermo_model"] != "nasa7":
            if thermo_config["thermo_model"] == "constant_gamma":
                raise NotImplementedError("constant_gamma thermo is not on the built path")
            raise KeyError("The thermo model you specified is not supported, only 'nasa7' "
                           "or 'constant_gamma' can be specified.")
        assert "mechanism_directory" in thermo_config, \
            "Please specify 'mechanism_directory' in your dict of settings"
        # reaction_model.py:41-46
        if not reaction_config["is_detailed_chemistry"]:
            raise NotImplementedError("only detailed chemistry is on the built path")
        assert "mechanism_directory" in reaction_config, \
            "You choosed detailed chemistry without specifying the directory of your " \
            "mechanism files"
        self.mech = Mechanism(reaction_config["mechanism_directory"], nondim_config)
        # flux.py:19-41 (set_flux_solver)
        st = flux_config["solver_type"]
        if st not in ("godunov", "flux_splitting"):
            raise KeyError("JANC only support 'godunov' and 'flux_splitting'")
        rec = flux_config["interface_reconstruction"]
        if rec not in _lib.RECONSTRUCTIONS:           # reconstruction/__init__.py:3-21
            raise KeyError(rec)
        if st == "godunov":
            rs = flux_config["riemann_solver"]
            if rs == "AUSMPW_plus":
                # AUSMPW_plus.py:63,67 averages the y faces' gamma with the x
                # faces' array; the shapes (nx, ny+1) and (nx+1, ny) do not
                # broadcast, so the reference cannot run it on a 2-D grid
                raise NotImplementedError(
                    "AUSMPW_plus: the reference's 2-D form mixes x- and y-face gamma "
                    "arrays of different shapes and cannot run")
            if rs not in _lib.RIEMANN_SOLVERS:        # riemann_solver/__init__.py:3-6
                raise KeyError(rs)
        else:
            rs = "HLLC"
        # split_method is read for flux splitting only (flux.py:33-35)
        sm = flux_config["split_method"] if st == "flux_splitting" else "LF"
        if sm not in _lib.SPLIT_METHODS:              # flux_splitting/__init__.py:3-6
            raise KeyError(sm)
        self.flux = (_lib.SOLVER_TYPES[st], _lib.SPLIT_METHODS[sm], _lib.RECONSTRUCTIONS[rec],
                     _lib.RIEMANN_SOLVERS[rs])
        if flux_config["viscosity"] == "on":
            raise NotImplementedError("viscous fluxes are not built")
        if time_control["temporal_evolution_scheme"] != "TVD_RK3":
            raise KeyError(time_control["temporal_evolution_scheme"])
        # boundary.py:51-118
        self.bc_kinds, self.user_bc = [], {}
        for side in _lib.SIDES:
            k = boundary_config[side + "_boundary"]
            if callable(k):
                self.user_bc[side] = k
                self.bc_kinds.append(_lib.BC_KINDS["user"])
            elif k in _lib.BC_KINDS and k != "user":
                self.bc_kinds.append(_lib.BC_KINDS[k])
            else:
                raise KeyError(k)
        self.dx, self.dy = float(grid["dx"]), float(grid["dy"])
        self.fixed_dt = float(time_control["dt"]) if "dt" in time_control else None
        self.cfl = float(time_control["CFL"]) if "CFL" in time_control else None
        if self.fixed_dt is None and self.cfl is None:
            raise KeyError("time_config needs 'dt' or 'CFL'")
        self._ctx = {}
        self.device = torch.device("cuda", torch.cuda.current_device())

    # ------------------------------------------------------------ helpers
    def _bc_params(self, theta):
        theta = theta if theta is not None else self.theta
        out = []
        for k in self.bc_kinds:
            if k == _lib.BC_KINDS["pressure_outlet"]:
                out.append(float(theta["Pb"]))
            elif k == _lib.BC_KINDS["injection_inlet"]:
                out.append(float(theta["P0_inj"]))
            else:
                out.append(0.0)
        return out

    def _inlet(self, theta):
        """(T0, Y_inj) of injection-inlet sides: theta['T0_inj'] (default
        300 K) and theta['Yinj'] (default: stoichiometric H2-air as the
        rotating-detonation example, janc_basic_example2.ipynb)."""
        theta = theta if theta is not None else self.theta
        theta = theta or {}
        T0 = float(theta.get("T0_inj", 300.0))
        Y = theta.get("Yinj")
        if Y is None:
            Y = self.mech.injection_Y()
        Y = np.asarray(Y, dtype=np.float64)
        if Y.ndim > 1:
            flat = Y.reshape(Y.shape[0], -1)
            if not np.all(flat == flat[:, :1]):
                raise NotImplementedError("injected composition must be uniform along the inlet")
            Y = flat[:, 0]
        if Y.shape != (self.mech.ns - 1,):
            raise ValueError("Yinj needs %d mass fractions" % (self.mech.ns - 1))
        return T0, Y

    def _context(self, U, theta, dx=None, dy=None):
        nx, ny = int(U.shape[1]), int(U.shape[2])
        dx = self.dx if dx is None else float(dx)
        dy = self.dy if dy is None else float(dy)
        key = (nx, ny, dx, dy)
        ctx = self._ctx.get(key)
        if ctx is None:
            ctx = _Context(nx, ny, dx, dy, self.bc_kinds, self._bc_params(theta),
                           self.precision, self.mech)
            _lib.check(ctx.lib.jh_ctx_set_flux(ctx.h, *self.flux), ctx.h)
            _lib.check(ctx.lib.jh_ctx_set_halley_stop(ctx.h, _lib.HALLEY_STOP[self.halley_stop]),
                       ctx.h)
            self._ctx[key] = ctx
        else:
            for s, v in enumerate(self._bc_params(theta)):
                _lib.check(ctx.lib.jh_ctx_set_bc_param(ctx.h, s, v), ctx.h)
        if _lib.BC_KINDS["injection_inlet"] in self.bc_kinds:
            T0, Y = self._inlet(theta)
            arr = (ctypes.c_double * len(Y))(*Y)
            _lib.check(ctx.lib.jh_ctx_set_inlet(ctx.h, T0, arr, len(Y)), ctx.h)
        return ctx

    def _state(self, U, aux):
        U = _as_device_tensor(U, self.dtype, self.device)
        aux = _as_device_tensor(aux, self.dtype, self.device)
        nv = 4 + self.mech.ns - 1
        if U.dim() != 3 or U.shape[0] != nv:
            raise ValueError("U must be (%d, nx, ny)" % nv)
        if aux.shape != (2,) + tuple(U.shape[1:]):
            raise ValueError("aux must be (2, nx, ny)")
        return U, aux

    # ------------------------------------------------------------ callables
    def _fill_user(self, ctx, U, aux, theta):
        """Call each side's boundary function on its three boundary-adjacent
        layers (boundary.py:52-57, 127-131: U_bd (12, 3, ny) for left/right,
        (12, nx, 3) for bottom/top) and copy the ghost layers it returns into
        the side's strip."""
        th = theta if theta is not None else self.theta
        for j, side in enumerate(_lib.SIDES):
            fn = self.user_bc.get(side)
            if fn is None:
                continue
            sl = {"left": (slice(None), slice(0, 3)), "right": (slice(None), slice(-3, None)),
                  "bottom": (slice(None), slice(None), slice(0, 3)),
                  "top": (slice(None), slice(None), slice(-3, None))}[side]
            Ug, ag = fn(U[sl], aux[sl], th)
            strip = torch.cat([torch.as_tensor(Ug), torch.as_tensor(ag)], 0).to(
                device=U.device, dtype=self.dtype)
            if side in ("bottom", "top"):
                strip = strip.permute(0, 2, 1)
            n = U.shape[2] if side in ("left", "right") else U.shape[1]
            if tuple(strip.shape) != (U.shape[0] + 2, 3, n):
                raise ValueError("boundary function for %s returned shapes %s, %s"
                                 % (side, tuple(Ug.shape), tuple(ag.shape)))
            strip = strip.contiguous()
            ctx.call("jh_set_ghost_strip", j, _ptr(strip))

    def _step_pieces(self, ctx, U, aux, Uo, ao, dt, cfl, dt_out, theta):
        """One step through jh_step_begin / _stage / _finish, calling the
        boundary functions before every RK stage."""
        U1, A1 = torch.empty_like(U), torch.empty_like(aux)
        U2, A2 = torch.empty_like(U), torch.empty_like(aux)
        ctx.call("jh_step_begin", _ptr(U), _ptr(aux), dt, cfl)
        for stage, (Us, As, Ud, Ad) in enumerate(((U, aux, U1, A1), (U1, A1, U2, A2),
                                                  (U2, A2, Uo, ao))):
            self._fill_user(ctx, Us, As, theta)
            ctx.call("jh_step_stage", stage, _ptr(Us), _ptr(As), _ptr(U), _ptr(Ud), _ptr(Ad),
                     dt, cfl)
        ctx.call("jh_step_finish", _ptr(Uo), _ptr(ao), dt, cfl,
                 _ptr(dt_out) if dt_out is not None else None)

    # ------------------------------------------------------------ step API
    def advance_one_step(self, U, aux, dx, dy, dt, theta=None):
        """set_advance_func.advance_one_step (set_simulator.py:117-122)."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, theta, dx, dy)
        Uo, ao = torch.empty_like(U), torch.empty_like(aux)
        if self.user_bc:
            self._step_pieces(ctx, U, aux, Uo, ao, float(dt), 0.0, None, theta)
            return Uo, ao
        ctx.call("jh_advance", _ptr(U), _ptr(aux), _ptr(Uo), _ptr(ao), float(dt), 0.0, 1, None)
        return Uo, ao

    def advance_func(self, U, aux, t, theta=None):
        """Simulator.advance_func (set_simulator.py:405-408 / :444-447)."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, theta)
        Uo, ao = torch.empty_like(U), torch.empty_like(aux)
        dt = torch.empty((1,), dtype=torch.float64, device=U.device)
        if self.user_bc:
            fixed = self.fixed_dt is not None
            self._step_pieces(ctx, U, aux, Uo, ao, self.fixed_dt if fixed else 0.0,
                              0.0 if fixed else self.cfl, dt, theta)
        elif self.fixed_dt is not None:
            ctx.call("jh_advance", _ptr(U), _ptr(aux), _ptr(Uo), _ptr(ao), self.fixed_dt, 0.0, 1,
                     _ptr(dt))
        else:
            ctx.call("jh_advance", _ptr(U), _ptr(aux), _ptr(Uo), _ptr(ao), 0.0, self.cfl, 1,
                     _ptr(dt))
        dt = dt[0]
        return Uo, ao, t + dt, dt

    def advance(self, U, aux, nsteps, theta=None, dt_hist=None, dt=None):
        """nsteps calls of advance_func fused into one library call, in place
        (dt: a fixed step instead of the configuration's dt or CFL).

        Returns the per-step dt as a device tensor."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, theta)
        if dt_hist is None:
            dt_hist = torch.empty((nsteps,), dtype=torch.float64, device=U.device)
        fixed_dt = self.fixed_dt if dt is None else float(dt)
        if self.user_bc:
            fixed = fixed_dt is not None
            for s in range(nsteps):
                self._step_pieces(ctx, U, aux, U, aux, fixed_dt if fixed else 0.0,
                                  0.0 if fixed else self.cfl, dt_hist[s:s + 1], theta)
            return U, aux, dt_hist
        if fixed_dt is not None:
            ctx.call("jh_advance", _ptr(U), _ptr(aux), _ptr(U), _ptr(aux), fixed_dt, 0.0,
                     int(nsteps), _ptr(dt_hist))
        else:
            ctx.call("jh_advance", _ptr(U), _ptr(aux), _ptr(U), _ptr(aux), 0.0, self.cfl,
                     int(nsteps), _ptr(dt_hist))
        return U, aux, dt_hist

    def run(self, U_init, aux_init, save_aux=False):
        """Simulator.run (set_simulator.py:457-494): advance to t_end, saving
        the state to results_path (janc_amd.h5.H5Saver, groups save_step{k})
        at t = 0, every save_dt and at the end.  Snapshots are written by a
        worker thread while the steps continue; ``save_aux`` also stores
        (gamma, T) for an exact restart (H5Saver.load_state)."""
        from .h5 import H5Saver
        U, aux = self._state(U_init, aux_init)
        U, aux = U.clone(), aux.clone()
        saver = H5Saver(self.results_path, "2D", simulator=self)
        extra = (lambda a: {"aux": a}) if save_aux else (lambda a: {})
        t = 0.0
        step = 0
        save_step = 1
        saver.save_async(0, t, step, u=U, **extra(aux))
        next_save_time = self.save_dt
        while t < self.t_end:
            U, aux, t_dev, dt = self.advance_func(U, aux, t, self.theta)
            t = float(t_dev)
            step += 1
            if t >= next_save_time or t >= self.t_end:
                saver.save_async(save_step, t, step, u=U, **extra(aux))
                next_save_time += self.save_dt
                save_step += 1
        saver.close()
        return U, aux, t

    # ------------------------------------------------------------ unit pieces
    def flux_func(self, U, aux, dx, dy, dt, theta=None):
        """set_rhs.flux_func (set_simulator.py:44-47): dt * total_flux(BC(U, aux))."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, theta, dx, dy)
        out = torch.empty_like(U)
        self._fill_user(ctx, U, aux, theta)
        ctx.call("jh_rhs", _ptr(U), _ptr(aux), float(dt), _ptr(out))
        return out

    def update_func(self, U, aux):
        """aux_func.update_aux (aux_func.py:21-29)."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, None)
        out = torch.empty_like(aux)
        ctx.call("jh_update_aux", _ptr(U), _ptr(aux), _ptr(out))
        return out

    def reaction_source_terms(self, U, aux, dt, theta=None):
        """reaction_model.reaction_source_terms (reaction_model.py:202-203)."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, theta)
        out = torch.empty_like(U)
        ctx.call("jh_reaction_source", _ptr(U), _ptr(aux), float(dt), _ptr(out))
        return out

    def CFL(self, U, aux, cfl=None):
        """CFL_2D (set_simulator.py:438-443) as a device scalar."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, None)
        out = torch.empty((1,), dtype=torch.float64, device=U.device)
        ctx.call("jh_cfl_dt", _ptr(U), _ptr(aux), float(cfl if cfl is not None else self.cfl),
                 _ptr(out))
        return out[0]

    def lf_alpha(self, U, aux, theta=None):
        """The two LF maxima of LF.py:10-11 over the ghost-padded field."""
        U, aux = self._state(U, aux)
        ctx = self._context(U, theta)
        out = torch.empty((2,), dtype=torch.float64, device=U.device)
        self._fill_user(ctx, U, aux, theta)
        ctx.call("jh_lf_alpha", _ptr(U), _ptr(aux), _ptr(out))
        return out

    def _adjoint_form(self):
        if self.precision != _lib.PREC_FP64:
            raise NotImplementedError("the adjoint is built for the FP64 form")
        if self.flux != (0, 0, 0, 0):
            raise NotImplementedError("the adjoint is built for LF flux splitting with WENO5_JS")

    def step_keep(self, U, aux, dt):
        """One fixed-dt FP64 step that also returns its residuals for
        reverse mode (jh_step_keep: the RK stage states and their LF maxima,
        what jax.vjp of set_simulator.py:117-122 keeps): (U_next, aux_next,
        res), res a flat float64 tensor for step_vjp(..., residuals=res)."""
        self._adjoint_form()
        U, aux = self._state(U, aux)
        ctx = self._context(U, None)
        cnt = ctypes.c_longlong(0)
        _lib.check(ctx.lib.jh_step_res_count(ctx.h, ctypes.byref(cnt)), ctx.h)
        res = torch.empty((cnt.value,), dtype=torch.float64, device=U.device)
        Uo, ao = torch.empty_like(U), torch.empty_like(aux)
        ctx.call("jh_step_keep", _ptr(U), _ptr(aux), float(dt), _ptr(Uo), _ptr(ao), _ptr(res))
        return Uo, ao, res

    def step_vjp(self, U, aux, dt, U_bar, aux_bar, with_output=False, known_output=None,
                 residuals=None):
        """Reverse-mode derivative of advance_one_step at fixed dt (FP64).

        Returns (U_in_bar, aux_in_bar), plus the step's output (U, aux) when
        ``with_output``: the VJP of set_simulator.py:117-122 under JAX's
        rules, get_T_nasa7's custom VJP (thermo_model.py:155-190) included.
        ``known_output`` = the step's output (U, aux) when the caller has it
        (jh_step_vjp_known: the chemistry's forward pass is not repeated);
        with it, ``residuals`` = step_keep's res (jh_step_vjp_res: nothing of
        the forward is recomputed)."""
        self._adjoint_form()
        U, aux = self._state(U, aux)
        Ub, ab = self._state(U_bar, aux_bar)
        ctx = self._context(U, None)
        gU, ga = torch.empty_like(U), torch.empty_like(aux)
        if residuals is not None:
            if known_output is None or with_output:
                raise ValueError("residuals go with known_output (and without with_output)")
            Un, an = self._state(*known_output)
            cnt = ctypes.c_longlong(0)
            _lib.check(ctx.lib.jh_step_res_count(ctx.h, ctypes.byref(cnt)), ctx.h)
            if residuals.dtype != torch.float64 or not residuals.is_contiguous() \
                    or residuals.device != U.device or residuals.numel() != cnt.value:
                raise ValueError("residuals must be step_keep's float64 tensor for this grid")
            ctx.call("jh_step_vjp_res", _ptr(U), _ptr(aux), float(dt), _ptr(residuals),
                     _ptr(Un), _ptr(an), _ptr(Ub), _ptr(ab), _ptr(gU), _ptr(ga))
            return gU, ga
        if known_output is not None and not with_output:
            Un, an = self._state(*known_output)
            ctx.call("jh_step_vjp_known", _ptr(U), _ptr(aux), float(dt), _ptr(Un), _ptr(an),
                     _ptr(Ub), _ptr(ab), _ptr(gU), _ptr(ga))
            return gU, ga
        Uo = torch.empty_like(U) if with_output else None
        ao = torch.empty_like(aux) if with_output else None
        ctx.call("jh_step_vjp", _ptr(U), _ptr(aux), float(dt), _ptr(Ub), _ptr(ab), _ptr(gU),
                 _ptr(ga), _ptr(Uo) if with_output else None,
                 _ptr(ao) if with_output else None)
        return (gU, ga, Uo, ao) if with_output else (gU, ga)

    def set_tile(self, U, variant, theta=None):
        """Select the stage-kernel variant (jh_ctx_set_tile) for U's grid."""
        ctx = self._context(U, theta)
        _lib.check(ctx.lib.jh_ctx_set_tile(ctx.h, int(variant)), ctx.h)

    def get_tile(self, U, theta=None):
        ctx = self._context(U, theta)
        return int(ctx.lib.jh_ctx_get_tile(ctx.h))

    def autotune(self, U, aux, steps=8, theta=None, force=False):
        """Time every stage-kernel variant on a copy of (U, aux); keep the fastest."""
        from .autotune import tune
        return tune(self, U, aux, steps=steps, theta=theta, force=force)

    def timing(self, U, enable=None):
        """Enable/disable per-launch HIP-event timing, or read (and reset)
        {kind: (total_ms, launches)} when ``enable`` is None.  ``enable`` may
        be a sampling period k: fused advance() calls then time the launches
        of every k-th step only."""
        ctx = self._context(U, None)
        if enable is not None:
            _lib.check(ctx.lib.jh_timing_enable(ctx.h, int(enable)), ctx.h)
            return None
        ms = (ctypes.c_double * 4)()
        cnt = (ctypes.c_int * 4)()
        _lib.check(ctx.lib.jh_timing_read(ctx.h, ms, cnt), ctx.h)
        return {k: (ms[i], cnt[i]) for i, k in enumerate(_lib.KERNEL_KINDS)}

    def trace(self, U, records=None):
        """Record, for each following fused step (jh_ctx_set_trace), where the
        LF maxima of every RK stage are attained and which chemistry clip
        branches are active and which cells are below the NASA-7 switch
        temperature after each temperature solve: ``records`` is an int64 CUDA
        tensor of shape (steps, 22) filled step by step; None stops tracing."""
        ctx = self._context(U, None)
        if records is None:
            _lib.check(ctx.lib.jh_ctx_set_trace(ctx.h, None, 0), ctx.h)
            return None
        if records.dtype != torch.int64 or not records.is_cuda or not records.is_contiguous() \
                or records.dim() != 2 or records.shape[1] != 22:
            raise ValueError("records must be a contiguous int64 CUDA tensor (steps, 22)")
        _lib.check(ctx.lib.jh_ctx_set_trace(ctx.h, _ptr(records), int(records.shape[0])), ctx.h)
        return records

    def clip_branches(self, U, masks=None, mode="off"):
        """Record ('record') or replay ('replay') the chemistry clip's branches
        of the following fused steps (jh_ctx_set_clip_branches); ``masks`` is
        an int16 CUDA tensor (steps, nx*ny); 'off' stops."""
        ctx = self._context(U, None)
        m = {"off": 0, "record": 1, "replay": 2}[mode]
        if m == 0:
            _lib.check(ctx.lib.jh_ctx_set_clip_branches(ctx.h, None, 0, 0), ctx.h)
            return None
        n = int(U.shape[1]) * int(U.shape[2])
        if masks.dtype != torch.int16 or not masks.is_cuda or not masks.is_contiguous() \
                or masks.dim() != 2 or masks.shape[1] != n:
            raise ValueError("masks must be a contiguous int16 CUDA tensor (steps, nx*ny)")
        _lib.check(ctx.lib.jh_ctx_set_clip_branches(ctx.h, _ptr(masks), int(masks.shape[0]), m),
                   ctx.h)
        return masks

    def health(self, U):
        """Device health flags since the last query (bits of
        ``_lib.HEALTH_FLAGS``; ``_lib.health_names`` spells them out)."""
        ctx = self._context(U, None)
        f = ctypes.c_uint()
        _lib.check(ctx.lib.jh_health(ctx.h, ctypes.byref(f)), ctx.h)
        return int(f.value)
