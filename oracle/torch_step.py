"""Differentiable FP64 twin of the oracle step in torch (test infrastructure).

Stands in for ``jax.grad`` of the reference, which cannot run here: the same
restatement as ``oracle/step.py`` written in torch ops so that
``torch.autograd`` differentiates it the way JAX differentiates
``src/janc_v2``:

* ``get_T_nasa7`` has a custom VJP (``model/thermo_model.py:135-190``): the
  implicit derivative at the converged T, dT/de = 1/cv,
  dT/dY_k = -(e_k - e_N2)/cv, and zero cotangent for the initial T;
* ``nanmax`` in LF (``LF.py:10-11``) and the clip of the chemistry increment
  (``reaction_model.py:191``) are written with amax / maximum / minimum, whose
  gradients split ties evenly like JAX's reductions;
* ``jnp.linalg.solve`` -> ``torch.linalg.solve``.

Only fixed-dt steps are differentiated (dt is a constant).  CPU only, small
grids.
"""
import numpy as np
import torch

from .boundary import Boundary
from .mech import load_mechanism
from .thermo import Thermo

T64 = torch.float64


def _t(x):
    return torch.as_tensor(np.asarray(x), dtype=T64)


class TorchOracle:
    def __init__(self, mech_path, boundary_config, dx, dy, nondim_config=None):
        rp, th, nd = load_mechanism(mech_path, nondim_config)
        self.rp, self.th, self.nd = rp, th, nd
        self.np_thermo = Thermo(th)
        self.dx, self.dy = float(dx), float(dy)
        self.nr = rp["num_of_species"] - rp["num_of_inert_species"]
        self.Mex = _t(th["species_M"])
        self.invM = 1.0 / self.Mex
        self.kinds = {}
        for s in ("left", "right", "bottom", "top"):
            k = boundary_config[s + "_boundary"]
            if callable(k) or k == "pressure_outlet":
                raise NotImplementedError("differentiable twin supports linear BC kinds")
            self.kinds[s] = k
        outer = self

        class GetT(torch.autograd.Function):
            """get_T_nasa7 with its custom VJP (thermo_model.py:135-190)."""

            @staticmethod
            def forward(ctx, e, Y, T0):
                a = outer.np_thermo.get_T(e.detach().numpy(), Y.detach().numpy(),
                                          T0.detach().numpy())
                aux = _t(a)
                ctx.save_for_backward(e, Y, aux)
                return aux

            @staticmethod
            def backward(ctx, g):
                e, Y, aux = ctx.saved_tensors
                de, dY, dT0 = outer.np_thermo.get_T_bwd(e.numpy(), Y.numpy(), aux.numpy(),
                                                        g.numpy())
                return _t(de), _t(dY), _t(dT0)

        self.GetT = GetT

    # ---------------------------------------------------------------- thermo
    def fill_Y(self, Y):
        return torch.cat([Y, 1.0 - Y.sum(0, keepdim=True)], 0)

    def get_R(self, Y):
        Y9 = self.fill_Y(Y)
        return (self.invM.reshape((-1,) + (1,) * (Y9.dim() - 1)) * Y9).sum(0, keepdim=True)

    def update_aux(self, U, aux):
        rho = U[0:1]
        u = U[1:2] / rho
        v = U[2:3] / rho
        e = U[3:4] / rho - 0.5 * (u ** 2 + v ** 2)
        Y = U[4:] / rho
        return self.GetT.apply(e, Y, aux[1:2])

    # ---------------------------------------------------------------- boundary
    def pad(self, U, aux):
        f = torch.cat([U, aux], 0)
        fx = torch.cat([f[:, -4:-3], f[:, -3:-2], f[:, -2:-1], f, f[:, 1:2], f[:, 2:3],
                        f[:, 3:4]], 1)
        fp = torch.cat([fx[:, :, -4:-3], fx[:, :, -3:-2], fx[:, :, -2:-1], fx, fx[:, :, 1:2],
                        fx[:, :, 2:3], fx[:, :, 3:4]], 2)
        nvar = U.shape[0]
        for side in ("left", "right", "bottom", "top"):
            k = self.kinds[side]
            if k == "periodic":
                continue
            xside = side in ("left", "right")
            low = side in ("left", "bottom")
            n_norm = U.shape[1] if xside else U.shape[2]
            if k == "slip_wall":
                src = [0, 0, 0] if low else [n_norm - 1] * 3
            else:
                src = [2, 1, 0] if low else [n_norm - 1, n_norm - 2, n_norm - 3]
            sign = torch.ones(nvar + 2, dtype=T64)
            if k == "slip_wall":
                sign[1 if xside else 2] = -1.0
            if k == "no_slip_wall":
                sign[1] = -1.0
                sign[2] = -1.0
            sign = sign.reshape(-1, 1, 1)
            if xside:
                ghosts = torch.stack([f[:, s, :] for s in src], 1) * sign
                rng = slice(0, 3) if low else slice(-3, None)
                fp = fp.clone()
                fp[:, rng, 3:-3] = ghosts
            else:
                ghosts = torch.stack([f[:, :, s] for s in src], 2) * sign
                rng = slice(0, 3) if low else slice(-3, None)
                fp = fp.clone()
                fp[:, 3:-3, rng] = ghosts
        return fp[0:nvar], fp[nvar:]

    # ---------------------------------------------------------------- flux
    def prim(self, U, aux):
        gamma, T = aux[0:1], aux[1:2]
        rho = U[0:1]
        u = U[1:2] / rho
        v = U[2:3] / rho
        Y = U[4:] / rho
        p = rho * self.get_R(Y) * T
        a = torch.sqrt(gamma * p / rho)
        return rho, u, v, Y, p, a

    @staticmethod
    def _nanmax(x):
        return torch.amax(torch.where(torch.isnan(x), torch.full_like(x, -float("inf")), x))

    def split(self, ixy, U, aux):
        rho, u, v, Y, p, a = self.prim(U, aux)
        rhoE = U[3:4]
        if ixy == 1:
            F = torch.cat([rho * u, rho * u ** 2 + p, rho * u * v, u * (rhoE + p), rho * u * Y], 0)
            th = self._nanmax(torch.abs(u) + a)
        else:
            F = torch.cat([rho * v, rho * u * v, rho * v ** 2 + p, v * (rhoE + p), rho * v * Y], 0)
            th = self._nanmax(torch.abs(v) + a)
        return 0.5 * (F + th * U), 0.5 * (F - th * U)

    @staticmethod
    def _weno(fm2, fm1, f0, fp1, fp2):
        IS1 = 0.25 * (fm2 - 4 * fm1 + 3 * f0) ** 2 + 13 / 12 * (fm2 - 2 * fm1 + f0) ** 2
        IS2 = 0.25 * (fm1 - fp1) ** 2 + 13 / 12 * (fm1 - 2 * f0 + fp1) ** 2
        IS3 = 0.25 * (3 * f0 - 4 * fp1 + fp2) ** 2 + 13 / 12 * (f0 - 2 * fp1 + fp2) ** 2
        a1 = 0.1 / (1e-6 + IS1) ** 2
        a2 = 0.6 / (1e-6 + IS2) ** 2
        a3 = 0.3 / (1e-6 + IS3) ** 2
        s = a1 + a2 + a3
        q1 = 1 / 3 * fm2 - 7 / 6 * fm1 + 11 / 6 * f0
        q2 = -1 / 6 * fm1 + 5 / 6 * f0 + 1 / 3 * fp1
        q3 = 1 / 3 * f0 + 5 / 6 * fp1 - 1 / 6 * fp2
        return a1 / s * q1 + a2 / s * q2 + a3 / s * q3

    def total_flux(self, Up, ap):
        Fp, Fm = self.split(1, Up, ap)
        Gp, Gm = self.split(2, Up, ap)
        w = self._weno
        F = w(Fp[:, 0:-5, 3:-3], Fp[:, 1:-4, 3:-3], Fp[:, 2:-3, 3:-3], Fp[:, 3:-2, 3:-3],
              Fp[:, 4:-1, 3:-3]) + \
            w(Fm[:, 5:, 3:-3], Fm[:, 4:-1, 3:-3], Fm[:, 3:-2, 3:-3], Fm[:, 2:-3, 3:-3],
              Fm[:, 1:-4, 3:-3])
        G = w(Gp[:, 3:-3, 0:-5], Gp[:, 3:-3, 1:-4], Gp[:, 3:-3, 2:-3], Gp[:, 3:-3, 3:-2],
              Gp[:, 3:-3, 4:-1]) + \
            w(Gm[:, 3:-3, 5:], Gm[:, 3:-3, 4:-1], Gm[:, 3:-3, 3:-2], Gm[:, 3:-3, 2:-3],
              Gm[:, 3:-3, 1:-4])
        return -((F[:, 1:] - F[:, :-1]) / self.dx + (G[:, :, 1:] - G[:, :, :-1]) / self.dy)

    def flux_func(self, U, aux, dt):
        Up, ap = self.pad(U, aux)
        return dt * self.total_flux(Up, ap)

    # ---------------------------------------------------------------- chemistry
    def _gibbs(self, T):
        th, nr, nd = self.th, self.nr, self.nd
        def pv(c, x):
            y = torch.zeros((c.shape[0],) + tuple(x.shape), dtype=T64)
            for j in range(c.shape[1]):
                y = y * x + _t(c[:, j]).reshape((-1,) + (1,) * x.dim())
            return y
        lo = T < float(th["Tcr"][0])
        h = torch.where(lo, pv(th["h_chem_low"][:nr], T), pv(th["h_chem_high"][:nr], T))
        ext = (-1,) + (1,) * T.dim()
        s = torch.where(lo, pv(th["s_low"][:nr], T) + _t(th["logcof_low"][:nr]).reshape(ext) * torch.log(nd.T0 * T),
                        pv(th["s_high"][:nr], T) + _t(th["logcof_high"][:nr]).reshape(ext) * torch.log(nd.T0 * T))
        return s - h / T

    def reaction(self, U, aux, dt):
        rp, nr, nd = self.rp, self.nr, self.nd
        rho = U[0:1]
        Y = self.fill_Y(U[4:] / rho)
        T = aux[1:2]
        X = rho * Y / self.Mex.reshape(-1, 1, 1)
        g = self._gibbs(T[0])
        Xc = torch.maximum(X, torch.full_like(X, 1e-50))
        logX = torch.log(Xc[0:nr])
        w = torch.zeros((nr,) + tuple(X.shape[1:]), dtype=T64)
        J = torch.zeros((nr, nr) + tuple(X.shape[1:]), dtype=T64)
        ext = (1,) * (X.dim() - 1)
        for i in range(rp["num_of_reactions"]):
            A, B, E = rp["A"][i], rp["B"][i], rp["Ea/Ru"][i]
            vf = _t(rp["vf"][i]).reshape((-1,) + ext)
            vb = _t(rp["vb"][i]).reshape((-1,) + ext)
            aij = _t(rp["third_body_coeffs"][i]).reshape((-1,) + ext)
            tb = rp["is_third_body"][i]
            kfi = A * torch.pow(T, B) * torch.exp(-E / T)
            aX = (aij * X).sum(0, keepdim=True)
            aX = tb * aX + (1 - tb)
            kf = kfi * torch.exp((vf * logX).sum(0, keepdim=True))
            kb = kfi / (torch.exp(((vb - vf) * g).sum(0, keepdim=True)) *
                        (101325 / nd.P0 / T) ** rp["vsum"][i]) * torch.exp((vb * logX).sum(0, keepdim=True))
            dnu = vb - vf
            w = w + dnu * aX * (kf - kb)
            ain = _t(rp["third_body_coeffs"][i, 0:nr]).reshape((-1,) + ext)
            Mn = self.Mex[0:nr].reshape((-1,) + ext)
            dvec = (kf - kb) * (ain / Mn) + 1 / (Mn * Xc[0:nr]) * aX * (vf * kf - vb * kb)
            J = J + dnu[:, None] * dvec[None, :]
        Mk = self.Mex[0:nr].reshape((-1,) + ext)
        w = Mk * w
        J = Mk[:, None] * J
        Amat = torch.eye(nr, dtype=T64) / dt - J.permute(2, 3, 0, 1)
        x = torch.linalg.solve(Amat, w.permute(1, 2, 0)[..., None])[..., 0].permute(2, 0, 1)
        dY = x / rho
        Yb = Y[0:-1]
        dY = torch.minimum(torch.maximum(dY, -Yb), 1 - Yb)
        return torch.cat([torch.zeros_like(U[:4]), rho * dY], 0)

    # ---------------------------------------------------------------- step
    def advance_one_step(self, U, aux, dt):
        """set_simulator.py:117-122 with TVD_RK3.py:5-12, fixed dt."""
        U1 = U + self.flux_func(U, aux, dt)
        aux1 = self.update_aux(U1, aux)
        U2 = 3 / 4 * U + 1 / 4 * (U1 + self.flux_func(U1, aux1, dt))
        aux2 = self.update_aux(U2, aux1)
        U3 = 1 / 3 * U + 2 / 3 * (U2 + self.flux_func(U2, aux2, dt))
        aux3 = self.update_aux(U3, aux2)
        U4 = U3 + self.reaction(U3, aux3, dt)
        aux4 = self.update_aux(U4, aux3)
        return U4, aux4
