"""Oracle restatement of the 2-D step closure (test infrastructure only).

Follows ``src/janc_v2/simulation/set_simulator.py:44-47`` (flux_func),
``solver_2D/flux.py:19-40`` (set_flux_solver: flux splitting or Godunov),
``:117-122`` (advance_one_step, detailed-chemistry branch), ``:438-447``
(CFL_2D / advance_func) and ``simulation/time_step/TVD_RK3.py:5-12``.
"""
import numpy as np

from . import flux as _flux
from .boundary import Boundary
from .mech import load_mechanism
from .reaction import Reaction
from .thermo import Thermo


class Oracle:
    """One configured solver: mechanism + grid + boundary kinds."""

    def __init__(self, mech_path, boundary_config, dx, dy, nondim_config=None,
                 per_cell_halley=False, flux_config=None):
        self.rp, self.th, self.nd = load_mechanism(mech_path, nondim_config)
        self.thermo = Thermo(self.th)
        self.reaction = Reaction(self.rp, self.th, self.nd)
        self.boundary = Boundary(boundary_config, self.thermo)
        self.dx, self.dy = dx, dy
        self.per_cell = per_cell_halley
        fc = flux_config or {}                     # flux.py:19-40
        self.solver_type = fc.get("solver_type", "flux_splitting")
        self.reconstruction = fc.get("interface_reconstruction", "WENO5_JS")
        self.riemann = fc.get("riemann_solver", "HLLC")
        self.split = fc.get("split_method", "LF")

    # set_simulator.py:44-47
    def flux_func(self, U, aux, dt, theta=None):
        Ug, ag = self.boundary(U, aux, theta)
        return dt * (_flux.scheme_total_flux(Ug, ag, self.dx, self.dy, self.thermo,
                                             self.solver_type, self.reconstruction,
                                             self.riemann, self.split))

    def update_func(self, U, aux):                 # aux_func.py:21-29
        return _flux.update_aux(U, aux, self.thermo, per_cell=self.per_cell)

    def reaction_source_terms(self, U, aux, dt):   # reaction_model.py:183-203
        return self.reaction.detailed_reaction(U, aux, dt, self.thermo.fill_Y)

    # TVD_RK3.py:5-12
    def advance_flux(self, U, aux, dt, theta=None):
        U1 = U + self.flux_func(U, aux, dt, theta)
        aux1 = self.update_func(U1, aux)
        U2 = 3 / 4 * U + 1 / 4 * (U1 + self.flux_func(U1, aux1, dt, theta))
        aux2 = self.update_func(U2, aux1)
        U3 = 1 / 3 * U + 2 / 3 * (U2 + self.flux_func(U2, aux2, dt, theta))
        aux3 = self.update_func(U3, aux2)
        return U3, aux3

    # set_simulator.py:117-122
    def advance_one_step(self, U, aux, dt, theta=None):
        U, aux = self.advance_flux(U, aux, dt, theta)
        dU = self.reaction_source_terms(U, aux, dt)
        U = U + dU
        aux = self.update_func(U, aux)
        return U, aux

    # set_simulator.py:438-443
    def cfl_dt(self, U, aux, cfl):
        _, u, v, _, _, a = _flux.U_to_prim(U, aux, self.thermo)
        sx = np.abs(u) + a
        sy = np.abs(v) + a
        return cfl * np.min(1 / (sx / self.dx + sy / self.dy))

    # set_simulator.py:444-447
    def advance_func(self, U, aux, t, cfl, theta=None):
        dt = self.cfl_dt(U, aux, cfl)
        U, aux = self.advance_one_step(U, aux, dt, theta)
        return U, aux, t + dt, dt

    # ---- helpers used by the tests ----
    def lf_alpha(self, U, aux, theta=None):
        """(alpha_x, alpha_y): the LF.py:10-11 maxima over the padded field."""
        Ug, ag = self.boundary(U, aux, theta)
        with np.errstate(invalid="ignore"):
            _, u, v, _, _, a = _flux.U_to_prim(Ug, ag, self.thermo)
        return np.nanmax(abs(u) + a), np.nanmax(abs(v) + a)
