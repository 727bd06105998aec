"""Oracle restatement of the 2-D ghost-cell fill (test infrastructure only).

Follows ``src/janc_v2/boundary/boundary.py:51-132``,
``boundary/boundary_padding.py:20-45`` and
``boundary/boundary_conditions_2D/*.py`` of the reference.
"""
import numpy as np


def pad_2D(U, aux):                                # boundary_padding.py:20-25
    """Periodic pre-pad.  Left ghosts are cells N-4..N-2, right ghosts 1..3."""
    field = np.concatenate([U, aux], axis=0)
    fx = np.concatenate([field[:, -4:-3], field[:, -3:-2], field[:, -2:-1], field,
                         field[:, 1:2], field[:, 2:3], field[:, 3:4]], axis=1)
    fp = np.concatenate([fx[:, :, -4:-3], fx[:, :, -3:-2], fx[:, :, -2:-1], fx,
                         fx[:, :, 1:2], fx[:, :, 2:3], fx[:, :, 3:4]], axis=2)
    return fp[0:-2], fp[-2:]


# ---- BC kinds: each takes the 3 boundary-adjacent interior layers ----------
# zero_gradient.py:3-21 (mirror)
def _zg(U_bd, aux_bd, side, theta=None):
    idx = [2, 1, 0] if side in ("left", "bottom") else [-1, -2, -3]
    ax = 1 if side in ("left", "right") else 2
    return np.take(U_bd, idx, axis=ax), np.take(aux_bd, idx, axis=ax)


# slip_wall.py:3-25 (3 copies of the wall cell, normal momentum negated)
def _slip(U_bd, aux_bd, side, theta=None):
    i0 = 0 if side in ("left", "bottom") else -1
    ax = 1 if side in ("left", "right") else 2
    Ug = np.take(U_bd, [i0, i0, i0], axis=ax).copy()
    ag = np.take(aux_bd, [i0, i0, i0], axis=ax)
    comp = 1 if ax == 1 else 2
    Ug[comp:comp + 1] *= -1
    return Ug, ag


# no_slip_wall.py:3-25 (mirror, both momenta negated)
def _noslip(U_bd, aux_bd, side, theta=None):
    Ug, ag = _zg(U_bd, aux_bd, side)
    Ug = Ug.copy()
    Ug[1:3] *= -1
    return Ug, ag


def _pressure_outlet_state(state_out, gamma_out, T_out, normal_vel, Pb, thermo):
    """pressure_outlet.py:5-22."""
    rho_out = state_out[0:1]
    u_out = state_out[1:2] / rho_out
    v_out = state_out[2:3] / rho_out
    Y_out = state_out[4:] / rho_out
    R_out = thermo.get_R(Y_out)
    p_out = rho_out * (R_out * T_out)
    a_out = np.sqrt(gamma_out * p_out / rho_out)
    mask = (normal_vel / a_out < 1)
    rho_cor = np.where(mask, Pb / (p_out / rho_out), rho_out)
    p_cor = np.where(mask, Pb * np.ones_like(p_out), p_out)
    T_cor = np.where(mask, p_cor / (rho_cor * R_out), T_out)
    _, gamma_o, h_out, _, _ = thermo.get_thermo(T_cor, Y_out)
    U_bd = np.concatenate([rho_cor, rho_cor * u_out, rho_cor * v_out,
                           rho_cor * h_out - p_cor + 0.5 * rho_cor * (u_out ** 2 + v_out ** 2),
                           rho_cor * Y_out], axis=0)
    aux_bd = np.concatenate([gamma_o, T_cor], axis=0)
    return U_bd, aux_bd


def _make_pressure_outlet(thermo):
    # pressure_outlet.py:24-66
    def bc(U_bd, aux_bd, side, theta):
        Pb = theta["Pb"]
        if side == "left":
            sl = (slice(None), slice(0, 1), slice(None))
            normal = lambda s: -(s[1:2] / s[0:1])
            ax = 1
        elif side == "right":
            sl = (slice(None), slice(-1, None), slice(None))
            normal = lambda s: s[1:2] / s[0:1]
            ax = 1
        elif side == "bottom":
            sl = (slice(None), slice(None), slice(0, 1))
            normal = lambda s: -(s[2:3] / s[0:1])
            ax = 2
        else:
            sl = (slice(None), slice(None), slice(-1, None))
            normal = lambda s: s[2:3] / s[0:1]
            ax = 2
        state = U_bd[sl]
        Ub, ab = _pressure_outlet_state(state, aux_bd[0:1][sl], aux_bd[1:2][sl],
                                        normal(state), Pb, thermo)
        return np.repeat(Ub, 3, axis=ax), np.repeat(ab, 3, axis=ax)
    return bc


class Boundary:
    """``set_boundary`` / ``boundary_conditions_2D`` (boundary.py:51-132)."""

    SIDES = ("left", "right", "bottom", "top")

    def __init__(self, boundary_config, thermo):
        self.kinds = {}
        po = _make_pressure_outlet(thermo)
        table = {"zero_gradient": _zg, "slip_wall": _slip, "no_slip_wall": _noslip,
                 "pressure_outlet": po}
        for s in self.SIDES:
            k = boundary_config[s + "_boundary"]
            if callable(k):
                self.kinds[s] = ("user", k)
            elif k == "periodic":
                self.kinds[s] = ("periodic", None)
            else:
                self.kinds[s] = ("builtin", table[k])

    def __call__(self, U, aux, theta=None):
        Up, ap = pad_2D(U, aux)
        Up = Up.copy()
        ap = ap.copy()
        strips = {
            "left": ((slice(None), slice(3, 6), slice(3, -3)), (slice(None), slice(0, 3), slice(3, -3))),
            "right": ((slice(None), slice(-6, -3), slice(3, -3)), (slice(None), slice(-3, None), slice(3, -3))),
            "bottom": ((slice(None), slice(3, -3), slice(3, 6)), (slice(None), slice(3, -3), slice(0, 3))),
            "top": ((slice(None), slice(3, -3), slice(-6, -3)), (slice(None), slice(3, -3), slice(-3, None))),
        }
        for s in self.SIDES:                      # boundary.py:127-131 order
            kind, fn = self.kinds[s]
            if kind == "periodic":
                continue
            src, dst = strips[s]
            U_bd, aux_bd = Up[src], ap[src]
            if kind == "user":
                Ug, ag = fn(U_bd, aux_bd, theta)
            else:
                Ug, ag = fn(U_bd, aux_bd, s, theta)
            Up[dst] = Ug
            ap[dst] = ag
        return Up, ap
