"""Oracle restatement of the NASA-7 thermo model (test infrastructure only).

Follows ``src/janc_v2/model/thermo_model.py:17-190`` of the reference.
Arrays are shaped like the reference's: ``Y`` is ``(8, ...)`` (N2 implied),
``T`` is ``(1, ...)``.
"""
import numpy as np

MAX_ITER = 5     # thermo_model.py:17
TOL = 5e-9       # thermo_model.py:18


class Thermo:
    def __init__(self, th):
        self.th = th
        self.Mex = th["species_M"]
        self.inv_M = 1.0 / self.Mex

    @staticmethod
    def _bc(a, ref):
        return a.reshape(a.shape + (1,) * (ref.ndim - 1)) if ref.ndim > 1 else a

    def fill_Y(self, Y):                           # thermo_model.py:80-82
        Y_last = 1.0 - np.sum(Y, axis=0, keepdims=True)
        return np.concatenate([Y, Y_last], axis=0)

    def get_R(self, Y):                            # thermo_model.py:87-92
        Y9 = self.fill_Y(Y)
        invM = self.inv_M.reshape((-1,) + (1,) * (Y9.ndim - 1))
        return np.sum(invM * Y9, axis=0, keepdims=True)

    def properties(self, T):                       # thermo_model.py:100-112
        """cp_i, dcp_i, h_i for every species at T (T without the leading axis)."""
        th = self.th
        mask = T < th["Tcr"].reshape((-1,) + (1,) * T.ndim)
        cp = np.where(mask, polyval(th["cp_low"], T), polyval(th["cp_high"], T))
        dcp = np.where(mask, polyval(th["dcp_low"], T), polyval(th["dcp_high"], T))
        h = np.where(mask, polyval(th["h_low"], T), polyval(th["h_high"], T))
        return cp, dcp, h

    def get_thermo(self, T, Y):                    # thermo_model.py:115-126
        R = self.get_R(Y)
        Y9 = self.fill_Y(Y)
        cp_i, dcp_i, h_i = self.properties(T[0])
        cp = np.sum(cp_i * Y9, axis=0, keepdims=True)
        h = np.sum(h_i * Y9, axis=0, keepdims=True)
        dcp = np.sum(dcp_i * Y9, axis=0, keepdims=True)
        gamma = cp / (cp - R)
        return cp, gamma, h, R, dcp

    def e_eqn(self, T, e, Y):                      # thermo_model.py:128-133
        cp, gamma, h, R, dcp = self.get_thermo(T, Y)
        res = ((h - R * T) - e)
        return res, (cp - R), dcp, gamma

    def get_T(self, e, Y, T0, per_cell=False):
        """Halley solve for T from e, thermo_model.py:135-153.

        The reference's loop runs while the *global* max|res| exceeds TOL and
        fewer than MAX_ITER iterations were made; ``per_cell=True`` applies the
        same test cell by cell (the form the GPU kernels use).
        """
        res, d1, d2, gamma = self.e_eqn(T0, e, Y)
        T = T0
        i = 0
        if not per_cell:
            while np.max(np.abs(res)) > TOL and i < MAX_ITER:
                dT = -2 * res * d1 / (2 * np.power(d1, 2) - res * d2)
                T = T + dT
                res, d1, d2, gamma = self.e_eqn(T, e, Y)
                i += 1
        else:
            active = np.abs(res) > TOL
            while active.any() and i < MAX_ITER:
                dT = -2 * res * d1 / (2 * np.power(d1, 2) - res * d2)
                Tn = T + dT
                rn, d1n, d2n, gn = self.e_eqn(Tn, e, Y)
                T = np.where(active, Tn, T)
                res = np.where(active, rn, res)
                d1 = np.where(active, d1n, d1)
                d2 = np.where(active, d2n, d2)
                gamma = np.where(active, gn, gamma)
                active = active & (np.abs(res) > TOL)
                i += 1
        return np.concatenate([gamma, T], axis=0)

    def get_T_bwd(self, e, Y, aux_new, g):         # thermo_model.py:159-188
        """VJP of get_T at the converged state (zero cotangent for initial T)."""
        T = aux_new[1:2]
        cp, _, h, R, dcp_dT = self.get_thermo(T, Y)
        cv = cp - R
        dcv_dT = dcp_dT
        dT_de = 1 / cv
        dgamma_dT = (dcp_dT * cv - dcv_dT * cp) / (cv ** 2)
        dgamma_de = dgamma_dT * dT_de
        cp_i, dcp_i_dT, h_i = self.properties(T[0])
        invM = self.inv_M.reshape((-1,) + (1,) * (T.ndim - 1))
        e_i = h_i - invM * T
        dT_dY = (-e_i[0:-1] + e_i[-1:]) / cv
        dR_dY = invM[0:-1] - invM[-1:]
        dcp_dY = cp_i[0:-1] - cp_i[-1:]
        dcv_dY = dcp_dY - dR_dY
        dgamma_dY = (dcp_dY * cv - dcv_dY * cp) / (cv ** 2)
        dgamma_dY = dgamma_dT * dT_dY + dgamma_dY
        dL_dgamma = g[0:1]
        dL_dT = g[1:2]
        dL_de = dL_dgamma * dgamma_de + dL_dT * dT_de
        dL_dY = dL_dgamma * dgamma_dY + dL_dT * dT_dY
        return dL_de, dL_dY, np.zeros_like(T)


def polyval(coefs, x):
    """Row-wise Horner evaluation: coefs (n_species, deg+1), x broadcast.

    Same recurrence as ``jnp.polyval`` (y <- y*x + c, starting from 0).
    """
    c = coefs.reshape(coefs.shape + (1,) * x.ndim)
    y = np.zeros((coefs.shape[0],) + x.shape)
    for j in range(coefs.shape[1]):
        y = y * x + c[:, j]
    return y
