"""Oracle restatement of the point-implicit detailed chemistry (test only).

Follows ``src/janc_v2/model/reaction_model.py:78-193`` of the reference.
The reference evaluates ``reactionConstant_i`` for every (k, n, i) through
``vmap``; the arithmetic per entry is identical here, vectorised over cells.
"""
import numpy as np

from .thermo import polyval


class Reaction:
    def __init__(self, rp, th, nd, n_y0=4):
        self.rp = rp
        self.th = th
        self.nd = nd
        self.nr = rp["num_of_species"] - rp["num_of_inert_species"]   # :60-63
        self.nY = n_y0                                                 # :70-71
        self.species_M = th["species_M"]
        dnu = np.abs(rp["vb"] - rp["vf"])
        assert np.all(np.isin(dnu, (0.0, 1.0, 2.0, 4.0))), \
            "stoichiometric differences must be exact scale factors"

    def get_gibbs(self, T):                         # reaction_model.py:78-89
        th, nr, nd = self.th, self.nr, self.nd
        Tcr = th["Tcr"][0:nr].reshape((-1,) + (1,) * T.ndim)
        mask = T < Tcr
        h = np.where(mask, polyval(th["h_chem_low"][0:nr], T),
                     polyval(th["h_chem_high"][0:nr], T))
        lcl = th["logcof_low"][0:nr].reshape((-1,) + (1,) * T.ndim)
        lch = th["logcof_high"][0:nr].reshape((-1,) + (1,) * T.ndim)
        s = np.where(mask, polyval(th["s_low"][0:nr], T) + lcl * np.log(nd.T0 * T),
                     polyval(th["s_high"][0:nr], T) + lch * np.log(nd.T0 * T))
        return s - h / T

    def rates_and_jacobian(self, T, X):
        """w_k (nr, ...) and dw_k/d(rho Y_n) (nr, nr, ...).

        ``reactionConstant_i`` (:91-122) summed over i by
        ``reaction_rate_with_derievative`` (:124-130); T is (1, ...), X is
        (ns, ...) of unclipped concentrations rho*Y/M.
        """
        rp, nr, nd = self.rp, self.nr, self.nd
        ext = (1,) * (X.ndim - 1)
        gibbs = self.get_gibbs(T[0])                      # (nr, ...)
        Xc = np.clip(X, 1e-50, None)                      # :108
        log_X = np.log(Xc[0:nr])                          # :109
        w = np.zeros((nr,) + X.shape[1:])
        J = np.zeros((nr, nr) + X.shape[1:])
        for i in range(rp["num_of_reactions"]):
            A, B, E = rp["A"][i], rp["B"][i], rp["Ea/Ru"][i]
            vf_i = rp["vf"][i].reshape((-1,) + ext)
            vb_i = rp["vb"][i].reshape((-1,) + ext)
            aij = rp["third_body_coeffs"][i].reshape((-1,) + ext)
            tb = rp["is_third_body"][i]
            kf_i = A * np.power(T, B) * np.exp(-E / T)                  # :105
            aX = np.sum(aij * X, axis=0, keepdims=True)                 # :106
            aX = tb * aX + (1 - tb)                                     # :107
            kf = kf_i * np.exp(np.sum(vf_i * log_X, axis=0, keepdims=True))   # :110
            kb = kf_i / (np.exp(np.sum((vb_i - vf_i) * gibbs, axis=0, keepdims=True))
                         * ((101325 / nd.P0 / T) ** rp["vsum"][i])) \
                * np.exp(np.sum(vb_i * log_X, axis=0, keepdims=True))  # :113
            # :115 and :121.  Every stoichiometric difference of the mechanisms
            # this oracle accepts is in {0, +-1, +-2}, so multiplying by dnu is
            # exact and may be factored out of the (k, n) products without
            # changing a single rounding.
            dnu = (rp["vb"][i, 0:nr] - rp["vf"][i, 0:nr]).reshape((-1,) + ext)
            w += dnu * aX * (kf - kb)
            ain = rp["third_body_coeffs"][i, 0:nr].reshape((-1,) + ext)
            Mn = self.species_M[0:nr].reshape((-1,) + ext)
            vfn, vbn = vf_i[0:nr], vb_i[0:nr]
            dvec = (kf - kb) * (ain / Mn) + 1 / (Mn * Xc[0:nr]) * aX * (vfn * kf - vbn * kb)
            J += dnu[:, None] * dvec[None, :]
        Mk = self.species_M[0:nr].reshape((-1,) + ext)
        w = Mk * w                                                      # :128
        J = Mk[:, None] * J                                             # :129
        return w, J

    def construct_matrix_equation(self, T, X, dt):  # reaction_model.py:147-159
        w, J = self.rates_and_jacobian(T, X)
        nr = self.nr
        S = np.moveaxis(w, 0, -1)[..., None]                      # (..., nr, 1)
        DSDU = np.moveaxis(np.moveaxis(J, 0, -1), 0, -1)          # (..., nr, nr)
        A = np.eye(nr) / dt - DSDU
        x = np.linalg.solve(A, S)
        return np.moveaxis(x[..., 0], -1, 0)

    def detailed_reaction(self, U, aux, dt, fill_Y):  # reaction_model.py:183-193
        nY = self.nY
        Mex = self.species_M.reshape((-1,) + (1,) * (U.ndim - 1))
        rho = U[0:1]
        Y = fill_Y(U[nY:] / rho)
        rhoY = rho * Y
        T = aux[1:2]
        X = rhoY / Mex
        drhoY = self.construct_matrix_equation(T, X, dt)
        dY = drhoY / rho
        dY = np.clip(dY, -Y[0:-1], 1 - Y[0:-1])
        return np.concatenate([np.zeros_like(U[:nY]), rho * dY], axis=0)
