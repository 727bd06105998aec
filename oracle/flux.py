"""Oracle restatement of the 2-D flux-splitting Euler flux (test only).

Follows ``src/janc_v2/solver_2D/aux_func.py:21-51``,
``solver_2D/flux_splitting/LF.py:4-16``,
``solver_2D/reconstruction/WENO5_JS.py:1-117`` and
``solver_2D/flux.py:56-68,123-130`` of the reference.
Arrays are (var, x, y) with 3 ghost cells on each side in the padded form.
"""
import numpy as np

P = 2              # WENO5_JS.py:3
EPS = 1e-6         # WENO5_JS.py:4
C1, C2, C3 = 1 / 10, 3 / 5, 3 / 10   # WENO5_JS.py:5-7


def U_to_prim(U, aux, thermo):                     # aux_func.py:39-51
    gamma, T = aux[0:1], aux[1:2]
    rho = U[0:1]
    u = U[1:2] / rho
    v = U[2:3] / rho
    Y = U[4:] / rho
    R = thermo.get_R(Y)
    p = (rho * R * T)
    a = np.sqrt(gamma * p / rho)
    return rho, u, v, Y, p, a


def update_aux(U, aux, thermo, per_cell=False):    # aux_func.py:21-29
    rho = U[0:1]
    u = U[1:2] / rho
    v = U[2:3] / rho
    e = U[3:4] / rho - 0.5 * (u ** 2 + v ** 2)
    Y = U[4:] / rho
    initial_T = aux[1:2]
    aux_new = thermo.get_T(e, Y, initial_T, per_cell=per_cell)
    out = aux.copy()
    out[0:2] = aux_new
    return out


def split_flux_lf(ixy, U, aux, thermo):            # LF.py:4-16
    with np.errstate(invalid="ignore"):
        rho, u, v, Y, p, a = U_to_prim(U, aux, thermo)
    rhoE = U[3:4]
    if ixy == 1:
        F = np.concatenate([rho * u, rho * u ** 2 + p, rho * u * v, u * (rhoE + p),
                            rho * u * Y], axis=0)
    else:
        F = np.concatenate([rho * v, rho * u * v, rho * v ** 2 + p, v * (rhoE + p),
                            rho * v * Y], axis=0)
    um = np.nanmax(abs(u) + a)
    vm = np.nanmax(abs(v) + a)
    theta = um if ixy == 1 else vm
    Hplus = 0.5 * (F + theta * U)
    Hminus = 0.5 * (F - theta * U)
    return Hplus, Hminus


def _weno_core(fjm2, fjm1, fj, fjp1, fjp2):
    """Left-biased WENO5-JS value at j+1/2 (WENO5_JS.py:17-33)."""
    IS1 = 1 / 4 * np.power((fjm2 - 4 * fjm1 + 3 * fj), 2) + 13 / 12 * np.power((fjm2 - 2 * fjm1 + fj), 2)
    IS2 = 1 / 4 * np.power((fjm1 - fjp1), 2) + 13 / 12 * np.power((fjm1 - 2 * fj + fjp1), 2)
    IS3 = 1 / 4 * np.power((3 * fj - 4 * fjp1 + fjp2), 2) + 13 / 12 * np.power((fj - 2 * fjp1 + fjp2), 2)
    alpha1 = C1 / np.power((EPS + IS1), P)
    alpha2 = C2 / np.power((EPS + IS2), P)
    alpha3 = C3 / np.power((EPS + IS3), P)
    w1 = alpha1 / (alpha1 + alpha2 + alpha3)
    w2 = alpha2 / (alpha1 + alpha2 + alpha3)
    w3 = alpha3 / (alpha1 + alpha2 + alpha3)
    f1 = 1 / 3 * fjm2 - 7 / 6 * fjm1 + 11 / 6 * fj
    f2 = -1 / 6 * fjm1 + 5 / 6 * fj + 1 / 3 * fjp1
    f3 = 1 / 3 * fj + 5 / 6 * fjp1 - 1 / 6 * fjp2
    return w1 * f1 + w2 * f2 + w3 * f3


def weno5_L_x(q):                                  # WENO5_JS.py:10-35
    return _weno_core(q[:, 0:-5, 3:-3], q[:, 1:-4, 3:-3], q[:, 2:-3, 3:-3],
                      q[:, 3:-2, 3:-3], q[:, 4:-1, 3:-3])


def weno5_L_y(q):                                  # WENO5_JS.py:37-62
    return _weno_core(q[:, 3:-3, 0:-5], q[:, 3:-3, 1:-4], q[:, 3:-3, 2:-3],
                      q[:, 3:-3, 3:-2], q[:, 3:-3, 4:-1])


def weno5_R_x(q):                                  # WENO5_JS.py:64-90 (mirror)
    return _weno_core(q[:, 5:, 3:-3], q[:, 4:-1, 3:-3], q[:, 3:-2, 3:-3],
                      q[:, 2:-3, 3:-3], q[:, 1:-4, 3:-3])


def weno5_R_y(q):                                  # WENO5_JS.py:92-117 (mirror)
    return _weno_core(q[:, 3:-3, 5:], q[:, 3:-3, 4:-1], q[:, 3:-3, 3:-2],
                      q[:, 3:-3, 2:-3], q[:, 3:-3, 1:-4])


def face_fluxes(U, aux, thermo):
    """(F, G): x-face fluxes (12, nx+1, ny) and y-face fluxes (12, nx, ny+1)."""
    Fplus, Fminus = split_flux_lf(1, U, aux, thermo)
    Gplus, Gminus = split_flux_lf(2, U, aux, thermo)
    F = weno5_L_x(Fplus) + weno5_R_x(Fminus)
    G = weno5_L_y(Gplus) + weno5_R_y(Gminus)
    return F, G


def total_flux(U, aux, dx, dy, thermo):            # flux.py:56-68 (+:123-130)
    """-(dF/dx + dG/dy) on the interior, from the ghost-padded U and aux."""
    F, G = face_fluxes(U, aux, thermo)
    dF = F[:, 1:, :] - F[:, :-1, :]
    dG = G[:, :, 1:] - G[:, :, :-1]
    net_flux = dF / dx + dG / dy
    return -net_flux
