"""Oracle restatement of the 2-D Euler fluxes (test only).

Follows ``src/janc_v2/solver_2D/aux_func.py:21-51``,
``solver_2D/flux_splitting/LF.py:4-16``, ``flux_splitting/SW.py:4-38``,
``solver_2D/reconstruction/WENO5_JS.py:1-117``, ``reconstruction/MUSCL.py:1-65``,
``solver_2D/riemann_solver/HLLC.py:5-77`` and
``solver_2D/flux.py:43-74,123-130`` of the reference.
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


SW_EPS = 1e-6      # SW.py:26


def split_flux_sw(ixy, U, aux, thermo):            # SW.py:4-38
    """Steger-Warming splitting with the eigenvalue smoothing
    lambda+- = (lambda +- sqrt(lambda^2 + eps^2)) / 2."""
    with np.errstate(invalid="ignore"):
        rho, u, v, Y, p, a = U_to_prim(U, aux, thermo)
    rhoE = U[3:4]
    gamma = aux[0:1]
    zx = (ixy == 1) * 1.0
    zy = (ixy == 2) * 1.0
    theta = zx * u + zy * v
    H1 = (1 / (2 * gamma)) * np.concatenate([rho, rho * u - rho * a * zx, rho * v - rho * a * zy,
                                             rhoE + p - rho * a * theta, rho * Y], axis=0)
    H2 = ((gamma - 1) / gamma) * np.concatenate(
        [rho, rho * u, rho * v, 0.5 * rho * (u ** 2 + v ** 2), rho * Y], axis=0)
    H4 = (1 / (2 * gamma)) * np.concatenate([rho, rho * u + rho * a * zx, rho * v + rho * a * zy,
                                             rhoE + p + rho * a * theta, rho * Y], axis=0)
    lambda1 = theta - a
    lambda2 = theta
    lambda4 = theta + a
    eps = SW_EPS
    with np.errstate(invalid="ignore"):
        lap1 = 0.5 * (lambda1 + np.sqrt(np.power(lambda1, 2) + eps ** 2))
        lam1 = 0.5 * (lambda1 - np.sqrt(np.power(lambda1, 2) + eps ** 2))
        lap2 = 0.5 * (lambda2 + np.sqrt(np.power(lambda2, 2) + eps ** 2))
        lam2 = 0.5 * (lambda2 - np.sqrt(np.power(lambda2, 2) + eps ** 2))
        lap4 = 0.5 * (lambda4 + np.sqrt(np.power(lambda4, 2) + eps ** 2))
        lam4 = 0.5 * (lambda4 - np.sqrt(np.power(lambda4, 2) + eps ** 2))
    Hplus = lap1 * H1 + lap2 * H2 + lap4 * H4
    Hminus = lam1 * H1 + lam2 * H2 + lam4 * H4
    return Hplus, Hminus


SPLIT = {"LF": split_flux_lf, "SW": split_flux_sw}   # flux_splitting/__init__.py:3-6


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


# ---------------------------------------------------------------- Godunov path
# solver_type='godunov' (flux.py:43-54): primitive reconstruction + an
# approximate Riemann solver at every face.

MUSCL_K = 1 / 3      # MUSCL.py:9 (k)
MUSCL_BETA = 1.0     # MUSCL.py:9 (beta)


def prim_q(U, aux, thermo):                        # flux.py:44-45
    """q = [rho, u, v, p, Y(8)] of the padded field."""
    rho, u, v, Y, p, a = U_to_prim(U, aux, thermo)
    return np.concatenate([rho, u, v, p, Y], axis=0)


def min_mod_limiter(x, y):                         # MUSCL.py:4-7
    mask1 = np.minimum(x * np.sign(y), y * np.sign(x))
    mask = np.maximum(np.zeros_like(mask1), mask1)
    return np.sign(x) * mask


def _muscl_slopes(Uc, axis):
    """(phi_fw, phi_bw) at every centre of Uc along ``axis`` (MUSCL.py:10-16)."""
    sl = lambda a, b: tuple(slice(a, b) if i == axis else slice(None) for i in range(3))
    d_minus = Uc[sl(1, -1)] - Uc[sl(None, -2)]
    d_plus = Uc[sl(2, None)] - Uc[sl(1, -1)]
    phi_fw = min_mod_limiter(d_plus, MUSCL_BETA * d_minus)
    phi_bw = min_mod_limiter(d_minus, MUSCL_BETA * d_plus)
    return phi_fw, phi_bw


def muscl_L_x(q):                                  # MUSCL.py:9-21
    U_x = q[:, 1:-1, 3:-3]
    phi_fw, phi_bw = _muscl_slopes(U_x, 1)
    s = 1 / 4 * ((1 - MUSCL_K) * phi_bw + (1 + MUSCL_K) * phi_fw)
    return U_x[:, 1:-2, :] + s[:, :-1, :]


def muscl_R_x(q):                                  # MUSCL.py:38-50
    U_x = q[:, 1:-1, 3:-3]
    phi_fw, phi_bw = _muscl_slopes(U_x, 1)
    s = 1 / 4 * ((1 - MUSCL_K) * phi_fw + (1 + MUSCL_K) * phi_bw)
    return U_x[:, 2:-1, :] - s[:, 1:, :]


def muscl_L_y(q):                                  # MUSCL.py:23-36
    U_y = q[:, 3:-3, 1:-1]
    phi_fw, phi_bw = _muscl_slopes(U_y, 2)
    s = 1 / 4 * ((1 - MUSCL_K) * phi_bw + (1 + MUSCL_K) * phi_fw)
    return U_y[:, :, 1:-2] + s[:, :, :-1]


def muscl_R_y(q):                                  # MUSCL.py:52-65
    U_y = q[:, 3:-3, 1:-1]
    phi_fw, phi_bw = _muscl_slopes(U_y, 2)
    s = 1 / 4 * ((1 - MUSCL_K) * phi_fw + (1 + MUSCL_K) * phi_bw)
    return U_y[:, :, 2:-1] - s[:, :, 1:]


RECONSTRUCTION = {                                 # reconstruction/__init__.py:3-21
    "WENO5_JS": (weno5_L_x, weno5_R_x, weno5_L_y, weno5_R_y),
    "MUSCL": (muscl_L_x, muscl_R_x, muscl_L_y, muscl_R_y),
}


def hllc_parts(qL, qR, thermo, ixy):
    """Wave speeds, states and fluxes of HLLC.py:5-77 on one face direction.

    qL, qR: reconstructed [rho, u, v, p, Y(8)] on either side of each face;
    ixy = 1 for x faces (normal velocity u), 2 for y faces (normal v).
    """
    rho_L, rho_R = qL[0:1], qR[0:1]
    p_L, p_R = qL[3:4], qR[3:4]
    Y_L, Y_R = qL[4:], qR[4:]
    R_L, R_R = thermo.get_R(Y_L), thermo.get_R(Y_R)
    T_L, T_R = p_L / (R_L * rho_L), p_R / (R_R * rho_R)
    _, gamma_L, h_L, _, _ = thermo.get_thermo(T_L, Y_L)
    _, gamma_R, h_R, _, _ = thermo.get_thermo(T_R, Y_R)
    a_L = np.sqrt(gamma_L * p_L / rho_L)
    a_R = np.sqrt(gamma_R * p_R / rho_R)
    u_L, u_R = qL[1:2], qR[1:2]
    v_L, v_R = qL[2:3], qR[2:3]
    E_L = rho_L * h_L - p_L + 0.5 * rho_L * (u_L ** 2 + v_L ** 2)
    E_R = rho_R * h_R - p_R + 0.5 * rho_R * (u_R ** 2 + v_R ** 2)
    n_L, n_R = (u_L, u_R) if ixy == 1 else (v_L, v_R)     # normal velocity
    s_L = np.minimum(n_L - a_L, n_R - a_R)
    s_R = np.maximum(n_L + a_L, n_R + a_R)
    s_M = ((p_R - p_L + rho_L * n_L * (s_L - n_L) - rho_R * n_R * (s_R - n_R))
           / (rho_L * (s_L - n_L) - rho_R * (s_R - n_R)))
    out = dict(s_L=s_L, s_R=s_R, s_M=s_M, a_L=a_L, a_R=a_R, E_L=E_L, E_R=E_R)
    for tag, (rho, u, v, p, E, Y, n, s) in (
            ("L", (rho_L, u_L, v_L, p_L, E_L, Y_L, n_L, s_L)),
            ("R", (rho_R, u_R, v_R, p_R, E_R, Y_R, n_R, s_R))):
        U = np.concatenate([rho, rho * u, rho * v, E, rho * Y], axis=0)
        mom = [s_M, v] if ixy == 1 else [u, s_M]
        U_M = rho * (s - n) / (s - s_M) * np.concatenate(
            [np.ones_like(s_M)] + mom + [E / rho + (s_M - n) * (s_M + p / (rho * (s - n))), Y],
            axis=0)
        if ixy == 1:
            F = np.concatenate([rho * u, rho * u ** 2 + p, rho * u * v, u * (E + p),
                                rho * u * Y], axis=0)
        else:
            F = np.concatenate([rho * v, rho * u * v, rho * v ** 2 + p, v * (E + p),
                                rho * v * Y], axis=0)
        out["U_" + tag], out["UM_" + tag], out["F_" + tag] = U, U_M, F
        out["FM_" + tag] = F + s * (U_M - U)
    return out


def hllc_dir(qL, qR, thermo, ixy):                 # HLLC.py:32-40 (x), 66-74 (y)
    h = hllc_parts(qL, qR, thermo, ixy)
    s_L, s_M, s_R = h["s_L"], h["s_M"], h["s_R"]
    cond1 = s_L >= 0
    cond2 = np.logical_and(s_L <= 0, s_M >= 0)
    cond3 = np.logical_and(s_M <= 0, s_R >= 0)
    F1 = np.where(cond1, h["F_L"], h["F_R"])
    F2 = np.where(cond2, h["FM_L"], F1)
    return np.where(cond3, h["FM_R"], F2)


RIEMANN = {"HLLC": hllc_dir}                       # riemann_solver/__init__.py:3-6


def godunov_face_fluxes(U, aux, thermo, reconstruction="WENO5_JS", riemann="HLLC"):
    """(F, G) of flux.py:43-54 from the ghost-padded U and aux."""
    with np.errstate(invalid="ignore", divide="ignore"):
        q = prim_q(U, aux, thermo)
        Lx, Rx, Ly, Ry = RECONSTRUCTION[reconstruction]
        F = RIEMANN[riemann](Lx(q), Rx(q), thermo, 1)
        G = RIEMANN[riemann](Ly(q), Ry(q), thermo, 2)
    return F, G


def split_face_fluxes(U, aux, thermo, reconstruction="WENO5_JS", split="LF"):
    """(F, G) of flux.py:56-68 with either reconstruction and split."""
    Lx, Rx, Ly, Ry = RECONSTRUCTION[reconstruction]
    Fplus, Fminus = SPLIT[split](1, U, aux, thermo)
    Gplus, Gminus = SPLIT[split](2, U, aux, thermo)
    return Lx(Fplus) + Rx(Fminus), Ly(Gplus) + Ry(Gminus)


def scheme_total_flux(U, aux, dx, dy, thermo, solver_type="flux_splitting",
                      reconstruction="WENO5_JS", riemann="HLLC", split="LF"):
    """-(dF/dx + dG/dy) for either solver_type (flux.py:43-74)."""
    if solver_type == "godunov":
        F, G = godunov_face_fluxes(U, aux, thermo, reconstruction, riemann)
    else:
        F, G = split_face_fluxes(U, aux, thermo, reconstruction, split)
    dF = F[:, 1:, :] - F[:, :-1, :]
    dG = G[:, :, 1:] - G[:, :, :-1]
    net_flux = dF / dx + dG / dy
    return -net_flux
