"""Oracle restatement of the injection inlet boundary (test infrastructure only).

The reference ships this boundary as user code in its rotating-detonation
example (``examples/janc_basic_example2.ipynb``, cell "Set boundary
conditions": ``bottom_boundary`` and ``inj_model``, the pressure-feedback
injector model of Fievisohn et al. 2017, doi:10.2514/1.B36103).  The example
runs nondimensionalised with the manifold state as the reference (p0 = T0 =
1); here the manifold total pressure P0 and temperature T0 appear explicitly
so the same model runs in SI units:

* model constants: throat area A1 = 1, chamber area A3 = 5, A2 = A3 - A1,
  gamma = 1.29, R = R(Y_inj);
* 20 bisection steps on the throat Mach number M in [0, 1]:
  P1 = P0 (1 + (g-1)/2 M^2)^(-g/(g-1)),  V1 = M (1 + (g-1)/2 M^2)^(-1/2) C0,
  C0 = sqrt(g R T0),  MFC = A1 P0 / sqrt(T0) sqrt(g/R) M (1 + (g-1)/2 M^2)^(-(g+1)/(2(g-1))),
  V3 = root of V^2/2 + B V + C = 0 with B = g/(g-1) P3 A3 / MFC, C = -g/(g-1) R T0,
  P2 = (MFC (V3 - V1) - P1 A1 + P3 A3) / A2;  M -> upper bound if P2 >= P1
  else lower bound;
* injected state: rho = MFC / V3 / A3, normal velocity V3 into the domain,
  T = p / (R rho) with p the wall cell's pressure;
* blocked injection where the wall cell's pressure is >= P0: the wall cell's
  own state with its normal velocity negated;
* three copies of the corrected wall-cell state fill the ghost layers.

The example defines the bottom side only; the same map is applied on any
side with the normal velocity taken into the domain.
"""
import numpy as np

A1, A3 = 1.0, 5.0
A2 = A3 - A1
GAMMA_INJ = 1.29
N_BISECT = 20


def inj_model(p, Y_inj, P0, T0, thermo):
    """Fievisohn injector model (example 2 ``inj_model``) in SI units.
    p: chamber pressure at the wall cell; Y_inj: (ns-1, ...) mass fractions.
    Returns (V3, T, h, gamma) of the injected stream."""
    g = GAMMA_INJ
    R = thermo.get_R(Y_inj)
    C0 = np.sqrt(g * R * T0)
    P3 = p
    M1 = np.zeros_like(p)
    M2 = np.ones_like(p)
    for _ in range(N_BISECT):
        M = 0.5 * (M1 + M2)
        f = 1 + (g - 1) / 2 * M ** 2
        P1 = P0 * f ** (-g / (g - 1))
        V1 = M * f ** (-0.5) * C0
        MFC = A1 * P0 / np.sqrt(T0) * np.sqrt(g / R) * M * f ** (-(g + 1) / 2 / (g - 1))
        A = 0.5
        B = g / (g - 1) * P3 * A3 / MFC
        C = -g / (g - 1) * R * T0
        V3 = (-B + np.sqrt(B ** 2 - 4 * A * C)) / (2 * A)
        P2 = (MFC * (V3 - V1) - P1 * A1 + P3 * A3) / A2
        up = P2 >= P1
        M2 = np.where(up, M, M2)
        M1 = np.where(~up, M, M1)
    rho = MFC / V3 / A3
    T = p / (R * rho)
    _, gamma, h, _, _ = thermo.get_thermo(T, Y_inj)
    return V3, T, h, gamma


def make_injection_inlet(thermo, side, P0, T0, Y_inj):
    """A boundary callable (U_bd, aux_bd, theta) -> (U_ghost, aux_ghost) for
    ``side``, following example 2's ``bottom_boundary``."""
    xside = side in ("left", "right")
    low = side in ("left", "bottom")
    ax = 1 if xside else 2
    Y_inj = np.asarray(Y_inj, dtype=np.float64)

    def bc(U_bd, aux_bd, theta=None):
        i0 = 0 if low else -1
        U_in = np.take(U_bd, [i0], axis=ax)
        aux_in = np.take(aux_bd, [i0], axis=ax)
        rho = U_in[0:1]
        u = U_in[1:2] / rho
        v = U_in[2:3] / rho
        Y = U_in[4:] / rho
        gamma_in, T_in = aux_in[0:1], aux_in[1:2]
        R_in = thermo.get_R(Y)
        p_in = rho * R_in * T_in
        _, _, h_in, _, _ = thermo.get_thermo(T_in, Y)
        Yi = np.broadcast_to(Y_inj.reshape((-1, 1, 1)), Y.shape)
        Vn, T_inj, h_inj, gamma_inj = inj_model(p_in, Yi, P0, T0, thermo)
        R_inj = thermo.get_R(Yi)
        rho_inj = p_in / (R_inj * T_inj)
        sgn = 1.0 if low else -1.0            # into the domain
        if xside:
            u_inj, v_inj = sgn * Vn, np.zeros_like(v)
        else:
            u_inj, v_inj = np.zeros_like(u), sgn * Vn
        blocked = p_in >= P0
        rho_c = np.where(blocked, rho, rho_inj)
        if xside:
            u_c = np.where(blocked, -u, u_inj)
            v_c = np.where(blocked, v, v_inj)
        else:
            u_c = np.where(blocked, u, u_inj)
            v_c = np.where(blocked, -v, v_inj)
        T_c = np.where(blocked, T_in, T_inj)
        h_c = np.where(blocked, h_in, h_inj)
        Y_c = np.where(blocked, Y, Yi)
        g_c = np.where(blocked, gamma_in, gamma_inj)
        p_c = p_in
        Ug = np.concatenate([rho_c, rho_c * u_c, rho_c * v_c,
                             rho_c * h_c - p_c + 0.5 * rho_c * (u_c ** 2 + v_c ** 2),
                             rho_c * Y_c], axis=0)
        ag = np.concatenate([g_c, T_c], axis=0)
        return np.repeat(Ug, 3, axis=ax), np.repeat(ag, 3, axis=ax)

    return bc


def stoichiometric_injection_Y(ns, ratio=1.0):
    """Example 2's injected mass fractions: H2 and O2 of H2-air at equivalence
    ``ratio`` (O2 0.232 by mass in air), 1e-20 for the other reacting species."""
    yH2 = (ratio / 8 * 0.232) / (1 + ratio / 8 * 0.232)
    yO2 = 0.232 / (1 + ratio / 8 * 0.232)
    return np.concatenate([[yH2, yO2], 1e-20 * np.ones(ns - 3)])


def attach_inlet(oracle, side, P0, T0, Y_inj):
    """Install the injection inlet on ``side`` of an oracle.step.Oracle (the
    reference passes the example's function as that side's boundary)."""
    oracle.boundary.kinds[side] = ("user", make_injection_inlet(oracle.thermo, side, P0, T0,
                                                                Y_inj))
    return oracle
