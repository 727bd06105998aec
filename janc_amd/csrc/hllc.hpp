// HLLC approximate Riemann solver at one face (solver_2D/riemann_solver/
// HLLC.py:5-77 of the reference), shared by the Godunov stage kernel and the
// host build that checks it against the oracle.  Tab is MixTables (an LDS
// copy in the kernels) or MechConst.
#pragma once
#include <math.h>

#include "fastmath.hpp"
#include "janc_types.hpp"
#include "thermo.hpp"

namespace jh {

constexpr int kNQ = JH_NS + 3;   // q = [rho, u, v, p, Y(NS-1)]

// Thermodynamic state of one reconstructed face value (HLLC.py:6-12):
// R from the filled mass fractions, T = p/(R rho), then gamma and h of the
// NASA-7 mixture at T.  Evaluated in FP64 whatever the storage precision.
template <typename W, typename Tab>
__host__ __device__ __forceinline__ void face_thermo(const MechConst& c, const Tab& mix,
                                                     const W (&q)[kNQ], double& gamma,
                                                     double& h) {
  double Y[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Y[k] = (double)q[4 + k];
  const double R = fill_Y_get_R(c, Y);
  const double T = (double)q[3] / (R * (double)q[0]);
  gamma = mix_gamma(mix, Y, R, T, &h);
}

// HLLC flux through one face with normal x (DIR 0) or y (DIR 1) from the
// left and right reconstructed primitive states (HLLC.py:14-40 for x,
// :42-74 for y).  Only the branch the reference's where() chain selects is
// evaluated: F_R* if s_M <= 0 <= s_R, else F_L* if s_L <= 0 <= s_M, else F_L
// if s_L >= 0, else F_R; the star flux of side K is F_K + s_K (U_K* - U_K).
template <typename W, int DIR, typename Tab>
__host__ __device__ __forceinline__ void hllc_flux(const MechConst& c, const Tab& mix,
                                                   const W (&qL)[kNQ], const W (&qR)[kNQ],
                                                   W (&F)[kNQ]) {
  double gLd, hLd, gRd, hRd;
  face_thermo(c, mix, qL, gLd, hLd);
  face_thermo(c, mix, qR, gRd, hRd);
  const W rL = qL[0], uL = qL[1], vL = qL[2], pL = qL[3];
  const W rR = qR[0], uR = qR[1], vR = qR[2], pR = qR[3];
  const W aL = sqrt((W)gLd * pL / rL);
  const W aR = sqrt((W)gRd * pR / rR);
  const W EL = rL * (W)hLd - pL + (W)0.5 * rL * (uL * uL + vL * vL);
  const W ER = rR * (W)hRd - pR + (W)0.5 * rR * (uR * uR + vR * vR);
  const W nL = DIR == 0 ? uL : vL, nR = DIR == 0 ? uR : vR;
  const W sL = fmin(nL - aL, nR - aR);
  const W sR = fmax(nL + aL, nR + aR);
  const W sM = (pR - pL + rL * nL * (sL - nL) - rR * nR * (sR - nR)) /
               (rL * (sL - nL) - rR * (sR - nR));
  const bool c1 = sL >= (W)0;
  const bool c2 = sL <= (W)0 && sM >= (W)0;
  const bool c3 = sM <= (W)0 && sR >= (W)0;
  const bool right = c3 || (!c2 && !c1);
  const bool star = c3 || c2;
  // the selected side
  const W r = right ? rR : rL, u = right ? uR : uL, v = right ? vR : vL;
  const W p = right ? pR : pL, E = right ? ER : EL, n = right ? nR : nL;
  const W s = right ? sR : sL;
  F[0] = r * n;
  F[1] = DIR == 0 ? r * (u * u) + p : r * u * v;
  F[2] = DIR == 0 ? r * u * v : r * (v * v) + p;
  F[3] = n * (E + p);
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) F[4 + k] = r * n * (right ? qR[4 + k] : qL[4 + k]);
  if (star) {
    const W fac = r * (s - n) / (s - sM);
    F[0] = F[0] + s * (fac - r);
    F[1] = F[1] + s * (fac * (DIR == 0 ? sM : u) - r * u);
    F[2] = F[2] + s * (fac * (DIR == 0 ? v : sM) - r * v);
    F[3] = F[3] + s * (fac * (E / r + (sM - n) * (sM + p / (r * (s - n)))) - E);
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) {
      const W Y = right ? qR[4 + k] : qL[4 + k];
      F[4 + k] = F[4 + k] + s * (fac * Y - r * Y);
    }
  }
}

}  // namespace jh
