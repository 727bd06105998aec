// NASA-7 mixture thermo and the Halley temperature solve, one cell at a time.
// Follows model/thermo_model.py:80-153 of the reference.  All arithmetic is
// FP64 whatever the storage precision of the state.
#pragma once
#include <math.h>

#include "fastmath.hpp"
#include "janc_types.hpp"

namespace jh {

constexpr int kHalleyMaxIter = 5;     // thermo_model.py:17
constexpr double kHalleyTol = 5e-9;   // thermo_model.py:18

// Mixture polynomial of one NASA-7 range: cp and h per mass (highest power
// first).  Every species shares the switch temperature, so the reference's
// per-species mask (thermo_model.py:101-104) selects one range per cell and
// sum_k Y_k P_k(T) can be evaluated as (sum_k Y_k c_k)(T).
struct MixPoly {
  double cp[5];
  double h[6];
};

// The NASA-7 tables the mixture routines read (cp and h of both ranges and
// the switch temperature).  MechConst has the same members; kernels that
// evaluate the mixture with a per-lane range index read a copy in LDS
// (load_mix_tables), where a per-lane address costs no vector-memory round
// trip.
struct MixTables {
  double cp[2][JH_NS][5];
  double h[2][JH_NS][6];
  double Tcr;
};

// The range index is a runtime (per-lane) value on purpose: with a
// compile-time index the compiler hoists both ranges' ~200 constants into
// registers wherever the mixture is evaluated more than once.
// The h polynomial's coefficients are the cp polynomial's divided by the
// power they multiply (h_j = cp_j / (5 - j), j < 5: the tables are built so),
// plus the constant term; the mixture's h_j are formed the same way from its
// cp_j, and only the constant term is summed over the species.
template <typename Tab>
__host__ __device__ __forceinline__ void mix_poly(const Tab& c, int r,
                                                  const double (&Y)[JH_NS],
                                                  MixPoly& m) {
#pragma unroll
  for (int j = 0; j < 5; ++j) m.cp[j] = 0.0;
  m.h[5] = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) {
#pragma unroll
    for (int j = 0; j < 5; ++j) m.cp[j] += Y[k] * c.cp[r][k][j];
    m.h[5] += Y[k] * c.h[r][k][5];
  }
  m.h[0] = m.cp[0] * (1.0 / 5.0);
  m.h[1] = m.cp[1] * (1.0 / 4.0);
  m.h[2] = m.cp[2] * (1.0 / 3.0);
  m.h[3] = m.cp[3] * (1.0 / 2.0);
  m.h[4] = m.cp[4];
}

// cp, dcp/dT and h of the mixture at T (thermo_model.py:115-126).
__host__ __device__ __forceinline__ void mix_eval(const MixPoly& m, double T,
                                                  double& cp, double& dcp,
                                                  double& h) {
  cp = m.cp[0];
#pragma unroll
  for (int j = 1; j < 5; ++j) cp = cp * T + m.cp[j];
  dcp = 4.0 * m.cp[0];
  dcp = dcp * T + 3.0 * m.cp[1];
  dcp = dcp * T + 2.0 * m.cp[2];
  dcp = dcp * T + m.cp[3];
  h = m.h[0];
#pragma unroll
  for (int j = 1; j < 6; ++j) h = h * T + m.h[j];
}

// fill_Y + get_R (thermo_model.py:80-92): Y[NS-1] = 1 - sum(Y[0:NS-1]).
__host__ __device__ __forceinline__ double fill_Y_get_R(const MechConst& c,
                                                        double (&Y)[JH_NS]) {
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) sum += Y[k];
  Y[JH_NS - 1] = 1.0 - sum;
  double R = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) R += c.R[k] * Y[k];
  return R;
}

// gamma = cp/(cp - R) of the mixture at T (get_thermo_nasa7).
template <typename Tab>
__host__ __device__ __forceinline__ double mix_gamma(const Tab& c,
                                                     const double (&Y)[JH_NS],
                                                     double R, double T,
                                                     double* h_out) {
  MixPoly m;
  mix_poly(c, T < c.Tcr ? 0 : 1, Y, m);
  double cp, dcp, h;
  mix_eval(m, T, cp, dcp, h);
  if (h_out) *h_out = h;
  return cp * rcp(cp - R);
}

// What a Halley solve reports besides (T, gamma) when the grid-wide stop is
// reproduced: n = the updates this cell made before its
// own residual was within tol (or kHalleyMaxIter); T1, g1 = the state after
// one more update (n < kHalleyMaxIter); ok1 = that update's residual is
// within tol too (a cell that converged stays converged); conv = the cell's
// final residual is within tol.
struct HalleyMore {
  int n;
  double T1, g1;
  bool ok1, conv;
};

// Halley iteration for T from e (get_T_nasa7, thermo_model.py:135-153),
// stopping per cell: each cell iterates while its own |res| > tol, at most
// kHalleyMaxIter times.  The reference stops when the grid-wide max|res| <=
// tol, so every cell makes N* = max over cells of n updates.  In SI units a
// converged cell's further updates move T by O(1e-16) relative; with
// nondimensional scales (tol is absolute in the energy unit) they can move it
// by ~1e-9, and the grid-wide stop is reproduced from HalleyMore: a cell with
// n < N* takes the state after one more update (the updates after that move
// it by rounding only, the residual being ~tol^3 after one Halley step).
// Returns whether the final residual is within tol.
template <typename Tab, bool MORE = false>
__host__ __device__ __forceinline__ bool halley_T(const Tab& c,
                                                  const double (&Y)[JH_NS],
                                                  double R, double e, double& T,
                                                  double& gamma, HalleyMore* more = nullptr) {
  int r = T < c.Tcr ? 0 : 1;
  MixPoly m;
  mix_poly(c, r, Y, m);
  double cp, dcp, h;
  mix_eval(m, T, cp, dcp, h);
  double res = (h - R * T) - e;
  double d1 = cp - R;
  gamma = cp * rcp(cp - R);
  int i = 0;
  for (; i < kHalleyMaxIter && fabs(res) > kHalleyTol; ++i) {
    const double dT = -2.0 * res * d1 * rcp(2.0 * (d1 * d1) - res * dcp);
    T = T + dT;
    const int rn = T < c.Tcr ? 0 : 1;
    if (rn != r) {
      r = rn;
      mix_poly(c, r, Y, m);
    }
    mix_eval(m, T, cp, dcp, h);
    res = (h - R * T) - e;
    d1 = cp - R;
    gamma = cp * rcp(cp - R);
  }
  if (MORE) {
    more->n = i;
    more->conv = !(fabs(res) > kHalleyTol);
    more->T1 = T;
    more->g1 = gamma;
    more->ok1 = more->conv;
    if (i < kHalleyMaxIter) {
      const double dT = -2.0 * res * d1 * rcp(2.0 * (d1 * d1) - res * dcp);
      const double T1 = T + dT;
      const int rn = T1 < c.Tcr ? 0 : 1;
      if (rn != r) mix_poly(c, rn, Y, m);
      mix_eval(m, T1, cp, dcp, h);
      const double res1 = (h - R * T1) - e;
      more->T1 = T1;
      more->g1 = cp * rcp(cp - R);
      more->ok1 = !(fabs(res1) > kHalleyTol);
    }
  }
  return !(fabs(res) > kHalleyTol);
}

}  // namespace jh

namespace jh {

// Reverse of get_T_nasa7 at the converged T (get_T_bwd,
// thermo_model.py:159-188): (gamma_bar, T_bar) -> (e_bar, Y_bar[0:NS-1]).
// Y is the filled mass-fraction vector, R its gas constant.
__host__ __device__ __forceinline__ void getT_vjp(const MechConst& c, const double (&Y)[JH_NS],
                                                  double R, double T, double gb, double Tbr,
                                                  double& eb, double (&Yb)[JH_NS - 1]) {
  const int r = T < c.Tcr ? 0 : 1;
  double cpk[JH_NS], ek[JH_NS];
  double cp = 0.0, dcp = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) {
    double a = c.cp[r][k][0];
#pragma unroll
    for (int j = 1; j < 5; ++j) a = a * T + c.cp[r][k][j];
    double d = 4.0 * c.cp[r][k][0];
    d = d * T + 3.0 * c.cp[r][k][1];
    d = d * T + 2.0 * c.cp[r][k][2];
    d = d * T + c.cp[r][k][3];
    double h = c.h[r][k][0];
#pragma unroll
    for (int j = 1; j < 6; ++j) h = h * T + c.h[r][k][j];
    cpk[k] = a;
    ek[k] = h - c.R[k] * T;
    cp += a * Y[k];
    dcp += d * Y[k];
  }
  const double cv = cp - R, icv = rcp(cv);
  const double dTde = icv;
  const double dgdT = (dcp * cv - dcp * cp) * icv * icv;
  eb = gb * dgdT * dTde + Tbr * dTde;
  constexpr int L = JH_NS - 1;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) {
    const double dTdY = (-ek[k] + ek[L]) * icv;
    const double dRdY = c.R[k] - c.R[L];
    const double dcpdY = cpk[k] - cpk[L];
    const double dcvdY = dcpdY - dRdY;
    const double dgdY = dgdT * dTdY + (dcpdY * cv - dcvdY * cp) * icv * icv;
    Yb[k] = gb * dgdY + Tbr * dTdY;
  }
}

}  // namespace jh
