// Point-implicit detailed chemistry for one cell (FP64).
// Follows model/reaction_model.py:78-193 of the reference:
//   X = rho*Y/M, (I/dt - dw/d(rhoY)) x = w, dY = clip(x/rho, -Y, 1-Y).
#pragma once
#include <math.h>

#include "fastmath.hpp"
#include "janc_types.hpp"
#include "mech_gen.inc"

namespace jh {

// g_k = s_k/R - h_k/(R T) for the active species (get_gibbs, :78-89).
__host__ __device__ __forceinline__ void gibbs(const MechConst& c, double T, double invT,
                                               double (&g)[JH_NA]) {
  const int r = T < c.Tcr ? 0 : 1;
  const double lT = log(c.T0 * T);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    double h = c.hc[r][k][0];
#pragma unroll
    for (int j = 1; j < 6; ++j) h = h * T + c.hc[r][k][j];
    double s = c.s[r][k][0];
#pragma unroll
    for (int j = 1; j < 5; ++j) s = s * T + c.s[r][k][j];
    s = s + c.logc[r][k] * lT;
    g[k] = s - h * invT;
  }
}

// Solve A x = b in place (b <- x) by LU with partial pivoting, the
// factorisation jnp.linalg.solve uses (LAPACK getrf/getrs).  Row swaps are
// predicated selects so every index stays compile-time and A lives in VGPRs.
template <int N>
__host__ __device__ __forceinline__ void lu_solve(double (&A)[N][N], double (&b)[N]) {
#pragma unroll
  for (int col = 0; col < N; ++col) {
    int p = col;
    double amax = fabs(A[col][col]);
#pragma unroll
    for (int r = col + 1; r < N; ++r) {
      const double v = fabs(A[r][col]);
      if (v > amax) {
        amax = v;
        p = r;
      }
    }
#pragma unroll
    for (int r = col + 1; r < N; ++r) {
      const bool s = (p == r);
#pragma unroll
      for (int j = col; j < N; ++j) {
        const double t = A[col][j];
        A[col][j] = s ? A[r][j] : t;
        A[r][j] = s ? t : A[r][j];
      }
      const double t = b[col];
      b[col] = s ? b[r] : t;
      b[r] = s ? t : b[r];
    }
    const double inv = rcp(A[col][col]);
#pragma unroll
    for (int r = col + 1; r < N; ++r) {
      const double l = A[r][col] * inv;
#pragma unroll
      for (int j = col + 1; j < N; ++j) A[r][j] -= l * A[col][j];
      b[r] -= l * b[col];
    }
  }
#pragma unroll
  for (int r = N - 1; r >= 0; --r) {
    double s = b[r];
#pragma unroll
    for (int j = r + 1; j < N; ++j) s -= A[r][j] * b[j];
    b[r] = s * rcp(A[r][r]);
  }
}

// Increment of rho*Y_k for the active species (detailed_reaction, :183-193).
// rhoY_in: the state's rho*Y_k for the NS-1 transported species.
__host__ __device__ __forceinline__ void chem_increment(const MechConst& c, double rho,
                                                        const double (&rhoY_in)[JH_NS - 1],
                                                        double T, double dt,
                                                        double (&drhoY)[JH_NA]) {
  const double irho = rcp(rho);
  double Y[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Y[k] = rhoY_in[k] * irho;
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) sum += Y[k];
  Y[JH_NS - 1] = 1.0 - sum;
  double X[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) X[k] = (rho * Y[k]) * c.R[k];
  double Xc[JH_NA], invMX[JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    Xc[k] = fmax(X[k], 1e-50);
    invMX[k] = rcp(c.M[k] * Xc[k]);
  }
  const double invT = rcp(T);
  double g[JH_NA];
  gibbs(c, T, invT, g);
  double w[JH_NA], J[JH_NA][JH_NA];
  mech_rates_jacobian(c, T, invT, log(T), X, Xc, invMX, g, w, J);
  const double idt = rcp(dt);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
#pragma unroll
    for (int n = 0; n < JH_NA; ++n) J[k][n] = ((k == n) ? idt : 0.0) - J[k][n];
  }
  lu_solve<JH_NA>(J, w);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    double dY = w[k] * irho;
    dY = fmin(fmax(dY, -Y[k]), 1.0 - Y[k]);
    drhoY[k] = rho * dY;
  }
}

}  // namespace jh
