// Point-implicit detailed chemistry for one cell (FP64).
// Follows model/reaction_model.py:78-193 of the reference:
//   X = rho*Y/M, (I/dt - dw/d(rhoY)) x = w, dY = clip(x/rho, -Y, 1-Y).
#pragma once
#include <math.h>

#include "fastmath.hpp"
#include "janc_types.hpp"
#include "mech_gen.inc"

namespace jh {

// The tables gibbs() reads.  MechConst has the same members; the chemistry
// kernel reads a copy in LDS (the range index is per lane).
struct GibbsTables {
  double hc[2][JH_NA][6];
  double s[2][JH_NA][5];
  double logc[2][JH_NA];
  double Tcr;
  double T0;
};

// g_k = s_k/R - h_k/(R T) for the active species (get_gibbs, :78-89).
template <typename Tab>
__host__ __device__ __forceinline__ void gibbs(const Tab& c, double T, double invT,
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
// Returns false when a pivot is zero or not finite (getrf's info > 0; the
// reference's solve returns non-finite values then, and so does this one).
template <int N>
__host__ __device__ __forceinline__ bool lu_solve(double (&A)[N][N], double (&b)[N]) {
  bool ok = true;
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
    // the predicated swap is skipped when no lane of the wave pivots
#if defined(__HIP_DEVICE_COMPILE__)
    const bool any_swap = __any(p != col);
#else
    const bool any_swap = p != col;
#endif
    if (any_swap) {
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
    }
    ok = ok && (amax > 0.0) && (amax < INFINITY);
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
  return ok;
}

// Increment of rho*Y_k for the active species (detailed_reaction, :183-193).
// ld.rho(), ld(k): the state's rho and rho*Y_k for the NS-1 transported
// species (register values, or loads from the state: they are read again
// after the solve, behind a compiler barrier, so the state is not held in
// registers across the Jacobian and its factorisation); gt: the Gibbs tables (c
// itself, or its copy in LDS); JT: the Jacobian's storage type (float in the
// mixed-precision form: assembled in FP32 from FP64 terms, factorised and
// solved in FP64).  Returns lu_solve's pivot check.  clip (optional): bit k
// set where the clip of species k takes its lower bound -Y_k, bit JH_NA + k
// where it takes 1 - Y_k.  branches (optional): take the clip's branches
// from these bits instead of comparing (a diagnostic that keeps a perturbed
// state on the branches of another).
template <typename JT = double, typename GTab, typename LoadRY>
__host__ __device__ __forceinline__ bool chem_increment_ld(const MechConst& c, const GTab& gt,
                                                           const LoadRY& ld, double T, double dt,
                                                           double (&drhoY)[JH_NA],
                                                           unsigned int* clip = nullptr,
                                                           const unsigned int* branches = nullptr) {
  double X[JH_NS];
  {
    const double rho = ld.rho(), irho = rcp(rho);
    double Y[JH_NS];
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = ld(k) * irho;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) sum += Y[k];
    Y[JH_NS - 1] = 1.0 - sum;
#pragma unroll
    for (int k = 0; k < JH_NS; ++k) X[k] = (rho * Y[k]) * c.R[k];
  }
  double Xc[JH_NA], invMX[JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    Xc[k] = fmax(X[k], 1e-50);
    invMX[k] = rcp(c.M[k] * Xc[k]);
  }
  const double invT = rcp(T);
  double g[JH_NA];
  gibbs(gt, T, invT, g);
  double eg[JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) eg[k] = exp(g[k]);
  double w[JH_NA];
  JT J[JH_NA][JH_NA];
  mech_rates_jacobian<JT>(c, T, invT, log(T), X, Xc, invMX, eg, w, J);
  const double idt = rcp(dt);
  double A[JH_NA][JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
#pragma unroll
    for (int n = 0; n < JH_NA; ++n) A[k][n] = ((k == n) ? idt : 0.0) - (double)J[k][n];
  }
  const bool ok = lu_solve<JH_NA>(A, w);
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ volatile("" ::: "memory");   // ld below reloads
#endif
  const double rho = ld.rho(), irho = rcp(rho);
  unsigned int cm = 0u;
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    const double Yk = ld(k) * irho;
    const double x = w[k] * irho;
    if (clip) cm |= (x < -Yk ? 1u << k : 0u) | (x > 1.0 - Yk ? 1u << (JH_NA + k) : 0u);
    double dY;
    if (branches) {
      const unsigned int b = *branches;
      dY = (b >> k) & 1u ? -Yk : (b >> (JH_NA + k)) & 1u ? 1.0 - Yk : x;
    } else {
      dY = fmin(fmax(x, -Yk), 1.0 - Yk);
    }
    drhoY[k] = rho * dY;
  }
  if (clip) *clip = cm;
  return ok;
}

// chem_increment_ld's state from register values.
struct RegState {
  double r;
  const double (&y)[JH_NS - 1];
  __host__ __device__ double rho() const { return r; }
  __host__ __device__ double operator()(int k) const { return y[k]; }
};

template <typename JT = double, typename GTab>
__host__ __device__ __forceinline__ bool chem_increment(const MechConst& c, const GTab& gt,
                                                        double rho,
                                                        const double (&rhoY_in)[JH_NS - 1],
                                                        double T, double dt,
                                                        double (&drhoY)[JH_NA],
                                                        unsigned int* clip = nullptr,
                                                        const unsigned int* branches = nullptr) {
  return chem_increment_ld<JT>(c, gt, RegState{rho, rhoY_in}, T, dt, drhoY, clip, branches);
}

__host__ __device__ __forceinline__ bool chem_increment(const MechConst& c, double rho,
                                                        const double (&rhoY_in)[JH_NS - 1],
                                                        double T, double dt,
                                                        double (&drhoY)[JH_NA]) {
  return chem_increment(c, c, rho, rhoY_in, T, dt, drhoY);
}

}  // namespace jh

namespace jh {

// ---------------------------------------------------------------- adjoint
// LU with partial pivoting kept for reuse (LAPACK getrf layout: unit-lower
// multipliers below the diagonal, U on and above it, piv[col] = pivot row).
template <int N>
__host__ __device__ __forceinline__ void lu_factor(double (&A)[N][N], int (&piv)[N]) {
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
    piv[col] = p;
#pragma unroll
    for (int r = col + 1; r < N; ++r) {
      const bool s = (p == r);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double t = A[col][j];
        A[col][j] = s ? A[r][j] : t;
        A[r][j] = s ? t : A[r][j];
      }
    }
    const double inv = rcp(A[col][col]);
#pragma unroll
    for (int r = col + 1; r < N; ++r) {
      const double l = A[r][col] * inv;
      A[r][col] = l;
#pragma unroll
      for (int j = col + 1; j < N; ++j) A[r][j] -= l * A[col][j];
    }
  }
}

// Swap b[col] <-> b[piv[col]] with compile-time indices.
template <int N>
__host__ __device__ __forceinline__ void swap_rows(double (&b)[N], int col, int p) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (r <= col) continue;
    const bool s = (p == r);
    const double t = b[col];
    b[col] = s ? b[r] : t;
    b[r] = s ? t : b[r];
  }
}

// A x = b with the factors of lu_factor (b <- x): P b (every row swap, in
// order, before the substitutions: the stored multipliers are in the final
// row order), L y = P b, U x = y.
template <int N>
__host__ __device__ __forceinline__ void lu_apply(const double (&A)[N][N], const int (&piv)[N],
                                                  double (&b)[N]) {
#pragma unroll
  for (int col = 0; col < N; ++col) swap_rows<N>(b, col, piv[col]);
#pragma unroll
  for (int col = 0; col < N; ++col) {
#pragma unroll
    for (int r = col + 1; r < N; ++r) b[r] -= A[r][col] * b[col];
  }
#pragma unroll
  for (int r = N - 1; r >= 0; --r) {
    double s = b[r];
#pragma unroll
    for (int j = r + 1; j < N; ++j) s -= A[r][j] * b[j];
    b[r] = s * rcp(A[r][r]);
  }
}

// A^T y = b with the factors of lu_factor (b <- y):  U^T z = b, L^T w = z,
// y = P^T w (the row swaps undone in reverse order).
template <int N>
__host__ __device__ __forceinline__ void lu_apply_T(const double (&A)[N][N], const int (&piv)[N],
                                                    double (&b)[N]) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    double s = b[r];
#pragma unroll
    for (int j = 0; j < r; ++j) s -= A[j][r] * b[j];
    b[r] = s * rcp(A[r][r]);
  }
#pragma unroll
  for (int r = N - 1; r >= 0; --r) {
    double s = b[r];
#pragma unroll
    for (int j = r + 1; j < N; ++j) s -= A[j][r] * b[j];
    b[r] = s;
  }
#pragma unroll
  for (int col = N - 1; col >= 0; --col) swap_rows<N>(b, col, piv[col]);
}

// Gradient of one branch of the reference's jnp.clip = min(max(v, lo), hi)
// with JAX's tie rule (a tie splits the cotangent evenly).
__host__ __device__ __forceinline__ void clip_vjp(double v, double lo, double hi, double gout,
                                                  double& gv, double& glo, double& ghi) {
  const double m = fmax(v, lo);
  const double gm = m < hi ? gout : (m == hi ? 0.5 * gout : 0.0);
  ghi = hi < m ? gout : (m == hi ? 0.5 * gout : 0.0);
  gv = v > lo ? gm : (v == lo ? 0.5 * gm : 0.0);
  glo = lo > v ? gm : (v == lo ? 0.5 * gm : 0.0);
}

// Reverse of chem_increment: cotangent Sb of drhoY -> cotangents of rho,
// rho*Y_k (the state's species densities) and T.  The forward quantities are
// recomputed; the solve is differentiated as lambda = A^-T x_bar,
// w_bar = lambda, J_bar = lambda x^T (model/reaction_model.py:147-193 under
// JAX's reverse mode).  Two halves, each with its own register budget:
// chem_vjp_solve (Jacobian, factorisation, x = A^-1 w and lambda) and
// chem_vjp_rates (the clip's and the rates' reverse given x and lambda).

// The species state shared by both halves: Y (all species), X = rho Y / M,
// Xc = max(X, 1e-50), invMX = 1 / (M Xc).
template <typename LD>
__host__ __device__ __forceinline__ void chem_vjp_state(const MechConst& c, const LD& ld,
                                                        double (&Y)[JH_NS], double (&X)[JH_NS],
                                                        double (&Xc)[JH_NA],
                                                        double (&invMX)[JH_NA]) {
  const double rho = ld.rho(), irho = rcp(rho);
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Y[k] = ld(k) * irho;
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) sum += Y[k];
  Y[JH_NS - 1] = 1.0 - sum;
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) X[k] = (rho * Y[k]) * c.R[k];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    Xc[k] = fmax(X[k], 1e-50);
    invMX[k] = rcp(c.M[k] * Xc[k]);
  }
}

// Gibbs terms at T: gp_k = dg_k/dT (= h_k / T^2) and exp(g_k); gt: the
// Gibbs tables (MechConst, or its copy in LDS: the range index is per lane).
template <typename Tab>
__host__ __device__ __forceinline__ void chem_vjp_gibbs(const Tab& c, double T, double invT,
                                                        double (&gp)[JH_NA],
                                                        double (&eg)[JH_NA]) {
  const int rr = T < c.Tcr ? 0 : 1;
  const double lT = log(c.T0 * T);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    double h = c.hc[rr][k][0];
#pragma unroll
    for (int j = 1; j < 6; ++j) h = h * T + c.hc[rr][k][j];
    double s = c.s[rr][k][0];
#pragma unroll
    for (int j = 1; j < 5; ++j) s = s * T + c.s[rr][k][j];
    s = s + c.logc[rr][k] * lT;
    const double g = s - h * invT;
    gp[k] = h * invT * invT;
    eg[k] = exp(g);
  }
}

// The state of both halves: ld.rho(), ld(k) = rho*Y_k (k < NS-1) and
// ld.sb(k) = the cotangent of drhoY_k.  Each half reads Y and the
// cotangents where it needs them and again after its long middle part,
// behind a compiler barrier, instead of holding them in registers.
// JH_READY(v): v is computed here (an empty asm that takes v in a
// register), so work whose results are tied this way stays ahead of a
// later JH_REREAD instead of being sunk past it.
#if defined(__HIP_DEVICE_COMPILE__)
#define JH_REREAD() __asm__ volatile("" ::: "memory")
#define JH_READY(v) __asm__ volatile("" : "+v"(v))
#else
#define JH_REREAD() ((void)0)
#define JH_READY(v) ((void)0)
#endif

// x = A^-1 w (the increment before the clip) and lambda = A^-T x_bar, x_bar
// the clip's cotangent of x.
template <typename LD, typename GTab>
__host__ __device__ __forceinline__ void chem_vjp_solve(const MechConst& c, const GTab& gt,
                                                        const LD& ld, double T, double dt,
                                                        double (&x)[JH_NA],
                                                        double (&lam)[JH_NA]) {
  double X[JH_NS], Xc[JH_NA], invMX[JH_NA];
  {
    double Y[JH_NS];
    chem_vjp_state(c, ld, Y, X, Xc, invMX);
  }
  const double invT = rcp(T), logT = log(T);
  double gp[JH_NA], eg[JH_NA];
  chem_vjp_gibbs(gt, T, invT, gp, eg);
  double J[JH_NA][JH_NA];
  mech_rates_jacobian(c, T, invT, logT, X, Xc, invMX, eg, x, J);
  const double idt = rcp(dt);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
#pragma unroll
    for (int n = 0; n < JH_NA; ++n) J[k][n] = ((k == n) ? idt : 0.0) - J[k][n];
  }
  int piv[JH_NA];
  lu_factor<JH_NA>(J, piv);
  lu_apply<JH_NA>(J, piv, x);          // x <- A^-1 w
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) JH_READY(x[k]);
  JH_REREAD();
  const double rho = ld.rho(), irho = rcp(rho);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {    // S = rho * clip(x/rho, -Y, 1-Y)
    const double Yk = ld(k) * irho;
    const double v = x[k] * irho;
    double gv, glo, ghi;
    clip_vjp(v, -Yk, 1.0 - Yk, ld.sb(k) * rho, gv, glo, ghi);
    lam[k] = gv * irho;
  }
  lu_apply_T<JH_NA>(J, piv, lam);      // lam <- A^-T x_bar
}

// The rest of the reverse given x and lambda: the rates' reverse (w_bar =
// lambda, J_bar = lambda x^T), the clip's cotangents of rho and Y, and the
// chain through X = rho Y / M.
template <typename LD, typename GTab>
__host__ __device__ __forceinline__ void chem_vjp_rates(const MechConst& c, const GTab& gt,
                                                        const LD& ld,
                                                        double T, const double (&x)[JH_NA],
                                                        const double (&lam)[JH_NA],
                                                        double& rhob,
                                                        double (&rYb)[JH_NS - 1],
                                                        double& Tb) {
  double X[JH_NS], Xc[JH_NA], invMX[JH_NA];
  {
    double Y[JH_NS];
    chem_vjp_state(c, ld, Y, X, Xc, invMX);
  }
  const double invT = rcp(T), logT = log(T);
  double gp[JH_NA], eg[JH_NA];
  chem_vjp_gibbs(gt, T, invT, gp, eg);
  double Xb[JH_NS], Xcb[JH_NA], iMXb[JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) Xb[k] = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    Xcb[k] = 0.0;
    iMXb[k] = 0.0;
  }
  Tb = 0.0;
  mech_rates_vjp(c, T, invT, logT, X, Xc, invMX, eg, gp, lam, x, Xb, Xcb, iMXb, Tb);
#pragma unroll
  for (int n = 0; n < JH_NA; ++n) {
    Xcb[n] -= iMXb[n] * invMX[n] * invMX[n] * c.M[n];
    JH_READY(Xcb[n]);
  }
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) JH_READY(Xb[k]);
  JH_READY(Tb);
  JH_REREAD();
  double Y[JH_NS];
  chem_vjp_state(c, ld, Y, X, Xc, invMX);
  const double rho = ld.rho(), irho = rcp(rho);
  double Yb[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) Yb[k] = 0.0;
  rhob = 0.0;
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    const double v = x[k] * irho, lo = -Y[k], hi = 1.0 - Y[k];
    const double out = fmin(fmax(v, lo), hi);
    const double sb = ld.sb(k);
    rhob += sb * out;
    double gv, glo, ghi;
    clip_vjp(v, lo, hi, sb * rho, gv, glo, ghi);
    Yb[k] -= glo + ghi;
    rhob -= gv * x[k] * irho * irho;
  }
#pragma unroll
  for (int n = 0; n < JH_NA; ++n)
    Xb[n] += X[n] > 1e-50 ? Xcb[n] : (X[n] == 1e-50 ? 0.5 * Xcb[n] : 0.0);
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) {
    Yb[k] += Xb[k] * rho * c.R[k];
    rhob += Xb[k] * Y[k] * c.R[k];
  }
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) {
    const double yk = Yb[k] - Yb[JH_NS - 1];
    rYb[k] = yk * irho;
    rhob -= yk * Y[k] * irho;
  }
}

// chem_vjp_*'s state from register values.
struct RegStateVjp {
  double r;
  const double (&y)[JH_NS - 1];
  const double (&s)[JH_NA];
  __host__ __device__ double rho() const { return r; }
  __host__ __device__ double operator()(int k) const { return y[k]; }
  __host__ __device__ double sb(int k) const { return s[k]; }
};

__host__ __device__ __forceinline__ void chem_increment_vjp(
    const MechConst& c, double rho, const double (&rhoY_in)[JH_NS - 1], double T, double dt,
    const double (&Sb)[JH_NA], double& rhob, double (&rYb)[JH_NS - 1], double& Tb) {
  double x[JH_NA], lam[JH_NA];
  const RegStateVjp ld{rho, rhoY_in, Sb};
  chem_vjp_solve(c, c, ld, T, dt, x, lam);
  chem_vjp_rates(c, c, ld, T, x, lam, rhob, rYb, Tb);
}

}  // namespace jh
