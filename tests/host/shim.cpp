// Host build of the per-cell device functions (they are __host__ __device__)
// for CPU tests against the oracle: compiled with g++ by tests/test_host_kernels.py.
#include <cmath>
#include <algorithm>
using std::fmax;
using std::fmin;
#include "../../janc_amd/csrc/chem.hpp"
#include "../../janc_amd/csrc/hllc.hpp"
#include "../../janc_amd/csrc/recon.hpp"
#include "../../janc_amd/csrc/thermo.hpp"
#include "../../janc_amd/csrc/weno.hpp"

extern "C" {

// chem_increment per cell: U (12, n), T (n) -> drhoY (8, n)
void host_chem(const double* mc, int n, const double* U, const double* T, double dt,
               double* d_out) {
  const MechConst& c = *reinterpret_cast<const MechConst*>(mc);
  for (int i = 0; i < n; ++i) {
    double rY[JH_NS - 1];
    for (int k = 0; k < JH_NS - 1; ++k) rY[k] = U[(4 + k) * n + i];
    double d[JH_NA];
    jh::chem_increment(c, U[i], rY, T[i], dt, d);
    for (int k = 0; k < JH_NA; ++k) d_out[k * n + i] = d[k];
  }
}

// chem_increment_vjp per cell: cotangent Sb (8, n) -> rho_bar (n), rhoY_bar (8, n), T_bar (n)
void host_chem_vjp(const double* mc, int n, const double* U, const double* T, double dt,
                   const double* Sb, double* rhob, double* rYb, double* Tb) {
  const MechConst& c = *reinterpret_cast<const MechConst*>(mc);
  for (int i = 0; i < n; ++i) {
    double rY[JH_NS - 1], sb[JH_NA], yb[JH_NS - 1], rb, tb;
    for (int k = 0; k < JH_NS - 1; ++k) rY[k] = U[(4 + k) * n + i];
    for (int k = 0; k < JH_NA; ++k) sb[k] = Sb[k * n + i];
    jh::chem_increment_vjp(c, U[i], rY, T[i], dt, sb, rb, yb, tb);
    rhob[i] = rb;
    for (int k = 0; k < JH_NS - 1; ++k) rYb[k * n + i] = yb[k];
    Tb[i] = tb;
  }
}

// getT_vjp per cell at the converged T: (gamma_bar, T_bar) -> (e_bar, Y_bar (8))
void host_getT_vjp(const double* mc, int n, const double* Y8, const double* T, const double* gb,
                   const double* Tbar, double* eb, double* Yb) {
  const MechConst& c = *reinterpret_cast<const MechConst*>(mc);
  for (int i = 0; i < n; ++i) {
    double Y[JH_NS];
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = Y8[k * n + i];
    const double R = jh::fill_Y_get_R(c, Y);
    double e, yb[JH_NS - 1];
    jh::getT_vjp(c, Y, R, T[i], gb[i], Tbar[i], e, yb);
    eb[i] = e;
    for (int k = 0; k < JH_NS - 1; ++k) Yb[k * n + i] = yb[k];
  }
}

// Halley solve per cell (update_aux): Y8 (8, n), e, initial T -> gamma, T
void host_halley(const double* mc, int n, const double* Y8, const double* e, const double* T0,
                 double* gamma, double* T) {
  const MechConst& c = *reinterpret_cast<const MechConst*>(mc);
  for (int i = 0; i < n; ++i) {
    double Y[JH_NS];
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = Y8[k * n + i];
    const double R = jh::fill_Y_get_R(c, Y);
    double t = T0[i], g;
    jh::halley_T(c, Y, R, e[i], t, g);
    gamma[i] = g;
    T[i] = t;
  }
}

// Halley solve per cell with the grid-wide-stop record: the per-cell (gamma,
// T), the update count n, the state after one more update (g1, T1) and the
// flags conv / ok1 (HalleyMore)
void host_halley_more(const double* mc, int n, const double* Y8, const double* e,
                      const double* T0, double* gamma, double* T, int* nit, double* g1,
                      double* T1, int* conv, int* ok1) {
  const MechConst& c = *reinterpret_cast<const MechConst*>(mc);
  for (int i = 0; i < n; ++i) {
    double Y[JH_NS];
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = Y8[k * n + i];
    const double R = jh::fill_Y_get_R(c, Y);
    double t = T0[i], g;
    jh::HalleyMore hm;
    const bool cv = jh::halley_T<MechConst, true>(c, Y, R, e[i], t, g, &hm);
    gamma[i] = g;
    T[i] = t;
    nit[i] = hm.n;
    g1[i] = hm.g1;
    T1[i] = hm.T1;
    conv[i] = cv && hm.conv;
    ok1[i] = hm.ok1;
  }
}

// lu_solve on one N x N system (row-major A, b <- x); returns the pivot check
int host_lu_solve8(double* A, double* b) {
  double M[JH_NA][JH_NA], v[JH_NA];
  for (int r = 0; r < JH_NA; ++r) {
    v[r] = b[r];
    for (int j = 0; j < JH_NA; ++j) M[r][j] = A[r * JH_NA + j];
  }
  const bool ok = jh::lu_solve<JH_NA>(M, v);
  for (int r = 0; r < JH_NA; ++r) b[r] = v[r];
  return ok ? 1 : 0;
}

// K = 4 FP32 faces from one 8-value window: L[i] left-, R[i] right-biased
void host_weno_run_f32(const float* f, float* L, float* R) {
  jh::weno5_run<4, false, false>(f, 1, L, 1);
  jh::weno5_run<4, true, false>(f, 1, R, 1);
}

double host_weno(const double* v) { return jh::weno5(v[0], v[1], v[2], v[3], v[4]); }
float host_weno_f32(const float* v) { return jh::weno5(v[0], v[1], v[2], v[3], v[4]); }
double host_weno_grad(const double* v, double* g) {
  double gg[5];
  const double q = jh::weno5_grad(v[0], v[1], v[2], v[3], v[4], gg);
  for (int k = 0; k < 5; ++k) g[k] = gg[k];
  return q;
}

// HLLC flux per face: qL, qR (12, n) -> F (12, n); dir 0 = x, 1 = y
void host_hllc(const double* mc, int n, const double* qL, const double* qR, int dir,
               double* F) {
  const MechConst& c = *reinterpret_cast<const MechConst*>(mc);
  for (int i = 0; i < n; ++i) {
    double l[jh::kNQ], r[jh::kNQ], f[jh::kNQ];
    for (int v = 0; v < jh::kNQ; ++v) {
      l[v] = qL[v * n + i];
      r[v] = qR[v * n + i];
    }
    if (dir == 0) jh::hllc_flux<double, 0>(c, c, l, r, f);
    else jh::hllc_flux<double, 1>(c, c, l, r, f);
    for (int v = 0; v < jh::kNQ; ++v) F[v * n + i] = f[v];
  }
}

// face values of one stencil v0..v5: out = (L, R); recon 0 = WENO5, 1 = MUSCL
void host_recon(int recon, const double* v, double* out) {
  if (recon == 1) {
    out[0] = jh::recon_L<jh::RECON_MUSCL>(v[0], v[1], v[2], v[3], v[4]);
    out[1] = jh::recon_R<jh::RECON_MUSCL>(v[1], v[2], v[3], v[4], v[5]);
  } else {
    out[0] = jh::recon_L<jh::RECON_WENO5>(v[0], v[1], v[2], v[3], v[4]);
    out[1] = jh::recon_R<jh::RECON_WENO5>(v[1], v[2], v[3], v[4], v[5]);
  }
}
}
