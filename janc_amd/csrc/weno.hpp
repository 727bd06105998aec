// WENO5-JS reconstruction of split fluxes at one face.
// Follows solver_2D/reconstruction/WENO5_JS.py:10-117 of the reference:
// smoothness IS_k = 1/4(.)^2 + 13/12(.)^2, alpha_k = C_k/(eps+IS_k)^2 with
// eps = 1e-6, C = (0.1, 0.6, 0.3), candidate stencils (1/3,-7/6,11/6),
// (-1/6,5/6,1/3), (1/3,5/6,-1/6).  The right-biased form is the same formula
// with the stencil mirrored, so one routine serves both.
#pragma once
#include <math.h>

#include "fastmath.hpp"
#include "janc_types.hpp"

namespace jh {

// FP64 form.  With the first differences d and the second differences b
// around f0 (d0 = f0 - fm1, d1 = fp1 - f0, b_k centred on fm1, f0, fp1), the
// WENO5_JS.py:17-33 smoothness indicators and candidates are
//   IS_1 = (b_-1/2 + d0)^2 + 13/12 b_-1^2,   q_1 = f0 + b_-1/3 + d0/2
//   IS_2 = ((d0 + d1)/2)^2 + 13/12 b_0^2,    q_2 = f0 - b_0/6 + d1/2
//   IS_3 = (b_+1/2 - d1)^2 + 13/12 b_+1^2,   q_3 = f0 - b_+1/6 + d1/2
// (a_1 = b_-1 + 2 d0, a_2 = -(d0 + d1), a_3 = b_+1 - 2 d1 in the reference's
// terms).  w_k = alpha_k / sum(alpha) is evaluated as
// C_k prod_{j!=k} e_j / sum_m C_m prod_{j!=m} e_j with e_k = (eps+IS_k)^2,
// the same ratio with one reciprocal instead of six divisions, and the value
// as f0 + sum_k w_k (q_k - f0).  44 FP64 operations per face.
__host__ __device__ __forceinline__ double weno5(double fm2, double fm1, double f0,
                                                 double fp1, double fp2) {
  const double dA = fm1 - fm2, dB = f0 - fm1, dC = fp1 - f0, dD = fp2 - fp1;
  const double bA = dB - dA, bB = dC - dB, bC = dD - dC;
  const double tA = fma((13.0 / 12.0) * bA, bA, 1e-6);
  const double tB = fma((13.0 / 12.0) * bB, bB, 1e-6);
  const double tC = fma((13.0 / 12.0) * bC, bC, 1e-6);
  const double hd0 = 0.5 * dB, hd1 = 0.5 * dC;
  const double A1 = fma(0.5, bA, dB), A2 = hd0 + hd1, A3 = fma(0.5, bC, -dC);
  const double e1 = fma(A1, A1, tA), e2 = fma(A2, A2, tB), e3 = fma(A3, A3, tC);
  const double s1 = e1 * e1, s2 = e2 * e2, s3 = e3 * e3;
  const double P1 = 0.1 * (s2 * s3), P2 = 0.6 * (s1 * s3), P3 = 0.3 * (s1 * s2);
  const double c1 = fma(1.0 / 3.0, bA, hd0);
  const double c2 = fma(-1.0 / 6.0, bB, hd1);
  const double c3 = fma(-1.0 / 6.0, bC, hd1);
  return fma(fma(P1, c1, fma(P2, c2, P3 * c3)), rcp(P1 + P2 + P3), f0);
}

// FP32 form for the mixed-precision step: the FP64 form's differences and
// candidates, with the weights normalised by the smallest e_k = eps + IS_k,
// alpha_k = C_k (e_min / e_k)^2 (ratios <= 1, the smallest weight's ratio 1).
// e_k itself is never squared: in SI units the split energy flux jumps by
// ~1e11 across a detonation front, so e_k reaches ~1e23 and e_k^2 would
// overflow FP32, while a ratio that underflows belongs to a weight that is
// zero at FP32 precision anyway.  ~45 FP32 operations per face.
__host__ __device__ __forceinline__ float weno5(float fm2, float fm1, float f0, float fp1,
                                                float fp2) {
  const float dA = fm1 - fm2, dB = f0 - fm1, dC = fp1 - f0, dD = fp2 - fp1;
  const float bA = dB - dA, bB = dC - dB, bC = dD - dC;
  const float hB = 0.5f * dB, hC = 0.5f * dC;
  const float A1 = fmaf(0.5f, bA, dB), A2 = hB + hC, A3 = fmaf(0.5f, bC, -dC);
  const float e1 = fmaf(A1, A1, fmaf((13.0f / 12.0f) * bA, bA, 1e-6f));
  const float e2 = fmaf(A2, A2, fmaf((13.0f / 12.0f) * bB, bB, 1e-6f));
  const float e3 = fmaf(A3, A3, fmaf((13.0f / 12.0f) * bC, bC, 1e-6f));
  const float em = fminf(e1, fminf(e2, e3));
  const float r1 = em * rcp(e1), r2 = em * rcp(e2), r3 = em * rcp(e3);
  const float a1 = 0.1f * (r1 * r1), a2 = 0.6f * (r2 * r2), a3 = 0.3f * (r3 * r3);
  const float c1 = fmaf(1.0f / 3.0f, bA, hB);
  const float c2 = fmaf(-1.0f / 6.0f, bB, hC);
  const float c3 = fmaf(-1.0f / 6.0f, bC, hC);
  return fmaf(fmaf(a1, c1, fmaf(a2, c2, a3 * c3)), rcp(a1 + a2 + a3), f0);
}

// The FP32 values at K consecutive faces of one line, read from memory (LDS)
// one value per face (f(j) = p[j * s]) and written to (ACC: added into)
// q[i * qs].  Neighbouring faces share the first and second differences and
// the 13/12 b^2 + eps terms, which the run carries; every operation is the
// one the FP32 weno5 performs for that face, so the values equal per-face
// calls bit for bit.  MIRROR = false: out_i = weno5(f(i), .., f(i+4)) (the
// face between f(i+2) and f(i+3)); MIRROR = true: out_i = weno5(f(i+4), ..,
// f(i)) (the face between f(i+1) and f(i+2)).
template <int K, bool MIRROR, bool ACC>
__host__ __device__ __forceinline__ void weno5_run(const float* p, int s, float* q, int qs) {
  // forward differences d_j = f(j+1) - f(j) and b_j = d_j - d_(j-1); the
  // mirrored stencil's differences are their exact negations
  float dB = p[2 * s] - p[s], dC = p[3 * s] - p[2 * s];
  float bA = dB - (p[s] - p[0]), bB = dC - dB;
  float tA = fmaf((13.0f / 12.0f) * bA, bA, 1e-6f), tB = fmaf((13.0f / 12.0f) * bB, bB, 1e-6f);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const float dD = p[(i + 4) * s] - p[(i + 3) * s];
    const float bC = dD - dC, tC = fmaf((13.0f / 12.0f) * bC, bC, 1e-6f);
    float A1, A2, A3, c1, c2, c3, t1, t3, f0;
    if (!MIRROR) {   // f0 = f(i+2): d0 = dB, d1 = dC; b_-1, b_0, b_+1 = bA, bB, bC
      const float hB = 0.5f * dB, hC = 0.5f * dC;
      A1 = fmaf(0.5f, bA, dB); A2 = hB + hC; A3 = fmaf(0.5f, bC, -dC);
      c1 = fmaf(1.0f / 3.0f, bA, hB); c2 = fmaf(-1.0f / 6.0f, bB, hC);
      c3 = fmaf(-1.0f / 6.0f, bC, hC);
      t1 = tA; t3 = tC;
      f0 = p[(i + 2) * s];
    } else {         // f0 = f(i+2): d0 = -dC, d1 = -dB; b_-1, b_0, b_+1 = bC, bB, bA
      const float hB = 0.5f * -dC, hC = 0.5f * -dB;
      A1 = fmaf(0.5f, bC, -dC); A2 = hB + hC; A3 = fmaf(0.5f, bA, dB);
      c1 = fmaf(1.0f / 3.0f, bC, hB); c2 = fmaf(-1.0f / 6.0f, bB, hC);
      c3 = fmaf(-1.0f / 6.0f, bA, hC);
      t1 = tC; t3 = tA;
      f0 = p[(i + 2) * s];
    }
    const float e1 = fmaf(A1, A1, t1), e2 = fmaf(A2, A2, tB), e3 = fmaf(A3, A3, t3);
    const float em = fminf(e1, fminf(e2, e3));
    const float r1 = em * rcp(e1), r2 = em * rcp(e2), r3 = em * rcp(e3);
    const float a1 = 0.1f * (r1 * r1), a2 = 0.6f * (r2 * r2), a3 = 0.3f * (r3 * r3);
    const float v = fmaf(fmaf(a1, c1, fmaf(a2, c2, a3 * c3)), rcp(a1 + a2 + a3), f0);
    if (ACC) q[i * qs] += v;
    else q[i * qs] = v;
    dB = dC; dC = dD;
    bA = bB; tA = tB; bB = bC; tB = tC;
  }
}

}  // namespace jh

namespace jh {

// Value and gradient of the FP64 weno5 with respect to its five arguments
// (the reverse mode of WENO5_JS.py:17-33 as JAX differentiates it):
//   dq/dv_j = sum_k w_k dq_k/dv_j + sum_k kappa_k de_k/dv_j,
//   kappa_k = -2 w_k (q_k - q) / e_k,  e_k = eps + IS_k,
// in weno5's difference form (q_k - q = c_k - sum_m w_m c_m), with the three
// 1/e_k from 1/(e_1 e_2 e_3) and the weights' normaliser from the same
// reciprocal (e^7 stays in FP64 range for any e = eps + IS_k the split
// fluxes produce: 1e-42 at eps, 1e154 at IS = 1e22).  With A_k the halved first-
// derivative terms (IS_k = A_k^2 + 13/12 b_k^2), de_k/dv = 2 A_k dA_k/dv +
// 13/6 b_k db_k/dv, where dA_1 = (1/2, -2, 3/2, 0, 0), dA_2 = (0, -1/2, 0,
// 1/2, 0), dA_3 = (0, 0, 3/2, -2, 1/2) and db_1 = (1, -2, 1, 0, 0), db_2 =
// (0, 1, -2, 1, 0), db_3 = (0, 0, 1, -2, 1).
__host__ __device__ __forceinline__ double weno5_grad(double v0, double v1, double v2, double v3,
                                                      double v4, double (&g)[5]) {
  const double dA = v1 - v0, dB = v2 - v1, dC = v3 - v2, dD = v4 - v3;
  const double bA = dB - dA, bB = dC - dB, bC = dD - dC;
  const double hB = 0.5 * dB, hC = 0.5 * dC;
  const double A1 = fma(0.5, bA, dB), A2 = hB + hC, A3 = fma(0.5, bC, -dC);
  const double e1 = fma(A1, A1, fma((13.0 / 12.0) * bA, bA, 1e-6));
  const double e2 = fma(A2, A2, fma((13.0 / 12.0) * bB, bB, 1e-6));
  const double e3 = fma(A3, A3, fma((13.0 / 12.0) * bC, bC, 1e-6));
  const double E12 = e1 * e2, E13 = e1 * e3, E23 = e2 * e3;
  const double P1 = 0.1 * (E23 * E23), P2 = 0.6 * (E13 * E13), P3 = 0.3 * (E12 * E12);
  // R = 1 / sum(P) and iE = 1 / (e1 e2 e3) from one reciprocal
  const double S = P1 + P2 + P3, E = E12 * e3, X = rcp(E * S);
  const double R = E * X, iE = S * X;
  const double w1 = P1 * R, w2 = P2 * R, w3 = P3 * R;
  const double c1 = fma(1.0 / 3.0, bA, hB);
  const double c2 = fma(-1.0 / 6.0, bB, hC);
  const double c3 = fma(-1.0 / 6.0, bC, hC);
  const double cb = fma(w1, c1, fma(w2, c2, w3 * c3));
  // 2 kappa_k A_k and 13/6 kappa_k b_k
  const double m1 = -2.0 * (w1 * (E23 * iE)) * (c1 - cb);
  const double m2 = -2.0 * (w2 * (E13 * iE)) * (c2 - cb);
  const double m3 = -2.0 * (w3 * (E12 * iE)) * (c3 - cb);
  const double a1 = 2.0 * m1 * A1, a2 = 2.0 * m2 * A2, a3 = 2.0 * m3 * A3;
  const double b1 = (13.0 / 6.0) * m1 * bA, b2 = (13.0 / 6.0) * m2 * bB;
  const double b3 = (13.0 / 6.0) * m3 * bC;
  g[0] = w1 * (1.0 / 3.0) + 0.5 * a1 + b1;
  g[1] = w1 * (-7.0 / 6.0) + w2 * (-1.0 / 6.0) - 2.0 * (a1 + b1) - 0.5 * a2 + b2;
  g[2] = w1 * (11.0 / 6.0) + w2 * (5.0 / 6.0) + w3 * (1.0 / 3.0) + 1.5 * (a1 + a3) + b1 -
         2.0 * b2 + b3;
  g[3] = w2 * (1.0 / 3.0) + w3 * (5.0 / 6.0) + 0.5 * a2 + b2 - 2.0 * (a3 + b3);
  g[4] = w3 * (-1.0 / 6.0) + 0.5 * a3 + b3;
  return v2 + cb;
}

}  // namespace jh
