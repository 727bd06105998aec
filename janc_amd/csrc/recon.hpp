// Interface reconstructions selectable by the flux configuration's
// 'interface_reconstruction' key (solver_2D/reconstruction/__init__.py:3-21):
// WENO5-JS (weno.hpp) and MUSCL with the minmod limiter
// (solver_2D/reconstruction/MUSCL.py:1-65, k = 1/3, beta = 1).
//
// Both take the six values v0..v5 of one face's stencil along the face
// normal, v2 | v3 being the two cells that share the face, and return the
// left-biased (L) and right-biased (R) face values.
#pragma once
#include <math.h>

#include "weno.hpp"

namespace jh {

enum Recon : int { RECON_WENO5 = 0, RECON_MUSCL = 1 };

// min_mod_limiter (MUSCL.py:4-7): sign(x) * max(0, min(x sign(y), y sign(x))).
template <typename W>
__host__ __device__ __forceinline__ W sgn(W x) {
  return x > (W)0 ? (W)1 : (x < (W)0 ? (W)-1 : (W)0);
}

template <typename W>
__host__ __device__ __forceinline__ W min_mod(W x, W y) {
  const W m = fmin(x * sgn(y), y * sgn(x));
  return sgn(x) * fmax((W)0, m);
}

// MUSCL face values (MUSCL.py:9-21 left, :38-50 right).  The left value
// extrapolates the centre cell v by its limited slope, from neighbours vm, vp;
// the right value of the same face is taken from the next cell the other way.
template <typename W>
__host__ __device__ __forceinline__ W muscl_L(W vm, W v, W vp) {
  // (1 - k) and (1 + k) with k = 1/3, rounded as the reference forms them
  constexpr W km = (W)(1.0 - 1.0 / 3.0), kp = (W)(1.0 + 1.0 / 3.0), beta = (W)1.0;
  const W dm = v - vm, dp = vp - v;
  const W fw = min_mod(dp, beta * dm), bw = min_mod(dm, beta * dp);
  return v + (W)0.25 * (km * bw + kp * fw);
}

template <typename W>
__host__ __device__ __forceinline__ W muscl_R(W vm, W v, W vp) {
  constexpr W km = (W)(1.0 - 1.0 / 3.0), kp = (W)(1.0 + 1.0 / 3.0), beta = (W)1.0;
  const W dm = v - vm, dp = vp - v;
  const W fw = min_mod(dp, beta * dm), bw = min_mod(dm, beta * dp);
  return v - (W)0.25 * (km * fw + kp * bw);
}

// Left-biased value at the face between v2 and v3 from v0..v4, and the
// right-biased value from v1..v5.
template <int RECON, typename W>
__host__ __device__ __forceinline__ W recon_L(W v0, W v1, W v2, W v3, W v4) {
  if (RECON == RECON_MUSCL) return muscl_L(v1, v2, v3);
  return weno5(v0, v1, v2, v3, v4);
}

template <int RECON, typename W>
__host__ __device__ __forceinline__ W recon_R(W v1, W v2, W v3, W v4, W v5) {
  if (RECON == RECON_MUSCL) return muscl_R(v2, v3, v4);
  return weno5(v5, v4, v3, v2, v1);
}

}  // namespace jh
