// Reverse-mode (VJP) kernels of one fixed-dt step, FP64.
//
// Semantics are those of JAX's reverse mode applied to the reference
// (src/janc_v2), with get_T's custom VJP (thermo_model.py:159-188): zero
// cotangent for the Halley initial guess, the LF maxima's cotangent shared
// evenly among tied maximisers of the padded field, the clip's cotangent
// routed to its active branch.
//
// Data path of the flux VJP (one RK stage): the stage input is padded with
// its ghost cells (pad_kernel); per x- and y-face, the WENO gradients times
// the face cotangent are written as ten coefficients per variable
// (face_vjp_kernel); per padded cell these are gathered into H+-_bar and
// pushed through the split flux and the primitive map (cell_vjp_kernel); the
// LF maxima's cotangent goes to the argmax cells (alpha_*); ghost-cell
// cotangents are folded back onto their source cells (unpad_kernel).
#pragma once
#include <hip/hip_runtime.h>

#include "stage.hpp"

namespace jh {

struct AdjArgs {
  int nx, ny;
  double dx, dy, dt;
  const MechConst* __restrict__ mc;
  int bc[4];
};

__device__ __forceinline__ int wrap_x(int i, int n) { return i < 0 ? n - 1 + i : 1 + (i - n); }

// Source interior cell of padded cell (P, Q) and the sign flips of the two
// momenta (bit 0: rho*u, bit 1: rho*v) -- boundary_padding.py:20-25 then the
// BC kinds of boundary_conditions_2D/*.py.  Returns false for a kind without
// a linear map (pressure outlet, user strips).
__device__ __forceinline__ bool pad_src(const AdjArgs& a, int P, int Q, int& sx, int& sy,
                                        int& flip) {
  const int nx = a.nx, ny = a.ny, i = P - 3, j = Q - 3;
  const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
  flip = 0;
  if (xin && yin) {
    sx = i;
    sy = j;
    return true;
  }
  if (!xin && !yin) {     // corners keep the periodic pre-pad
    sx = wrap_x(i, nx);
    sy = wrap_x(j, ny);
    return true;
  }
  const bool xs = !xin;
  const int side = xs ? (i < 0 ? JH_LEFT : JH_RIGHT) : (j < 0 ? JH_BOTTOM : JH_TOP);
  const int N = xs ? nx : ny;
  const int g = xs ? i : j;          // index along the normal, outside [0, N)
  const bool low = g < 0;
  int s;
  switch (a.bc[side]) {
    case JH_BC_PERIODIC: s = wrap_x(g, N); break;
    case JH_BC_ZERO_GRADIENT: s = low ? -1 - g : 2 * N - 1 - g; break;
    case JH_BC_NO_SLIP_WALL:
      s = low ? -1 - g : 2 * N - 1 - g;
      flip = 3;
      break;
    case JH_BC_SLIP_WALL:
      s = low ? 0 : N - 1;
      flip = xs ? 1 : 2;
      break;
    default: return false;
  }
  if (xs) {
    sx = s;
    sy = j;
  } else {
    sx = i;
    sy = s;
  }
  return true;
}

// Ghost-padded copy of (U, aux): (12, nx+6, ny+6) and (2, nx+6, ny+6).
__global__ __launch_bounds__(256) void pad_kernel(AdjArgs a, const double* __restrict__ U,
                                                  const double* __restrict__ aux,
                                                  double* __restrict__ Up,
                                                  double* __restrict__ ap) {
  const int PX = a.nx + 6, PY = a.ny + 6, np = PX * PY, n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  const int P = t / PY, Q = t - P * PY;
  int sx, sy, flip;
  if (!pad_src(a, P, Q, sx, sy, flip)) return;
  const int s = sx * a.ny + sy;
#pragma unroll
  for (int v = 0; v < NU; ++v) {
    double val = U[v * n + s];
    if ((v == 1 && (flip & 1)) || (v == 2 && (flip & 2))) val = -val;
    Up[v * np + t] = val;
  }
  ap[t] = aux[s];
  ap[np + t] = aux[n + s];
}

struct PrimD {
  double rho, ir, u, v, p, R, T, gamma;
  double Y[JH_NS];
};

__device__ __forceinline__ PrimD prim_at(const MechConst& c, const double* __restrict__ Up,
                                         const double* __restrict__ ap, int np, int t) {
  PrimD q;
  q.rho = Up[t];
  q.ir = rcp(q.rho);
  q.u = Up[np + t] * q.ir;
  q.v = Up[2 * np + t] * q.ir;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) q.Y[k] = Up[(4 + k) * np + t] * q.ir;
  q.R = fill_Y_get_R(c, q.Y);
  q.gamma = ap[t];
  q.T = ap[np + t];
  q.p = q.rho * q.R * q.T;
  return q;
}

// Physical flux of variable cv in direction dir at a cell (LF.py:12).
__device__ __forceinline__ double phys_flux_d(int cv, int dir, const PrimD& q, double ucv) {
  const double un = dir == 0 ? q.u : q.v;
  switch (cv) {
    case 0: return q.rho * un;
    case 1: return dir == 0 ? q.rho * (q.u * q.u) + q.p : (q.rho * q.u) * q.v;
    case 2: return dir == 0 ? (q.rho * q.u) * q.v : q.rho * (q.v * q.v) + q.p;
    case 3: return un * (ucv + q.p);
    default: return un * ucv;
  }
}

// Face cotangent coefficients.  Faces along DIR (0: x-faces (nx+1) x ny,
// 1: y-faces nx x (ny+1)); coefficient plane (cv*10 + m): m < 5 left-biased
// WENO of H+ at stencil slot m, m >= 5 right-biased WENO of H- at slot m-5,
// each multiplied by the face cotangent F_bar = dt/dh (rbar_f - rbar_{f-1}).
template <int DIR>
__global__ __launch_bounds__(256) void face_vjp_kernel(AdjArgs a, const double* __restrict__ Up,
                                                       const double* __restrict__ ap,
                                                       const double* __restrict__ rbar,
                                                       const unsigned long long* alpha_bits,
                                                       double* __restrict__ coef) {
  const MechConst& c = *a.mc;
  const int nx = a.nx, ny = a.ny, n = nx * ny, PY = ny + 6, np = (nx + 6) * PY;
  const int nf = DIR == 0 ? (nx + 1) * ny : nx * (ny + 1);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf) return;
  int f, row;   // face index along DIR and the cross index
  if (DIR == 0) {
    f = t / ny;
    row = t - f * ny;
  } else {
    row = t / (ny + 1);
    f = t - row * (ny + 1);
  }
  const double alpha = bitsd(alpha_bits[DIR]);
  // six cells of the stencil: interior index f-3 .. f+2 along DIR
  PrimD q[6];
  int pt[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    pt[m] = DIR == 0 ? (f + m) * PY + (row + 3) : (row + 3) * PY + (f + m);
    q[m] = prim_at(c, Up, ap, np, pt[m]);
  }
  const double dtdh = a.dt / (DIR == 0 ? a.dx : a.dy);
  for (int cv = 0; cv < NU; ++cv) {
    // cotangent of rhs at the cells either side of the face
    const int cl = DIR == 0 ? (f - 1) * ny + row : row * ny + (f - 1);
    const int cr = DIR == 0 ? f * ny + row : row * ny + f;
    const double rl = (f - 1 >= 0) ? rbar[cv * n + cl] : 0.0;
    const double rr = (f < (DIR == 0 ? nx : ny)) ? rbar[cv * n + cr] : 0.0;
    const double Fb = dtdh * (rr - rl);
    double hp[6], hm[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double ucv = Up[cv * np + pt[m]];
      const double F = phys_flux_d(cv, DIR, q[m], ucv);
      hp[m] = 0.5 * (F + alpha * ucv);
      hm[m] = 0.5 * (F - alpha * ucv);
    }
    double gL[5], gR[5];
    weno5_grad(hp[0], hp[1], hp[2], hp[3], hp[4], gL);
    weno5_grad(hm[5], hm[4], hm[3], hm[2], hm[1], gR);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      coef[(size_t)(cv * 10 + m) * nf + t] = Fb * gL[m];
      coef[(size_t)(cv * 10 + 5 + m) * nf + t] = Fb * gR[m];
    }
  }
}

// Cotangents of the primitive map at one cell -> cotangents of U and aux.
// Inputs: rho_b, u_b, v_b, p_b, a_b (sound speed), Y_b[0:NS-1] and direct
// U_b contributions; see aux_func.py:39-51 for the forward map.
__device__ __forceinline__ void prim_to_U_vjp(const MechConst& c, const PrimD& q, double rhob,
                                              double ub, double vb, double pb, double ab,
                                              double (&Yb)[JH_NS - 1], double (&Ub)[NU],
                                              double& Tb, double& gb) {
  // a = sqrt(gamma p / rho)
  if (ab != 0.0) {
    const double snd = sqrt(q.gamma * q.p * q.ir);
    const double k = 0.5 * ab * rcp(snd);
    gb += k * q.p * q.ir;
    pb += k * q.gamma * q.ir;
    rhob -= k * q.gamma * q.p * q.ir * q.ir;
  }
  // p = rho R T, R = sum R_k Y_k (Y_N2 = 1 - sum)
  rhob += pb * q.R * q.T;
  Tb += pb * q.rho * q.R;
  const double Rb = pb * q.rho * q.T;
  constexpr int L = JH_NS - 1;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) {
    const double yb = Yb[k] + Rb * (c.R[k] - c.R[L]);
    Ub[4 + k] += yb * q.ir;
    rhob -= yb * q.Y[k] * q.ir;
  }
  Ub[1] += ub * q.ir;
  rhob -= ub * q.u * q.ir;
  Ub[2] += vb * q.ir;
  rhob -= vb * q.v * q.ir;
  Ub[0] += rhob;
}

// Per padded cell: gather H+-_bar from the face coefficients, apply the split
// flux and primitive VJPs; also this cell's share of the LF-alpha cotangent
// sums (per-block partials, reduced in a fixed order afterwards).
__global__ __launch_bounds__(256) void cell_vjp_kernel(AdjArgs a, const double* __restrict__ Up,
                                                       const double* __restrict__ ap,
                                                       const double* __restrict__ cx,
                                                       const double* __restrict__ cy,
                                                       const unsigned long long* alpha_bits,
                                                       double* __restrict__ Ubp,
                                                       double* __restrict__ abp,
                                                       double* __restrict__ part) {
  __shared__ double red[2][4];
  const MechConst& c = *a.mc;
  const int nx = a.nx, ny = a.ny, PY = ny + 6, np = (nx + 6) * PY;
  const int nfx = (nx + 1) * ny, nfy = nx * (ny + 1);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double sx_part = 0.0, sy_part = 0.0;
  if (t < np) {
    const int P = t / PY, Q = t - P * PY, i = P - 3, j = Q - 3;
    const bool xrow = j >= 0 && j < ny, ycol = i >= 0 && i < nx;
    int sxx, syy, flip;
    const bool linear = pad_src(a, P, Q, sxx, syy, flip);
    const double ax = bitsd(alpha_bits[0]), ay = bitsd(alpha_bits[1]);
    double Ub[NU];
#pragma unroll
    for (int v = 0; v < NU; ++v) Ub[v] = 0.0;
    double Tb = 0.0, gb = 0.0;
    if (linear && (xrow || ycol)) {
      const PrimD q = prim_at(c, Up, ap, np, t);
      double rhob = 0.0, ub = 0.0, vb = 0.0, pb = 0.0;
      double Yb[JH_NS - 1];
#pragma unroll
      for (int k = 0; k < JH_NS - 1; ++k) Yb[k] = 0.0;
      for (int cv = 0; cv < NU; ++cv) {
        double hpx = 0.0, hmx = 0.0, hpy = 0.0, hmy = 0.0;
        if (xrow) {
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            const int fL = i + 3 - m, fR = i - 2 + m;
            if (fL >= 0 && fL <= nx) hpx += cx[(size_t)(cv * 10 + m) * nfx + fL * ny + j];
            if (fR >= 0 && fR <= nx) hmx += cx[(size_t)(cv * 10 + 5 + m) * nfx + fR * ny + j];
          }
        }
        if (ycol) {
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            const int fL = j + 3 - m, fR = j - 2 + m;
            if (fL >= 0 && fL <= ny) hpy += cy[(size_t)(cv * 10 + m) * nfy + i * (ny + 1) + fL];
            if (fR >= 0 && fR <= ny)
              hmy += cy[(size_t)(cv * 10 + 5 + m) * nfy + i * (ny + 1) + fR];
          }
        }
        const double ucv = Up[cv * np + t];
        Ub[cv] += 0.5 * ax * (hpx - hmx) + 0.5 * ay * (hpy - hmy);
        sx_part += 0.5 * ucv * (hpx - hmx);
        sy_part += 0.5 * ucv * (hpy - hmy);
        const double Fx = 0.5 * (hpx + hmx), Fy = 0.5 * (hpy + hmy);
        // VJP of phys_flux for both directions
        switch (cv) {
          case 0:   // rho*u, rho*v
            rhob += Fx * q.u + Fy * q.v;
            ub += Fx * q.rho;
            vb += Fy * q.rho;
            break;
          case 1:   // rho*u^2 + p, rho*u*v
            rhob += Fx * q.u * q.u + Fy * q.u * q.v;
            ub += Fx * 2.0 * q.rho * q.u + Fy * q.rho * q.v;
            vb += Fy * q.rho * q.u;
            pb += Fx;
            break;
          case 2:   // rho*u*v, rho*v^2 + p
            rhob += Fx * q.u * q.v + Fy * q.v * q.v;
            ub += Fx * q.rho * q.v;
            vb += Fx * q.rho * q.u + Fy * 2.0 * q.rho * q.v;
            pb += Fy;
            break;
          case 3:   // u (rhoE + p), v (rhoE + p)
            ub += Fx * (ucv + q.p);
            vb += Fy * (ucv + q.p);
            Ub[3] += Fx * q.u + Fy * q.v;
            pb += Fx * q.u + Fy * q.v;
            break;
          default:  // rho*u*Y = u * rhoY
            ub += Fx * ucv;
            vb += Fy * ucv;
            Ub[cv] += Fx * q.u + Fy * q.v;
        }
      }
      prim_to_U_vjp(c, q, rhob, ub, vb, pb, 0.0, Yb, Ub, Tb, gb);
    }
#pragma unroll
    for (int v = 0; v < NU; ++v) Ubp[v * np + t] = Ub[v];
    abp[t] = gb;
    abp[np + t] = Tb;
  }
  // fixed-order block sums of the alpha cotangent
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sx_part += __shfl_xor(sx_part, off, 64);
    sy_part += __shfl_xor(sy_part, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = sx_part;
    red[1][wid] = sy_part;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// alpha_bar = sum of the partials (one block, fixed order), and the number
// of padded cells attaining each LF maximum.  scal[0..1] = alpha_bar,
// cnt[0..1] = tie counts.
// Exact tie test against the maxima alpha_max_kernel reduced from the same
// arithmetic: JAX's max VJP splits the cotangent among the elements equal to
// the maximum.
__device__ __forceinline__ bool is_max(double val, double alpha) { return val == alpha; }

// LF maxima of the padded field recomputed with the arithmetic of
// alpha_count_kernel / alpha_apply_kernel (NaN ignored, nanmax of LF.py:10-11).
__global__ __launch_bounds__(256) void alpha_max_kernel(AdjArgs a, const double* __restrict__ Up,
                                                        const double* __restrict__ ap,
                                                        unsigned long long* amx) {
  __shared__ double red[4];
  const MechConst& c = *a.mc;
  const int PY = a.ny + 6, np = (a.nx + 6) * PY;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double mx = 0.0, my = 0.0;
  int sx, sy, flip;
  if (t < np && pad_src(a, t / PY, t % PY, sx, sy, flip)) {
    const PrimD q = prim_at(c, Up, ap, np, t);
    const double snd = sqrt(q.gamma * q.p * q.ir);
    mx = nan_to_zero(fabs(q.u) + snd);
    my = nan_to_zero(fabs(q.v) + snd);
  }
  block_max_atomic<256>(mx, &amx[0], red);
  block_max_atomic<256>(my, &amx[1], red);
}

__global__ __launch_bounds__(256) void alpha_reduce_kernel(const double* __restrict__ part,
                                                           int nblk, double* scal) {
  __shared__ double red[2][256];
  double sx = 0.0, sy = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    sx += part[2 * b];
    sy += part[2 * b + 1];
  }
  red[0][threadIdx.x] = sx;
  red[1][threadIdx.x] = sy;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    scal[0] = red[0][0];
    scal[1] = red[1][0];
  }
}

__global__ __launch_bounds__(256) void alpha_count_kernel(AdjArgs a, const double* __restrict__ Up,
                                                          const double* __restrict__ ap,
                                                          const unsigned long long* alpha_bits,
                                                          unsigned int* cnt) {
  const MechConst& c = *a.mc;
  const int PY = a.ny + 6, np = (a.nx + 6) * PY;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  int sx, sy, flip;
  const int P = t / PY, Q = t - P * PY;
  if (!pad_src(a, P, Q, sx, sy, flip)) return;
  const PrimD q = prim_at(c, Up, ap, np, t);
  const double snd = sqrt(q.gamma * q.p * q.ir);
  if (is_max(fabs(q.u) + snd, bitsd(alpha_bits[0]))) atomicAdd(&cnt[0], 1u);
  if (is_max(fabs(q.v) + snd, bitsd(alpha_bits[1]))) atomicAdd(&cnt[1], 1u);
}

__global__ __launch_bounds__(256) void alpha_apply_kernel(AdjArgs a, const double* __restrict__ Up,
                                                          const double* __restrict__ ap,
                                                          const unsigned long long* alpha_bits,
                                                          const double* scal,
                                                          const unsigned int* cnt,
                                                          double* __restrict__ Ubp,
                                                          double* __restrict__ abp) {
  const MechConst& c = *a.mc;
  const int PY = a.ny + 6, np = (a.nx + 6) * PY;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  int sx, sy, flip;
  const int P = t / PY, Q = t - P * PY;
  if (!pad_src(a, P, Q, sx, sy, flip)) return;
  const PrimD q = prim_at(c, Up, ap, np, t);
  const double snd = sqrt(q.gamma * q.p * q.ir);
  const bool mx = cnt[0] > 0 && is_max(fabs(q.u) + snd, bitsd(alpha_bits[0]));
  const bool my = cnt[1] > 0 && is_max(fabs(q.v) + snd, bitsd(alpha_bits[1]));
  if (!mx && !my) return;
  const double wx = mx ? scal[0] / cnt[0] : 0.0, wy = my ? scal[1] / cnt[1] : 0.0;
  // d(|u| + a) and d(|v| + a); sign(0) = 0 as in JAX
  const double su = q.u > 0 ? 1.0 : (q.u < 0 ? -1.0 : 0.0);
  const double sv = q.v > 0 ? 1.0 : (q.v < 0 ? -1.0 : 0.0);
  double Ub[NU];
#pragma unroll
  for (int v = 0; v < NU; ++v) Ub[v] = 0.0;
  double Yb[JH_NS - 1];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Yb[k] = 0.0;
  double Tb = 0.0, gb = 0.0;
  prim_to_U_vjp(c, q, 0.0, wx * su, wy * sv, 0.0, wx + wy, Yb, Ub, Tb, gb);
#pragma unroll
  for (int v = 0; v < NU; ++v) Ubp[v * np + t] += Ub[v];
  abp[t] += gb;
  abp[np + t] += Tb;
}

// Fold padded cotangents back onto the interior cells they were copied from
// (transpose of pad_kernel), added into Ub / ab.
__global__ __launch_bounds__(256) void unpad_kernel(AdjArgs a, const double* __restrict__ Ubp,
                                                    const double* __restrict__ abp,
                                                    double* __restrict__ Ub,
                                                    double* __restrict__ ab) {
  const int nx = a.nx, ny = a.ny, n = nx * ny, PY = ny + 6, np = (nx + 6) * PY;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int x = t / ny, y = t - x * ny;
  int Pc[7], Qc[7], npc = 0, nqc = 0;
  Pc[npc++] = x + 3;
  if (x <= 3 || x >= nx - 4) {
    for (int g = 0; g < 3; ++g) {
      Pc[npc++] = g;
      Pc[npc++] = nx + 3 + g;
    }
  }
  Qc[nqc++] = y + 3;
  if (y <= 3 || y >= ny - 4) {
    for (int g = 0; g < 3; ++g) {
      Qc[nqc++] = g;
      Qc[nqc++] = ny + 3 + g;
    }
  }
  double acc[NU + 2];
#pragma unroll
  for (int v = 0; v < NU + 2; ++v) acc[v] = 0.0;
  for (int ip = 0; ip < npc; ++ip) {
    for (int iq = 0; iq < nqc; ++iq) {
      int sx, sy, flip;
      if (!pad_src(a, Pc[ip], Qc[iq], sx, sy, flip)) continue;
      if (sx != x || sy != y) continue;
      const int pt = Pc[ip] * PY + Qc[iq];
#pragma unroll
      for (int v = 0; v < NU; ++v) {
        double val = Ubp[v * np + pt];
        if ((v == 1 && (flip & 1)) || (v == 2 && (flip & 2))) val = -val;
        acc[v] += val;
      }
      acc[NU] += abp[pt];
      acc[NU + 1] += abp[np + pt];
    }
  }
#pragma unroll
  for (int v = 0; v < NU; ++v) Ub[v * n + t] += acc[v];
  ab[t] += acc[NU];
  ab[n + t] += acc[NU + 1];
}

// Cotangent of (e, Y) -> U: e = U3/rho - (u^2+v^2)/2, Y = U_k/rho.
__device__ __forceinline__ void eY_to_U(const double (&Uc)[NU], double eb,
                                        const double (&Yb)[JH_NS - 1], double (&Ub)[NU]) {
  const double rho = Uc[0], ir = rcp(rho);
  const double u = Uc[1] * ir, v = Uc[2] * ir;
  double rhob = -eb * Uc[3] * ir * ir;
  Ub[3] += eb * ir;
  const double ub = -eb * u, vb = -eb * v;
  Ub[1] += ub * ir;
  rhob -= ub * u * ir;
  Ub[2] += vb * ir;
  rhob -= vb * v * ir;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) {
    Ub[4 + k] += Yb[k] * ir;
    rhob -= Yb[k] * Uc[4 + k] * ir * ir;
  }
  Ub[0] += rhob;
}

// Ub += VJP of aux = get_T(U) at the converged aux (gamma_bar, T_bar = ab).
__global__ __launch_bounds__(256) void getT_vjp_kernel(AdjArgs a, const double* __restrict__ U,
                                                       const double* __restrict__ aux,
                                                       const double* __restrict__ ab,
                                                       double* __restrict__ Ub) {
  const MechConst& c = *a.mc;
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double Uc[NU];
#pragma unroll
  for (int v = 0; v < NU; ++v) Uc[v] = U[v * n + t];
  const double ir = rcp(Uc[0]);
  double Y[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Y[k] = Uc[4 + k] * ir;
  const double R = fill_Y_get_R(c, Y);
  double eb, Yb[JH_NS - 1];
  getT_vjp(c, Y, R, aux[n + t], ab[t], ab[n + t], eb, Yb);
  double Ubc[NU];
#pragma unroll
  for (int v = 0; v < NU; ++v) Ubc[v] = 0.0;
  eY_to_U(Uc, eb, Yb, Ubc);
#pragma unroll
  for (int v = 0; v < NU; ++v) Ub[v * n + t] += Ubc[v];
}

// U4 = U3 + S(U3, T3): Ub3 = Ub4 + S_vjp, Tb3 = S_vjp_T (model/reaction_model.py:183-193).
// Ub3 may alias Ub4 (each thread reads its cell before writing it).
__global__ __launch_bounds__(256) void chem_vjp_kernel(AdjArgs a, const double* __restrict__ U3,
                                                       const double* __restrict__ aux3,
                                                       const double* Ub4, double* Ub3,
                                                       double* __restrict__ ab3) {
  const MechConst& c = *a.mc;
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double rho = U3[t];
  double rY[JH_NS - 1], Sb[JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) rY[k] = U3[(4 + k) * n + t];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) Sb[k] = Ub4[(4 + k) * n + t];
  double rhob, rYb[JH_NS - 1], Tb;
  chem_increment_vjp(c, rho, rY, aux3[n + t], a.dt, Sb, rhob, rYb, Tb);
#pragma unroll
  for (int v = 0; v < NU; ++v) {
    double g = Ub4[v * n + t];
    if (v == 0) g += rhob;
    if (v >= 4) g += rYb[v - 4];
    Ub3[v * n + t] = g;
  }
  ab3[t] = 0.0;
  ab3[n + t] = Tb;
}

// y = a*x + b*y over m doubles (b = 0 overwrites).
__global__ __launch_bounds__(256) void axpby_kernel(int m, double a, const double* __restrict__ x,
                                                    double b, double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  y[t] = b == 0.0 ? a * x[t] : a * x[t] + b * y[t];
}

}  // namespace jh
