// Reverse-mode (VJP) kernels of one fixed-dt step, FP64.
//
// Semantics are those of JAX's reverse mode applied to the reference
// (src/janc_v2), with get_T's custom VJP (thermo_model.py:159-188): zero
// cotangent for the Halley initial guess, the LF maxima's cotangent shared
// evenly among tied maximisers of the padded field, the clip's cotangent
// routed to its active branch.
//
// Data path of the flux VJP (one RK stage), over the reference's ghost-padded
// field without storing it: a padded cell reads its source interior cell
// through the boundary map (pad_src, prim_src).  Per padded cell, H+-_bar is
// gathered from the WENO gradients of the faces that reach it and pushed
// through the split flux and the primitive map (flux_vjp_tile_kernel); an
// interior cell's cotangent is added into the output, a ghost cell's is kept
// in a padded buffer; the LF maxima's cotangent goes to the argmax cells
// (alpha_*); the ghost-cell cotangents are then folded back onto their source
// cells (fold_ghosts_kernel).
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

__host__ __device__ inline int wrap_x(int i, int n) { return i < 0 ? n - 1 + i : 1 + (i - n); }

// Source interior cell of padded cell (P, Q) and the sign flips of the two
// momenta (bit 0: rho*u, bit 1: rho*v) -- boundary_padding.py:20-25 then the
// BC kinds of boundary_conditions_2D/*.py.  Returns false for a kind without
// a linear map (pressure outlet, user strips).
__host__ __device__ inline bool pad_src(const AdjArgs& a, int P, int Q, int& sx, int& sy,
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

struct PrimD {
  double rho, ir, u, v, p, R, T, gamma;
  double Y[JH_NS];
};

// Primitives of a padded cell, read from its source interior cell s with the
// momenta's signs flipped per flip (pad_src).  The flips are applied to the
// quotients: -(m * ir) equals (-m) * ir exactly, so these are the values the
// reference computes on its ghost-padded copy.
__device__ __forceinline__ PrimD prim_src(const MechConst& c, const double* __restrict__ U,
                                          const double* __restrict__ aux, int n, int s,
                                          int flip) {
  PrimD q;
  q.rho = U[s];
  q.ir = rcp(q.rho);
  q.u = U[n + s] * q.ir;
  q.v = U[2 * n + s] * q.ir;
  if (flip & 1) q.u = -q.u;
  if (flip & 2) q.v = -q.v;
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) q.Y[k] = U[(4 + k) * n + s] * q.ir;
  q.R = fill_Y_get_R(c, q.Y);
  q.gamma = aux[s];
  q.T = aux[n + s];
  q.p = q.rho * q.R * q.T;
  return q;
}

// Value of variable v at a padded cell whose source cell holds val.
__device__ __forceinline__ double flip_var(int v, int flip, double val) {
  return (v == 1 && (flip & 1)) || (v == 2 && (flip & 2)) ? -val : val;
}

// The padded cell (P, Q) is an interior cell (its own source, no flips).
__host__ __device__ inline bool pad_interior(const AdjArgs& a, int P, int Q) {
  return P >= 3 && P < a.nx + 3 && Q >= 3 && Q < a.ny + 3;
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

// ------------------------------------------------------------------
// Flux VJP of one RK stage, tiled: no face quantities are staged in HBM.
// Tile = TXA padded rows x 64 padded columns, one cell per thread.  Padded
// cell P receives, per direction and variable, H+_bar = sum_m F_bar(f = P - m)
// dq_L/dv_m(face f) and H-_bar = sum_m F_bar(f = P - 5 + m) dq_R/dv_m(face f).
// Per variable: H+- of the tile and its 4-cell halo and the face cotangents
// F_bar go to LDS; then, one direction at a time, every face that reaches
// the tile evaluates its WENO gradient once and stores F_bar * dq/dv_m in LDS,
// and each cell gathers its five left- and five right-biased terms (in the
// order m = 0..4); then the split-flux and primitive VJPs of the reference's
// forward map.
// Tiles are aligned to the interior: origins (X0, Y0) = (3 + TXA a, 3 + 64 b)
// in padded coordinates, a >= -1, b >= -1, so an nx x ny interior whose
// sides are multiples of (TXA, 64) splits into whole tiles and the 3-cell
// ghost frame falls into thin tiles around them, whose face work is packed
// into their first waves (a padded grid tiled from its corner puts a
// partial tile into every band: at 1024 x 256, 516 full and 129 thin tiles
// instead of 512 and 268).  Measured at 1024 x 256 the two tilings take the
// same time (the kernel is bound by latency at two waves per SIMD, not by
// its tile count).
// Tile counts (bands a = -1 .. , b = -1 ..) and the origin of tile blk.
template <int TXA>
__host__ __device__ inline int adj_tiles_x(int nx) { return (nx + 3 + TXA - 1) / TXA + 1; }
__host__ __device__ inline int adj_tiles_y(int ny) { return (ny + 3 + TY - 1) / TY + 1; }
template <int TXA>
__device__ inline void adj_tile_origin(int blk, int ny, int& X0, int& Y0) {
  const int ty = adj_tiles_y(ny), a = blk / ty;
  X0 = 3 + TXA * (a - 1);
  Y0 = 3 + TY * (blk - a * ty - 1);
}

template <int TXA>
struct AdjTile {
  static constexpr int NT = 64 * TXA;          // threads (one padded cell each)
  static constexpr int NW = TXA;               // waves
  static constexpr int RX = TXA + 8;           // x-region rows (4-cell halo)
  static constexpr int RY = TY + 8;            // y-region columns
  static constexpr int NFX = TXA + 5;          // x-face rows reaching the tile
  static constexpr int NFY = TY + 5;           // y-face columns reaching the tile
  static constexpr int NHS = 8 / NW;           // x-halo rows per wave
};

// SEQ: the x faces' products are gathered before the y faces' are formed,
// through one buffer (two more barriers per variable, ~2/3 of the LDS).
template <int TXA, bool SEQ>
struct AdjLds {
  static constexpr int NCX = 2 * (TXA + 4) * 5 * TY, NCY = 2 * TXA * 5 * (TY + 4);
  double Hxp[TXA + 8][TY], Hxm[TXA + 8][TY];
  double Hyp[TXA][TY + 8], Hym[TXA][TY + 8];
  double Fbx[TXA + 5][TY];
  double Fby[TXA][TY + 5];
  // F_bar * dq/dv_m of the faces reaching the tile: x faces
  // [side][k][m][TY] (k < TXA+4), y faces [side][row][m][TY+4]
  double C[SEQ ? (NCX > NCY ? NCX : NCY) : NCX + NCY];
  double red[4][TXA];
  unsigned int ties[2][TXA];
};

// Exact tie test of the LF maxima: JAX's max VJP splits the cotangent
// among the elements equal to the maximum.
__device__ __forceinline__ bool is_max(double val, double alpha) { return val == alpha; }

struct PrimL {   // the primitives the physical flux needs
  double rho, u, v, p;
};

__device__ __forceinline__ double phys_flux_l(int cv, int dir, const PrimL& q, double ucv) {
  const double un = dir == 0 ? q.u : q.v;
  switch (cv) {
    case 0: return q.rho * un;
    case 1: return dir == 0 ? q.rho * (q.u * q.u) + q.p : (q.rho * q.u) * q.v;
    case 2: return dir == 0 ? (q.rho * q.u) * q.v : q.rho * (q.v * q.v) + q.p;
    case 3: return un * (ucv + q.p);
    default: return un * ucv;
  }
}

template <int TXA, bool SEQ>
__global__ __launch_bounds__(AdjTile<TXA>::NT) __attribute__((amdgpu_waves_per_eu(2))) void flux_vjp_tile_kernel(
    AdjArgs a, const double* __restrict__ U, const double* __restrict__ aux,
    const double* __restrict__ rbar, const unsigned long long* alpha_bits,
    double* __restrict__ Ut, double* __restrict__ At, double* __restrict__ Ubp,
    double* __restrict__ abp, double* __restrict__ spd, double* __restrict__ part) {
  using G = AdjTile<TXA>;
  __shared__ AdjLds<TXA, SEQ> L;
  double* Cx = L.C;
  double* Cy = SEQ ? L.C : L.C + AdjLds<TXA, SEQ>::NCX;
  const MechConst& c = *a.mc;
  const int nx = a.nx, ny = a.ny, n = nx * ny, PX = nx + 6, PY = ny + 6, np = PX * PY;
  const int t = threadIdx.x, ly = t & 63, w = t >> 6;
  int X0, Y0;
  adj_tile_origin<TXA>(blockIdx.x, ny, X0, Y0);
  const double ax = bitsd(alpha_bits[0]), ay = bitsd(alpha_bits[1]);
  const double dtdx = a.dt / a.dx, dtdy = a.dt / a.dy;

  // region slots: 0 = own cell (x and y regions), 1..NHS = x-halo rows,
  // NHS+1 = y-halo cell (t < 8 * TXA)
  constexpr int NS_ = G::NHS + 2;
  int px[NS_], py[NS_], xr[NS_], yc[NS_];
  bool in[NS_];
  px[0] = X0 + w;
  py[0] = Y0 + ly;
  xr[0] = w + 4;
  yc[0] = ly + 4;
#pragma unroll
  for (int h = 0; h < G::NHS; ++h) {
    const int hr = w + h * G::NW;                  // halo row 0..7
    const int r = hr < 4 ? hr : TXA + hr;          // region rows 0..3, TXA+4..TXA+7
    xr[1 + h] = r;
    px[1 + h] = X0 - 4 + r;
    py[1 + h] = Y0 + ly;
    yc[1 + h] = -1;
  }
  {
    constexpr int s = NS_ - 1;
    const int row = t >> 3, k = t & 7;
    yc[s] = t < 8 * TXA ? (k < 4 ? k : TY + k) : -1;
    xr[s] = -1;
    px[s] = X0 + row;
    py[s] = Y0 - 4 + (k < 4 ? k : TY + k);
  }
  PrimL q[NS_];
  int src[NS_], fl[NS_];   // source interior cell and momentum flips (pad_src)
#pragma unroll
  for (int s = 0; s < NS_; ++s) {
    in[s] = px[s] >= 0 && px[s] < PX && py[s] >= 0 && py[s] < PY &&
            (s < NS_ - 1 || t < 8 * TXA);
    int sx = 0, sy = 0, f = 0;
    // the adjoint admits linear boundary kinds only, so pad_src holds here
    if (in[s]) in[s] = pad_src(a, px[s], py[s], sx, sy, f);
    src[s] = sx * ny + sy;
    fl[s] = f;
    if (in[s]) {
      const PrimD d = prim_src(c, U, aux, n, src[s], f);
      q[s] = PrimL{d.rho, d.u, d.v, d.p};
    } else {
      q[s] = PrimL{1.0, 0.0, 0.0, 0.0};
    }
  }
  // own cell: which face families reach it
  const int P = px[0], Q = py[0], i = P - 3, j = Q - 3;
  const bool own = in[0];
  const bool xrow = own && j >= 0 && j < ny, ycol = own && i >= 0 && i < nx;
  const bool work = own && (xrow || ycol);
  double Ub[NU];
#pragma unroll
  for (int v = 0; v < NU; ++v) Ub[v] = 0.0;
  double rhob = 0.0, ub = 0.0, vb = 0.0, pb = 0.0;
  double Yb[JH_NS - 1];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Yb[k] = 0.0;
  double sx_part = 0.0, sy_part = 0.0;

  // This thread's face-cotangent entries: x faces e = t + r NT (r < NEX) of
  // the (NFX, TY) table, y faces e = t + r NT (r < NEY) of the (TXA, NFY)
  // table; F_bar = dt/dx (rbar_f - rbar_{f-1}) with rbar zero outside the
  // grid.  The rbar and state loads of variable cv + 1 are issued while
  // variable cv's gradients run; they are unconditional (clamped offsets,
  // zeroed afterwards) so no load sits on a divergent branch.
  constexpr int NEX = (G::NFX * TY + G::NT - 1) / G::NT;
  constexpr int NEY = (TXA * G::NFY + G::NT - 1) / G::NT;
  int xoff[NEX], yoff[NEY];
  unsigned xfl[NEX], yfl[NEY];   // bit 0: entry exists, 1: face valid, 2: left, 3: right
#pragma unroll
  for (int r = 0; r < NEX; ++r) {
    const int e = t + r * G::NT;
    const int k = e >> 6, col = e & 63;
    const int f = X0 - 5 + k, jj = Y0 + col - 3;
    const bool ex = e < G::NFX * TY;
    const bool ok = ex && f >= 0 && f <= nx && jj >= 0 && jj < ny;
    const bool lo = ok && f - 1 >= 0, hi = ok && f < nx;
    xfl[r] = (ex ? 1u : 0u) | (ok ? 2u : 0u) | (lo ? 4u : 0u) | (hi ? 8u : 0u);
    xoff[r] = ok ? (f - 1) * ny + jj : 0;          // right value at + ny
  }
#pragma unroll
  for (int r = 0; r < NEY; ++r) {
    const int e = t + r * G::NT;
    const int row = e / G::NFY, k = e - row * G::NFY;
    const int gg = Y0 - 5 + k, ii = X0 + row - 3;
    const bool ex = e < TXA * G::NFY;
    const bool ok = ex && gg >= 0 && gg <= ny && ii >= 0 && ii < nx;
    const bool lo = ok && gg - 1 >= 0, hi = ok && gg < ny;
    yfl[r] = (ex ? 1u : 0u) | (ok ? 2u : 0u) | (lo ? 4u : 0u) | (hi ? 8u : 0u);
    yoff[r] = ok ? ii * ny + gg - 1 : 0;           // right value at + 1
  }
  double uraw[NS_], rxl[NEX], rxr[NEX], ryl[NEY], ryr[NEY];
  auto load_raw = [&](int v) {
    const size_t o = (size_t)v * n;
#pragma unroll
    for (int s = 0; s < NS_; ++s) uraw[s] = U[o + (in[s] ? src[s] : 0)];
#pragma unroll
    for (int r = 0; r < NEX; ++r) {
      const int b = xoff[r];
      rxl[r] = rbar[o + (xfl[r] & 4u ? b : 0)];
      rxr[r] = rbar[o + (xfl[r] & 8u ? b + ny : 0)];
    }
#pragma unroll
    for (int r = 0; r < NEY; ++r) {
      const int b = yoff[r];
      ryl[r] = rbar[o + (yfl[r] & 4u ? b : 0)];
      ryr[r] = rbar[o + (yfl[r] & 8u ? b + 1 : 0)];
    }
  };
  load_raw(0);
  // x faces of this tile inside the grid lie in columns [cx0, cx0 + nxv)
  // (jj = Y0 + col - 3 in [0, ny)).  A tile with at most half its columns
  // there (the ghost-frame tile columns, small grids) is thin: its face
  // work is packed into the first waves.  Its cells (columns [l0, l1) of
  // the padded grid) read the y faces kk = ky0 .. ky1 (a cell at column ly
  // reads ly .. ly + 4).
  const int cx0 = Y0 < 3 ? 3 - Y0 : 0;
  const int nxv = max(0, min(TY, ny + 3 - Y0) - cx0);
  const bool thin = 2 * nxv <= TY;
  const int ky0 = max(0, -Y0), ky1 = min(TY + 3, min(TY, PY - Y0) + 3);
  const int nky = max(0, ky1 - ky0 + 1);

  for (int cv = 0; cv < NU; ++cv) {
    // ---- H+- of the region and the face cotangents ----
    double ucv_own = 0.0;
#pragma unroll
    for (int s = 0; s < NS_; ++s) {
      const double ucv = in[s] ? flip_var(cv, fl[s], uraw[s]) : 0.0;
      if (s == 0) ucv_own = ucv;
      if (xr[s] >= 0) {
        double hp = 0.0, hm = 0.0;
        if (in[s]) {
          const double hF = 0.5 * phys_flux_l(cv, 0, q[s], ucv);   // as the forward
          hp = fma(0.5 * ax, ucv, hF);
          hm = fma(-0.5 * ax, ucv, hF);
        }
        L.Hxp[xr[s]][ly] = hp;
        L.Hxm[xr[s]][ly] = hm;
      }
      if (yc[s] >= 0) {
        const int row = s == 0 ? w : (t >> 3);
        double hp = 0.0, hm = 0.0;
        if (in[s]) {
          const double hG = 0.5 * phys_flux_l(cv, 1, q[s], ucv);
          hp = fma(0.5 * ay, ucv, hG);
          hm = fma(-0.5 * ay, ucv, hG);
        }
        L.Hyp[row][yc[s]] = hp;
        L.Hym[row][yc[s]] = hm;
      }
    }
#pragma unroll
    for (int r = 0; r < NEX; ++r) {
      if (!(xfl[r] & 1u)) continue;
      const int e = t + r * G::NT;
      const double rl = xfl[r] & 4u ? rxl[r] : 0.0, rr = xfl[r] & 8u ? rxr[r] : 0.0;
      L.Fbx[e >> 6][e & 63] = xfl[r] & 2u ? dtdx * (rr - rl) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NEY; ++r) {
      if (!(yfl[r] & 1u)) continue;
      const int e = t + r * G::NT;
      const int row = e / G::NFY;
      const double rl = yfl[r] & 4u ? ryl[r] : 0.0, rr = yfl[r] & 8u ? ryr[r] : 0.0;
      L.Fby[row][e - row * G::NFY] = yfl[r] & 2u ? dtdy * (rr - rl) : 0.0;
    }
    __syncthreads();
    if (cv + 1 < NU) load_raw(cv + 1);
    // ---- x and y faces: F_bar * dq/dv_m into LDS ----
    // x faces kk = 0..TXA+3 per side and column; with 8 rows a thread takes
    // a run of three consecutive faces of one column (2 sides x 4 runs x 64
    // columns = the 512 threads) and reads the run's seven stencil values
    // once.
    const auto xface = [&](int side, int kk, int col, const double* h) {
      const int f = side == 0 ? X0 - 4 + kk : X0 - 5 + kk;   // face index
      const int jj = Y0 + col - 3;
      double* out = &Cx[((side * (TXA + 4) + kk) * 5) * TY + col];
      if (f >= 0 && f <= nx && jj >= 0 && jj < ny) {
        double g[5];
        if (side == 0)   // left-biased, cells f..f+4 = region rows kk..kk+4
          weno5_grad(h[0], h[1], h[2], h[3], h[4], g);
        else             // right-biased, cells f+5..f+1 = region rows kk+4..kk
          weno5_grad(h[4], h[3], h[2], h[1], h[0], g);
        const double fb = L.Fbx[side == 0 ? kk + 1 : kk][col];
#pragma unroll
        for (int m = 0; m < 5; ++m) out[m * TY] = fb * g[m];
      } else {         // a face outside the grid contributes nothing
#pragma unroll
        for (int m = 0; m < 5; ++m) out[m * TY] = 0.0;
      }
    };
    if constexpr (TXA == 8) {
      const auto xrun = [&](int side, int col, int kk0) {
        double h[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) h[j] = side == 0 ? L.Hxp[kk0 + j][col] : L.Hxm[kk0 + j][col];
#pragma unroll
        for (int i = 0; i < 3; ++i) xface(side, kk0 + i, col, h + i);
      };
      if (!thin) {
        xrun(t >> 8, t & 63, 3 * ((t >> 6) & 3));
      } else if (t < 8 * nxv) {
        // a thin tile packs the runs of its columns with x faces into the
        // first waves instead of giving every wave a few working lanes
        const int q = t / nxv;
        xrun(q >> 2, cx0 + (t - q * nxv), 3 * (q & 3));
      }
    } else {
      for (int e = t; e < 2 * (TXA + 4) * TY; e += G::NT) {
        const int side = e / ((TXA + 4) * TY), rem = e - side * ((TXA + 4) * TY);
        const int kk = rem >> 6, col = rem & 63;
        double h[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) h[j] = side == 0 ? L.Hxp[kk + j][col] : L.Hxm[kk + j][col];
        xface(side, kk, col, h);
      }
    }
    // ---- the own cell's sums over the faces that reach it ----
    // (every face a cell reads was written this variable, faces outside the
    // grid as zeros, so the reads are unconditional and the ten reads of a
    // direction are in flight together)
    double hpx = 0.0, hmx = 0.0, hpy = 0.0, hmy = 0.0;
    const auto gather_x = [&]() {
      if (work && xrow) {
#pragma unroll
        for (int m = 0; m < 5; ++m) hpx += Cx[((w + 4 - m) * 5 + m) * TY + ly];
#pragma unroll
        for (int m = 0; m < 5; ++m) hmx += Cx[(((TXA + 4) + w + m) * 5 + m) * TY + ly];
      }
    };
    if constexpr (SEQ) {
      __syncthreads();
      gather_x();
      __syncthreads();
    }
    // y faces kk = 0..TY+3 per side and row: with 8 rows the 512 threads
    // take the pairs (kk, kk+1), kk < TY, of every side and row, and the
    // first 64 threads the last four faces of each
    const auto yface = [&](int side, int row, int kk, const double* h) {
      const int gf = side == 0 ? Y0 - 4 + kk : Y0 - 5 + kk;
      const int ii = X0 + row - 3;
      double* out = &Cy[((side * TXA + row) * 5) * (TY + 4) + kk];
      if (gf >= 0 && gf <= ny && ii >= 0 && ii < nx) {
        double g[5];
        if (side == 0)
          weno5_grad(h[0], h[1], h[2], h[3], h[4], g);
        else
          weno5_grad(h[4], h[3], h[2], h[1], h[0], g);
        const double fb = L.Fby[row][side == 0 ? kk + 1 : kk];
#pragma unroll
        for (int m = 0; m < 5; ++m) out[m * (TY + 4)] = fb * g[m];
      } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) out[m * (TY + 4)] = 0.0;
      }
    };
    if constexpr (TXA == 8) {
      const auto ypair = [&](int side, int row, int kk0) {
        double h[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) h[j] = side == 0 ? L.Hyp[row][kk0 + j] : L.Hym[row][kk0 + j];
        yface(side, row, kk0, h);
        yface(side, row, kk0 + 1, h + 1);
      };
      if (!thin) {
        ypair(t >> 8, (t >> 5) & 7, 2 * (t & 31));
      } else {         // the faces a thin tile's cells read, packed
        for (int e = t; e < 2 * TXA * nky; e += G::NT) {
          const int q = e / nky, side = q / TXA, row = q - side * TXA;
          const int kk = ky0 + (e - q * nky);
          double h[5];
#pragma unroll
          for (int j = 0; j < 5; ++j) h[j] = side == 0 ? L.Hyp[row][kk + j] : L.Hym[row][kk + j];
          yface(side, row, kk, h);
        }
      }
      if (!thin && t < 64) {
        const int side = t >> 5, row = (t >> 2) & 7, kk = TY + (t & 3);
        double h[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) h[j] = side == 0 ? L.Hyp[row][kk + j] : L.Hym[row][kk + j];
        yface(side, row, kk, h);
      }
    } else {
      for (int e = t; e < 2 * TXA * (TY + 4); e += G::NT) {
        const int side = e / (TXA * (TY + 4)), rem = e - side * (TXA * (TY + 4));
        const int row = rem / (TY + 4), kk = rem - row * (TY + 4);
        double h[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) h[j] = side == 0 ? L.Hyp[row][kk + j] : L.Hym[row][kk + j];
        yface(side, row, kk, h);
      }
    }
    __syncthreads();
    if constexpr (!SEQ) gather_x();
    if (work && ycol) {
#pragma unroll
      for (int m = 0; m < 5; ++m) hpy += Cy[(w * 5 + m) * (TY + 4) + ly + 4 - m];
#pragma unroll
      for (int m = 0; m < 5; ++m) hmy += Cy[((TXA + w) * 5 + m) * (TY + 4) + ly + m];
    }
    if (work) {
      const PrimL& qo = q[0];
      const double ucv = ucv_own;
      Ub[cv] += 0.5 * ax * (hpx - hmx) + 0.5 * ay * (hpy - hmy);
      sx_part += 0.5 * ucv * (hpx - hmx);
      sy_part += 0.5 * ucv * (hpy - hmy);
      const double Fx = 0.5 * (hpx + hmx), Fy = 0.5 * (hpy + hmy);
      switch (cv) {
        case 0:
          rhob += Fx * qo.u + Fy * qo.v;
          ub += Fx * qo.rho;
          vb += Fy * qo.rho;
          break;
        case 1:
          rhob += Fx * qo.u * qo.u + Fy * qo.u * qo.v;
          ub += Fx * 2.0 * qo.rho * qo.u + Fy * qo.rho * qo.v;
          vb += Fy * qo.rho * qo.u;
          pb += Fx;
          break;
        case 2:
          rhob += Fx * qo.u * qo.v + Fy * qo.v * qo.v;
          ub += Fx * qo.rho * qo.v;
          vb += Fx * qo.rho * qo.u + Fy * 2.0 * qo.rho * qo.v;
          pb += Fy;
          break;
        case 3:
          ub += Fx * (ucv + qo.p);
          vb += Fy * (ucv + qo.p);
          Ub[3] += Fx * qo.u + Fy * qo.v;
          pb += Fx * qo.u + Fy * qo.v;
          break;
        default:
          ub += Fx * ucv;
          vb += Fy * ucv;
          Ub[cv] += Fx * qo.u + Fy * qo.v;
      }
    }
    // No barrier before the next variable: its H and F_bar writes only
    // conflict with this variable's face loops (finished before the last
    // barrier above), and its face loops' product writes come after its own
    // first barrier, which every thread reaches after these reads.
  }
  // the own cell's |u| + a and |v| + a: the padded field's LF maxima are
  // reduced from these (per block here, then alpha_reduce_kernel), with the
  // arithmetic that alpha_apply_kernel's tie test repeats
  double rxo = 0.0, ryo = 0.0;   // raw values, for the tie test
  if (own) {
    double Tb = 0.0, gb = 0.0;
    const PrimD qf = prim_src(c, U, aux, n, src[0], fl[0]);
    const double snd = sqrt(qf.gamma * qf.p * qf.ir);
    rxo = fabs(qf.u) + snd;
    ryo = fabs(qf.v) + snd;
    spd[P * PY + Q] = rxo;         // alpha_apply_kernel's tie test reads these
    spd[np + P * PY + Q] = ryo;
    if (work) prim_to_U_vjp(c, qf, rhob, ub, vb, pb, 0.0, Yb, Ub, Tb, gb);
    if (pad_interior(a, P, Q)) {   // its own source: straight into the output
      const int s = src[0];
#pragma unroll
      for (int v = 0; v < NU; ++v) Ut[(size_t)v * n + s] += Ub[v];
      At[s] += gb;
      At[n + s] += Tb;
    } else {                       // a ghost: kept for fold_ghosts_kernel
      const int pt = P * PY + Q;
#pragma unroll
      for (int v = 0; v < NU; ++v) Ubp[(size_t)v * np + pt] = Ub[v];
      abp[pt] = gb;
      abp[np + pt] = Tb;
    }
  }
  // fixed-order block sums of the alpha cotangent, the block maxima, and
  // the block's count of cells attaining them (exact: a block whose maximum
  // is the field's holds all of its cells that attain it)
  double mxo = nan_to_zero(rxo), myo = nan_to_zero(ryo);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sx_part += __shfl_xor(sx_part, off, 64);
    sy_part += __shfl_xor(sy_part, off, 64);
    mxo = fmax(mxo, __shfl_xor(mxo, off, 64));
    myo = fmax(myo, __shfl_xor(myo, off, 64));
  }
  if (ly == 0) {
    L.red[0][w] = sx_part;
    L.red[1][w] = sy_part;
    L.red[2][w] = mxo;
    L.red[3][w] = myo;
  }
  __syncthreads();
  double mx = 0.0, my = 0.0;
#pragma unroll
  for (int k = 0; k < TXA; ++k) {
    mx = fmax(mx, L.red[2][k]);
    my = fmax(my, L.red[3][k]);
  }
  const unsigned long long tx = __ballot(own && is_max(rxo, mx));
  const unsigned long long ty = __ballot(own && is_max(ryo, my));
  if (ly == 0) {
    L.ties[0][w] = (unsigned int)__popcll(tx);
    L.ties[1][w] = (unsigned int)__popcll(ty);
  }
  __syncthreads();
  if (t == 0) {
    double sx = 0.0, sy = 0.0;
    unsigned int cx = 0, cy = 0;
#pragma unroll
    for (int k = 0; k < TXA; ++k) {
      sx += L.red[0][k];
      sy += L.red[1][k];
      cx += L.ties[0][k];
      cy += L.ties[1][k];
    }
    double* o = part + 6 * blockIdx.x;
    o[0] = sx;
    o[1] = sy;
    o[2] = mx;
    o[3] = my;
    o[4] = (double)cx;
    o[5] = (double)cy;
  }
}

// alpha_bar = sum of the per-block partials, in a fixed order; the LF maxima
// of the padded field = max of the block maxima; the tie counts = sum of the
// counts of the blocks whose maximum is the field's (one block): scal[0..1],
// amx[0..1] (bit patterns) and cnt[0..1] for alpha_apply_kernel, which hands
// each padded cell attaining a maximum its share of alpha_bar.
__global__ __launch_bounds__(256) void alpha_reduce_kernel(const double* __restrict__ part,
                                                           int nblk, double* scal,
                                                           unsigned long long* amx,
                                                           unsigned int* cnt) {
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  double sx = 0.0, sy = 0.0, mx = 0.0, my = 0.0;
  for (int b = t; b < nblk; b += 256) {
    sx += part[6 * b];
    sy += part[6 * b + 1];
    mx = fmax(mx, part[6 * b + 2]);
    my = fmax(my, part[6 * b + 3]);
  }
  red[0][t] = sx;
  red[1][t] = sy;
  red[2][t] = mx;
  red[3][t] = my;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] += red[0][t + s];
      red[1][t] += red[1][t + s];
      red[2][t] = fmax(red[2][t], red[2][t + s]);
      red[3][t] = fmax(red[3][t], red[3][t + s]);
    }
    __syncthreads();
  }
  const double gx = red[2][0], gy = red[3][0];
  if (t == 0) {
    scal[0] = red[0][0];
    scal[1] = red[1][0];
    amx[0] = dbits(gx);
    amx[1] = dbits(gy);
  }
  __syncthreads();                 // red[0..1][0] read before reuse
  double cx = 0.0, cy = 0.0;
  for (int b = t; b < nblk; b += 256) {
    if (part[6 * b + 2] == gx) cx += part[6 * b + 4];
    if (part[6 * b + 3] == gy) cy += part[6 * b + 5];
  }
  red[0][t] = cx;
  red[1][t] = cy;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] += red[0][t + s];
      red[1][t] += red[1][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    cnt[0] = (unsigned int)red[0][0];
    cnt[1] = (unsigned int)red[1][0];
  }
}

// One block per flux_vjp_tile_kernel tile (same cells): a tile whose block
// maxima (part) are not the field's holds no maximiser and exits at once.
template <int TXA>
__global__ __launch_bounds__(AdjTile<TXA>::NT) void alpha_apply_kernel(
    AdjArgs a, const double* __restrict__ U, const double* __restrict__ aux,
    const double* __restrict__ spd, const double* __restrict__ part,
    const unsigned long long* alpha_bits, const double* scal, const unsigned int* cnt,
    double* __restrict__ Ut, double* __restrict__ At, double* __restrict__ Ubp,
    double* __restrict__ abp) {
  const double gx = bitsd(alpha_bits[0]), gy = bitsd(alpha_bits[1]);
  const double* pb = part + 6 * blockIdx.x;
  if (!(cnt[0] > 0 && pb[2] == gx) && !(cnt[1] > 0 && pb[3] == gy)) return;
  const MechConst& c = *a.mc;
  const int PX = a.nx + 6, PY = a.ny + 6, np = PX * PY;
  int X0, Y0;
  adj_tile_origin<TXA>(blockIdx.x, a.ny, X0, Y0);
  const int P = X0 + (threadIdx.x >> 6), Q = Y0 + (threadIdx.x & 63);
  if (P < 0 || P >= PX || Q < 0 || Q >= PY) return;
  const int t = P * PY + Q;
  int sx, sy, flip;
  if (!pad_src(a, P, Q, sx, sy, flip)) return;
  // the tie test compares the very values the maxima were reduced from
  // (stored by flux_vjp_tile_kernel), not a recomputation
  const bool mx = cnt[0] > 0 && is_max(spd[t], gx);
  const bool my = cnt[1] > 0 && is_max(spd[np + t], gy);
  if (!mx && !my) return;
  const PrimD q = prim_src(c, U, aux, a.nx * a.ny, sx * a.ny + sy, flip);
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
  if (pad_interior(a, P, Q)) {     // as in flux_vjp_tile_kernel
    const int n = a.nx * a.ny, s = sx * a.ny + sy;
#pragma unroll
    for (int v = 0; v < NU; ++v) Ut[v * n + s] += Ub[v];
    At[s] += gb;
    At[n + s] += Tb;
  } else {
#pragma unroll
    for (int v = 0; v < NU; ++v) Ubp[v * np + t] += Ub[v];
    abp[t] += gb;
    abp[np + t] += Tb;
  }
}

// Fold the ghost cells' cotangents back onto the interior cells they were
// copied from (the transpose of pad_src on the ghosts), added into Ut / At.
// Target k is interior cell tgt[k]; its ghost copies are ent[beg[k] ..
// beg[k+1]), each the padded position * 4 + the momentum flips.
__global__ __launch_bounds__(256) void fold_ghosts_kernel(AdjArgs a, const int* __restrict__ tgt,
                                                          const int* __restrict__ beg,
                                                          const int* __restrict__ ent, int m,
                                                          const double* __restrict__ Ubp,
                                                          const double* __restrict__ abp,
                                                          double* __restrict__ Ut,
                                                          double* __restrict__ At) {
  // one thread per (target, variable): blockIdx.y = variable (NU + 2 of
  // them), so the few thousand targets still fill the GPU
  const int n = a.nx * a.ny, np = (a.nx + 6) * (a.ny + 6);
  const int k = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
  if (k >= m) return;
  const double* src = v < NU ? Ubp + (size_t)v * np : abp + (size_t)(v - NU) * np;
  double acc = 0.0;
  const int e1 = beg[k + 1];
  for (int e = beg[k]; e < e1; ++e) {
    const int pt = ent[e] >> 2, flip = ent[e] & 3;
    acc += v < NU ? flip_var(v, flip, src[pt]) : src[pt];
  }
  const int t = tgt[k];
  double* dst = v < NU ? Ut + (size_t)v * n : At + (size_t)(v - NU) * n;
  dst[t] += acc;
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
// The U cotangent that (gamma_bar, T_bar) = (ab[t], ab[n + t]) of cell t
// sends through get_T's VJP and (e, Y)(U).
__device__ __forceinline__ void getT_vjp_cell(const MechConst& c, const double* __restrict__ U,
                                              const double* __restrict__ aux,
                                              const double* __restrict__ ab, int n, int t,
                                              double (&Ubc)[NU]) {
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
#pragma unroll
  for (int v = 0; v < NU; ++v) Ubc[v] = 0.0;
  eY_to_U(Uc, eb, Yb, Ubc);
}

// Ub = Ubar + get_T's VJP of ab at the state (U, aux): the start of the
// reverse sweep, at the step's output.
__global__ __launch_bounds__(256) void getT_vjp_kernel(AdjArgs a, const double* __restrict__ U,
                                                       const double* __restrict__ aux,
                                                       const double* __restrict__ ab,
                                                       const double* __restrict__ Ubar,
                                                       double* __restrict__ Ub) {
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double Ubc[NU];
  getT_vjp_cell(*a.mc, U, aux, ab, n, t, Ubc);
#pragma unroll
  for (int v = 0; v < NU; ++v) Ub[v * n + t] = Ubar[v * n + t] + Ubc[v];
}

// U4 = U3 + S(U3, T3): Ub3 = Ub4 + S_vjp, Tb3 = S_vjp_T (model/reaction_model.py:183-193),
// in two kernels: chem_vjp_solve_kernel writes x = A^-1 w and lambda =
// A^-T x_bar (8 planes each), chem_vjp_rates_kernel the cotangents.  Split,
// each half fits two waves per SIMD without scratch (one kernel held 512
// registers per lane and still spilled).
// chem_vjp_*'s state of cell t read from the stage state U and the
// cotangent Ub of the step's output.
// (plane base + lane index: the base is wave-uniform, so every plane is
// addressed from scalar registers and one shared lane offset)
struct AdjCellState {
  const double* U;
  const double* Ub;
  int n, t;
  __device__ double rho() const { return U[t]; }
  __device__ double operator()(int k) const { return (U + (size_t)(4 + k) * n)[t]; }
  __device__ double sb(int k) const { return (Ub + (size_t)(4 + k) * n)[t]; }
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void chem_vjp_solve_kernel(
    AdjArgs a, const double* __restrict__ U3, const double* __restrict__ aux3,
    const double* __restrict__ Ub4, double* __restrict__ xs, double* __restrict__ ls) {
  __shared__ GibbsTables gt;
  const MechConst& c = *a.mc;
  load_gibbs_tables(c, gt);
  __syncthreads();
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double x[JH_NA], lam[JH_NA];
  chem_vjp_solve(c, gt, AdjCellState{U3, Ub4, n, t}, aux3[n + t], a.dt, x, lam);
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    xs[(size_t)k * n + t] = x[k];
    ls[(size_t)k * n + t] = lam[k];
  }
}

// Ub3 may alias Ub4 (each thread reads its cell before writing it).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void chem_vjp_rates_kernel(
    AdjArgs a, const double* __restrict__ U3, const double* __restrict__ aux3, const double* Ub4,
    const double* __restrict__ xs, const double* __restrict__ ls, double* Ub3,
    double* __restrict__ ab3) {
  __shared__ GibbsTables gt;
  const MechConst& c = *a.mc;
  load_gibbs_tables(c, gt);
  __syncthreads();
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double x[JH_NA], lam[JH_NA];
#pragma unroll
  for (int k = 0; k < JH_NA; ++k) {
    x[k] = xs[(size_t)k * n + t];
    lam[k] = ls[(size_t)k * n + t];
  }
  double rhob, rYb[JH_NS - 1], Tb;
  chem_vjp_rates(c, gt, AdjCellState{U3, Ub4, n, t}, aux3[n + t], x, lam, rhob, rYb, Tb);
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

// getT_vjp_kernel followed by the reverse of one Shu-Osher combine
// U' = ... + cr (U_s + r(U_s)), in one pass over the cells: with x = Ub +
// get_T's VJP of ab (the value getT_vjp_kernel would store), Uacc =
// aacc*x + bacc*Uacc (bacc = 0 overwrites), Ub = x (cr = 0) or
// rb = Ub = cr*x, and ab zeroed.
__global__ __launch_bounds__(256) void getT_split_kernel(AdjArgs a, const double* __restrict__ U,
                                                         const double* __restrict__ aux,
                                                         double aacc, double bacc, double cr,
                                                         double* __restrict__ Ub,
                                                         double* __restrict__ Uacc,
                                                         double* __restrict__ rb,
                                                         double* __restrict__ ab) {
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double Ubc[NU];
  getT_vjp_cell(*a.mc, U, aux, ab, n, t, Ubc);
#pragma unroll
  for (int v = 0; v < NU; ++v) {
    const size_t i = (size_t)v * n + t;
    const double x = Ub[i] + Ubc[v];
    Uacc[i] = bacc == 0.0 ? aacc * x : aacc * x + bacc * Uacc[i];
    if (cr != 0.0) {
      const double r = cr * x;
      rb[i] = r;
      Ub[i] = r;
    } else {
      Ub[i] = x;
    }
  }
  ab[t] = 0.0;
  ab[n + t] = 0.0;
}

}  // namespace jh
