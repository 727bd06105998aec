// Reverse-mode (VJP) kernels of one fixed-dt step, FP64.
//
// Semantics are those of JAX's reverse mode applied to the reference
// (src/janc_v2), with get_T's custom VJP (thermo_model.py:159-188): zero
// cotangent for the Halley initial guess, the LF maxima's cotangent shared
// evenly among tied maximisers of the padded field, the clip's cotangent
// routed to its active branch.
//
// Data path of the flux VJP (one RK stage), over the reference's ghost-padded
// field without storing the padded state: a padded cell reads its source
// interior cell through the boundary map (pad_src, prim_src).  Per padded
// cell, H+-_bar is summed from the WENO gradients of the faces that reach it
// (the x faces' sums in flux_vjp_x_kernel, kept per padded cell; the y
// faces' in flux_vjp_y_kernel) and pushed through the split flux and the
// primitive map (flux_vjp_y_kernel); an interior cell's cotangent is added
// into the output, a ghost cell's is kept in a padded buffer; the LF
// maxima's cotangent goes to the argmax cells (alpha_*); the ghost-cell
// cotangents are then folded back onto their source cells
// (fold_ghosts_kernel).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

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
// Flux VJP of one RK stage, one direction per pass.  Padded cell P receives,
// per direction and variable, H+_bar = sum_m F_bar(f = P - m) dq_L/dv_m(f)
// and H-_bar = sum_m F_bar(f = P - 5 + m) dq_R/dv_m(f), with the face
// cotangent F_bar(f) = dt/dx (rbar_f - rbar_{f-1}) (rbar zero outside the
// grid) and dq/dv_m the WENO gradient of face f with respect to its m-th
// stencil value.  Everything after H+-_bar (the split flux's and the
// primitive map's VJPs) is linear in it, so the x and y contributions add:
//   adj_prim_kernel    (rho, u, v, p) of every interior cell, for the x pass;
//   flux_vjp_x_kernel  H+-_bar of the x faces of every padded cell of every
//                      interior column, swept along x in registers;
//   flux_vjp_y_kernel  the y faces' H+-_bar, exchanged between the lanes of
//                      a wave, then each padded cell's sums pushed through
//                      the split flux and the primitive map, and the block
//                      partials of the LF maxima's cotangent.
// No pass uses a block barrier inside its variable loop and none keeps face
// quantities in LDS.

// x pass: one thread per (interior column, variable, side, segment of
// kXSeg padded rows), the 64 lanes of a wave on consecutive columns.
constexpr int kXSeg = 32;
__host__ __device__ inline int xs_segments(int nx) { return (nx + 6 + kXSeg - 1) / kXSeg; }
// (variable, side) instances x groups of four tiles, one block (4 waves)
// each; the launch rounds it up to a multiple of the 8 XCDs
__host__ __device__ inline int xs_work_blocks(int nx, int ny) {
  return 2 * NU * ((((ny + 63) / 64) * xs_segments(nx) + 3) / 4);
}
__host__ __device__ inline int xs_blocks(int nx, int ny) {
  return (xs_work_blocks(nx, ny) + 7) / 8 * 8;
}

// y pass: a wave covers 64 consecutive padded columns of one padded row and
// owns the 56 cells on lanes 4..59 (their stencils and faces lie within the
// wave); kYRows waves (rows) per block.
constexpr int kYRows = 4;
constexpr int kYCells = 56;
__host__ __device__ inline int ys_segments(int ny) { return (ny + 6 + kYCells - 1) / kYCells; }
__host__ __device__ inline int ys_blocks(int nx, int ny) {
  return ((nx + 6 + kYRows - 1) / kYRows) * ys_segments(ny);
}
// Padded cell of thread t of y-pass block blk, and whether the lane owns it.
__host__ __device__ inline void ys_cell(int blk, int ny, int t, int& P, int& Q, bool& mine) {
  const int ns = ys_segments(ny), rg = blk / ns, sg = blk - rg * ns, l = t & 63;
#if defined(__HIP_DEVICE_COMPILE__)
  P = rg * kYRows + __builtin_amdgcn_readfirstlane(t >> 6);   // wave-uniform row
#else
  P = rg * kYRows + (t >> 6);
#endif
  Q = sg * kYCells - 4 + l;
  mine = l >= 4 && l < 4 + kYCells;
}

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

__global__ __launch_bounds__(256) void adj_prim_kernel(AdjArgs a, const double* __restrict__ U,
                                                       const double* __restrict__ aux,
                                                       double* __restrict__ qb) {
  const int n = a.nx * a.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const PrimD q = prim_src(*a.mc, U, aux, n, t, 0);
  qb[t] = q.rho;
  qb[n + t] = q.u;
  qb[2 * n + t] = q.v;
  qb[3 * n + t] = q.p;
}

// x faces f = 0..nx of column j; the stencil of face f is padded rows
// b .. b+4 with b = f + side, and face f's m-th gradient entry belongs to row
// b + m (side 0) or b + 4 - m (side 1).  A thread sweeps the faces reaching
// its rows [c0, c0 + kXSeg), holding the five stencil values and the five
// open row sums: row b is complete once face f = b - side is done.
// hx[side][cv][P][j]: H+_bar (side 0) and H-_bar (side 1) of padded row P.
// CLS = min(cv, 4) selects the physical flux and the momentum flips at
// compile time.
template <int CLS, int SIDE>
__device__ __forceinline__ void xs_sweep(const AdjArgs& a, const double* __restrict__ U,
                                         const double* __restrict__ qb,
                                         const double* __restrict__ rbar, double ax,
                                         double* __restrict__ hx, int cv, int j, int seg) {
  const int nx = a.nx, ny = a.ny, n = nx * ny, PX = nx + 6;
  constexpr int side = SIDE;
  const double hal = (side == 0 ? 0.5 : -0.5) * ax;
  const double dtdx = a.dt / a.dx;
  const int c0 = seg * kXSeg, c1 = min(c0 + kXSeg, PX);
  const int fA = max(0, c0 - 4 - side), fB = min(nx, c0 + kXSeg - 1 - side);
  const double* Uv = U + (size_t)cv * n;
  const double* rv = rbar + (size_t)cv * n + j;
  double* out = hx + (size_t)(side * NU + cv) * PX * ny + j;
  // H+ (side 0) or H- (side 1) of padded row P of this column, read through
  // the boundary map as the forward pads it (LF.py:10-12); the loads of a
  // row are issued two faces before its value is needed
  struct Raw {
    double rho, u, v, p, uc;
    int fl;   // momentum flips, -1: no source (H = 0)
  };
  const auto load = [&](int P, Raw& r) {
    const int i = P - 3;
    int s = i * ny + j, fl = 0;
    if (i < 0 || i >= nx) {
      int sx, sy;
      if (pad_src(a, P, j + 3, sx, sy, fl)) {
        s = sx * ny + sy;
      } else {
        s = 0;
        fl = -1;
      }
    }
    r.rho = qb[s];
    r.u = qb[n + s];
    r.v = qb[2 * n + s];
    r.p = qb[3 * n + s];
    r.uc = Uv[s];
    r.fl = fl;
  };
  const auto hval = [&](const Raw& r) -> double {
    if (r.fl < 0) return 0.0;
    const PrimL q{r.rho, (r.fl & 1) ? -r.u : r.u, (r.fl & 2) ? -r.v : r.v, r.p};
    const double ucv = flip_var(CLS, r.fl, r.uc);
    return fma(hal, ucv, 0.5 * phys_flux_l(CLS, 0, q, ucv));
  };
  const auto rload = [&](int f) -> double {   // rbar of interior row f, zero outside
    const double v = rv[(size_t)(f >= 0 && f < nx ? f : 0) * ny];
    return f >= 0 && f < nx ? v : 0.0;
  };
  // rows before the first stencil get no face of this side
  for (int P = c0; P < min(c1, fA + side); ++P) out[(size_t)P * ny] = 0.0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int bB = fA + side - 1;   // last row whose stencil was swept
  if (fA <= fB) {
    const int b0 = fA + side, Pmax = PX - 1;
    Raw r0, r1, r2;
    load(b0, r0);
    load(b0 + 1, r1);
    load(b0 + 2, r2);
    double h0 = hval(r0), h1 = hval(r1), h2 = hval(r2);
    load(b0 + 3, r0);
    load(b0 + 4, r1);
    load(min(b0 + 5, Pmax), r2);
    double h3 = hval(r0), h4 = hval(r1);
    load(min(b0 + 6, Pmax), r0);
    // r2, r0: raw rows b+5, b+6 of the current face's base b
    double rprev = rload(fA - 1), rc0 = rload(fA), rc1 = rload(fA + 1);
    // two faces per trip: the second face's loads and gradient overlap the
    // first's
#pragma unroll 2
    for (int f = fA; f <= fB; ++f) {
      const int b = f + side;
      const double fb = dtdx * (rc0 - rprev);
      rprev = rc0;
      rc0 = rc1;
      rc1 = rload(f + 2);
      load(min(b + 7, Pmax), r1);
      double g[5];
      if constexpr (side == 0) {
        weno5_grad(h0, h1, h2, h3, h4, g);
      } else {
        double gr[5];
        weno5_grad(h4, h3, h2, h1, h0, gr);
#pragma unroll
        for (int m = 0; m < 5; ++m) g[m] = gr[4 - m];
      }
      const double d0 = s0 + fb * g[0];
      if (b >= c0 && b < c1) out[(size_t)b * ny] = d0;
      s0 = s1 + fb * g[1];
      s1 = s2 + fb * g[2];
      s2 = s3 + fb * g[3];
      s3 = fb * g[4];
      h0 = h1;
      h1 = h2;
      h2 = h3;
      h3 = h4;
      h4 = hval(r2);   // row b + 5 (beyond the padded grid only after the last face)
      r2 = r0;
      r0 = r1;
    }
    bB = fB + side;
  }
  // rows after the last stencil: complete when no later face exists (fB =
  // nx), else outside [c0, c1)
  const double tail[4] = {s0, s1, s2, s3};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int P = bB + 1 + k;
    if (P >= c0 && P < c1) out[(size_t)P * ny] = tail[k];
  }
  for (int P = max(c0, bB + 5); P < c1; ++P) out[(size_t)P * ny] = 0.0;
}


__global__ __launch_bounds__(256) void flux_vjp_x_kernel(AdjArgs a, const double* __restrict__ U,
                                                         const double* __restrict__ qb,
                                                         const double* __restrict__ rbar,
                                                         const unsigned long long* alpha_bits,
                                                         double* __restrict__ hx) {
  const int nx = a.nx, ny = a.ny;
  const int njb = (ny + 63) >> 6, nt = njb * xs_segments(nx);
  // Blocks go round-robin to the 8 XCDs; logical block lb gives XCD x the
  // contiguous run lb = x * (gridDim.x / 8) + k, and consecutive logical
  // blocks are the 2 NU (variable, side) instances of one group of four
  // tiles (a block's four waves share one instance), so a group's
  // primitives, state and rhs cotangent are fetched into one L2 and read
  // there by all of them.  gridDim.x is xs_blocks (a multiple of 8); the
  // logical blocks past xs_work_blocks have no work.
  const int lb = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (lb >= xs_work_blocks(nx, ny)) return;
  // the wave index through readfirstlane: every quantity derived from it
  // (segment, rows, faces, loop bounds) is then wave-uniform to the compiler
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int sc = lb % (2 * NU), tile = (lb / (2 * NU)) * 4 + wv;
  if (tile >= nt) return;
  const int side = sc & 1, cv = sc >> 1;
  const int jb = tile % njb, seg = tile / njb;
  const int j = (jb << 6) + (threadIdx.x & 63);
  if (j >= ny) return;
  const double ax = bitsd(alpha_bits[0]);
  switch (2 * min(cv, 4) + side) {
    case 0: xs_sweep<0, 0>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 1: xs_sweep<0, 1>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 2: xs_sweep<1, 0>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 3: xs_sweep<1, 1>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 4: xs_sweep<2, 0>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 5: xs_sweep<2, 1>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 6: xs_sweep<3, 0>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 7: xs_sweep<3, 1>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    case 8: xs_sweep<4, 0>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
    default: xs_sweep<4, 1>(a, U, qb, rbar, ax, hx, cv, j, seg); break;
  }
}

// Whole-wave lane shifts of a double (DPP wave_shr:1 / wave_shl:1, two
// 32-bit moves): lane l receives lane l-1 (shr1) or lane l+1 (shl1); the lane
// shifted in receives 0.
template <int CTRL>
__device__ __forceinline__ double wave_shift1(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double shr1(double x) { return wave_shift1<0x138>(x); }
__device__ __forceinline__ double shl1(double x) { return wave_shift1<0x130>(x); }

// y pass and the per-cell reverse of the split flux and the primitive map.
// Lane l at padded column Q evaluates the left-biased y face g = Q and the
// right-biased face g = Q - 1 of its row (both stencils: columns Q .. Q+4,
// taken from lanes l .. l+4 by DPP whole-wave shifts, no LDS); the cell on
// lane l sums the left-biased products of lanes l-4 .. l and the
// right-biased ones of lanes l-4 .. l.
__global__ __launch_bounds__(64 * kYRows) void flux_vjp_y_kernel(
    AdjArgs a, const double* __restrict__ U, const double* __restrict__ aux,
    const double* __restrict__ rbar, const unsigned long long* alpha_bits,
    const double* __restrict__ hx, double* __restrict__ Ut, double* __restrict__ At,
    double* __restrict__ Ubp, double* __restrict__ abp, double* __restrict__ spd,
    double* __restrict__ part) {
  __shared__ double red[4][kYRows];
  __shared__ unsigned int ties[2][kYRows];
  const MechConst& c = *a.mc;
  const int nx = a.nx, ny = a.ny, n = nx * ny, PX = nx + 6, PY = ny + 6, np = PX * PY;
  const int t = threadIdx.x, w = t >> 6, ly = t & 63;
  int P, Q;
  bool mine;
  ys_cell(blockIdx.x, ny, t, P, Q, mine);
  const double ax = bitsd(alpha_bits[0]), ay = bitsd(alpha_bits[1]);
  const double dtdy = a.dt / a.dy;
  int sx = 0, sy = 0, fl = 0;
  bool in = P < PX && Q >= 0 && Q < PY;
  // the adjoint admits linear boundary kinds only, so pad_src holds here
  if (in) in = pad_src(a, P, Q, sx, sy, fl);
  const int src = sx * ny + sy;
  const bool own = in && mine;
  const int i = P - 3, j = Q - 3;
  const bool rowin = i >= 0 && i < nx;          // wave-uniform: the row has y faces
  const bool xrow = own && j >= 0 && j < ny, ycol = own && rowin;
  const bool work = own && (xrow || ycol);
  PrimL q{1.0, 0.0, 0.0, 0.0};
  if (in) {
    const PrimD d = prim_src(c, U, aux, n, src, fl);
    q = PrimL{d.rho, d.u, d.v, d.p};
  }
  // faces g = Q (left-biased) and g = Q - 1 (right-biased) exist for
  // 0 <= g <= ny; rbar of row i at columns Q, Q-1, Q-2 (zero outside [0, ny))
  const bool v0 = rowin && Q >= 0 && Q <= ny, v1 = rowin && Q >= 1 && Q <= ny + 1;
  const bool okA = rowin && Q >= 0 && Q < ny, okB = rowin && Q >= 1 && Q <= ny,
             okC = rowin && Q >= 2 && Q <= ny + 1;
  const int rrow = (rowin ? i : 0) * ny;
  const int oA = okA ? rrow + Q : 0, oB = okB ? rrow + Q - 1 : 0, oC = okC ? rrow + Q - 2 : 0;
  const size_t ox0 = xrow ? (size_t)P * ny + j : 0;
  double Ub[NU];
#pragma unroll
  for (int v = 0; v < NU; ++v) Ub[v] = 0.0;
  double rhob = 0.0, ub = 0.0, vb = 0.0, pb = 0.0;
  double Yb[JH_NS - 1];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Yb[k] = 0.0;
  double sx_part = 0.0, sy_part = 0.0;

  // the loads of variable cv + 1 are issued before variable cv's gradients;
  // they are unconditional (clamped offsets, masked afterwards)
  double nu_, nxp, nxm, nrA, nrB, nrC;
  const auto load = [&](int cv) {
    const size_t o = (size_t)cv * n;
    nu_ = U[o + (in ? src : 0)];
    nxp = hx[(size_t)cv * PX * ny + ox0];
    nxm = hx[(size_t)(NU + cv) * PX * ny + ox0];
    nrA = rbar[o + oA];
    nrB = rbar[o + oB];
    nrC = rbar[o + oC];
  };
  // one variable; CLS = min(cv, 4) selects the flux, the flips and the
  // split flux's reverse at compile time, and cv is a constant after
  // unrolling, so Ub[cv] stays in registers
  const auto step = [&](auto K, int cv) {
    constexpr int CLS = decltype(K)::value;
    const double ucv = in ? flip_var(CLS, fl, nu_) : 0.0;
    const double xp = nxp, xm = nxm, rA = nrA, rB = nrB, rC = nrC;
    if (cv + 1 < NU) load(cv + 1);
    double hpy = 0.0, hmy = 0.0;
    if (rowin) {   // wave-uniform
      double hp = 0.0, hm = 0.0;
      if (in) {
        const double hG = 0.5 * phys_flux_l(CLS, 1, q, ucv);   // as the forward
        hp = fma(0.5 * ay, ucv, hG);
        hm = fma(-0.5 * ay, ucv, hG);
      }
      const double fb0 = v0 ? dtdy * ((okA ? rA : 0.0) - (okB ? rB : 0.0)) : 0.0;
      const double fb1 = v1 ? dtdy * ((okB ? rB : 0.0) - (okC ? rC : 0.0)) : 0.0;
      double ep[5], em[5];
      ep[0] = hp;
      em[0] = hm;
#pragma unroll
      for (int d = 1; d < 5; ++d) {   // lane l + d's values, one shift at a time
        ep[d] = shl1(ep[d - 1]);
        em[d] = shl1(em[d - 1]);
      }
      double g0[5], g1[5];
      weno5_grad(ep[0], ep[1], ep[2], ep[3], ep[4], g0);
      weno5_grad(em[4], em[3], em[2], em[1], em[0], g1);
      // left-biased: entry m of lane l belongs to lane l + m; right-biased:
      // entry m (argument em[4-m]) to lane l + 4 - m.  The cell on lane l
      // gathers them Horner-fashion, one shift per term:
      //   hpy(l) = p0(l) + p1(l-1) + p2(l-2) + p3(l-3) + p4(l-4),
      //   hmy(l) = q4(l) + q3(l-1) + q2(l-2) + q1(l-3) + q0(l-4).
      hpy = fb0 * g0[0] +
            shr1(fb0 * g0[1] + shr1(fb0 * g0[2] + shr1(fb0 * g0[3] + shr1(fb0 * g0[4]))));
      hmy = fb1 * g1[4] +
            shr1(fb1 * g1[3] + shr1(fb1 * g1[2] + shr1(fb1 * g1[1] + shr1(fb1 * g1[0]))));
    }
    if (work) {
      const double hpx = xrow ? xp : 0.0, hmx = xrow ? xm : 0.0;
      const PrimL& qo = q;
      Ub[cv] += 0.5 * ax * (hpx - hmx) + 0.5 * ay * (hpy - hmy);
      sx_part += 0.5 * ucv * (hpx - hmx);
      sy_part += 0.5 * ucv * (hpy - hmy);
      const double Fx = 0.5 * (hpx + hmx), Fy = 0.5 * (hpy + hmy);
      if constexpr (CLS == 0) {
        rhob += Fx * qo.u + Fy * qo.v;
        ub += Fx * qo.rho;
        vb += Fy * qo.rho;
      } else if constexpr (CLS == 1) {
        rhob += Fx * qo.u * qo.u + Fy * qo.u * qo.v;
        ub += Fx * 2.0 * qo.rho * qo.u + Fy * qo.rho * qo.v;
        vb += Fy * qo.rho * qo.u;
        pb += Fx;
      } else if constexpr (CLS == 2) {
        rhob += Fx * qo.u * qo.v + Fy * qo.v * qo.v;
        ub += Fx * qo.rho * qo.v;
        vb += Fx * qo.rho * qo.u + Fy * 2.0 * qo.rho * qo.v;
        pb += Fy;
      } else if constexpr (CLS == 3) {
        ub += Fx * (ucv + qo.p);
        vb += Fy * (ucv + qo.p);
        Ub[3] += Fx * qo.u + Fy * qo.v;
        pb += Fx * qo.u + Fy * qo.v;
      } else {
        ub += Fx * ucv;
        vb += Fy * ucv;
        Ub[cv] += Fx * qo.u + Fy * qo.v;
      }
    }
  };
  load(0);
  step(std::integral_constant<int, 0>{}, 0);
  step(std::integral_constant<int, 1>{}, 1);
  step(std::integral_constant<int, 2>{}, 2);
  step(std::integral_constant<int, 3>{}, 3);
#pragma unroll
  for (int cv = 4; cv < NU; ++cv) step(std::integral_constant<int, 4>{}, cv);
  // the own cell's |u| + a and |v| + a: the padded field's LF maxima are
  // reduced from these (per block here, then alpha_reduce_kernel), with the
  // arithmetic that alpha_apply_kernel's tie test repeats
  double rxo = 0.0, ryo = 0.0;   // raw values, for the tie test
  if (own) {
    double Tb = 0.0, gb = 0.0;
    const PrimD qf = prim_src(c, U, aux, n, src, fl);
    const double snd = sqrt(qf.gamma * qf.p * qf.ir);
    rxo = fabs(qf.u) + snd;
    ryo = fabs(qf.v) + snd;
    spd[P * PY + Q] = rxo;         // alpha_apply_kernel's tie test reads these
    spd[np + P * PY + Q] = ryo;
    if (work) prim_to_U_vjp(c, qf, rhob, ub, vb, pb, 0.0, Yb, Ub, Tb, gb);
    if (pad_interior(a, P, Q)) {   // its own source: straight into the output
#pragma unroll
      for (int v = 0; v < NU; ++v) Ut[(size_t)v * n + src] += Ub[v];
      At[src] += gb;
      At[n + src] += Tb;
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
    red[0][w] = sx_part;
    red[1][w] = sy_part;
    red[2][w] = mxo;
    red[3][w] = myo;
  }
  __syncthreads();
  double mx = 0.0, my = 0.0;
#pragma unroll
  for (int k = 0; k < kYRows; ++k) {
    mx = fmax(mx, red[2][k]);
    my = fmax(my, red[3][k]);
  }
  const unsigned long long tx = __ballot(own && is_max(rxo, mx));
  const unsigned long long ty = __ballot(own && is_max(ryo, my));
  if (ly == 0) {
    ties[0][w] = (unsigned int)__popcll(tx);
    ties[1][w] = (unsigned int)__popcll(ty);
  }
  __syncthreads();
  if (t == 0) {
    double sxs = 0.0, sys = 0.0;
    unsigned int cx = 0, cy = 0;
#pragma unroll
    for (int k = 0; k < kYRows; ++k) {
      sxs += red[0][k];
      sys += red[1][k];
      cx += ties[0][k];
      cy += ties[1][k];
    }
    double* o = part + 6 * blockIdx.x;
    o[0] = sxs;
    o[1] = sys;
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

// One block per flux_vjp_y_kernel block (same cells): a block whose maxima
// (part) are not the field's holds no maximiser and exits at once.
__global__ __launch_bounds__(64 * kYRows) void alpha_apply_kernel(
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
  int P, Q;
  bool mine;
  ys_cell(blockIdx.x, a.ny, threadIdx.x, P, Q, mine);
  if (!mine || P >= PX || Q < 0 || Q >= PY) return;
  const int t = P * PY + Q;
  int sx, sy, flip;
  if (!pad_src(a, P, Q, sx, sy, flip)) return;
  // the tie test compares the very values the maxima were reduced from
  // (stored by flux_vjp_y_kernel), not a recomputation
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
  if (pad_interior(a, P, Q)) {     // as in flux_vjp_y_kernel
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
