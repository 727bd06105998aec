// Same-size AMR blocks advanced together: the AMR branch of the reference's
// 2-D step (simulation/set_simulator.py:37-42, :107-111, and
// time_step/TVD_RK3.py:15-37), where flux_func, update_func and the
// chemistry are vmap-ed over a leading block axis.
//
// Layout: a batch of blocks is the reference's blk_data, (n_blk, 14, PX, PY)
// = U (12 planes) and aux (gamma, T) of each block, C-order; the stage
// entry's blocks carry their 3 ghost layers (PX = bnx + 6, PY = bny + 6,
// filled by the caller: jaxamr's get_ghost_block_data), the chemistry's do
// not (PX = bnx, PY = bny).  Each block has its own StepScalars: its LF
// maxima (vmap gives every block the nanmax over its own padded field) and
// health flags.  The temperature solves stop per cell.
#pragma once
#include <hip/hip_runtime.h>

#include "stage.hpp"

namespace jh {

// Rows of the block stage kernel's tile: AMR blocks are small (tens of cells
// per side), so the short tile wastes the fewest rows.
constexpr int kBlkRows = 8;

// LF maxima of stage `stage` of every block over its padded field
// (LF.py:10-11 under vmap): one workgroup per block, stored (not folded) into
// the block's StepScalars::amax[0][stage].
template <typename Real>
__global__ __launch_bounds__(256) void blocks_alpha_kernel(StepArgs<Real> a, int stage,
                                                           long long bs,
                                                           const Real* __restrict__ B) {
  __shared__ double red[8];
  const int b = blockIdx.x;
  const int P = (a.nx + 6) * (a.ny + 6);
  const Real* U = B + (size_t)b * bs;
  const Real* aux = U + (size_t)NU * P;
  const MechConst& c = *a.mc;
  double mx = 0.0, my = 0.0;
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    const Prim<double> q = cell_prim<double, true>(c, U + i, aux + i, P);
    mx = fmax(mx, nan_to_zero(fabs(q.u) + q.a));
    my = fmax(my, nan_to_zero(fabs(q.v) + q.a));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, off, 64));
    my = fmax(my, __shfl_xor(my, off, 64));
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[wid] = mx;
    red[4 + wid] = my;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double m = red[4 * threadIdx.x];
    for (int w = 1; w < 4; ++w) m = fmax(m, red[4 * threadIdx.x + w]);
    a.sc[b].amax[0][stage][threadIdx.x] = dbits(m);
  }
}

// One RK stage (MODE_RK) or the flux_func alone (MODE_RHS) of every block's
// interior: the stage kernel body on a block's padded planes, the block's
// own maxima.  Jobs (block, tile) in XCD-contiguous order.  bs / bo: block
// strides of the inputs and of the output (12 planes per block for MODE_RHS).
template <typename Real, int MODE, int TX, int WPE>
__global__ __launch_bounds__(Tile<TX>::NT) __attribute__((amdgpu_waves_per_eu(WPE)))
void blocks_stage_kernel(StepArgs<Real> a, int stage, int tiles_per_blk, long long bs,
                         long long bo, const Real* Bs, const Real* Bn, Real* Bo) {
  __shared__ StageLds<Real, TX> L;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int b = t / tiles_per_blk, tile = t - b * tiles_per_blk;
  const size_t P = (size_t)(a.nx + 6) * (a.ny + 6);
  StepArgs<Real> ab = a;
  ab.sc = a.sc + b;
  ab.parity = 0;
  const Real* Us = Bs + (size_t)b * bs;
  const Real* Un = Bn + (size_t)b * bs;
  Real* Uo = Bo + (size_t)b * bo;
  stage_body<Real, MODE, TX, RECON_WENO5, SPLIT_LF, Real, false, true>(
      ab, stage, Us, Us + NU * P, Un, Uo, Uo + NU * P, tile, L);
}

// The ghost layers of a stage's output blocks, as the reference computes
// them over the padded arrays (its rhs is zero there, TVD_RK3.py:21-35):
// U_out = the stage's combination of U_n and U_s, aux_out = update_aux from
// the stage input's T; MODE_RHS writes zeros (flux_func's jnp.pad).  Grid:
// (ring cells / 256, n_blk).
template <typename Real, int MODE>
__global__ __launch_bounds__(256) void blocks_ring_kernel(StepArgs<Real> a, int stage,
                                                          long long bs, long long bo,
                                                          const Real* __restrict__ Bs,
                                                          const Real* __restrict__ Bn,
                                                          Real* __restrict__ Bo) {
  __shared__ MixTables mix;
  const MechConst& c = *a.mc;
  if (MODE == MODE_RK) load_mix_tables(c, mix);
  __syncthreads();
  const int b = blockIdx.y;
  const int nx = a.nx, ny = a.ny, PY = ny + 6, P = (nx + 6) * PY;
  const int nring = P - nx * ny;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int fl = 0u;
  if (r < nring) {
    // rows 0-2 and nx+3..nx+5 whole, then the 6 ghost columns of rows 3..nx+2
    int px, py;
    if (r < 6 * PY) {
      const int row = r / PY;
      px = row < 3 ? row : nx + row;
      py = r - row * PY;
    } else {
      const int k = r - 6 * PY, row = k / 6, col = k - row * 6;
      px = 3 + row;
      py = col < 3 ? col : ny + col;
    }
    const int o = px * PY + py;
    Real* Uo = Bo + (size_t)b * bo;
    if (MODE == MODE_RHS) {
#pragma unroll
      for (int v = 0; v < NU; ++v) Uo[(size_t)v * P + o] = (Real)0;
    } else {
      const Real* Us = Bs + (size_t)b * bs;
      const Real* Un = Bn + (size_t)b * bs;
      double un[NU];
#pragma unroll
      for (int v = 0; v < NU; ++v) {
        const double us = (double)Us[(size_t)v * P + o];
        if (stage == 0) {
          un[v] = us + 0.0;
        } else {
          const double u0 = (double)Un[(size_t)v * P + o];
          if (stage == 1) un[v] = (3.0 / 4.0) * u0 + (1.0 / 4.0) * (us + 0.0);
          else un[v] = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (us + 0.0);
        }
        Uo[(size_t)v * P + o] = (Real)un[v];
        un[v] = (double)(Real)un[v];
      }
      const double rho = un[0], ir = rcp(rho);
      const double u = un[1] * ir, v = un[2] * ir;
      const double e = un[3] * ir - 0.5 * (u * u + v * v);
      double Y[JH_NS];
#pragma unroll
      for (int k = 0; k < JH_NS - 1; ++k) Y[k] = un[4 + k] * ir;
      const double R = fill_Y_get_R(c, Y);
      double T = (double)Us[(size_t)(NU + 1) * P + o];
      double gamma;
      const bool conv = halley_T(mix, Y, R, e, T, gamma);
      Uo[(size_t)NU * P + o] = (Real)gamma;
      Uo[(size_t)(NU + 1) * P + o] = (Real)T;
      fl = halley_flags(rho, e, R, T, gamma, conv);
    }
  }
  raise_flags(a.sc + b, fl);
}

// Chemistry and final update_aux of every block's interior, in place
// (set_simulator.py:107-111): blocks (n_blk, 14, nx, ny).  Grid: (cells /
// 256, n_blk).
template <typename Real>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void blocks_chem_kernel(StepArgs<Real> a, long long bs, Real* B) {
  __shared__ double red[4];
  __shared__ unsigned long long redb[4];
  __shared__ unsigned long long redg[kGridSlotsChem * 4];
  __shared__ MixTables mix;
  __shared__ GibbsTables gt;
  const int b = blockIdx.y;
  StepArgs<Real> ab = a;
  ab.sc = a.sc + b;
  ab.parity = 0;
  Real* U = B + (size_t)b * bs;
  chem_body<Real, CHEM_STEP>(ab, U, U + (size_t)NU * a.nx * a.ny, nullptr, nullptr, blockIdx.x,
                             red, redb, redg, mix, gt);
}

// Ghost refill of same-size blocks tiling one uniform level: out[b][v][p]
// (padded blocks, P = (bnx + 6)(bny + 6) cells each) = the interior cell
// idx[b P + p] of the level (flat index block * bnx * bny + x * bny + y) of
// plane v, times the cell's momentum sign (su for v = 1, sv for v = 2).
// The source is addressed by strides (block, plane, row; columns
// contiguous), so a padded stage output's interior view needs no copy.
// Grid: (ceil(n_blk P / 256), NV).
template <typename Real>
__global__ __launch_bounds__(256) void blocks_fill_kernel(
    const Real* __restrict__ src, long long sb, long long sv, long long sr, int bnx, int bny,
    const int* __restrict__ idx, const Real* __restrict__ su, const Real* __restrict__ svg,
    Real* __restrict__ out, int n_blk, int P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (i >= (long long)n_blk * P) return;
  const long long b = i / P, p = i - b * P;
  const int col = idx[i], B = bnx * bny;
  const int cb = col / B, loc = col - cb * B, lx = loc / bny, ly = loc - lx * bny;
  Real x = src[cb * sb + v * sv + lx * sr + ly];
  if (v == 1) x *= su[i];
  else if (v == 2) x *= svg[i];
  out[(b * NV + v) * P + p] = x;
}

}  // namespace jh
