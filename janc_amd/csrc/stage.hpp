// Device kernels of the SSP-RK3 reacting-Euler step (gfx950).
//
// Layout in HBM: the state is SoA, C-order (var, x, y) with y contiguous,
// exactly the reference's (12, nx, ny) / (2, nx, ny) arrays.  Ghost cells are
// not stored in the state: four ghost strips (left/right: (14, 3, ny),
// bottom/top: (14, 3, nx); vars 0-11 = U, 12-13 = aux) are filled once per RK
// stage by bc_kernel and read by the stage kernel's halo loads.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#include "chem.hpp"
#include "janc_types.hpp"
#include "recon.hpp"
#include "thermo.hpp"
#include "weno.hpp"

namespace jh {

constexpr int NU = 4 + (JH_NS - 1);   // conserved variables (12)
constexpr int kBcThreads = 64;        // bc_kernel block size
constexpr int kInletTeam = 16;        // bc_kernel lanes per injection-inlet line
constexpr int NV = NU + 2;            // + aux (gamma, T) in ghost strips

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double bitsd(unsigned long long b) {
  return __longlong_as_double((long long)b);
}

// Wave-then-block max of non-negative doubles (NaN handling by the caller),
// then one 64-bit atomicMax per block.  Every thread of the block must call.
// A slot only grows while kernels fold into it, so the atomic is skipped when
// a relaxed load already shows a value at least as large (a stale load is a
// lower bound): blocks that finish together would otherwise queue their
// atomics on the one address.
template <int NT>
__device__ __forceinline__ void block_max_atomic(double v, unsigned long long* dst,
                                                 double* red_lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_lds[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red_lds[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) m = fmax(m, red_lds[i]);
    const unsigned long long mb = dbits(m);
    if (mb > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(dst, mb);
  }
}

// Same reduction that lets a NaN through (CFL_2D uses jnp.min, which
// propagates NaN).  Values are reduced as bit patterns.
template <int NT>
__device__ __forceinline__ void block_max_bits_atomic(unsigned long long v,
                                                      unsigned long long* dst,
                                                      unsigned long long* red_lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_lds[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red_lds[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) m = red_lds[i] > m ? red_lds[i] : m;
    if (m > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(dst, m);
  }
}

// K maxima at once (bit patterns / counts), one atomicMax per value and
// block into dst[0..K): the grid-wide-stop slots of one producer.  red holds
// K * NT/64 values.  Every thread of the block must call.
template <int NT, int K>
__device__ __forceinline__ void block_max_bits_multi(unsigned long long (&v)[K],
                                                     unsigned long long* dst,
                                                     unsigned long long* red) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(v[k], off, 64);
      v[k] = o > v[k] ? o : v[k];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * NW + wid] = v[k];
  }
  __syncthreads();
  if ((int)threadIdx.x < K) {
    const int k = threadIdx.x;
    unsigned long long m = red[k * NW];
#pragma unroll
    for (int i = 1; i < NW; ++i) m = red[k * NW + i] > m ? red[k * NW + i] : m;
    if (m > __hip_atomic_load(dst + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(dst + k, m);
  }
}

__device__ __forceinline__ double nan_to_zero(double v) { return v == v ? v : 0.0; }

// Health flags of one temperature solve and the state it solved for
// (JH_FLAG_*): e and R carry a non-finite conserved value (a NaN residual
// ends the Halley loop at once and leaves T, gamma finite).
__device__ __forceinline__ unsigned int halley_flags(double rho, double e, double R, double T,
                                                     double gamma, bool conv) {
  unsigned int f = 0u;
  if (rho <= 0.0) f |= JH_FLAG_NEGATIVE_RHO;
  if (!(isfinite(e) && isfinite(R) && isfinite(T) && isfinite(gamma))) f |= JH_FLAG_NONFINITE;
  else if (!(T > 0.0)) f |= JH_FLAG_NONPOSITIVE_T;
  if (!conv) f |= JH_FLAG_HALLEY_NOCONV;
  return f;
}

// One atomicOr per wave that has flags to raise (flags are rare).
__device__ __forceinline__ void raise_flags(StepScalars* sc, unsigned int f) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off, 64);
  if (f && (threadIdx.x & 63) == 0) atomicOr(&sc->flags, f);
}

// Per-candidate contribution of one cell to the grid-wide-stop maxima
// v[off .. off + kNstar): the cell's value at its own stop (x, for N* = n) or
// after one more update (x1, for N* > n); nothing for N* < n (N* >= n for
// every cell).  off is a compile-time constant at every call site.
template <int K>
__device__ __forceinline__ void cand_max(unsigned long long (&v)[K], int off, int n, double x,
                                         double x1) {
  const unsigned long long b = dbits(x), b1 = dbits(x1);
#pragma unroll
  for (int j = 0; j < kNstar; ++j) {
    const unsigned long long c = j == n ? b : j > n ? b1 : 0ull;
    v[off + j] = c > v[off + j] ? c : v[off + j];
  }
}

// v[i] of a 4-entry per-side array with a runtime side index, as selects: a
// kernel that holds its arguments in registers (the batched kernels copy
// them from memory) would otherwise index them through scratch.
template <typename T>
__device__ __forceinline__ T pick4(const T (&v)[4], int i) {
  return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

// Block-cooperative copy of MechConst's Gibbs tables into t (LDS); the
// caller synchronises the block before reading t.
__device__ __forceinline__ void load_gibbs_tables(const MechConst& c, GibbsTables& t) {
  constexpr int n = 2 * JH_NA * 12;   // hc[2][NA][6], s[2][NA][5], logc[2][NA], contiguous
  static_assert(sizeof(GibbsTables) == (n + 2) * sizeof(double), "GibbsTables layout");
  const double* src = &c.hc[0][0][0];
  double* dst = &t.hc[0][0][0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  if (threadIdx.x == 0) {
    t.Tcr = c.Tcr;
    t.T0 = c.T0;
  }
}

// Block-cooperative copy of MechConst's mixture tables into t (LDS); the
// caller synchronises the block before reading t.
__device__ __forceinline__ void load_mix_tables(const MechConst& c, MixTables& t) {
  constexpr int n = 2 * JH_NS * 11;   // cp[2][NS][5] then h[2][NS][6], contiguous in both
  static_assert(sizeof(MixTables) == (n + 1) * sizeof(double), "MixTables layout");
  const double* src = &c.cp[0][0][0];
  double* dst = &t.cp[0][0][0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  if (threadIdx.x == 0) t.Tcr = c.Tcr;
}


// Primitive state of one cell (aux_func.py:39-51), in compute type W.
template <typename W>
struct Prim {
  W rho, u, v, p, a, g;       // g: gamma from aux (set with a)
};

template <typename W, bool WITH_A, typename Real>
__device__ __forceinline__ Prim<W> cell_prim(const MechConst& c, const Real* __restrict__ Up,
                                             const Real* __restrict__ ap, int stride) {
  Prim<W> q;
  q.rho = (W)Up[0];
  const W ir = rcp(q.rho);
  q.u = (W)Up[stride] * ir;
  q.v = (W)Up[2 * stride] * ir;
  W sum = (W)0, R = (W)0;
  W Y[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) {
    Y[k] = (W)Up[(4 + k) * stride] * ir;
    sum += Y[k];
  }
  Y[JH_NS - 1] = (W)1 - sum;
#pragma unroll
  for (int k = 0; k < JH_NS; ++k) R += (W)c.R[k] * Y[k];
  const W T = (W)ap[stride];
  q.p = (q.rho * R * T);
  q.a = (W)0;
  q.g = (W)0;
  if (WITH_A) {
    q.g = (W)ap[0];
    q.a = sqrt(q.g * q.p * ir);
  }
  return q;
}

// ------------------------------------------------------------- arguments

template <typename Real>
struct StepArgs {
  int nx, ny;
  double dx, dy;
  double cfl, dt_fixed;       // dt_fixed > 0 selects a fixed dt
  int parity;                 // step parity (slot set of StepScalars)
  const MechConst* __restrict__ mc;
  StepScalars* sc;
  Real* g[4];                 // ghost strips L, R, B, T (NV planes each)
  // host side: 1 selects the stage kernel that reads each halo cell from
  // its source cell (locate_src) instead of the strips (every side a copy kind)
  int mirror;
  int bc[4];
  double bc_p[4];             // pressure_outlet back pressure / inlet P0 per side
  double inj_T0;              // injection inlet: manifold total temperature
  double inj_Y[JH_NS - 1];    // injection inlet: injected mass fractions
  // Grid-wide Halley stop (null: each cell stops on its own residual).  A
  // producer of (gamma, T) writes, per cell, its update count (hnit) and the
  // (gamma, T) after one more update (hside, two planes), and reduces its N*
  // and per-candidate maxima into StepScalars::gs; halley_fix_kernel then
  // applies N*.
  Real* hside;
  unsigned char* hnit;
  // Chemistry clip branches of this step, one mask per cell (clip bits of
  // chem_increment): written by clip_mask_kernel, read by the replaying
  // chemistry kernel (jh_ctx_set_clip_branches); null otherwise.
  unsigned short* clipm;
};

template <typename Real>
__device__ __forceinline__ double step_dt(const StepArgs<Real>& a) {
  return a.dt_fixed > 0.0 ? a.dt_fixed
                          : a.cfl * (1.0 / bitsd(a.sc->cmax[a.parity]));
}

// ----------------------------------------------------- reset / reductions

// Zero the reduction slots of one step parity (blockDim.x >= 7).
__device__ __forceinline__ void zero_slots(StepScalars* sc, int parity) {
  const int t = threadIdx.x;
  if (t < 6) (&sc->amax[parity][0][0])[t] = 0ull;
  if (t == 6) sc->cmax[parity] = 0ull;
  constexpr int ng = (int)(sizeof(GridSlots) / 8) * JH_PRODUCERS;
  for (int i = t; i < ng; i += blockDim.x) (&sc->gs[parity][0].nstar)[i] = 0ull;
}

__global__ void reset_kernel(StepScalars* sc, int parity) { zero_slots(sc, parity); }

// Zero one producer's grid-wide-stop slots (update_aux alone).
__global__ void reset_grid_slot_kernel(StepScalars* sc, int parity, int prod) {
  constexpr int m = (int)(sizeof(GridSlots) / 8);
  if ((int)threadIdx.x < m) (&sc->gs[parity][prod].nstar)[threadIdx.x] = 0ull;
}

// Interior part of the stage-1 LF maxima and the CFL maximum of a state
// handed in by the caller (set_simulator.py:438-443, LF.py:10-11).
// Block blk of nblk (grid-stride over the cells).
template <typename Real>
__device__ __forceinline__ void reduce_body(const StepArgs<Real>& a, const Real* __restrict__ U,
                                            const Real* __restrict__ aux, int slot_parity,
                                            int with_cfl, int blk, int nblk, double* red,
                                            unsigned long long* redb) {
  const MechConst& c = *a.mc;
  const int n = a.nx * a.ny;
  double mx = 0.0, my = 0.0;
  unsigned long long cm = 0ull;
  for (int i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
    const Prim<double> q = cell_prim<double, true>(c, U + i, aux + i, n);
    const double sx = fabs(q.u) + q.a, sy = fabs(q.v) + q.a;
    mx = fmax(mx, nan_to_zero(sx));
    my = fmax(my, nan_to_zero(sy));
    const unsigned long long b = dbits(sx / a.dx + sy / a.dy);
    cm = b > cm ? b : cm;
  }
  block_max_atomic<256>(mx, &a.sc->amax[slot_parity][0][0], red);
  block_max_atomic<256>(my, &a.sc->amax[slot_parity][0][1], red);
  if (with_cfl) block_max_bits_atomic<256>(cm, &a.sc->cmax[slot_parity], redb);
}

template <typename Real>
__global__ __launch_bounds__(256) void reduce_kernel(StepArgs<Real> a, const Real* __restrict__ U,
                                                     const Real* __restrict__ aux, int slot_parity,
                                                     int with_cfl) {
  __shared__ double red[4];
  __shared__ unsigned long long redb[4];
  reduce_body<Real>(a, U, aux, slot_parity, with_cfl, blockIdx.x, gridDim.x, red, redb);
}

// The same maxima in two passes without same-address atomics: every block
// of a short kernel finishes at once, so a thousand blocks' atomicMax on
// three addresses serialise (~25 us at 1024 blocks).  Pass 1 writes each
// block's (max|u|+a, max|v|+a, CFL max) bit patterns to part[3 blk ..];
// reduce_final_kernel (one block) folds them into the slots.
template <typename Real>
__global__ __launch_bounds__(256) void reduce_part_kernel(StepArgs<Real> a,
                                                          const Real* __restrict__ U,
                                                          const Real* __restrict__ aux,
                                                          unsigned long long* __restrict__ part) {
  __shared__ unsigned long long red[3][4];
  const MechConst& c = *a.mc;
  const int n = a.nx * a.ny;
  unsigned long long v[3] = {0ull, 0ull, 0ull};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Prim<double> q = cell_prim<double, true>(c, U + i, aux + i, n);
    const double sx = fabs(q.u) + q.a, sy = fabs(q.v) + q.a;
    const unsigned long long bx = dbits(nan_to_zero(sx)), by = dbits(nan_to_zero(sy));
    const unsigned long long bc = dbits(sx / a.dx + sy / a.dy);
    v[0] = bx > v[0] ? bx : v[0];
    v[1] = by > v[1] ? by : v[1];
    v[2] = bc > v[2] ? bc : v[2];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(v[k], off, 64);
      v[k] = o > v[k] ? o : v[k];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) red[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    unsigned long long m = red[k][0];
#pragma unroll
    for (int w = 1; w < 4; ++w) m = red[k][w] > m ? red[k][w] : m;
    part[3 * blockIdx.x + k] = m;
  }
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const unsigned long long* __restrict__ part,
                                                           int nblk, StepScalars* sc, int parity,
                                                           int with_cfl) {
  __shared__ unsigned long long red[3][4];
  unsigned long long v[3] = {0ull, 0ull, 0ull};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = part[3 * b + k] > v[k] ? part[3 * b + k] : v[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(v[k], off, 64);
      v[k] = o > v[k] ? o : v[k];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) red[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    unsigned long long m = red[k][0];
#pragma unroll
    for (int w = 1; w < 4; ++w) m = red[k][w] > m ? red[k][w] : m;
    unsigned long long* dst = k < 2 ? &sc->amax[parity][0][k] : &sc->cmax[parity];
    if (k < 2 || with_cfl) atomicMax(dst, m);
  }
}

// ------------------------------------------------------------ ghost cells

// Pressure-feedback injector model (examples/janc_basic_example2.ipynb,
// inj_model; Fievisohn et al. 2017) with the manifold total state (P0, T0)
// explicit instead of the example's nondimensional 1.0: 20 bisection steps
// on the throat Mach number, then the chamber-side velocity V3 and density.
__device__ __forceinline__ void inj_model(double p, double R, double P0, double T0, double& V3,
                                          double& rho) {
  constexpr double g = 1.29, A1 = 1.0, A3 = 5.0, A2 = A3 - A1;
  const double C0 = sqrt(g * R * T0);
  const double P3 = p;
  double M1 = 0.0, M2 = 1.0, MFC = 0.0, V = 0.0;
  for (int i = 0; i < 20; ++i) {
    const double M = 0.5 * (M1 + M2);
    const double f = 1 + (g - 1) / 2 * (M * M);
    const double lf = log(f);           // the three powers of f from one log
    const double P1 = P0 * exp(-g / (g - 1) * lf);
    const double V1 = M * rsqrt(f) * C0;
    MFC = A1 * P0 / sqrt(T0) * sqrt(g / R) * M * exp(-(g + 1) / 2 / (g - 1) * lf);
    const double A = 0.5;
    const double B = g / (g - 1) * P3 * A3 / MFC;
    const double C = -g / (g - 1) * R * T0;
    V = (-B + sqrt(B * B - 4 * A * C)) / (2 * A);
    const double P2 = (MFC * (V - V1) - P1 * A1 + P3 * A3) / A2;
    if (P2 >= P1) M2 = M;
    else M1 = M;
  }
  V3 = V;
  rho = MFC / V / A3;
}

// inj_model for a team of TL lanes (a power of two, at most 64) sharing one
// boundary line.  The bisection's midpoints are dyadic rationals (exact in
// FP64), so the TL - 1 nodes of the next log2(TL) levels of the bisection
// tree below the current interval are evaluated at once, one per lane (team
// lane j = heap node j: level l = floor(log2 j), midpoint
// M1 + (M2 - M1)(2k + 1)/2^(l+1), k = j - 2^l), and the sequential decisions
// are replayed from the lanes' P2 >= P1 bits: the same midpoints and
// decisions as inj_model in 20 / log2(TL) passes instead of 20 steps.  V and
// MFC come from the lane that evaluated the last midpoint.  Every lane of a
// team calls it with the same arguments; other teams of the wave may not.
template <int TL>
__device__ __forceinline__ void inj_model_team(double p, double R, double P0, double T0,
                                               double& V3, double& rho) {
  static_assert(TL >= 2 && TL <= 64 && (TL & (TL - 1)) == 0, "team of 2..64 lanes");
  constexpr int LV = TL == 2 ? 1 : TL == 4 ? 2 : TL == 8 ? 3 : TL == 16 ? 4 : TL == 32 ? 5 : 6;
  constexpr double g = 1.29, A1 = 1.0, A3 = 5.0, A2 = A3 - A1;
  const double C0 = sqrt(g * R * T0);
  const double P3 = p;
  const int lane = threadIdx.x & 63, tl = lane & (TL - 1), base = lane - tl;
  double M1 = 0.0, M2 = 1.0, MFC = 0.0, V = 0.0;
  for (int done = 0; done < 20;) {
    const int lv = 20 - done < LV ? 20 - done : LV;   // levels in this pass
    bool ge = false;
    double Vl = 0.0, MFCl = 0.0;
    if (tl >= 1 && tl < (1 << lv)) {
      const int l = 31 - __clz(tl), k = tl - (1 << l);
      const double M = M1 + (M2 - M1) * ((double)(2 * k + 1) * ldexp(1.0, -(l + 1)));
      const double f = 1 + (g - 1) / 2 * (M * M);
      const double lf = log(f);
      const double P1 = P0 * exp(-g / (g - 1) * lf);
      const double V1 = M * rsqrt(f) * C0;
      MFCl = A1 * P0 / sqrt(T0) * sqrt(g / R) * M * exp(-(g + 1) / 2 / (g - 1) * lf);
      const double A = 0.5;
      const double B = g / (g - 1) * P3 * A3 / MFCl;
      const double C = -g / (g - 1) * R * T0;
      Vl = (-B + sqrt(B * B - 4 * A * C)) / (2 * A);
      const double P2 = (MFCl * (Vl - V1) - P1 * A1 + P3 * A3) / A2;
      ge = P2 >= P1;
    }
    const unsigned long long bits = __ballot(ge) >> base;
    int node = 1, last = 1;
    for (int l = 0; l < lv; ++l) {
      const double M = 0.5 * (M1 + M2);
      last = node;
      if ((bits >> node) & 1ull) {
        M2 = M;
        node = 2 * node;
      } else {
        M1 = M;
        node = 2 * node + 1;
      }
    }
    V = __shfl(Vl, base + last);
    MFC = __shfl(MFCl, base + last);
    done += lv;
  }
  V3 = V;
  rho = MFC / V / A3;
}

// Ghost state of an injection-inlet side from its wall cell ci
// (bottom_boundary of the same example): injected stream where the wall
// cell's pressure is below P0, else the wall cell with its normal velocity
// negated; the pressure is the wall cell's in both cases.
template <typename Real, bool TEAM = false>
__device__ __forceinline__ void inlet_ghost(const MechConst& c, const StepArgs<Real>& a,
                                            const Real* __restrict__ U,
                                            const Real* __restrict__ aux, int n, int ci,
                                            bool xside, bool low, double P0, double (&G)[NV]) {
  double Uc[NU];
#pragma unroll
  for (int v = 0; v < NU; ++v) Uc[v] = (double)U[v * n + ci];
  const double g_in = (double)aux[ci], T_in = (double)aux[n + ci];
  const double rho = Uc[0], u = Uc[1] / rho, vv = Uc[2] / rho;
  double Y[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Y[k] = Uc[4 + k] / rho;
  const double R = fill_Y_get_R(c, Y);
  const double p = rho * R * T_in;
  double Yi[JH_NS];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) Yi[k] = a.inj_Y[k];
  const double Ri = fill_Y_get_R(c, Yi);
  const bool blocked = p >= P0;
  double rho_c, un_c, ut_c, T_c, gam, h;
  const double un = xside ? u : vv, ut = xside ? vv : u;   // normal, tangential
  if (blocked) {
    rho_c = rho;
    un_c = -un;
    ut_c = ut;
    T_c = T_in;
    gam = g_in;
    (void)mix_gamma(c, Y, R, T_in, &h);
  } else {
    double V3, rho_m;
    if (TEAM) inj_model_team<kInletTeam>(p, Ri, P0, a.inj_T0, V3, rho_m);
    else inj_model(p, Ri, P0, a.inj_T0, V3, rho_m);
    T_c = p / (Ri * rho_m);
    gam = mix_gamma(c, Yi, Ri, T_c, &h);
    rho_c = p / (Ri * T_c);
    un_c = low ? V3 : -V3;
    ut_c = 0.0;
  }
  const double uc = xside ? un_c : ut_c, vc = xside ? ut_c : un_c;
  G[0] = rho_c;
  G[1] = rho_c * uc;
  G[2] = rho_c * vc;
  G[3] = rho_c * h - p + 0.5 * rho_c * (uc * uc + vc * vc);
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) G[4 + k] = rho_c * (blocked ? Y[k] : Yi[k]);
  G[NU] = gam;
  G[NU + 1] = T_c;
}

// LF maxima of a computed ghost state G (U, gamma, T) as the strip stores it
// (the arithmetic of cell_prim on the stored values).
template <typename Real>
__device__ __forceinline__ void ghost_maxima(const MechConst& c, const double (&G)[NV],
                                             double& mx, double& my) {
  Real Gr[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) Gr[v] = (Real)G[v];
  const Prim<double> q = cell_prim<double, true>(c, Gr, Gr + NU, 1);
  mx = fmax(mx, nan_to_zero(fabs(q.u) + q.a));
  my = fmax(my, nan_to_zero(fabs(q.v) + q.a));
}

// Ghost strips for one RK stage (boundary.py:126-132 with the BC kinds of
// boundary_conditions_2D/*.py and the periodic pre-pad of
// boundary_padding.py:20-25), plus the ghost cells' share of the LF maxima.
// One thread per boundary-normal line: threads [0, ny) left, [ny, 2ny) right,
// [2ny, 2ny+nx) bottom, [2ny+nx, 2ny+2nx) top.  blk = block index.
// TEAM_INLET: the lines of injection-inlet sides are left to the blocks from
// bc_line_blocks(nx, ny) on, kInletTeam lanes per line (inj_model_team),
// lines in side order left, right, bottom, top.
// Only computed ghost states (injection inlet, pressure outlet, callable
// strips) add to the maxima: a copied ghost (periodic, walls, zero gradient)
// has |u| + a and |v| + a of an interior cell, already in them.
__host__ __device__ __forceinline__ int bc_line_blocks(int nx, int ny) {
  return (2 * nx + 2 * ny + kBcThreads - 1) / kBcThreads;
}

template <typename Real, bool TEAM_INLET = false>
__device__ __forceinline__ void bc_body(const StepArgs<Real>& a, const Real* __restrict__ U,
                                        const Real* __restrict__ aux, int stage, int blk,
                                        double* red) {
  const MechConst& c = *a.mc;
  const int nx = a.nx, ny = a.ny, n = nx * ny;
  if (stage == 0 && blk == 0) {
    // the other parity's slots are written from here on in this step
    zero_slots(a.sc, a.parity ^ 1);
  }
  const int tid = blk * blockDim.x + threadIdx.x;
  double mx = 0.0, my = 0.0;
  int side = -1, pos = 0;
  const int line_blocks = bc_line_blocks(nx, ny);
  const bool team = TEAM_INLET && blk >= line_blocks;
  if (team) {
    int j = (blk - line_blocks) * (kBcThreads / kInletTeam) + (int)threadIdx.x / kInletTeam;
#pragma unroll
    for (int sd = 0; sd < 4; ++sd) {        // compile-time side index
      if (side < 0 && a.bc[sd] == JH_BC_INJECTION_INLET) {
        const int len = sd <= JH_RIGHT ? ny : nx;
        if (j < len) {
          side = sd;
          pos = j;
        } else {
          j -= len;
        }
      }
    }
  } else {
    if (tid < ny) { side = JH_LEFT; pos = tid; }
    else if (tid < 2 * ny) { side = JH_RIGHT; pos = tid - ny; }
    else if (tid < 2 * ny + nx) { side = JH_BOTTOM; pos = tid - 2 * ny; }
    else if (tid < 2 * ny + 2 * nx) { side = JH_TOP; pos = tid - 2 * ny - nx; }
    if (TEAM_INLET && side >= 0 && pick4(a.bc, side) == JH_BC_INJECTION_INLET) side = -1;
  }
  if (side >= 0) {
    const bool xside = side <= JH_RIGHT;
    const int len = xside ? ny : nx;          // strip row length
    const int N = xside ? nx : ny;            // cells along the normal
    const int gstride = 3 * len;              // plane stride in the strip
    Real* gs = pick4(a.g, side);
    const int kind = pick4(a.bc, side);
    // interior cell index (x, y) -> flat
    auto cell = [&](int along) -> int { return xside ? along * ny + pos : pos * ny + along; };
    if (kind != JH_BC_USER) {
      if (kind == JH_BC_PRESSURE_OUTLET) {
        // pressure_outlet.py:5-66: corrected wall-cell state, 3 copies
        const int w = (side == JH_LEFT || side == JH_BOTTOM) ? 0 : N - 1;
        const int ci = cell(w);
        double Uc[NU];
#pragma unroll
        for (int v = 0; v < NU; ++v) Uc[v] = (double)U[v * n + ci];
        const double g_out = (double)aux[ci], T_out = (double)aux[n + ci];
        const double rho = Uc[0], u = Uc[1] / rho, vv = Uc[2] / rho;
        double Y[JH_NS];
#pragma unroll
        for (int k = 0; k < JH_NS - 1; ++k) Y[k] = Uc[4 + k] / rho;
        const double R = fill_Y_get_R(c, Y);
        const double p_out = rho * (R * T_out);
        const double a_out = sqrt(g_out * p_out / rho);
        const double vn = side == JH_LEFT ? -(Uc[1] / Uc[0])
                        : side == JH_RIGHT ? Uc[1] / Uc[0]
                        : side == JH_BOTTOM ? -(Uc[2] / Uc[0]) : Uc[2] / Uc[0];
        const bool mask = (vn / a_out < 1.0);
        const double Pb = pick4(a.bc_p, side);
        const double rho_c = mask ? Pb / (p_out / rho) : rho;
        const double p_c = mask ? Pb : p_out;
        const double T_c = mask ? p_c / (rho_c * R) : T_out;
        double h;
        const double gam = mix_gamma(c, Y, R, T_c, &h);
        double G[NV];
        G[0] = rho_c;
        G[1] = rho_c * u;
        G[2] = rho_c * vv;
        G[3] = rho_c * h - p_c + 0.5 * rho_c * (u * u + vv * vv);
#pragma unroll
        for (int k = 0; k < JH_NS - 1; ++k) G[4 + k] = rho_c * Y[k];
        G[NU] = gam;
        G[NU + 1] = T_c;
        for (int gi = 0; gi < 3; ++gi) {
#pragma unroll
          for (int v = 0; v < NV; ++v) gs[v * gstride + gi * len + pos] = (Real)G[v];
        }
        ghost_maxima<Real>(c, G, mx, my);
      } else if (kind == JH_BC_INJECTION_INLET) {
        const bool low = (side == JH_LEFT || side == JH_BOTTOM);
        double G[NV];
        // team: the kInletTeam lanes of this line compute G together
        inlet_ghost<Real, TEAM_INLET>(c, a, U, aux, n, cell(low ? 0 : N - 1), xside, low,
                                      pick4(a.bc_p, side), G);
        if (!team || (threadIdx.x & (kInletTeam - 1)) == 0) {
          for (int gi = 0; gi < 3; ++gi) {
#pragma unroll
            for (int v = 0; v < NV; ++v) gs[v * gstride + gi * len + pos] = (Real)G[v];
          }
        }
        ghost_maxima<Real>(c, G, mx, my);
      } else {
        // every layer's values are loaded before any is stored, so the 3 x 14
        // loads are in flight together (interleaved with the strip stores,
        // which the compiler cannot prove disjoint from the state, each load
        // would wait for the one before)
        Real val[3][NV];
#pragma unroll
        for (int gi = 0; gi < 3; ++gi) {
          // source cell along the normal for ghost layer gi (gi = 0 is the
          // outermost on the low side, the innermost on the high side)
          const bool low = (side == JH_LEFT || side == JH_BOTTOM);
          int src;
          if (kind == JH_BC_PERIODIC) src = low ? N - 4 + gi : 1 + gi;
          else if (kind == JH_BC_SLIP_WALL) src = low ? 0 : N - 1;
          else src = low ? 2 - gi : N - 1 - gi;   // zero_gradient, no_slip_wall
          const int ci = cell(src);
#pragma unroll
          for (int v = 0; v < NU; ++v) {
            Real x = U[v * n + ci];
            if (kind == JH_BC_SLIP_WALL && v == (xside ? 1 : 2)) x = -x;
            if (kind == JH_BC_NO_SLIP_WALL && (v == 1 || v == 2)) x = -x;
            val[gi][v] = x;
          }
          val[gi][NU] = aux[ci];
          val[gi][NU + 1] = aux[n + ci];
        }
#pragma unroll
        for (int gi = 0; gi < 3; ++gi) {
#pragma unroll
          for (int v = 0; v < NV; ++v) gs[v * gstride + gi * len + pos] = val[gi][v];
        }
      }
    }
    if (kind == JH_BC_USER) {
      for (int gi = 0; gi < 3; ++gi) {
        const Real* gp = gs + gi * len + pos;
        const Prim<double> q = cell_prim<double, true>(c, gp, gp + NU * gstride, gstride);
        mx = fmax(mx, nan_to_zero(fabs(q.u) + q.a));
        my = fmax(my, nan_to_zero(fabs(q.v) + q.a));
      }
    }
  }
  block_max_atomic<kBcThreads>(mx, &a.sc->amax[a.parity][stage][0], red);
  block_max_atomic<kBcThreads>(my, &a.sc->amax[a.parity][stage][1], red);
}

template <typename Real>
__global__ __launch_bounds__(kBcThreads) void bc_kernel(StepArgs<Real> a, const Real* __restrict__ U,
                                                 const Real* __restrict__ aux, int stage) {
  __shared__ double red[4];
  bc_body<Real, true>(a, U, aux, stage, blockIdx.x, red);
}

// ------------------------------------------------------------ stage kernel

constexpr int TY = 64;            // tile width along y (one wave)

// Tile geometry for TX rows along x (TX = 8, 16 or 32; 32*TX threads).  The
// kernel's WPE template argument is the occupancy (waves per SIMD) the
// register allocation must allow: at most 512/WPE VGPRs per lane.
template <int TX>
struct Tile {
  static constexpr int NT = 32 * TX;             // threads per block
  static constexpr int NW = NT / 64;             // waves per block
  static constexpr int RX = TX + 6;              // x-extended region rows
  static constexpr int RY = TY + 6;              // y-extended region columns
  static constexpr int NXF = (TX + 1) * TY;      // x faces per tile
  static constexpr int NYF = TX * (TY + 1);      // y faces per tile
  static constexpr int NHS = (6 + NW - 1) / NW;  // x-halo rows per wave
  static constexpr int NSL = 2 + NHS + 1;        // region cells per thread
};

enum StageMode : int { MODE_RK = 0, MODE_RHS = 1 };

// Blocks go round-robin to the 8 XCDs; give each XCD a contiguous run of
// tiles (y fastest, then x) so a tile's x-neighbours, whose rows it re-reads
// as halo, are usually served by the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int xcd = b & 7, k = b >> 3, q = nb >> 3, r = nb & 7;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

// Address of one halo/interior cell: U base, aux base and plane stride.
template <typename Real>
struct CellRef {
  const Real* u;
  const Real* x;
  int stride;
  bool ok;
};

template <typename Real>
__device__ __forceinline__ CellRef<Real> locate(const StepArgs<Real>& a, const Real* __restrict__ U,
                                                const Real* __restrict__ aux, int gx, int gy) {
  const int nx = a.nx, ny = a.ny;
  CellRef<Real> r;
  r.ok = true;
  const bool inx = gx >= 0 && gx < nx, iny = gy >= 0 && gy < ny;
  if (inx && iny) {
    const int o = gx * ny + gy;
    r.u = U + o;
    r.x = aux + o;
    r.stride = nx * ny;
  } else if (iny && gx < 0 && gx >= -3) {
    r.stride = 3 * ny;
    r.u = a.g[JH_LEFT] + (gx + 3) * ny + gy;
    r.x = r.u + NU * r.stride;
  } else if (iny && gx >= nx && gx < nx + 3) {
    r.stride = 3 * ny;
    r.u = a.g[JH_RIGHT] + (gx - nx) * ny + gy;
    r.x = r.u + NU * r.stride;
  } else if (inx && gy < 0 && gy >= -3) {
    r.stride = 3 * nx;
    r.u = a.g[JH_BOTTOM] + (gy + 3) * nx + gx;
    r.x = r.u + NU * r.stride;
  } else if (inx && gy >= ny && gy < ny + 3) {
    r.stride = 3 * nx;
    r.u = a.g[JH_TOP] + (gy - ny) * nx + gx;
    r.x = r.u + NU * r.stride;
  } else {
    r.ok = false;
    r.u = U;
    r.x = aux;
    r.stride = 0;
  }
  return r;
}

// A cell of a block that carries its own 3 ghost layers (the AMR branch's
// blocks, (nx + 6) x (ny + 6) planes, interior at offset (3, 3)): cells
// within the ghost layers are read in place, cells beyond them do not exist.
// A halo cell of a copy-kind side (periodic with the reference's
// off-by-one, slip wall, zero gradient, no-slip wall) read from the interior
// cell bc_body copies into its ghost slot: the source cell of ghost layer gi
// (gi = 0 outermost on the low side, innermost on the high side) and, in
// flip, the momentum components the ghost negates (bit 0: rho u, bit 1:
// rho v).  Negation is exact, so the values equal the strip's.
__device__ __forceinline__ int ghost_src(int kind, bool low, int N, int gi) {
  if (kind == JH_BC_PERIODIC) return low ? N - 4 + gi : 1 + gi;
  if (kind == JH_BC_SLIP_WALL) return low ? 0 : N - 1;
  return low ? 2 - gi : N - 1 - gi;   // zero gradient, no-slip wall
}

__device__ __forceinline__ unsigned int ghost_flip(int kind, bool xside) {
  if (kind == JH_BC_SLIP_WALL) return xside ? 1u : 2u;
  return kind == JH_BC_NO_SLIP_WALL ? 3u : 0u;
}

template <typename Real>
__device__ __forceinline__ CellRef<Real> locate_src(const StepArgs<Real>& a,
                                                    const Real* __restrict__ U,
                                                    const Real* __restrict__ aux, int gx, int gy,
                                                    unsigned int& flip) {
  const int nx = a.nx, ny = a.ny;
  const bool inx = gx >= 0 && gx < nx, iny = gy >= 0 && gy < ny;
  CellRef<Real> r;
  r.ok = (inx || (gx >= -3 && gx < nx + 3)) && (iny || (gy >= -3 && gy < ny + 3)) &&
         (inx || iny);
  r.stride = nx * ny;
  flip = 0u;
  int sx = gx, sy = gy;
  if (!r.ok) {
    sx = 0;
    sy = 0;
  } else if (!inx) {
    const bool low = gx < 0;
    const int kind = low ? a.bc[JH_LEFT] : a.bc[JH_RIGHT];
    sx = ghost_src(kind, low, nx, low ? gx + 3 : gx - nx);
    flip = ghost_flip(kind, true);
  } else if (!iny) {
    const bool low = gy < 0;
    const int kind = low ? a.bc[JH_BOTTOM] : a.bc[JH_TOP];
    sy = ghost_src(kind, low, ny, low ? gy + 3 : gy - ny);
    flip = ghost_flip(kind, false);
  }
  const int o = sx * ny + sy;
  r.u = U + o;
  r.x = aux + o;
  return r;
}

template <typename Real>
__device__ __forceinline__ CellRef<Real> locate_padded(const StepArgs<Real>& a,
                                                       const Real* __restrict__ U,
                                                       const Real* __restrict__ aux, int gx,
                                                       int gy) {
  const int py = a.ny + 6;
  CellRef<Real> r;
  r.ok = gx >= -3 && gx < a.nx + 3 && gy >= -3 && gy < a.ny + 3;
  const int o = r.ok ? (gx + 3) * py + gy + 3 : 0;
  r.u = U + o;
  r.x = aux + o;
  r.stride = (a.nx + 6) * py;
  return r;
}

// Physical flux of component cv in direction x (dir 0) or y (dir 1) at a
// cell, LF.py:12 (rho*u, rho*u^2+p, rho*u*v, u*(rhoE+p), rho*u*Y).
template <typename W>
__device__ __forceinline__ W phys_flux(int cv, int dir, const Prim<W>& q, W ucv) {
  const W un = dir == 0 ? q.u : q.v;
  switch (cv) {
    case 0: return q.rho * un;
    case 1: return dir == 0 ? q.rho * (q.u * q.u) + q.p : (q.rho * q.u) * q.v;
    case 2: return dir == 0 ? (q.rho * q.u) * q.v : q.rho * (q.v * q.v) + q.p;
    case 3: return un * (ucv + q.p);
    default: return un * ucv;   // rho*u*Y with Y = rhoY/rho
  }
}

enum Split : int { SPLIT_LF = 0, SPLIT_SW = 1 };

// Steger-Warming split fluxes of component cv in direction x (dir 0) or y
// (SW.py:4-38): H+- = sum_k lambda_k+- H_k over the eigenvalues
// lambda = (un - a, un, un + a), lambda+- = (lambda +- sqrt(lambda^2 + eps^2))/2,
// eps = 1e-6, with H1, H4 = (1/(2 gamma)) [rho, rho u -+ rho a zx, rho v -+ rho a zy,
// rho E + p -+ rho a un, rho Y] and H2 = ((gamma-1)/gamma) [rho, rho u, rho v,
// rho |u|^2 / 2, rho Y]; gamma is the aux value, a = sqrt(gamma p / rho).
template <typename W>
__device__ __forceinline__ void sw_split(int cv, int dir, const Prim<W>& q, W ucv, W& hp,
                                         W& hm) {
  constexpr W eps2 = (W)(1e-6 * 1e-6);
  const W un = dir == 0 ? q.u : q.v;
  const W l1 = un - q.a, l2 = un, l4 = un + q.a;
  const W s1 = sqrt(l1 * l1 + eps2), s2 = sqrt(l2 * l2 + eps2), s4 = sqrt(l4 * l4 + eps2);
  const W ig = (W)1 / ((W)2 * q.g), g2 = (q.g - (W)1) / q.g;
  const W ra = q.rho * q.a;
  W h1, h2, h4;
  switch (cv) {
    case 0:
      h1 = h2 = h4 = q.rho;
      break;
    case 1: {
      const W ru = q.rho * q.u, d = dir == 0 ? ra : (W)0;
      h1 = ru - d; h2 = ru; h4 = ru + d;
      break;
    }
    case 2: {
      const W rv = q.rho * q.v, d = dir == 1 ? ra : (W)0;
      h1 = rv - d; h2 = rv; h4 = rv + d;
      break;
    }
    case 3:
      h1 = ucv + q.p - ra * un;
      h2 = (W)0.5 * q.rho * (q.u * q.u + q.v * q.v);
      h4 = ucv + q.p + ra * un;
      break;
    default:   // rho * Y with Y = rhoY / rho
      h1 = h2 = h4 = q.rho * (ucv / q.rho);
  }
  const W H1 = ig * h1, H2 = g2 * h2, H4 = ig * h4;
  hp = (W)0.5 * (l1 + s1) * H1 + (W)0.5 * (l2 + s2) * H2 + (W)0.5 * (l4 + s4) * H4;
  hm = (W)0.5 * (l1 - s1) * H1 + (W)0.5 * (l2 - s2) * H2 + (W)0.5 * (l4 - s4) * H4;
}

// One SSP-RK3 stage (TVD_RK3.py:5-12) for a 64 x TX tile:
//   rhs = dt * -(dF/dx + dG/dy)   (set_simulator.py:44-47, flux.py:56-68)
//   stage 0: U1 = U + rhs; stage 1: U2 = 3/4 U + 1/4 (U1 + rhs);
//   stage 2: U3 = 1/3 U + 2/3 (U2 + rhs)
// followed by update_aux (Halley) of the new state and the next stage's LF
// maxima.  MODE_RHS writes rhs only (the reference's flux_func).
//
// Work split per conserved variable: phase 1 evaluates the split fluxes
// H+- = (F +- alpha U)/2 of the tile + halo into LDS, phase 2 reconstructs the
// (TX+1)*64 x-faces and TX*65 y-faces (L + R WENO each), phase 3 takes the
// divergence and the RK combine for the thread's two cells.  The next
// variable's loads are issued before phase 2 so their latency is hidden.
//
// Region cells of thread t (wave w, lane ly): slots 0,1 = interior rows w and
// w+NW; slots 2..2+NHS-1 = x-halo rows {0,1,2,TX+3,TX+4,TX+5}[w + h*NW];
// last slot = y-halo cell t (t < 6*TX).
// LDS of one stage-kernel block.
template <typename Real, int TX>
struct StageLds {
  Real Hxp[TX + 6][TY], Hxm[TX + 6][TY];
  Real Hyp[TX][TY + 6], Hym[TX][TY + 6];
  Real Fx[TX + 1][TY];
  Real Fy[TX][TY + 1];
  double red[Tile<TX>::NW];
  MixTables mix;              // NASA-7 tables for the Halley epilogue
};

// One tile (index `tile` in y-fastest order) of one RK stage; RECON selects
// the reconstruction of the split fluxes (recon.hpp), SPLIT the splitting,
// W the arithmetic type of the split fluxes, reconstructions, divergence and
// RK combine (Real, the storage type, unless stated); GRID the grid-wide
// temperature-solve stop (StepArgs::hside set), compiled separately so the
// per-cell kernels keep their register budget; PAD the state planes are one
// AMR block with its own ghost layers ((nx + 6) x (ny + 6), no ghost strips,
// no next-stage maxima: the next stage's ghosts come from the caller).
template <typename Real, int MODE, int TX, int RECON = RECON_WENO5, int SPLIT = SPLIT_LF,
          typename W = Real, bool GRID = false, bool PAD = false, bool MIRROR = false>
__device__ __forceinline__ void stage_body(const StepArgs<Real>& a, int stage,
                                           const Real* __restrict__ Us,
                                           const Real* __restrict__ auxs, const Real* Un,
                                           Real* Uo, Real* __restrict__ auxo, int tile,
                                           StageLds<W, TX>& L) {
  using G = Tile<TX>;
  constexpr int NT = G::NT, NW = G::NW;
  constexpr int NXF = G::NXF, NYF = G::NYF, NHS = G::NHS, NSL = G::NSL;
  auto& sHxp = L.Hxp;
  auto& sHxm = L.Hxm;
  auto& sHyp = L.Hyp;
  auto& sHym = L.Hym;
  auto& sFx = L.Fx;
  auto& sFy = L.Fy;
  double* red = L.red;
  (void)NW;
  // face runs for the FP32 WENO5 (the FP64 form's carried differences would
  // push the kernel past its register budget)
  constexpr bool kRuns = RECON == RECON_WENO5 && std::is_same<W, float>::value;
  static_assert(!kRuns || (TX % 4 == 0 && TX <= 64), "face runs need TX a multiple of 4");

  const MechConst& c = *a.mc;
  if (MODE != MODE_RHS) load_mix_tables(c, L.mix);   // read after the phase barriers
  const int nx = a.nx, ny = a.ny, n = PAD ? (nx + 6) * (ny + 6) : nx * ny;
  // the wave index through readfirstlane: the tile rows, slots and face
  // runs derived from it are wave-uniform to the compiler (scalar registers
  // and branches)
  const int t = threadIdx.x, ly = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  if constexpr (MIRROR) {
    // with the halo read from the source cells no bc_kernel runs, so the
    // first stage zeroes the other parity's slots in its place
    if (stage == 0 && tile == 0) zero_slots(a.sc, a.parity ^ 1);
  }
  const int tiles_y = (ny + TY - 1) / TY;
  const int x0 = (tile / tiles_y) * TX;
  const int y0 = (tile % tiles_y) * TY;
  const double dt = step_dt(a);
  const double idx = 1.0 / a.dx, idy = 1.0 / a.dy;
  const W ax = (W)(Real)bitsd(a.sc->amax[a.parity][stage][0]);
  const W ay = (W)(Real)bitsd(a.sc->amax[a.parity][stage][1]);
  // H+- = (F +- alpha U) / 2 as fma(+-alpha/2, U, F/2): halving is exact, so
  // this is the same value with one operation fewer per pair
  const W hax = (W)0.5 * ax, hay = (W)0.5 * ay;

  int rrow[NSL], rcol[NSL];       // LDS coordinates: x-region row / y-region col
  int gxs[NSL], gys[NSL];
  bool has[NSL];
  rrow[0] = w + 3;
  rrow[1] = w + NW + 3;
#pragma unroll
  for (int h = 0; h < NHS; ++h) {
    const int hi = w + h * NW;
    rrow[2 + h] = hi < 6 ? (hi < 3 ? hi : TX + hi) : -1;
  }
#pragma unroll
  for (int s = 0; s < NSL - 1; ++s) {
    has[s] = rrow[s] >= 0;
    gxs[s] = x0 - 3 + rrow[s];
    gys[s] = y0 + ly;
    rcol[s] = ly + 3;
  }
  {
    constexpr int s = NSL - 1;
    has[s] = t < 6 * TX;
    const int xr = t / 6, pos = t % 6;
    const int col = pos < 3 ? pos : TY + pos;   // 0,1,2 or 67,68,69
    rrow[s] = xr;                               // y-region row (interior x)
    rcol[s] = col;
    gxs[s] = x0 + xr;
    gys[s] = y0 - 3 + col;
  }
  CellRef<Real> ref[NSL];
  Prim<W> q[NSL];
  Real nxt[NSL];
  // bits 2s, 2s + 1: slot s is a ghost whose rho u, rho v its side negates
  // (halo read from the source cells)
  unsigned int flips = 0u;
#pragma unroll
  for (int s = 0; s < NSL; ++s) {
    unsigned int f = 0u;
    if constexpr (PAD) ref[s] = locate_padded<Real>(a, Us, auxs, gxs[s], gys[s]);
    else if constexpr (MIRROR) ref[s] = locate_src<Real>(a, Us, auxs, gxs[s], gys[s], f);
    else ref[s] = locate<Real>(a, Us, auxs, gxs[s], gys[s]);
    if (!has[s]) ref[s].ok = false;
    if (ref[s].ok) {
      q[s] = cell_prim<W, SPLIT == SPLIT_SW, Real>(c, ref[s].u, ref[s].x, ref[s].stride);
      if (f & 1u) q[s].u = -q[s].u;
      if (f & 2u) q[s].v = -q[s].v;
      flips |= f << (2 * s);
      nxt[s] = ref[s].u[0];
    } else {
      q[s].rho = (W)1; q[s].u = (W)0; q[s].v = (W)0; q[s].p = (W)0; q[s].a = (W)0;
      q[s].g = (W)1;
      nxt[s] = (Real)0;
    }
  }
  bool own[2];
  int oidx[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    own[s] = gxs[s] < nx && gys[s] < ny;
    oidx[s] = PAD ? (gxs[s] + 3) * (ny + 6) + gys[s] + 3 : gxs[s] * ny + gys[s];
  }

  // One conserved variable; CLS = min(cv, 4) selects the physical flux
  // (and the Steger-Warming split) at compile time, so phase 1 carries no
  // per-slot selection on cv.  Variables 0-3 are instantiated once each,
  // the species share one loop body.
  const auto var_step = [&](auto K, const int cv) {
    constexpr int CLS = decltype(K)::value;
    // ---- phase 1: split fluxes into LDS ----
    W uc_own[2];
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      if (!has[s]) continue;
      W ucv = (W)nxt[s];
      if constexpr (CLS == 1 || CLS == 2) {
        if ((flips >> (2 * s + CLS - 1)) & 1u) ucv = -ucv;
      }
      if (s < 2) uc_own[s] = ucv;
      const bool xreg = s < NSL - 1, yreg = (s < 2) || (s == NSL - 1);
      if (xreg) {
        W hp = (W)0, hm = (W)0;
        if (ref[s].ok) {
          if (SPLIT == SPLIT_SW) {
            sw_split<W>(CLS, 0, q[s], ucv, hp, hm);
          } else {
            const W hF = (W)0.5 * phys_flux<W>(CLS, 0, q[s], ucv);
            hp = fma(hax, ucv, hF);
            hm = fma(-hax, ucv, hF);
          }
        }
        sHxp[rrow[s]][ly] = hp;
        sHxm[rrow[s]][ly] = hm;
      }
      if (yreg) {
        W hp = (W)0, hm = (W)0;
        if (ref[s].ok) {
          if (SPLIT == SPLIT_SW) {
            sw_split<W>(CLS, 1, q[s], ucv, hp, hm);
          } else {
            const W hG = (W)0.5 * phys_flux<W>(CLS, 1, q[s], ucv);
            hp = fma(hay, ucv, hG);
            hm = fma(-hay, ucv, hG);
          }
        }
        const int yr = s < 2 ? rrow[s] - 3 : rrow[s];
        sHyp[yr][rcol[s]] = hp;
        sHym[yr][rcol[s]] = hm;
      }
    }
    __syncthreads();
    // prefetch the next variable (consumed after two barriers) and this
    // variable's U^n for the RK combine (consumed in phase 3)
    if (cv + 1 < NU) {
#pragma unroll
      for (int s = 0; s < NSL; ++s)
        if (ref[s].ok) nxt[s] = ref[s].u[(cv + 1) * ref[s].stride];
    }
    W u0n[2] = {(W)0, (W)0};
    if (MODE != MODE_RHS && stage != 0) {
#pragma unroll
      for (int s = 0; s < 2; ++s)
        if (own[s]) u0n[s] = (W)Un[cv * n + oidx[s]];
    }
    // ---- phase 2: face reconstructions ----
    if constexpr (kRuns) {
      // FP32 WENO5: runs of consecutive faces (weno5_run).  The first half
      // of the block takes the x faces: lane = column, wave = run (TX / 4
      // runs of four per column, the last a run of five through face row
      // TX); its first wave also takes the y faces of column 64 (lane =
      // row), so no y wave runs a single-face pass for a few of its lanes.
      // The second half the y faces 0..63: 16 runs of four per row, run =
      // u % 16, row = u / 16.
      constexpr int NXT = NT / 2;
      static_assert(NXT % 64 == 0, "the x-face half of the block is whole waves");
      if (w < NXT / 64) {
        const int col = ly, fr0 = w * 4;
        if (fr0 + 4 == TX) {   // the last run also takes face row TX (wave-uniform)
          weno5_run<5, false, false>(&sHxp[fr0][col], TY, &sFx[fr0][col], TY);
          weno5_run<5, true, true>(&sHxm[fr0 + 1][col], TY, &sFx[fr0][col], TY);
        } else {
          weno5_run<4, false, false>(&sHxp[fr0][col], TY, &sFx[fr0][col], TY);
          weno5_run<4, true, true>(&sHxm[fr0 + 1][col], TY, &sFx[fr0][col], TY);
        }
        if (fr0 == 0 && col < TX) {
          const int row = col;
          sFy[row][TY] = weno5(sHyp[row][TY], sHyp[row][TY + 1], sHyp[row][TY + 2],
                               sHyp[row][TY + 3], sHyp[row][TY + 4]) +
                         weno5(sHym[row][TY + 5], sHym[row][TY + 4], sHym[row][TY + 3],
                               sHym[row][TY + 2], sHym[row][TY + 1]);
        }
      } else {
        const int u = t - NXT, row = u >> 4, fc0 = (u & 15) * 4;
        weno5_run<4, false, false>(&sHyp[row][fc0], 1, &sFy[row][fc0], 1);
        weno5_run<4, true, true>(&sHym[row][fc0 + 1], 1, &sFy[row][fc0], 1);
      }
    } else
    for (int it = t; it < NXF + NYF; it += NT) {
      if (it < NXF) {
        const int fr = it >> 6, col = it & 63;
        const W L = recon_L<RECON, W>(sHxp[fr][col], sHxp[fr + 1][col], sHxp[fr + 2][col],
                                            sHxp[fr + 3][col], sHxp[fr + 4][col]);
        const W R = recon_R<RECON, W>(sHxm[fr + 1][col], sHxm[fr + 2][col],
                                            sHxm[fr + 3][col], sHxm[fr + 4][col],
                                            sHxm[fr + 5][col]);
        sFx[fr][col] = L + R;
      } else {
        const int j = it - NXF;
        const int xr = j / (TY + 1), fc = j - xr * (TY + 1);
        const W L = recon_L<RECON, W>(sHyp[xr][fc], sHyp[xr][fc + 1], sHyp[xr][fc + 2],
                                            sHyp[xr][fc + 3], sHyp[xr][fc + 4]);
        const W R = recon_R<RECON, W>(sHym[xr][fc + 1], sHym[xr][fc + 2],
                                            sHym[xr][fc + 3], sHym[xr][fc + 4],
                                            sHym[xr][fc + 5]);
        sFy[xr][fc] = L + R;
      }
    }
    __syncthreads();
    // ---- phase 3: divergence + RK combine ----
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ir = rrow[s] - 3;
      // FP64 whatever W: in FP32 the stage increment, often below half an
      // ulp of the state, would be lost to each intermediate rounding and
      // the mixed form's error would grow systematically
      const double dF = (double)sFx[ir + 1][ly] - (double)sFx[ir][ly];
      const double dG = (double)sFy[ir][ly + 1] - (double)sFy[ir][ly];
      const double rhs = dt * (-(dF * idx + dG * idy));
      double un;
      if (MODE == MODE_RHS) {
        un = rhs;
      } else if (stage == 0) {
        un = (double)uc_own[s] + rhs;
      } else {
        const double u0 = (double)u0n[s], uc = (double)uc_own[s];
        if (stage == 1) un = (3.0 / 4.0) * u0 + (1.0 / 4.0) * (uc + rhs);
        else un = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (uc + rhs);
      }
      if (own[s]) Uo[cv * n + oidx[s]] = (Real)un;
    }
    // (no explicit drain of the memory counters here: with the loop
    // specialised per variable class the compiler's own waits are as good,
    // measured 0.8 % faster without one)
  };
  var_step(std::integral_constant<int, 0>{}, 0);
  var_step(std::integral_constant<int, 1>{}, 1);
  var_step(std::integral_constant<int, 2>{}, 2);
  var_step(std::integral_constant<int, 3>{}, 3);
#pragma unroll 1
  for (int cv = 4; cv < NU; ++cv) var_step(std::integral_constant<int, 4>{}, cv);
  if (MODE == MODE_RHS) return;

  // ---- update_aux (aux_func.py:21-29) + next stage's LF maxima ----
  // The new state is read back from this thread's own stores (a runtime
  // component index into a register array would live in scratch).
  constexpr bool grid = GRID;
  double mx = 0.0, my = 0.0;
  unsigned long long gv[kGridSlotsStage];   // grid-wide stop: N*, x and y candidates
#pragma unroll
  for (int k = 0; k < kGridSlotsStage; ++k) gv[k] = 0ull;
  unsigned int fl = 0u;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    if (!own[s]) continue;
    const Real* un = Uo + oidx[s];
    const double rho = (double)un[0], ir = rcp(rho);
    const double u = (double)un[n] * ir, v = (double)un[2 * n] * ir;
    const double e = (double)un[3 * n] * ir - 0.5 * (u * u + v * v);
    double Y[JH_NS];
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = (double)un[(4 + k) * n] * ir;
    const double R = fill_Y_get_R(c, Y);
    double T = (double)auxs[n + oidx[s]];
    double gamma;
    HalleyMore hm{0, 0.0, 0.0, true, true};
    bool conv;
    if constexpr (grid) {
      conv = halley_T<MixTables, true>(L.mix, Y, R, e, T, gamma, &hm);
      a.hside[oidx[s]] = (Real)hm.g1;
      a.hside[n + oidx[s]] = (Real)hm.T1;
      a.hnit[oidx[s]] = (unsigned char)hm.n;
      if (hm.n < kHalleyMaxIter && !hm.ok1) fl |= JH_FLAG_HALLEY_REGRESS;
      gv[0] = (unsigned long long)hm.n > gv[0] ? (unsigned long long)hm.n : gv[0];
    } else {
      conv = halley_T(L.mix, Y, R, e, T, gamma);
    }
    auxo[oidx[s]] = (Real)gamma;
    auxo[n + oidx[s]] = (Real)T;
    fl |= halley_flags(rho, e, R, T, gamma, conv);
    if (!PAD && stage < 2) {
      // next stage's alpha from the stored (rounded) values
      const double gs = (double)(Real)gamma, Ts = (double)(Real)T;
      const double p = rho * R * Ts;
      const double snd = sqrt(gs * p * ir);
      const double sx = nan_to_zero(fabs(u) + snd), sy = nan_to_zero(fabs(v) + snd);
      if constexpr (grid) {
        const double g1 = (double)(Real)hm.g1, T1 = (double)(Real)hm.T1;
        const double p1 = rho * R * T1;
        const double snd1 = sqrt(g1 * p1 * ir);
        cand_max(gv, 1, hm.n, sx, nan_to_zero(fabs(u) + snd1));
        cand_max(gv, 1 + kNstar, hm.n, sy, nan_to_zero(fabs(v) + snd1));
      } else {
        mx = fmax(mx, sx);
        my = fmax(my, sy);
      }
    }
  }
  raise_flags(a.sc, fl);
  if constexpr (grid) {
    // the H arrays are free once every thread has passed the last phase 3
    // (the reduction's first barrier)
    block_max_bits_multi<NT, kGridSlotsStage>(gv, &a.sc->gs[a.parity][stage].nstar,
                                              reinterpret_cast<unsigned long long*>(&L.Hxp[0][0]));
  } else if (!PAD && stage < 2) {
    block_max_atomic<NT>(mx, &a.sc->amax[a.parity][stage + 1][0], red);
    block_max_atomic<NT>(my, &a.sc->amax[a.parity][stage + 1][1], red);
  }
}

template <typename Real, int MODE, int TX, int WPE, int RECON, int SPLIT = SPLIT_LF,
          typename W = Real, bool GRID = false, bool MIRROR = false>
__global__ __launch_bounds__(Tile<TX>::NT) __attribute__((amdgpu_waves_per_eu(WPE)))
void stage_kernel(StepArgs<Real> a, int stage, const Real* __restrict__ Us,
                  const Real* __restrict__ auxs, const Real* Un, Real* Uo,
                  Real* __restrict__ auxo) {
  __shared__ StageLds<W, TX> L;
  stage_body<Real, MODE, TX, RECON, SPLIT, W, GRID, false, MIRROR>(a, stage, Us, auxs, Un, Uo, auxo,
                                                    xcd_tile(blockIdx.x, gridDim.x), L);
}

// ---------------------------------------------------------- chemistry

enum ChemMode : int { CHEM_STEP = 0, CHEM_SOURCE = 1 };

// chem_increment_ld's state read from the conserved planes of cell i.
template <typename Real>
struct MemState {
  const Real* U;
  int n, i;
  __device__ double rho() const { return (double)U[i]; }
  __device__ double operator()(int k) const { return (double)(U + (size_t)(4 + k) * n)[i]; }
};

// Point-implicit chemistry after RK3 (set_simulator.py:117-122):
//   U <- U + reaction_source_terms(U, aux, dt); aux <- update_aux(U, aux)
// plus the next step's CFL and stage-1 LF maxima.  CHEM_SOURCE writes the
// source S (12 planes, first 4 zero) instead.  GRID: the grid-wide Halley
// stop's epilogue (a.hside set), compiled apart so that the per-cell form
// keeps its registers.
template <typename Real, int MODE, bool REPLAY = false, bool GRID = false>
__device__ __forceinline__ void chem_body(const StepArgs<Real>& a, Real* U, Real* aux, Real* S,
                                          double* dt_hist, int blk, double* red,
                                          unsigned long long* redb, unsigned long long* redg,
                                          MixTables& mix, GibbsTables& gt) {
  const MechConst& c = *a.mc;
  if (MODE == CHEM_STEP) load_mix_tables(c, mix);
  load_gibbs_tables(c, gt);
  __syncthreads();
  const int n = a.nx * a.ny;
  const double dt = step_dt(a);
  constexpr bool grid = GRID;
  double mx = 0.0, my = 0.0;
  unsigned long long cm = 0ull;
  unsigned long long gv[kGridSlotsChem];   // grid-wide stop: N*, x, y and CFL candidates
#pragma unroll
  for (int k = 0; k < kGridSlotsChem; ++k) gv[k] = 0ull;
  unsigned int fl = 0u;
  // one cell per thread: a grid-stride loop would let the compiler keep every
  // mechanism constant in registers across iterations
  const int i = blk * blockDim.x + threadIdx.x;
  if (i < n) {
    // the state is read before the Jacobian and again after the solve
    // instead of being held across it
    const MemState<Real> ld{U, n, i};
    double d[JH_NA];
    bool ok;
    // Jacobian storage: FP32 in the mixed-precision form (Real = float)
    if constexpr (REPLAY) {
      const unsigned int br = a.clipm[i];
      ok = chem_increment_ld<Real>(c, gt, ld, (double)aux[n + i], dt, d, nullptr, &br);
    } else {
      ok = chem_increment_ld<Real>(c, gt, ld, (double)aux[n + i], dt, d);
    }
    double T = (double)aux[n + i];
    if (MODE == CHEM_SOURCE) {
#pragma unroll
      for (int v = 0; v < 4; ++v) S[v * n + i] = (Real)0;
#pragma unroll
      for (int k = 0; k < JH_NA; ++k) S[(4 + k) * n + i] = (Real)d[k];
      if (!ok) atomicOr(&a.sc->flags, (unsigned)JH_FLAG_SINGULAR_PIVOT);
      return;
    }
    if (!ok) fl |= JH_FLAG_SINGULAR_PIVOT;
    Real Un[NU];
#pragma unroll
    for (int v = 0; v < 4; ++v) Un[v] = U[v * n + i];
#pragma unroll
    for (int k = 0; k < JH_NA; ++k) {
      Un[4 + k] = (Real)((double)U[(4 + k) * n + i] + d[k]);
      U[(4 + k) * n + i] = Un[4 + k];
    }
    const double r2 = (double)Un[0], ir = rcp(r2);
    const double u = (double)Un[1] * ir, v = (double)Un[2] * ir;
    const double e = (double)Un[3] * ir - 0.5 * (u * u + v * v);
    double Y[JH_NS];
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = (double)Un[4 + k] * ir;
    const double R = fill_Y_get_R(c, Y);
    double gamma;
    HalleyMore hm{0, 0.0, 0.0, true, true};
    bool conv;
    if constexpr (grid) {
      conv = halley_T<MixTables, true>(mix, Y, R, e, T, gamma, &hm);
      a.hside[i] = (Real)hm.g1;
      a.hside[n + i] = (Real)hm.T1;
      a.hnit[i] = (unsigned char)hm.n;
      if (hm.n < kHalleyMaxIter && !hm.ok1) fl |= JH_FLAG_HALLEY_REGRESS;
      gv[0] = (unsigned long long)hm.n;
    } else {
      conv = halley_T(mix, Y, R, e, T, gamma);
    }
    aux[i] = (Real)gamma;
    aux[n + i] = (Real)T;
    fl |= halley_flags(r2, e, R, T, gamma, conv);
    // next step: CFL_2D and stage-1 LF maxima from the stored values
    const double gs = (double)(Real)gamma, Ts = (double)(Real)T;
    const double p = r2 * R * Ts;
    const double snd = sqrt(gs * p * ir);
    const double sx = fabs(u) + snd, sy = fabs(v) + snd;
    if constexpr (grid) {
      const double g1 = (double)(Real)hm.g1, T1 = (double)(Real)hm.T1;
      const double p1 = r2 * R * T1;
      const double snd1 = sqrt(g1 * p1 * ir);
      const double sx1 = fabs(u) + snd1, sy1 = fabs(v) + snd1;
      cand_max(gv, 1, hm.n, nan_to_zero(sx), nan_to_zero(sx1));
      cand_max(gv, 1 + kNstar, hm.n, nan_to_zero(sy), nan_to_zero(sy1));
      cand_max(gv, 1 + 2 * kNstar, hm.n, sx / a.dx + sy / a.dy, sx1 / a.dx + sy1 / a.dy);
    } else {
      mx = fmax(mx, nan_to_zero(sx));
      my = fmax(my, nan_to_zero(sy));
      const unsigned long long b = dbits(sx / a.dx + sy / a.dy);
      cm = b > cm ? b : cm;
    }
  }
  if (MODE == CHEM_STEP) {
    raise_flags(a.sc, fl);
    const int o = a.parity ^ 1;
    if constexpr (grid) {
      block_max_bits_multi<256, kGridSlotsChem>(gv, &a.sc->gs[a.parity][JH_PROD_CHEM].nstar,
                                                redg);
    } else {
      block_max_atomic<256>(mx, &a.sc->amax[o][0][0], red);
      block_max_atomic<256>(my, &a.sc->amax[o][0][1], red);
      block_max_bits_atomic<256>(cm, &a.sc->cmax[o], redb);
    }
    // the step's dt, stored last: a global store ahead of the rates would
    // make the compiler treat every mechanism constant read after it as
    // possibly clobbered, and fetch them with vector loads into VGPRs
    // instead of scalar loads
    if (blk == 0 && threadIdx.x == 0) {
      a.sc->dt_last = dt;
      if (dt_hist) *dt_hist = dt;
    }
  }
}

// Two waves per SIMD: the point-implicit solve holds the 8x8 Jacobian and
// the rates in registers (256 VGPRs at one wave).
template <typename Real> struct ChemWaves { static constexpr int value = 2; };
template <> struct ChemWaves<float> { static constexpr int value = 2; };
template <typename Real, int MODE, bool REPLAY = false, bool GRID = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ChemWaves<Real>::value))) void chem_kernel(StepArgs<Real> a, Real* U, Real* aux,
                                                   Real* S, double* dt_hist) {
  __shared__ double red[4];
  __shared__ unsigned long long redb[4];
  __shared__ unsigned long long redg[GRID ? kGridSlotsChem * 4 : 1];
  __shared__ MixTables mix;
  __shared__ GibbsTables gt;
  chem_body<Real, MODE, REPLAY, GRID>(a, U, aux, S, dt_hist, blockIdx.x, red, redb, redg, mix,
                                      gt);
}

// update_aux alone (aux_func.py:21-29), for the unit entry point; in grid
// mode followed by halley_fix_kernel (producer JH_PROD_UPDATE).
template <typename Real>
__global__ __launch_bounds__(256) void update_aux_kernel(StepArgs<Real> a, const Real* __restrict__ U,
                                                         const Real* __restrict__ aux_in,
                                                         Real* __restrict__ aux_out) {
  const MechConst& c = *a.mc;
  __shared__ unsigned long long redg[4];
  const int n = a.nx * a.ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nb[1] = {0ull};
  unsigned int fl = 0u;
  if (i < n) {
    const double rho = (double)U[i], ir = rcp(rho);
    const double u = (double)U[n + i] * ir, v = (double)U[2 * n + i] * ir;
    const double e = (double)U[3 * n + i] * ir - 0.5 * (u * u + v * v);
    double Y[JH_NS];
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = (double)U[(4 + k) * n + i] * ir;
    const double R = fill_Y_get_R(c, Y);
    double T = (double)aux_in[n + i];
    double gamma;
    bool conv;
    if (a.hside) {
      HalleyMore hm;
      conv = halley_T<MechConst, true>(c, Y, R, e, T, gamma, &hm);
      a.hside[i] = (Real)hm.g1;
      a.hside[n + i] = (Real)hm.T1;
      a.hnit[i] = (unsigned char)hm.n;
      if (hm.n < kHalleyMaxIter && !hm.ok1) fl |= JH_FLAG_HALLEY_REGRESS;
      nb[0] = (unsigned long long)hm.n;
    } else {
      conv = halley_T(c, Y, R, e, T, gamma);
    }
    aux_out[i] = (Real)gamma;
    aux_out[n + i] = (Real)T;
    fl |= halley_flags(rho, e, R, T, gamma, conv);
  }
  raise_flags(a.sc, fl);
  if (a.hside)
    block_max_bits_multi<256, 1>(nb, &a.sc->gs[a.parity][JH_PROD_UPDATE].nstar, redg);
}

// ------------------------------------------------------------ tracing
// Per step, where each stage's LF maxima are attained and which chemistry
// clip branches are active (jh_ctx_set_trace).  The step is only piecewise
// smooth: reverse mode (and the adjoint kernels) differentiate the branch
// taken, so a finite difference agrees with the adjoint only across
// perturbations under which none of these switch.  A record is 22 int64:
// per stage s, (first padded index attaining max|u|+a, number of padded cells
// attaining it, the same for max|v|+a); then a hash and the count of the
// (cell, species, bound) triples whose clip is active; then, after each of
// the step's four temperature solves (the three stages, the chemistry), a
// hash and the count of the cells below the NASA-7 switch temperature (the
// two fits differ there by up to ~4e-7 in cp: a cell that changes side
// changes its state discontinuously).

struct TraceScratch {
  unsigned long long mx[2];    // maxima (bit patterns) of the stage being traced
  unsigned long long idx[2];   // first padded index attaining each
  unsigned long long cnt[2];   // cells attaining each
};
constexpr int kTraceRecord = 22;

// |u| + a and |v| + a of padded cell P (interior cell or non-corner ghost
// cell from the strips), from the stored values.
template <typename Real>
__device__ __forceinline__ bool trace_speeds(const StepArgs<Real>& a, const Real* U,
                                             const Real* aux, int P, double& sx, double& sy) {
  const int py = a.ny + 6;
  const int gx = P / py - 3, gy = P - (P / py) * py - 3;
  const CellRef<Real> r = locate<Real>(a, U, aux, gx, gy);
  if (!r.ok) return false;
  const Prim<double> q = cell_prim<double, true>(*a.mc, r.u, r.x, r.stride);
  sx = nan_to_zero(fabs(q.u) + q.a);
  sy = nan_to_zero(fabs(q.v) + q.a);
  return true;
}

// Pass 1: the maxima over the padded field (grid-stride).
template <typename Real>
__global__ __launch_bounds__(256) void trace_max_kernel(StepArgs<Real> a, const Real* U,
                                                        const Real* aux, TraceScratch* ts) {
  __shared__ double red[4];
  const int np = (a.nx + 6) * (a.ny + 6);
  double mx = 0.0, my = 0.0;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < np; P += gridDim.x * blockDim.x) {
    double sx, sy;
    if (trace_speeds(a, U, aux, P, sx, sy)) {
      mx = fmax(mx, sx);
      my = fmax(my, sy);
    }
  }
  block_max_atomic<256>(mx, &ts->mx[0], red);
  block_max_atomic<256>(my, &ts->mx[1], red);
}

// Pass 2: the first padded index attaining each maximum and their number.
template <typename Real>
__global__ __launch_bounds__(256) void trace_arg_kernel(StepArgs<Real> a, const Real* U,
                                                        const Real* aux, TraceScratch* ts) {
  const int np = (a.nx + 6) * (a.ny + 6);
  const unsigned long long m0 = ts->mx[0], m1 = ts->mx[1];
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < np; P += gridDim.x * blockDim.x) {
    double sx, sy;
    if (!trace_speeds(a, U, aux, P, sx, sy)) continue;
    if (dbits(sx) == m0) {
      atomicMin(&ts->idx[0], (unsigned long long)P);
      atomicAdd(&ts->cnt[0], 1ull);
    }
    if (dbits(sy) == m1) {
      atomicMin(&ts->idx[1], (unsigned long long)P);
      atomicAdd(&ts->cnt[1], 1ull);
    }
  }
}

// Moves one stage's result into its record slots (rec: 4 values) and resets
// the scratch for the next stage; rec == nullptr only resets.
__global__ void trace_store_kernel(TraceScratch* ts, long long* rec) {
  if (threadIdx.x != 0) return;
  if (rec) {
    rec[0] = (long long)ts->idx[0];
    rec[1] = (long long)ts->cnt[0];
    rec[2] = (long long)ts->idx[1];
    rec[3] = (long long)ts->cnt[1];
  }
  ts->mx[0] = ts->mx[1] = 0ull;
  ts->idx[0] = ts->idx[1] = ~0ull;
  ts->cnt[0] = ts->cnt[1] = 0ull;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// The chemistry's active clip branches at the state the chemistry kernel
// will react: rec[0] += a hash of every (cell, bit), rec[1] += their count
// (rec zeroed beforehand).  One cell per thread.
template <typename Real>
__global__ __launch_bounds__(256) void trace_clip_kernel(StepArgs<Real> a, const Real* U,
                                                         const Real* aux, long long* rec) {
  const MechConst& c = *a.mc;
  const int n = a.nx * a.ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long h = 0ull, cnt = 0ull;
  if (i < n) {
    const double rho = (double)U[i];
    double rY[JH_NS - 1];
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) rY[k] = (double)U[(4 + k) * n + i];
    double d[JH_NA];
    unsigned int cm = 0u;
    (void)chem_increment<Real>(c, c, rho, rY, (double)aux[n + i], step_dt(a), d, &cm);
    for (int b = 0; b < 2 * JH_NA; ++b)
      if ((cm >> b) & 1u) {
        h += mix64(((unsigned long long)i << 5) | (unsigned long long)b);
        cnt += 1ull;
      }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    h += __shfl_xor(h, off, 64);
    cnt += __shfl_xor(cnt, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && cnt) {
    atomicAdd(reinterpret_cast<unsigned long long*>(rec), h);
    atomicAdd(reinterpret_cast<unsigned long long*>(rec + 1), cnt);
  }
}

// The clip branches the chemistry kernel will take at this state, one mask
// per cell, into a.clipm (recording for jh_ctx_set_clip_branches).
template <typename Real>
__global__ __launch_bounds__(256) void clip_mask_kernel(StepArgs<Real> a, const Real* U,
                                                        const Real* aux) {
  const MechConst& c = *a.mc;
  const int n = a.nx * a.ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double rY[JH_NS - 1];
#pragma unroll
  for (int k = 0; k < JH_NS - 1; ++k) rY[k] = (double)U[(4 + k) * n + i];
  double d[JH_NA];
  unsigned int cm = 0u;
  (void)chem_increment<Real>(c, c, (double)U[i], rY, (double)aux[n + i], step_dt(a), d, &cm);
  a.clipm[i] = (unsigned short)cm;
}

// Cells of aux's temperature field below the NASA-7 switch: rec[0] += a
// hash of their indices, rec[1] += their count (rec zeroed beforehand).
template <typename Real>
__global__ __launch_bounds__(256) void trace_range_kernel(StepArgs<Real> a, const Real* aux,
                                                          long long* rec) {
  const int n = a.nx * a.ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long h = 0ull, cnt = 0ull;
  if (i < n && (double)aux[n + i] < a.mc->Tcr) {
    h = mix64((unsigned long long)i);
    cnt = 1ull;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    h += __shfl_xor(h, off, 64);
    cnt += __shfl_xor(cnt, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && cnt) {
    atomicAdd(reinterpret_cast<unsigned long long*>(rec), h);
    atomicAdd(reinterpret_cast<unsigned long long*>(rec + 1), cnt);
  }
}

__global__ void zero_i64_kernel(long long* p, int m) {
  if ((int)threadIdx.x < m) p[threadIdx.x] = 0;
}

// Grid-wide Halley stop (thermo_model.py:140-142), after a producer of
// (gamma, T) with StepArgs::hside set: N* = max over cells of the per-cell
// update count n; a cell with n < N* takes its state after one more update
// (halley_T's HalleyMore).  Block 0 also folds the producer's maxima for N*
// into the slots the next kernels read: the LF maxima of stage lf_stage of
// parity lf_parity (lf_stage < 0: none) and, with with_cfl, that parity's CFL
// maximum.  Grid-stride over the cells.
template <typename Real>
__global__ __launch_bounds__(256) void halley_fix_kernel(StepScalars* sc, int parity, int prod,
                                                         int lf_parity, int lf_stage,
                                                         int with_cfl, Real* __restrict__ aux,
                                                         const Real* __restrict__ side,
                                                         const unsigned char* __restrict__ nit,
                                                         int n) {
  const GridSlots& g = sc->gs[parity][prod];
  const unsigned long long ns64 = g.nstar;
  const int ns = ns64 < (unsigned long long)kNstar ? (int)ns64 : kNstar - 1;
  if (blockIdx.x == 0) {
    const int t = threadIdx.x;
    if (t < 2 && lf_stage >= 0) atomicMax(&sc->amax[lf_parity][lf_stage][t], g.lf[t][ns]);
    if (t == 2 && with_cfl) atomicMax(&sc->cmax[lf_parity], g.cfl[ns]);
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if ((int)nit[i] < ns) {
      aux[i] = side[i];
      aux[n + i] = side[n + i];
    }
  }
}

}  // namespace jh
