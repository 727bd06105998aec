// Godunov form of the 2-D flux (solver_type = 'godunov'): reconstruction of
// the primitive variables q = [rho, u, v, p, Y(8)] at every face, then the
// HLLC approximate Riemann solver (solver_2D/flux.py:43-54,
// solver_2D/riemann_solver/HLLC.py:5-77), inside one SSP-RK3 stage kernel.
#pragma once
#include "hllc.hpp"
#include "recon.hpp"
#include "stage.hpp"

namespace jh {

// Tile of the Godunov stage kernel: TX rows along x by TY = 64 columns,
// TX + 1 waves.  x pass: wave w reconstructs x-face row w (TX + 1 rows of 64
// faces).  y pass: wave w < TX reconstructs the 64 y faces of row w that lie
// on each cell's low-y side, wave TX the tile's last y face (column 64) of
// every row.
template <int TX>
struct GTile {
  static constexpr int NW = TX + 1;
  static constexpr int NT = 64 * NW;
  static constexpr int RX = TX + 6;               // staged rows (x pass)
  static constexpr int NSX = (RX + NW - 1) / NW;  // staged cells per thread (x pass)
};

template <typename Real, int TX>
struct GodLds {
  Real Q[2][TX + 6][TY + 6];   // one primitive component, double-buffered
  Real Fx[NU][TX + 1][TY];     // x-face fluxes of the tile
  Real Ge[NU][TX];             // y flux at the tile's last face of each row
  MixTables mix;               // NASA-7 tables (face thermo, Halley epilogue)
};

// Cell of the staged region: its address, density, 1/rho, pressure, and the
// conserved value of the next component to stage.
template <typename Real>
struct GCell {
  CellRef<Real> ref;
  Real rho, ir, p, nxt;
};

template <typename Real>
__device__ __forceinline__ void gcell_init(const StepArgs<Real>& a, const Real* __restrict__ U,
                                           const Real* __restrict__ aux, int gx, int gy, bool has,
                                           GCell<Real>& g) {
  g.ref = locate<Real>(a, U, aux, gx, gy);
  if (!has) g.ref.ok = false;
  if (g.ref.ok) {
    const Prim<Real> q = cell_prim<Real, false>(*a.mc, g.ref.u, g.ref.x, g.ref.stride);
    g.rho = q.rho;
    g.ir = rcp(q.rho);
    g.p = q.p;
    g.nxt = g.ref.u[g.ref.stride];          // component 1 (rho u), staged second
  } else {
    // benign state for cells no owned face reads
    g.rho = (Real)1; g.ir = (Real)1; g.p = (Real)1; g.nxt = (Real)0;
  }
}

// q component cv of a staged cell (aux_func.py:39-51 order rho, u, v, p, Y).
template <typename Real>
__device__ __forceinline__ Real gcell_q(const GCell<Real>& g, int cv) {
  if (cv == 0) return g.rho;
  if (cv == 3) return g.p;
  return g.ref.ok ? g.nxt * g.ir : (Real)0;
}

// After component cv is staged, load the conserved value the next staged
// component that needs one will read (components 0 and 3 need none).
template <typename Real>
__device__ __forceinline__ void gcell_advance(GCell<Real>& g, int cv) {
  if (cv == 0 || cv == 3 || cv + 1 >= NU || !g.ref.ok) return;
  const int nc = cv == 2 ? 4 : cv + 1;
  g.nxt = g.ref.u[nc * g.ref.stride];
}

// One SSP-RK3 stage (TVD_RK3.py:5-12) with the Godunov right-hand side
//   rhs = dt * -(dF/dx + dG/dy), F, G from HLLC on reconstructed q
// for one tile, followed by update_aux (Halley) of the new state.
// MODE_RHS writes rhs only (the reference's flux_func).
template <typename Real, int MODE, int TX, int RECON>
__device__ __forceinline__ void godunov_stage_body(const StepArgs<Real>& a, int stage,
                                                   const Real* __restrict__ Us,
                                                   const Real* __restrict__ auxs, const Real* Un,
                                                   Real* Uo, Real* __restrict__ auxo, int tile,
                                                   GodLds<Real, TX>& L) {
  using G = GTile<TX>;
  constexpr int NW = G::NW, NSX = G::NSX, RX = G::RX;
  const MechConst& c = *a.mc;
  load_mix_tables(c, L.mix);                 // read after the first barrier
  const int nx = a.nx, ny = a.ny, n = nx * ny;
  const int t = threadIdx.x, ly = t & 63, w = t >> 6;
  const int tiles_y = (ny + TY - 1) / TY;
  const int x0 = (tile / tiles_y) * TX;
  const int y0 = (tile % tiles_y) * TY;

  // ---------------- x pass: q on rows x0-3 .. x0+TX+2, columns y0 .. y0+63
  Real qL[NU], qR[NU], Fl[NU];
  {
    GCell<Real> cell[NSX];
    int row[NSX];
#pragma unroll
    for (int h = 0; h < NSX; ++h) {
      row[h] = w + h * NW;
      gcell_init<Real>(a, Us, auxs, x0 - 3 + row[h], y0 + ly, row[h] < RX, cell[h]);
    }
#pragma unroll
    for (int cv = 0; cv < NU; ++cv) {
      auto& Q = L.Q[cv & 1];
#pragma unroll
      for (int h = 0; h < NSX; ++h)
        if (row[h] < RX) Q[row[h]][ly + 3] = gcell_q(cell[h], cv);
#pragma unroll
      for (int h = 0; h < NSX; ++h) gcell_advance(cell[h], cv);
      __syncthreads();
      const int col = ly + 3;
      qL[cv] = recon_L<RECON, Real>(Q[w][col], Q[w + 1][col], Q[w + 2][col], Q[w + 3][col],
                                    Q[w + 4][col]);
      qR[cv] = recon_R<RECON, Real>(Q[w + 1][col], Q[w + 2][col], Q[w + 3][col],
                                    Q[w + 4][col], Q[w + 5][col]);
    }
  }
  hllc_flux<Real, 0>(c, L.mix, qL, qR, Fl);
#pragma unroll
  for (int cv = 0; cv < NU; ++cv) L.Fx[cv][w][ly] = Fl[cv];

  // ---------------- y pass: q on rows x0 .. x0+TX-1, columns y0-3 .. y0+66
  {
    GCell<Real> cell;
    int qr, qc;
    bool has;
    if (w < TX) {
      qr = w + 3;
      qc = ly + 3;
      has = true;
    } else {
      const int xr = ly / 6, pos = ly % 6;
      qr = xr + 3;
      qc = pos < 3 ? pos : TY + pos;
      has = ly < 6 * TX;
    }
    gcell_init<Real>(a, Us, auxs, x0 + qr - 3, y0 + qc - 3, has, cell);
    // rows of the y faces this thread reconstructs
    const int fr = w < TX ? w + 3 : (ly < TX ? ly : 0) + 3;
    const int fc = w < TX ? ly : TY;
#pragma unroll
    for (int cv = 0; cv < NU; ++cv) {
      auto& Q = L.Q[cv & 1];
      if (has) Q[qr][qc] = gcell_q(cell, cv);
      gcell_advance(cell, cv);
      __syncthreads();
      qL[cv] = recon_L<RECON, Real>(Q[fr][fc], Q[fr][fc + 1], Q[fr][fc + 2], Q[fr][fc + 3],
                                    Q[fr][fc + 4]);
      qR[cv] = recon_R<RECON, Real>(Q[fr][fc + 1], Q[fr][fc + 2], Q[fr][fc + 3],
                                    Q[fr][fc + 4], Q[fr][fc + 5]);
    }
  }
  hllc_flux<Real, 1>(c, L.mix, qL, qR, Fl);
  if (w == TX && ly < TX) {
#pragma unroll
    for (int cv = 0; cv < NU; ++cv) L.Ge[cv][ly] = Fl[cv];
  }
  __syncthreads();
  if (w >= TX) return;

  // ---------------- divergence + RK combine for cell (x0 + w, y0 + ly)
  const int gx = x0 + w, gy = y0 + ly;
  const bool own = gx < nx && gy < ny;
  const int o = gx * ny + gy;
  const double dt = step_dt(a);
  const double idx = 1.0 / a.dx, idy = 1.0 / a.dy;
  double un[NU];       // FP64 whatever Real (see stage_body)
#pragma unroll
  for (int cv = 0; cv < NU; ++cv) {
    Real Gn = __shfl_down(Fl[cv], 1);
    if (ly == TY - 1) Gn = L.Ge[cv][w];
    const double dF = (double)L.Fx[cv][w + 1][ly] - (double)L.Fx[cv][w][ly];
    const double dG = (double)Gn - (double)Fl[cv];
    const double rhs = dt * (-(dF * idx + dG * idy));
    if (MODE == MODE_RHS) {
      un[cv] = rhs;
    } else {
      const double uc = own ? (double)Us[cv * n + o] : 0.0;
      if (stage == 0) {
        un[cv] = uc + rhs;
      } else {
        const double u0 = own ? (double)Un[cv * n + o] : 0.0;
        if (stage == 1) un[cv] = (3.0 / 4.0) * u0 + (1.0 / 4.0) * (uc + rhs);
        else un[cv] = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (uc + rhs);
      }
    }
    if (own) Uo[cv * n + o] = (Real)un[cv];
  }
  if (MODE == MODE_RHS) return;

  // ---------------- update_aux (aux_func.py:21-29) of the new state
  // the Halley solve sees the stored (rounded) state, as the next stage does
  unsigned long long nb = 0ull;     // grid-wide stop: this lane's update count
  unsigned int fl = 0u;
  if (own) {
    double us[NU];
#pragma unroll
    for (int cv = 0; cv < NU; ++cv) us[cv] = (double)(Real)un[cv];
    const double rho = us[0], ir = rcp(rho);
    const double u = us[1] * ir, v = us[2] * ir;
    const double e = us[3] * ir - 0.5 * (u * u + v * v);
    double Y[JH_NS];
#pragma unroll
    for (int k = 0; k < JH_NS - 1; ++k) Y[k] = us[4 + k] * ir;
    const double R = fill_Y_get_R(c, Y);
    double T = (double)auxs[n + o];
    double gamma;
    bool conv;
    if (a.hside) {
      HalleyMore hm;
      conv = halley_T<MixTables, true>(L.mix, Y, R, e, T, gamma, &hm);
      a.hside[o] = (Real)hm.g1;
      a.hside[n + o] = (Real)hm.T1;
      a.hnit[o] = (unsigned char)hm.n;
      if (hm.n < kHalleyMaxIter && !hm.ok1) fl |= JH_FLAG_HALLEY_REGRESS;
      nb = (unsigned long long)hm.n;
    } else {
      conv = halley_T(L.mix, Y, R, e, T, gamma);
    }
    auxo[o] = (Real)gamma;
    auxo[n + o] = (Real)T;
    fl |= halley_flags(rho, e, R, T, gamma, conv);
  }
  // every lane of the wave reaches here (waves w >= TX left above)
  raise_flags(a.sc, fl);
  if (a.hside) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long x = __shfl_xor(nb, off, 64);
      nb = x > nb ? x : nb;
    }
    unsigned long long* dst = &a.sc->gs[a.parity][stage].nstar;
    if ((threadIdx.x & 63) == 0 &&
        nb > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(dst, nb);
  }
}

template <typename Real, int MODE, int TX, int RECON>
__global__ __launch_bounds__(GTile<TX>::NT) void godunov_stage_kernel(
    StepArgs<Real> a, int stage, const Real* __restrict__ Us, const Real* __restrict__ auxs,
    const Real* Un, Real* Uo, Real* __restrict__ auxo) {
  __shared__ GodLds<Real, TX> L;
  godunov_stage_body<Real, MODE, TX, RECON>(a, stage, Us, auxs, Un, Uo, auxo,
                                            xcd_tile(blockIdx.x, gridDim.x), L);
}

}  // namespace jh
