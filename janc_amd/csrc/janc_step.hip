// C ABI of the gfx950 JANC step library: context management and the launch
// sequence of one SSP-RK3 + point-implicit-chemistry step.  Declarations and
// the reference entry points each call replaces are in include/janc_step.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/janc_step.h"
#include "adjoint.hpp"
#include "batch.hpp"
#include "blocks.hpp"
#include "godunov.hpp"
#include "stage.hpp"

using namespace jh;

struct jh_ctx {
  int nx = 0, ny = 0;
  double dx = 0, dy = 0;
  int precision = JH_PREC_FP64;
  int bc[4] = {0, 0, 0, 0};
  double bc_p[4] = {0, 0, 0, 0};
  double inj_T0 = 300.0;
  double inj_Y[JH_NS - 1] = {};
  hipStream_t stream = nullptr;
  MechConst* mc = nullptr;        // device
  StepScalars* sc = nullptr;      // device
  void* ghost[4] = {nullptr, nullptr, nullptr, nullptr};
  void* tmpU[2] = {nullptr, nullptr};
  void* tmpA[2] = {nullptr, nullptr};
  int parity = 0;
  bool user_filled[4] = {false, false, false, false};   // JH_BC_USER strips written
  int tile = 0;                   // stage-kernel variant (kTileRows / kTileWaves)
  int solver = JH_SOLVER_FLUX_SPLITTING;   // jh_ctx_set_flux
  int split = JH_SPLIT_LF;
  int recon = JH_RECON_WENO5_JS;
  // temperature-solve stop (jh_ctx_set_halley_stop); the grid-wide form's
  // per-cell buffers: (gamma, T) after one more update, update counts
  int halley = JH_HALLEY_PER_CELL;
  void* hside = nullptr;
  unsigned char* hnit = nullptr;
  // step tracing (jh_ctx_set_trace): records of kTraceRecord int64 per step
  long long* trace = nullptr;
  int trace_cap = 0, trace_step = 0;
  TraceScratch* trs = nullptr;
  // per-block maxima of the prologue's reduction (3 x grid_cells blocks)
  unsigned long long* rpart = nullptr;
  // chemistry clip branches recorded / replayed per step
  // (jh_ctx_set_clip_branches): (max steps, nx*ny) masks
  unsigned short* clipm = nullptr;
  int clip_mode = 0, clip_cap = 0, clip_step = 0;
  std::string err;
  // kernel timing with HIP events on the context stream (jh_timing_*)
  bool timing = false;             // record events around the current launches
  int timing_every = 0;            // jh_advance: time the launches of every k-th step
  long long timed_steps = 0;       // jh_advance steps since jh_timing_enable
  std::vector<hipEvent_t> ev_free;
  struct Mark { int kind; hipEvent_t a, b; };
  std::vector<Mark> marks;
  // adjoint workspace (allocated on the first jh_step_vjp call)
  double* adj = nullptr;
  size_t adj_len = 0;
  // its ghost-fold lists were built for these boundary kinds (-1: not built)
  int adj_fold_bc[4] = {-1, -1, -1, -1};
  int adj_fold_m = 0;              // interior cells with ghost copies
};

namespace {

thread_local std::string g_create_err;

int fail(jh_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg; else g_create_err = msg;
  return code;
}

#define HIP_TRY(ctx, call)                                                     \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(ctx, JH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t elem_size(const jh_ctx* c) { return c->precision == JH_PREC_FP64 ? 8 : 4; }

int check_launch(jh_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, JH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return JH_OK;
}

hipEvent_t take_event(jh_ctx* c) {
  if (!c->ev_free.empty()) {
    hipEvent_t e = c->ev_free.back();
    c->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with events when timing is on.
struct Timed {
  jh_ctx* c;
  int kind;
  hipEvent_t a = nullptr;
  Timed(jh_ctx* c_, int k) : c(c_), kind(k) {
    if (c->timing && (a = take_event(c))) (void)hipEventRecord(a, c->stream);
  }
  ~Timed() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->marks.push_back({kind, a, b});
  }
};

template <typename Real>
StepArgs<Real> make_args(jh_ctx* c, double dt_fixed, double cfl) {
  StepArgs<Real> a;
  a.nx = c->nx;
  a.ny = c->ny;
  a.dx = c->dx;
  a.dy = c->dy;
  a.cfl = cfl;
  a.dt_fixed = dt_fixed;
  a.parity = c->parity;
  a.mc = c->mc;
  a.sc = c->sc;
  for (int s = 0; s < 4; ++s) {
    a.g[s] = static_cast<Real*>(c->ghost[s]);
    a.bc[s] = c->bc[s];
    a.bc_p[s] = c->bc_p[s];
  }
  a.mirror = 0;
  a.inj_T0 = c->inj_T0;
  for (int k = 0; k < JH_NS - 1; ++k) a.inj_Y[k] = c->inj_Y[k];
  const bool grid = c->halley == JH_HALLEY_GRID;
  a.hside = grid ? static_cast<Real*>(c->hside) : nullptr;
  a.hnit = grid ? c->hnit : nullptr;
  a.clipm = nullptr;
  return a;
}

// Grid-wide temperature-solve stop: after a producer of (gamma, T) written
// to aux, apply N* and fold the producer's maxima for N* into the LF slot
// (lf_parity, lf_stage) (lf_stage < 0: none) and, with with_cfl, the CFL
// slot of lf_parity.  No launch in the per-cell form.
template <typename Real>
void halley_fix(jh_ctx* c, int prod, int lf_parity, int lf_stage, int with_cfl, Real* aux) {
  if (c->halley != JH_HALLEY_GRID) return;
  const int n = c->nx * c->ny;
  const int nb = std::min((n + 255) / 256, 2048);
  halley_fix_kernel<Real><<<nb, 256, 0, c->stream>>>(c->sc, c->parity, prod, lf_parity, lf_stage,
                                                     with_cfl, aux,
                                                     static_cast<const Real*>(c->hside), c->hnit,
                                                     n);
}

int grid_cells(const jh_ctx* c) {
  const int n = c->nx * c->ny;
  int b = (n + 255) / 256;
  return b < 2048 ? b : 2048;
}

// one thread per cell (pointwise kernels without a grid-stride loop)
int grid_all(const jh_ctx* c) { return (c->nx * c->ny + 255) / 256; }

// The step's chemistry launch: the grid-wide Halley stop's epilogue (a.hside
// set) and the clip-branch replay are separate instantiations.
template <typename Real>
void launch_chem_step(const jh_ctx* c, const StepArgs<Real>& a, hipStream_t st, Real* U,
                      Real* aux, double* dt_hist, bool replay) {
  const int g = grid_all(c);
  if (a.hside) {
    if (replay)
      chem_kernel<Real, CHEM_STEP, true, true><<<g, 256, 0, st>>>(a, U, aux, nullptr, dt_hist);
    else
      chem_kernel<Real, CHEM_STEP, false, true><<<g, 256, 0, st>>>(a, U, aux, nullptr, dt_hist);
  } else {
    if (replay)
      chem_kernel<Real, CHEM_STEP, true, false><<<g, 256, 0, st>>>(a, U, aux, nullptr, dt_hist);
    else
      chem_kernel<Real, CHEM_STEP, false, false><<<g, 256, 0, st>>>(a, U, aux, nullptr, dt_hist);
  }
}

int tile_rows(const jh_ctx* c);

// Rows of the Godunov stage kernel's tile (measured on the 4000 x 400 FP64
// step: with WENO5 the 4-wave 64 x 3 tile lets two blocks share a CU, 13%
// faster than 64 x 7; the lighter MUSCL kernel is faster with 64 x 7).
constexpr int kGodRowsWeno = 3;
constexpr int kGodRowsMuscl = 7;

int god_rows(const jh_ctx* c) {
  return c->recon == JH_RECON_MUSCL ? kGodRowsMuscl : kGodRowsWeno;
}

// Steger-Warming flux splitting: one tile shape, 64 x 16 at two waves per
// SIMD (its per-cell eigenvalue splits need ~250 VGPRs).
constexpr int kSwRows = 16;

int stage_blocks(const jh_ctx* c) {
  const int tx = c->solver == JH_SOLVER_GODUNOV ? god_rows(c)
                 : c->split == JH_SPLIT_SW     ? kSwRows
                                               : tile_rows(c);
  return ((c->nx + tx - 1) / tx) * ((c->ny + TY - 1) / TY);
}

// Stage-kernel variants (jh_ctx_set_tile): X(index, rows along x, waves per
// SIMD the register budget is sized for).
#define JH_TILE_VARIANTS(X) X(0, 8, 3) X(1, 8, 4) X(2, 16, 4) X(3, 16, 5) X(4, 32, 4)
constexpr int kTileVariants = 5;
constexpr int kTileRows[kTileVariants] = {8, 8, 16, 16, 32};
constexpr int kTileWaves[kTileVariants] = {3, 4, 4, 5, 4};

// With the grid-wide temperature-solve stop the flux-splitting stage kernel
// runs one tile shape (its epilogue carries the per-candidate maxima): 16
// rows at a 4-wave register budget, whatever the context's variant.
constexpr int kGridRows = 16;

template <typename Real, int MODE, int RECON>
void launch_split_stage(jh_ctx* c, const StepArgs<Real>& a, int stage, const Real* Us,
                        const Real* As, const Real* Un, Real* Uo, Real* Ao) {
  if (MODE == MODE_RK && c->halley == JH_HALLEY_GRID) {
    const int nbg = ((c->nx + kGridRows - 1) / kGridRows) * ((c->ny + TY - 1) / TY);
    stage_kernel<Real, MODE, kGridRows, 4, RECON, SPLIT_LF, Real, true>
        <<<nbg, Tile<kGridRows>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
    return;
  }
  const int nb = stage_blocks(c);
  switch (c->tile) {
#define JH_LAUNCH(V, TX, WPE)                                                          \
    case V:                                                                            \
      if constexpr (MODE == MODE_RK && RECON == RECON_WENO5) {                         \
        if (a.mirror) {                                                                \
          stage_kernel<Real, MODE, TX, WPE, RECON, SPLIT_LF, Real, false, true>        \
              <<<nb, Tile<TX>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);      \
          break;                                                                       \
        }                                                                              \
      }                                                                                \
      stage_kernel<Real, MODE, TX, WPE, RECON><<<nb, Tile<TX>::NT, 0, c->stream>>>(   \
          a, stage, Us, As, Un, Uo, Ao);                                               \
      break;
    JH_TILE_VARIANTS(JH_LAUNCH)
#undef JH_LAUNCH
    default:
      break;
  }
}

template <typename Real, int MODE>
void launch_stage(jh_ctx* c, const StepArgs<Real>& a, int stage, const Real* Us, const Real* As,
                  const Real* Un, Real* Uo, Real* Ao) {
  const int nb = stage_blocks(c);
  if (c->solver == JH_SOLVER_GODUNOV) {
    if (c->recon == JH_RECON_MUSCL)
      godunov_stage_kernel<Real, MODE, kGodRowsMuscl, RECON_MUSCL>
          <<<nb, GTile<kGodRowsMuscl>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
    else
      godunov_stage_kernel<Real, MODE, kGodRowsWeno, RECON_WENO5>
          <<<nb, GTile<kGodRowsWeno>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
  } else if (c->split == JH_SPLIT_SW) {
    const bool grid = MODE == MODE_RK && c->halley == JH_HALLEY_GRID;
    if (c->recon == JH_RECON_MUSCL) {
      if (grid)
        stage_kernel<Real, MODE, kSwRows, 2, RECON_MUSCL, SPLIT_SW, Real, true>
            <<<nb, Tile<kSwRows>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
      else
        stage_kernel<Real, MODE, kSwRows, 2, RECON_MUSCL, SPLIT_SW>
            <<<nb, Tile<kSwRows>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
    } else {
      if (grid)
        stage_kernel<Real, MODE, kSwRows, 2, RECON_WENO5, SPLIT_SW, Real, true>
            <<<nb, Tile<kSwRows>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
      else
        stage_kernel<Real, MODE, kSwRows, 2, RECON_WENO5, SPLIT_SW>
            <<<nb, Tile<kSwRows>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
    }
  } else if (c->recon == JH_RECON_MUSCL) {
    launch_split_stage<Real, MODE, RECON_MUSCL>(c, a, stage, Us, As, Un, Uo, Ao);
  } else {
    launch_split_stage<Real, MODE, RECON_WENO5>(c, a, stage, Us, As, Un, Uo, Ao);
  }
}

int tile_rows(const jh_ctx* c) { return kTileRows[c->tile]; }

// bc_kernel: one thread per boundary line in blocks of kBcThreads, then the
// lines of injection-inlet sides, kInletTeam lanes each, for their bisection
int bc_blocks(const jh_ctx* c) {
  int lines = 0;
  for (int s = 0; s < 4; ++s)
    if (c->bc[s] == JH_BC_INJECTION_INLET) lines += s <= JH_RIGHT ? c->ny : c->nx;
  constexpr int per_block = kBcThreads / kInletTeam;
  return bc_line_blocks(c->nx, c->ny) + (lines + per_block - 1) / per_block;
}

// Prologue for a state handed in by the caller: zero this parity's slots and
// reduce the interior maxima (stage-1 LF alpha, CFL).
template <typename Real>
int prologue(jh_ctx* c, const Real* U, const Real* aux, double dt, double cfl) {
  Timed tm(c, JH_KERNEL_REDUCE);
  reset_kernel<<<1, 64, 0, c->stream>>>(c->sc, c->parity);
  auto a = make_args<Real>(c, dt, cfl);
  const int nb = grid_cells(c);
  reduce_part_kernel<Real><<<nb, 256, 0, c->stream>>>(a, U, aux, c->rpart);
  reduce_final_kernel<<<1, 256, 0, c->stream>>>(c->rpart, nb, c->sc, c->parity, 1);
  return check_launch(c, "reduce_kernel");
}

// Tracing around one RK stage's input (its ghost strips filled): where the
// stage's LF maxima are attained, into rec (4 values).
template <typename Real>
void trace_range(jh_ctx* c, const StepArgs<Real>& a, const Real* aux, long long* rec) {
  zero_i64_kernel<<<1, 64, 0, c->stream>>>(rec, 2);
  trace_range_kernel<Real><<<grid_all(c), 256, 0, c->stream>>>(a, aux, rec);
}

template <typename Real>
void trace_stage(jh_ctx* c, const StepArgs<Real>& a, const Real* Us, const Real* As,
                 long long* rec) {
  const int nb = std::min(((c->nx + 6) * (c->ny + 6) + 255) / 256, 1024);
  trace_max_kernel<Real><<<nb, 256, 0, c->stream>>>(a, Us, As, c->trs);
  trace_arg_kernel<Real><<<nb, 256, 0, c->stream>>>(a, Us, As, c->trs);
  trace_store_kernel<<<1, 64, 0, c->stream>>>(c->trs, rec);
}

// Whether the RK stages can read their halo cells from the source cells
// (StepArgs::mirror) instead of ghost strips filled by bc_kernel: every side
// a copy kind, the LF-split WENO5 stage kernel, each cell stopping its own
// Halley solve, no tracing (whose kernels read the strips).
bool mirror_halo(const jh_ctx* c) {
  if (c->solver == JH_SOLVER_GODUNOV || c->split == JH_SPLIT_SW || c->recon != JH_RECON_WENO5_JS ||
      c->halley == JH_HALLEY_GRID || c->trace)
    return false;
  for (int s = 0; s < 4; ++s) {
    const int k = c->bc[s];
    if (k != JH_BC_PERIODIC && k != JH_BC_ZERO_GRADIENT && k != JH_BC_SLIP_WALL &&
        k != JH_BC_NO_SLIP_WALL)
      return false;
  }
  return true;
}

template <typename Real>
int advance_impl(jh_ctx* c, const Real* U, const Real* aux, Real* Uo, Real* auxo, double dt,
                 double cfl, int nsteps, double* dt_hist) {
  const size_t n = (size_t)c->nx * c->ny;
  Real* U1 = static_cast<Real*>(c->tmpU[0]);
  Real* U2 = static_cast<Real*>(c->tmpU[1]);
  Real* A1 = static_cast<Real*>(c->tmpA[0]);
  Real* A2 = static_cast<Real*>(c->tmpA[1]);
  int rc = prologue<Real>(c, U, aux, dt, cfl);
  if (rc) return rc;
  const Real* Un = U;
  const Real* An = aux;
  const bool timing = c->timing;
  const bool mir = mirror_halo(c);
  for (int s = 0; s < nsteps; ++s) {
    // with a sampling period, only every k-th step's launches are timed
    c->timing = timing && (c->timing_every <= 1 || c->timed_steps % c->timing_every == 0);
    ++c->timed_steps;
    auto a = make_args<Real>(c, dt, cfl);
    a.mirror = mir ? 1 : 0;
    const int bb = bc_blocks(c);
    long long* rec = (c->trace && c->trace_step < c->trace_cap)
                         ? c->trace + (size_t)kTraceRecord * c->trace_step++ : nullptr;
    // stage 1: (Un, An) -> (U1, A1)
    if (!mir) { Timed t(c, JH_KERNEL_BC); bc_kernel<Real><<<bb, kBcThreads, 0, c->stream>>>(a, Un, An, 0); }
    if (rec) trace_stage<Real>(c, a, Un, An, rec);
    { Timed t(c, JH_KERNEL_STAGE);
      launch_stage<Real, MODE_RK>(c, a, 0, Un, An, Un, U1, A1);
      halley_fix<Real>(c, 0, c->parity, 1, 0, A1); }
    // stage 2: (U1, A1) -> (U2, A2)
    if (!mir) { Timed t(c, JH_KERNEL_BC); bc_kernel<Real><<<bb, kBcThreads, 0, c->stream>>>(a, U1, A1, 1); }
    if (rec) {
      trace_range<Real>(c, a, A1, rec + 14);
      trace_stage<Real>(c, a, U1, A1, rec + 4);
    }
    { Timed t(c, JH_KERNEL_STAGE);
      launch_stage<Real, MODE_RK>(c, a, 1, U1, A1, Un, U2, A2);
      halley_fix<Real>(c, 1, c->parity, 2, 0, A2); }
    // stage 3: (U2, A2) -> (Uo, auxo); Uo may alias Un (each thread reads
    // Un at its own cell before writing it)
    if (!mir) { Timed t(c, JH_KERNEL_BC); bc_kernel<Real><<<bb, kBcThreads, 0, c->stream>>>(a, U2, A2, 2); }
    if (rec) {
      trace_range<Real>(c, a, A2, rec + 16);
      trace_stage<Real>(c, a, U2, A2, rec + 8);
    }
    { Timed t(c, JH_KERNEL_STAGE);
      launch_stage<Real, MODE_RK>(c, a, 2, U2, A2, Un, Uo, auxo);
      halley_fix<Real>(c, 2, c->parity, -1, 0, auxo); }
    // chemistry + final update_aux in place, next step's maxima
    if (rec) {
      trace_range<Real>(c, a, auxo, rec + 18);
      zero_i64_kernel<<<1, 64, 0, c->stream>>>(rec + 12, 2);
      trace_clip_kernel<Real><<<grid_all(c), 256, 0, c->stream>>>(a, Uo, auxo, rec + 12);
    }
    const bool replay = c->clip_mode == JH_CLIP_REPLAY && c->clip_step < c->clip_cap;
    if (c->clip_mode && c->clip_step < c->clip_cap) {
      a.clipm = c->clipm + (size_t)n * c->clip_step++;
      if (c->clip_mode == JH_CLIP_RECORD)
        clip_mask_kernel<Real><<<grid_all(c), 256, 0, c->stream>>>(a, Uo, auxo);
    }
    { Timed t(c, JH_KERNEL_CHEM);
      launch_chem_step<Real>(c, a, c->stream, Uo, auxo, dt_hist ? dt_hist + s : nullptr,
                             replay);
      halley_fix<Real>(c, JH_PROD_CHEM, c->parity ^ 1, 0, 1, auxo); }
    if (rec) trace_range<Real>(c, a, auxo, rec + 20);
    rc = check_launch(c, "step kernels");
    if (rc) {
      c->timing = timing;
      return rc;
    }
    c->parity ^= 1;
    Un = Uo;
    An = auxo;
  }
  c->timing = timing;
  (void)n;
  return JH_OK;
}

template <typename Real>
int rhs_impl(jh_ctx* c, const Real* U, const Real* aux, double dt, Real* out) {
  int rc = prologue<Real>(c, U, aux, dt, 0.0);
  if (rc) return rc;
  auto a = make_args<Real>(c, dt, 0.0);
  bc_kernel<Real><<<bc_blocks(c), kBcThreads, 0, c->stream>>>(a, U, aux, 0);
  launch_stage<Real, MODE_RHS>(c, a, 0, U, aux, U, out, nullptr);
  return check_launch(c, "rhs kernels");
}

// ------------------------------------------------------------------ adjoint

__global__ void fill_kernel(int m, double val, double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) y[t] = val;
}

__global__ void zero_u32_kernel(unsigned int* p, int m) {
  if ((int)threadIdx.x < m) p[threadIdx.x] = 0u;
}

int blocks_for(size_t m) { return (int)((m + 255) / 256); }

// Blocks of the adjoint's y pass (and of alpha_apply_kernel) over the
// ghost-padded grid.
int adj_blocks(const jh_ctx* c) { return ys_blocks(c->nx, c->ny); }

// JH_BC_USER sides must have been written (jh_set_ghost_strip or through
// jh_ghost_strip's pointer) since their last use; consumes the marks.
int take_user_strips(jh_ctx* c) {
  for (int s = 0; s < 4; ++s)
    if (c->bc[s] == JH_BC_USER && !c->user_filled[s])
      return fail(c, JH_ERR_ARG, "user boundary strip not filled for this stage");
  for (int s = 0; s < 4; ++s) c->user_filled[s] = false;
  return JH_OK;
}

struct AdjWork {
  double *U[5], *A[5];          // forward states of the step: 0 = input, 1-3 RK, 4 = output
  double *Ub, *Uacc, *ab, *rb;  // cotangents
  double *Ubp, *abp;            // ghost cells' cotangents, on the padded grid
  double* spd;                  // |u| + a, |v| + a of every padded cell (2 planes)
  double* qb;                   // rho, u, v, p of the stage state (4 planes)
  double* hx;                   // x faces' H+-_bar, [2][NU][nx + 6][ny]
  double *part, *scal;
  unsigned int* cnt;
  unsigned long long* amx;      // recomputed LF maxima (bit patterns)
  int* fold;                    // ghost-fold lists: tgt[m], beg[m + 1], ent[np - n]
  int npad, nblk;
};

int adj_work(jh_ctx* c, const double* U0, const double* A0, AdjWork& w) {
  const size_t n = (size_t)c->nx * c->ny;
  const size_t np = (size_t)(c->nx + 6) * (c->ny + 6);
  const int nblk = adj_blocks(c);
  const size_t ng = np - n;      // ghost cells
  const size_t nhx = 2 * NU * (size_t)(c->nx + 6) * c->ny;
  const size_t len = 4 * (NU + 2) * n + (3 * NU + 2) * n + (NU + 4) * np + 4 * n + nhx +
                     6 * (size_t)nblk + 2 + 2 + 2 + (3 * ng + 2) / 2;
  if (c->adj_len < len) {
    if (c->adj) (void)hipFree(c->adj);
    c->adj = nullptr;
    c->adj_len = 0;
    c->adj_fold_bc[0] = -1;
    HIP_TRY(c, hipMalloc(&c->adj, len * sizeof(double)));
    c->adj_len = len;
  }
  double* p = c->adj;
  w.U[0] = const_cast<double*>(U0);
  w.A[0] = const_cast<double*>(A0);
  for (int i = 1; i <= 4; ++i) {
    w.U[i] = p; p += NU * n;
    w.A[i] = p; p += 2 * n;
  }
  w.Ub = p; p += NU * n;
  w.Uacc = p; p += NU * n;
  w.rb = p; p += NU * n;
  w.ab = p; p += 2 * n;
  w.Ubp = p; p += NU * np;
  w.abp = p; p += 2 * np;
  w.spd = p; p += 2 * np;
  w.qb = p; p += 4 * n;
  w.hx = p; p += nhx;
  w.part = p; p += 6 * nblk;
  w.scal = p; p += 2;
  w.cnt = reinterpret_cast<unsigned int*>(p);
  p += 2;
  w.amx = reinterpret_cast<unsigned long long*>(p);
  p += 2;
  w.fold = reinterpret_cast<int*>(p);
  w.npad = (int)np;
  w.nblk = nblk;
  return JH_OK;
}

AdjArgs adj_args(const jh_ctx* c, double dt) {
  AdjArgs a;
  a.nx = c->nx;
  a.ny = c->ny;
  a.dx = c->dx;
  a.dy = c->dy;
  a.dt = dt;
  a.mc = c->mc;
  for (int s = 0; s < 4; ++s) a.bc[s] = c->bc[s];
  return a;
}

// Ghost-fold lists of fold_ghosts_kernel for the context's grid and
// boundary kinds: the ghost cells grouped by source cell, in padded order
// within a group.  Built on the host and copied once per set of kinds.
int adj_fold(jh_ctx* c, const AdjArgs& a, AdjWork& w) {
  if (std::equal(c->bc, c->bc + 4, c->adj_fold_bc)) return JH_OK;
  const int nx = c->nx, ny = c->ny, PX = nx + 6, PY = ny + 6;
  if ((long long)PX * PY > (1LL << 29))   // entries hold position * 4 in an int
    return fail(c, JH_ERR_ARG, "grid too large for the adjoint");
  std::vector<std::pair<int, int>> g;   // (source cell, padded position * 4 + flips)
  g.reserve((size_t)PX * PY - (size_t)nx * ny);
  for (int P = 0; P < PX; ++P)
    for (int Q = 0; Q < PY; ++Q) {
      if (pad_interior(a, P, Q)) continue;
      int sx, sy, flip;
      if (!pad_src(a, P, Q, sx, sy, flip)) continue;
      g.emplace_back(sx * ny + sy, (P * PY + Q) * 4 + flip);
    }
  std::stable_sort(g.begin(), g.end(),
                   [](const std::pair<int, int>& l, const std::pair<int, int>& r) {
                     return l.first < r.first;
                   });
  std::vector<int> tgt, beg, ent;
  for (size_t e = 0; e < g.size(); ++e) {
    if (e == 0 || g[e].first != g[e - 1].first) {
      tgt.push_back(g[e].first);
      beg.push_back((int)e);
    }
    ent.push_back(g[e].second);
  }
  beg.push_back((int)g.size());
  const int m = (int)tgt.size();
  std::vector<int> buf;
  buf.reserve(tgt.size() + beg.size() + ent.size());
  buf.insert(buf.end(), tgt.begin(), tgt.end());
  buf.insert(buf.end(), beg.begin(), beg.end());
  buf.insert(buf.end(), ent.begin(), ent.end());
  HIP_TRY(c, hipMemcpyAsync(w.fold, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));   // buf is freed on return
  std::copy(c->bc, c->bc + 4, c->adj_fold_bc);
  c->adj_fold_m = m;
  return JH_OK;
}

// Flux VJP of one RK stage with input (Us, As) and rhs cotangent rb; the
// results are added into (Ut, At).
void flux_vjp(jh_ctx* c, const AdjArgs& a, AdjWork& w, const double* Us, const double* As,
              const double* rb, const unsigned long long* alpha, double* Ut, double* At) {
  hipStream_t st = c->stream;
  const int n = a.nx * a.ny;
  adj_prim_kernel<<<blocks_for(n), 256, 0, st>>>(a, Us, As, w.qb);
  flux_vjp_x_kernel<<<xs_blocks(a.nx, a.ny), 256, 0, st>>>(a, Us, w.qb, rb, alpha, w.hx);
  flux_vjp_y_kernel<<<w.nblk, 64 * kYRows, 0, st>>>(a, Us, As, rb, alpha, w.hx, Ut, At, w.Ubp,
                                                    w.abp, w.spd, w.part);
  alpha_reduce_kernel<<<1, 256, 0, st>>>(w.part, w.nblk, w.scal, w.amx, w.cnt);
  alpha_apply_kernel<<<w.nblk, 64 * kYRows, 0, st>>>(a, Us, As, w.spd, w.part, w.amx, w.scal,
                                                     w.cnt, Ut, At, w.Ubp, w.abp);
  const int m = c->adj_fold_m;
  const int* fl = w.fold;
  if (m > 0)
    fold_ghosts_kernel<<<dim3(blocks_for(m), NU + 2), 256, 0, st>>>(a, fl, fl + m, fl + 2 * m + 1,
                                                                    m, w.Ubp, w.abp, Ut, At);
}

// Forward pass of one step keeping every RK stage: w.U[1..3], w.A[1..3] the
// stage states (the third before the chemistry), w.U[4], w.A[4] the output
// (a copy of stage 3 put through the chemistry), or the caller's output
// Uk/Ak, whose chemistry is then not repeated.  The stages' LF maxima are
// left in c->sc's slots of the parity the step ran with (returned in *P).
int vjp_forward(jh_ctx* c, AdjWork& w, const double* U0, const double* A0, double dt,
                const double* Uk, const double* Ak, int* Pout) {
  const size_t n = (size_t)c->nx * c->ny;
  hipStream_t st = c->stream;
  int rc;
  // ---- forward recompute, keeping every stage ----
  rc = prologue<double>(c, U0, A0, dt, 0.0);
  if (rc) return rc;
  const int P = c->parity;
  auto a = make_args<double>(c, dt, 0.0);
  a.mirror = mirror_halo(c) ? 1 : 0;
  const int bb = bc_blocks(c);
  for (int s = 0; s < 3; ++s) {
    if (!a.mirror) bc_kernel<double><<<bb, kBcThreads, 0, st>>>(a, w.U[s], w.A[s], s);
    launch_stage<double, MODE_RK>(c, a, s, w.U[s], w.A[s], U0, w.U[s + 1], w.A[s + 1]);
    halley_fix<double>(c, s, P, s < 2 ? s + 1 : -1, 0, w.A[s + 1]);
  }
  if (Uk && Ak) {
    w.U[4] = const_cast<double*>(Uk);
    w.A[4] = const_cast<double*>(Ak);
  } else {
    HIP_TRY(c, hipMemcpyAsync(w.U[4], w.U[3], NU * n * sizeof(double), hipMemcpyDeviceToDevice,
                              st));
    HIP_TRY(c, hipMemcpyAsync(w.A[4], w.A[3], 2 * n * sizeof(double), hipMemcpyDeviceToDevice,
                              st));
    launch_chem_step<double>(c, a, st, w.U[4], w.A[4], nullptr, false);
    halley_fix<double>(c, JH_PROD_CHEM, P ^ 1, 0, 1, w.A[4]);
  }
  rc = check_launch(c, "vjp forward");
  if (rc) return rc;
  c->parity ^= 1;
  *Pout = P;
  return JH_OK;
}

// Reverse sweep of one step from the forward's states in w and the stages'
// LF maxima alpha[s] (2 bit patterns each); accumulates into w.Uacc, w.ab.
int vjp_reverse(jh_ctx* c, AdjWork& w, double dt, const unsigned long long* const alpha[3],
                const double* Ubar, const double* Abar) {
  const size_t n = (size_t)c->nx * c->ny;
  const int bc1 = blocks_for(n);
  hipStream_t st = c->stream;
  int rc;
  const AdjArgs aa = adj_args(c, dt);
  rc = adj_fold(c, aa, w);
  if (rc) return rc;
  // ---- reverse sweep ----
  // final update_aux and chemistry: Ub = U4_bar -> U3_bar
  getT_vjp_kernel<<<bc1, 256, 0, st>>>(aa, w.U[4], w.A[4], Abar, Ubar, w.Ub);
  // x and lambda pass through buffers that are idle until the flux VJPs
  // (Ubp: 12 padded planes, rb: 12 planes)
  chem_vjp_solve_kernel<<<bc1, 256, 0, st>>>(aa, w.U[3], w.A[3], w.Ub, w.Ubp, w.rb);
  chem_vjp_rates_kernel<<<bc1, 256, 0, st>>>(aa, w.U[3], w.A[3], w.Ub, w.Ubp, w.rb, w.Ub, w.ab);
  // each getT_split_kernel: get_T's VJP at the stage state, then the
  // reverse of the stage's combine
  // stage 3: U3 = 1/3 U + 2/3 (U2 + r(U2))
  getT_split_kernel<<<bc1, 256, 0, st>>>(aa, w.U[3], w.A[3], 1.0 / 3.0, 0.0, 2.0 / 3.0, w.Ub,
                                         w.Uacc, w.rb, w.ab);
  flux_vjp(c, aa, w, w.U[2], w.A[2], w.rb, alpha[2], w.Ub, w.ab);
  // stage 2: U2 = 3/4 U + 1/4 (U1 + r(U1))
  getT_split_kernel<<<bc1, 256, 0, st>>>(aa, w.U[2], w.A[2], 3.0 / 4.0, 1.0, 1.0 / 4.0, w.Ub,
                                         w.Uacc, w.rb, w.ab);
  flux_vjp(c, aa, w, w.U[1], w.A[1], w.rb, alpha[1], w.Ub, w.ab);
  // stage 1: U1 = U + r(U)
  getT_split_kernel<<<bc1, 256, 0, st>>>(aa, w.U[1], w.A[1], 1.0, 1.0, 0.0, w.Ub, w.Uacc,
                                         nullptr, w.ab);
  flux_vjp(c, aa, w, w.U[0], w.A[0], w.Ub, alpha[0], w.Uacc, w.ab);
  return check_launch(c, "vjp kernels");
}

// Uk/Ak: the step's output when the caller already has it (the RK stages are
// still recomputed; only the chemistry's forward pass, whose output the
// reverse sweep starts from, is skipped).
int step_vjp_impl(jh_ctx* c, const double* U0, const double* A0, double dt, const double* Ubar,
                  const double* Abar, double* Ub0, double* Ab0, double* Uo, double* Ao,
                  const double* Uk = nullptr, const double* Ak = nullptr) {
  AdjWork w;
  int rc = adj_work(c, U0, A0, w);
  if (rc) return rc;
  // the sweep accumulates straight into the outputs: the input cotangents
  // are read by the first kernel (getT_vjp_kernel) only, before either
  // output is written, so the outputs may alias them
  w.Uacc = static_cast<double*>(Ub0);
  w.ab = static_cast<double*>(Ab0);
  const size_t n = (size_t)c->nx * c->ny;
  hipStream_t st = c->stream;
  int P = 0;
  rc = vjp_forward(c, w, U0, A0, dt, Uk, Ak, &P);
  if (rc) return rc;
  if (Uo) HIP_TRY(c, hipMemcpyAsync(Uo, w.U[4], NU * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (Ao) HIP_TRY(c, hipMemcpyAsync(Ao, w.A[4], 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  const unsigned long long* const alpha[3] = {&c->sc->amax[P][0][0], &c->sc->amax[P][1][0],
                                              &c->sc->amax[P][2][0]};
  return vjp_reverse(c, w, dt, alpha, Ubar, Abar);
}

// Residuals of one FP64 step (jh_step_keep, jh_step_vjp_res): the three RK
// stage states, each NU + 2 planes (U then aux; the third before the
// chemistry), then the three stages' LF maxima (x, y) as 6 bit patterns.
size_t res_count(const jh_ctx* c) { return 3 * (size_t)(NU + 2) * c->nx * c->ny + 6; }

void res_views(const jh_ctx* c, double* res, AdjWork& w, const unsigned long long** alpha) {
  const size_t n = (size_t)c->nx * c->ny;
  for (int s = 1; s <= 3; ++s) {
    w.U[s] = res + (size_t)(s - 1) * (NU + 2) * n;
    w.A[s] = w.U[s] + NU * n;
  }
  const auto* bits = reinterpret_cast<const unsigned long long*>(res + 3 * (NU + 2) * n);
  for (int s = 0; s < 3; ++s) alpha[s] = bits + 2 * s;
}

int step_keep_impl(jh_ctx* c, const double* U0, const double* A0, double dt, double* Uo,
                   double* Ao, double* res) {
  AdjWork w{};
  const unsigned long long* alpha[3];
  res_views(c, res, w, alpha);
  w.U[0] = const_cast<double*>(U0);
  w.A[0] = const_cast<double*>(A0);
  w.U[4] = Uo;
  w.A[4] = Ao;
  int P = 0;
  int rc = vjp_forward(c, w, U0, A0, dt, nullptr, nullptr, &P);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(const_cast<unsigned long long*>(alpha[0]), &c->sc->amax[P][0][0],
                            6 * sizeof(unsigned long long), hipMemcpyDeviceToDevice, c->stream));
  return JH_OK;
}

int step_vjp_res_impl(jh_ctx* c, const double* U0, const double* A0, double dt,
                      const double* res, const double* Uk, const double* Ak, const double* Ubar,
                      const double* Abar, double* Ub0, double* Ab0) {
  AdjWork w;
  int rc = adj_work(c, U0, A0, w);
  if (rc) return rc;
  const unsigned long long* alpha[3];
  res_views(c, const_cast<double*>(res), w, alpha);
  w.U[4] = const_cast<double*>(Uk);
  w.A[4] = const_cast<double*>(Ak);
  w.Uacc = Ub0;
  w.ab = Ab0;
  return vjp_reverse(c, w, dt, alpha, Ubar, Abar);
}

}  // namespace

// Device-side copy of the few scalars that a host-side query needs.
__global__ void copy_scalars_kernel(const StepScalars* sc, int parity, double cfl, double* dt,
                                    double* alpha) {
  if (threadIdx.x != 0) return;
  if (dt) *dt = cfl * (1.0 / bitsd(sc->cmax[parity]));
  if (alpha) {
    alpha[0] = bitsd(sc->amax[parity][0][0]);
    alpha[1] = bitsd(sc->amax[parity][0][1]);
  }
}


// ------------------------------------------------------------------ batch

struct jh_batch {
  int ncase = 0;
  int precision = JH_PREC_FP64;
  std::vector<int> nx, ny, bc;        // bc: 4 per case
  bool mirror = false;                // every side of every case a copy kind
  std::vector<double> dx, dy, bcp;    // bcp: 4 per case
  std::vector<double> inj;            // per case: T0, Y_inj[NS-1]
  hipStream_t user = nullptr;         // the caller's stream
  hipStream_t own = nullptr;          // launches and graph capture
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  MechConst* mc = nullptr;
  StepScalars* sc = nullptr;          // one per case
  void* bufs = nullptr;               // stage buffers and ghost strips of all cases
  std::vector<size_t> buf_off;        // per case: byte offset of its block in bufs
  void* desc = nullptr;               // device CaseDesc[ncase]
  int* off = nullptr;                 // device block tables
  std::vector<int> h_off;
  int t_stage[kTileVariants] = {}, t_bc = 0, t_red = 0, t_chem = 0;   // table starts in off
  int persist[kTileVariants] = {};    // persistent grid of each stage variant
  int tile = 0;
  std::vector<void*> U, aux;
  bool bound = false;
  unsigned gen = 0;                   // bumped by every bind
  hipGraphExec_t gexec = nullptr;
  struct Key {
    int nsteps = -1;
    double dt = 0, cfl = 0;
    double* hist = nullptr;
    int tile = -1;
    unsigned gen = 0;
  } key;
  std::string err;
};

namespace {

thread_local std::string g_batch_err;

int bfail(jh_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg; else g_batch_err = msg;
  return code;
}

#define BHIP_TRY(b, call)                                                      \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return bfail(b, JH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Per-case buffer block: U1, U2 (NU planes), A1, A2 (2 planes), 4 strips.
size_t case_bytes(size_t es, int nx, int ny) {
  const size_t n = (size_t)nx * ny;
  size_t t = 2 * align256(es * NU * n) + 2 * align256(es * 2 * n);
  for (int s = 0; s < 4; ++s) t += align256(es * NV * 3 * (s < 2 ? ny : nx));
  return t;
}

template <typename Real>
void fill_desc(const jh_batch* b, int k, CaseDesc<Real>& d) {
  std::memset(&d, 0, sizeof(d));
  char* base = static_cast<char*>(b->bufs) + b->buf_off[k];
  const size_t es = sizeof(Real), n = (size_t)b->nx[k] * b->ny[k];
  StepArgs<Real>& a = d.a;
  a.nx = b->nx[k];
  a.ny = b->ny[k];
  a.dx = b->dx[k];
  a.dy = b->dy[k];
  a.mc = b->mc;
  a.sc = b->sc + k;
  d.U1 = reinterpret_cast<Real*>(base);
  base += align256(es * NU * n);
  d.U2 = reinterpret_cast<Real*>(base);
  base += align256(es * NU * n);
  d.A1 = reinterpret_cast<Real*>(base);
  base += align256(es * 2 * n);
  d.A2 = reinterpret_cast<Real*>(base);
  base += align256(es * 2 * n);
  for (int s = 0; s < 4; ++s) {
    a.g[s] = reinterpret_cast<Real*>(base);
    base += align256(es * NV * 3 * (s < 2 ? b->ny[k] : b->nx[k]));
    a.bc[s] = b->bc[4 * k + s];
    a.bc_p[s] = b->bcp[4 * k + s];
  }
  a.inj_T0 = b->inj[(size_t)k * JH_NS];
  for (int j = 0; j < JH_NS - 1; ++j) a.inj_Y[j] = b->inj[(size_t)k * JH_NS + 1 + j];
  d.U = b->bound ? static_cast<Real*>(b->U[k]) : nullptr;
  d.aux = b->bound ? static_cast<Real*>(b->aux[k]) : nullptr;
}

template <typename Real>
int upload_desc(jh_batch* b) {
  std::vector<CaseDesc<Real>> h(b->ncase);
  for (int k = 0; k < b->ncase; ++k) fill_desc<Real>(b, k, h[k]);
  BHIP_TRY(b, hipStreamSynchronize(b->own));
  BHIP_TRY(b, hipMemcpy(b->desc, h.data(), sizeof(CaseDesc<Real>) * b->ncase,
                        hipMemcpyHostToDevice));
  return JH_OK;
}

template <typename Real, int TX, int WPE>
int persistent_grid(jh_batch* b, int v, int cus) {
  int per_cu = 0;
  BHIP_TRY(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, stage_batch_kernel<Real, TX, WPE>, Tile<TX>::NT, 0));
  if (per_cu < 1) per_cu = 1;
  int g = (per_cu * cus) & ~7;
  if (g < 8) g = 8;
  b->persist[v] = g;
  return JH_OK;
}

template <typename Real>
int setup_grids(jh_batch* b) {
  int dev = 0, cus = 0;
  BHIP_TRY(b, hipGetDevice(&dev));
  BHIP_TRY(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int rc = JH_OK;
#define JH_GRID(V, TX, WPE) \
  if (!rc) rc = persistent_grid<Real, TX, WPE>(b, V, cus);
  JH_TILE_VARIANTS(JH_GRID)
#undef JH_GRID
  return rc;
}

template <typename Real>
void launch_stage_batch(jh_batch* b, hipStream_t st, int stage, int parity, double dt,
                        double cfl) {
  const auto* d = static_cast<const CaseDesc<Real>*>(b->desc);
  const int v = b->tile;
  const int total = b->h_off[b->t_stage[v] + b->ncase];
  const int g = std::min(b->persist[v], total);
  const int* off = b->off + b->t_stage[v];
  switch (v) {
#define JH_LAUNCH(V, TX, WPE)                                                           \
    case V:                                                                             \
      if (b->mirror)                                                                    \
        stage_batch_kernel<Real, TX, WPE, true><<<g, Tile<TX>::NT, 0, st>>>(            \
            d, off, b->ncase, stage, parity, dt, cfl);                                  \
      else                                                                              \
        stage_batch_kernel<Real, TX, WPE><<<g, Tile<TX>::NT, 0, st>>>(                  \
            d, off, b->ncase, stage, parity, dt, cfl);                                  \
      break;
    JH_TILE_VARIANTS(JH_LAUNCH)
#undef JH_LAUNCH
    default:
      break;
  }
}

// The launch sequence of nsteps batched steps (also what a graph captures).
template <typename Real>
int batch_enqueue(jh_batch* b, hipStream_t st, double dt, double cfl, int nsteps,
                  double* hist) {
  const auto* d = static_cast<const CaseDesc<Real>*>(b->desc);
  const int nc = b->ncase;
  const int n_red = b->h_off[b->t_red + nc], n_bc = b->h_off[b->t_bc + nc];
  const int n_chem = b->h_off[b->t_chem + nc];
  reset_batch_kernel<Real><<<(nc + 63) / 64, 64, 0, st>>>(d, nc, 0);
  reduce_batch_kernel<Real><<<n_red, 256, 0, st>>>(d, b->off + b->t_red, nc, 0, dt, cfl);
  for (int s = 0; s < nsteps; ++s) {
    const int parity = s & 1;
    for (int stage = 0; stage < 3; ++stage) {
      // copy-kind sides only: the stage kernel reads the halo from the
      // source cells and zeroes the other parity's slots at stage 0
      if (!b->mirror)
        bc_batch_kernel<Real><<<n_bc, kBcThreads, 0, st>>>(d, b->off + b->t_bc, nc, stage, parity,
                                                           dt, cfl);
      launch_stage_batch<Real>(b, st, stage, parity, dt, cfl);
    }
    chem_batch_kernel<Real><<<n_chem, 256, 0, st>>>(d, b->off + b->t_chem, nc, parity, dt, cfl,
                                                     hist, nsteps, s);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, JH_ERR_HIP, std::string("batch kernels: ") +
                                                       hipGetErrorString(e));
  return JH_OK;
}

int batch_enqueue_any(jh_batch* b, hipStream_t st, double dt, double cfl, int nsteps,
                      double* hist) {
  return b->precision == JH_PREC_FP64 ? batch_enqueue<double>(b, st, dt, cfl, nsteps, hist)
                                      : batch_enqueue<float>(b, st, dt, cfl, nsteps, hist);
}

void drop_graph(jh_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->key = jh_batch::Key();
}

}  // namespace

// ------------------------------------------------------------------ AMR blocks

struct jh_blocks {
  int nx = 0, ny = 0;                 // block interior
  double dx = 0, dy = 0;
  int precision = JH_PREC_FP64;
  hipStream_t stream = nullptr;
  MechConst* mc = nullptr;
  StepScalars* sc = nullptr;          // one per block
  int cap = 0;                        // blocks sc holds
  StepScalars* scratch = nullptr;     // device address of `scratch` on this handle's GPU
  std::string err;
};

// Per-block scalars of jh_blocks_stage_noalloc: a file-scope device array,
// so that entry point allocates nothing.  Every GPU the library runs on holds
// its own copy (kNoallocBlocks x sizeof(StepScalars) = 3.4 MB) until the
// process ends; calls that overlap on one GPU share it and race.
constexpr int kNoallocBlocks = 2048;
static __device__ __align__(256) StepScalars scratch[kNoallocBlocks];

namespace {

thread_local std::string g_blocks_err;

// scratch's device address per GPU, taken once per device
// (hipGetSymbolAddress with that device current); the host-side name
// `scratch` is not a device address and is never passed to a kernel.
std::mutex g_scratch_mu;
std::vector<StepScalars*> g_scratch_addr;

hipError_t scratch_address(StepScalars** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[dev] = static_cast<StepScalars*>(p);
  }
  *out = g_scratch_addr[dev];
  return hipSuccess;
}

int kfail(jh_blocks* b, int code, const std::string& msg) {
  if (b) b->err = msg; else g_blocks_err = msg;
  return code;
}

#define KHIP_TRY(b, call)                                                      \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return kfail(b, JH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Room for n_blk blocks' scalars (flags start at zero).
int blocks_reserve(jh_blocks* b, int n_blk) {
  if (n_blk <= b->cap) return JH_OK;
  KHIP_TRY(b, hipStreamSynchronize(b->stream));
  StepScalars* sc = nullptr;
  KHIP_TRY(b, hipMalloc(&sc, sizeof(StepScalars) * n_blk));
  KHIP_TRY(b, hipMemset(sc, 0, sizeof(StepScalars) * n_blk));
  if (b->sc) (void)hipFree(b->sc);
  b->sc = sc;
  b->cap = n_blk;
  return JH_OK;
}

// sc: the per-block scalars (the handle's, or the no-allocation scratch)
template <typename Real>
StepArgs<Real> blocks_args(const jh_blocks* b, double dt, StepScalars* sc) {
  StepArgs<Real> a;
  std::memset(&a, 0, sizeof(a));
  a.nx = b->nx;
  a.ny = b->ny;
  a.dx = b->dx;
  a.dy = b->dy;
  a.dt_fixed = dt;
  a.mc = b->mc;
  a.sc = sc;
  return a;
}

template <typename Real, int MODE>
int blocks_stage_impl(jh_blocks* b, int n_blk, int stage, const Real* Bs, const Real* Bn,
                      Real* Bo, double dt, StepScalars* sc) {
  const long long P = (long long)(b->nx + 6) * (b->ny + 6);
  const long long bs = (NU + 2) * P, bo = MODE == MODE_RHS ? NU * P : (NU + 2) * P;
  auto a = blocks_args<Real>(b, dt, sc);
  hipStream_t st = b->stream;
  blocks_alpha_kernel<Real><<<n_blk, 256, 0, st>>>(a, stage, bs, Bs);
  const int tiles = ((b->nx + kBlkRows - 1) / kBlkRows) * ((b->ny + TY - 1) / TY);
  blocks_stage_kernel<Real, MODE, kBlkRows, 4>
      <<<n_blk * tiles, Tile<kBlkRows>::NT, 0, st>>>(a, stage, tiles, bs, bo, Bs, Bn, Bo);
  const int nring = (int)(P - (long long)b->nx * b->ny);
  blocks_ring_kernel<Real, MODE>
      <<<dim3((nring + 255) / 256, n_blk), 256, 0, st>>>(a, stage, bs, bo, Bs, Bn, Bo);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kfail(b, JH_ERR_HIP, std::string("block kernels: ") +
                                                       hipGetErrorString(e));
  return JH_OK;
}

// reserve: grow the handle's per-block scalars (allocates); the
// no-allocation entry point checks sizes only
int blocks_checks(jh_blocks* b, int n_blk, bool reserve = true) {
  if (n_blk < 1) return kfail(b, JH_ERR_ARG, "n_blk must be >= 1");
  if ((long long)n_blk * (NU + 2) * (b->nx + 6) * (b->ny + 6) >= (1ll << 31))
    return kfail(b, JH_ERR_ARG, "block batch too large for 32-bit offsets");
  return reserve ? blocks_reserve(b, n_blk) : JH_OK;
}

int stage_args_check(jh_blocks* b, int stage, const void* blk_s, const void* blk_n, double dt,
                     const void* blk_out) {
  if (!blk_s || !blk_n || !blk_out) return kfail(b, JH_ERR_ARG, "NULL pointer");
  if (blk_out == blk_s || blk_out == blk_n)
    return kfail(b, JH_ERR_ARG, "blk_out must not alias the inputs");
  if (stage < 0 || stage > 2) return kfail(b, JH_ERR_ARG, "stage must be 0, 1 or 2");
  if (!(dt > 0)) return kfail(b, JH_ERR_ARG, "dt must be > 0");
  return JH_OK;
}

}  // namespace

extern "C" {

int jh_mech_const_count(void) { return (int)(sizeof(MechConst) / sizeof(double)); }
const char* jh_mech_digest(void) { return JH_MECH_DIGEST; }
int jh_num_species(void) { return JH_NS; }

int jh_ctx_create(int nx, int ny, double dx, double dy, const int* bc_kind,
                  const double* bc_param, int precision, const double* mech, int n_mech,
                  void* stream, jh_ctx** out) {
  if (!out) return fail(nullptr, JH_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (nx < 4 || ny < 4) return fail(nullptr, JH_ERR_ARG, "nx and ny must be >= 4");
  if ((long long)nx * ny * NU >= (1ll << 31))
    return fail(nullptr, JH_ERR_ARG, "grid too large for 32-bit plane offsets");
  if (!(dx > 0) || !(dy > 0)) return fail(nullptr, JH_ERR_ARG, "dx, dy must be > 0");
  if (precision != JH_PREC_FP64 && precision != JH_PREC_MIXED)
    return fail(nullptr, JH_ERR_ARG, "unknown precision");
  if (!bc_kind || !mech) return fail(nullptr, JH_ERR_ARG, "bc_kind or mech is NULL");
  if (n_mech != jh_mech_const_count())
    return fail(nullptr, JH_ERR_MECH, "mechanism constant count mismatch");
  for (int s = 0; s < 4; ++s)
    if (bc_kind[s] < JH_BC_PERIODIC || bc_kind[s] > JH_BC_INJECTION_INLET)
      return fail(nullptr, JH_ERR_ARG, "unknown boundary kind");
  jh_ctx* c = new jh_ctx();
  c->nx = nx;
  c->ny = ny;
  c->dx = dx;
  c->dy = dy;
  c->precision = precision;
  // default stage-kernel variant (the autotuner's usual pick on MI355X):
  // 16 rows with the 5-wave register budget for FP64, 16 rows at 4 waves/SIMD
  // for the mixed form
  c->tile = precision == JH_PREC_FP64 ? 3 : 2;
  c->stream = static_cast<hipStream_t>(stream);
  for (int s = 0; s < 4; ++s) {
    c->bc[s] = bc_kind[s];
    c->bc_p[s] = bc_param ? bc_param[s] : 0.0;
  }
  const size_t es = elem_size(c), n = (size_t)nx * ny;
  hipError_t e = hipSuccess;
  e = e ? e : hipMalloc(&c->mc, sizeof(MechConst));
  e = e ? e : hipMalloc(&c->sc, sizeof(StepScalars));
  e = e ? e : hipMalloc(&c->rpart, 3 * sizeof(unsigned long long) * grid_cells(c));
  for (int s = 0; s < 4 && !e; ++s) {
    const size_t len = (s < 2) ? ny : nx;
    e = hipMalloc(&c->ghost[s], es * NV * 3 * len);
  }
  for (int i = 0; i < 2 && !e; ++i) e = hipMalloc(&c->tmpU[i], es * NU * n);
  for (int i = 0; i < 2 && !e; ++i) e = hipMalloc(&c->tmpA[i], es * 2 * n);
  if (!e) e = hipMemcpy(c->mc, mech, sizeof(MechConst), hipMemcpyHostToDevice);
  if (!e) e = hipMemset(c->sc, 0, sizeof(StepScalars));
  for (int s = 0; s < 4 && !e; ++s) {
    const size_t len = (s < 2) ? ny : nx;
    e = hipMemset(c->ghost[s], 0, es * NV * 3 * len);
  }
  if (e) {
    g_create_err = std::string("allocation: ") + hipGetErrorString(e);
    jh_ctx_destroy(c);
    return JH_ERR_HIP;
  }
  *out = c;
  return JH_OK;
}

int jh_ctx_destroy(jh_ctx* c) {
  if (!c) return JH_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream); else (void)hipDeviceSynchronize();
  for (auto& m : c->marks) {
    c->ev_free.push_back(m.a);
    c->ev_free.push_back(m.b);
  }
  for (auto e : c->ev_free) (void)hipEventDestroy(e);
  if (c->adj) (void)hipFree(c->adj);
  if (c->trs) (void)hipFree(c->trs);
  if (c->hside) (void)hipFree(c->hside);
  if (c->hnit) (void)hipFree(c->hnit);
  (void)hipFree(c->mc);
  (void)hipFree(c->sc);
  if (c->rpart) (void)hipFree(c->rpart);
  for (int s = 0; s < 4; ++s) (void)hipFree(c->ghost[s]);
  for (int i = 0; i < 2; ++i) {
    (void)hipFree(c->tmpU[i]);
    (void)hipFree(c->tmpA[i]);
  }
  delete c;
  return JH_OK;
}

int jh_ctx_set_bc_param(jh_ctx* c, int side, double value) {
  if (!c) return JH_ERR_ARG;
  if (side < 0 || side > 3) return fail(c, JH_ERR_ARG, "side out of range");
  c->bc_p[side] = value;
  return JH_OK;
}

int jh_ctx_set_inlet(jh_ctx* c, double T0, const double* Y, int n) {
  if (!c) return JH_ERR_ARG;
  if (!Y || n != JH_NS - 1) return fail(c, JH_ERR_ARG, "Y_inj needs num_species-1 values");
  if (!(T0 > 0)) return fail(c, JH_ERR_ARG, "T0 must be > 0");
  c->inj_T0 = T0;
  for (int k = 0; k < n; ++k) c->inj_Y[k] = Y[k];
  return JH_OK;
}

int jh_ctx_set_tile(jh_ctx* c, int variant) {
  if (!c) return JH_ERR_ARG;
  if (variant < 0 || variant >= jh_tile_variants()) return fail(c, JH_ERR_ARG, "unknown tile variant");
  c->tile = variant;
  return JH_OK;
}

int jh_tile_variants(void) { return kTileVariants; }

int jh_ctx_set_halley_stop(jh_ctx* c, int mode) {
  if (!c) return JH_ERR_ARG;
  if (mode != JH_HALLEY_PER_CELL && mode != JH_HALLEY_GRID)
    return fail(c, JH_ERR_ARG, "unknown temperature-solve stop mode");
  if (mode == JH_HALLEY_GRID && !c->hside) {
    const size_t n = (size_t)c->nx * c->ny;
    HIP_TRY(c, hipMalloc(&c->hside, elem_size(c) * 2 * n));
    HIP_TRY(c, hipMalloc(&c->hnit, n));
  }
  c->halley = mode;
  return JH_OK;
}

int jh_ctx_get_halley_stop(const jh_ctx* c) { return c ? c->halley : -1; }

int jh_ctx_set_clip_branches(jh_ctx* c, unsigned short* masks, int max_steps, int mode) {
  if (!c) return JH_ERR_ARG;
  if (mode != JH_CLIP_OFF && mode != JH_CLIP_RECORD && mode != JH_CLIP_REPLAY)
    return fail(c, JH_ERR_ARG, "unknown clip-branch mode");
  if (mode != JH_CLIP_OFF && (!masks || max_steps < 1))
    return fail(c, JH_ERR_ARG, "clip-branch masks need a buffer of max_steps >= 1 steps");
  c->clipm = mode ? masks : nullptr;
  c->clip_mode = mode;
  c->clip_cap = mode ? max_steps : 0;
  c->clip_step = 0;
  return JH_OK;
}

int jh_ctx_set_trace(jh_ctx* c, long long* records, int max_steps) {
  if (!c) return JH_ERR_ARG;
  if (records && max_steps < 1) return fail(c, JH_ERR_ARG, "max_steps must be >= 1");
  if (records && !c->trs) {
    HIP_TRY(c, hipMalloc(&c->trs, sizeof(TraceScratch)));
    trace_store_kernel<<<1, 64, 0, c->stream>>>(c->trs, nullptr);
    const int rc = check_launch(c, "trace_store_kernel");
    if (rc) return rc;
  }
  c->trace = records;
  c->trace_cap = records ? max_steps : 0;
  c->trace_step = 0;
  return JH_OK;
}

int jh_ctx_set_flux(jh_ctx* c, int solver_type, int split_method, int reconstruction,
                    int riemann_solver) {
  if (!c) return JH_ERR_ARG;
  if (solver_type != JH_SOLVER_FLUX_SPLITTING && solver_type != JH_SOLVER_GODUNOV)
    return fail(c, JH_ERR_ARG, "solver_type must be flux_splitting or godunov");
  if (split_method != JH_SPLIT_LF && split_method != JH_SPLIT_SW)
    return fail(c, JH_ERR_ARG, "split_method must be LF or SW");
  if (reconstruction != JH_RECON_WENO5_JS && reconstruction != JH_RECON_MUSCL)
    return fail(c, JH_ERR_ARG, "interface_reconstruction must be WENO5_JS or MUSCL");
  if (riemann_solver != JH_RIEMANN_HLLC)
    return fail(c, JH_ERR_ARG, "riemann_solver must be HLLC");
  c->solver = solver_type;
  c->split = split_method;
  c->recon = reconstruction;
  return JH_OK;
}

int jh_tile_info(int variant, int* rows, int* threads, int* waves) {
  if (variant < 0 || variant >= kTileVariants || !rows || !threads || !waves) return JH_ERR_ARG;
  *rows = kTileRows[variant];
  *threads = 32 * kTileRows[variant];
  *waves = kTileWaves[variant];
  return JH_OK;
}

int jh_ctx_get_tile(const jh_ctx* c) { return c ? c->tile : -1; }

int jh_ctx_set_stream(jh_ctx* c, void* stream) {
  if (!c) return JH_ERR_ARG;
  c->stream = static_cast<hipStream_t>(stream);
  return JH_OK;
}

const char* jh_last_error(const jh_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int jh_advance(jh_ctx* c, const void* U, const void* aux, void* U_out, void* aux_out, double dt,
               double cfl, int nsteps, double* dt_hist) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux || !U_out || !aux_out) return fail(c, JH_ERR_ARG, "NULL state pointer");
  if (nsteps < 0) return fail(c, JH_ERR_ARG, "nsteps < 0");
  if (!(dt > 0) && !(cfl > 0)) return fail(c, JH_ERR_ARG, "need dt > 0 or cfl > 0");
  for (int s = 0; s < 4; ++s)
    if (c->bc[s] == JH_BC_USER)
      return fail(c, JH_ERR_ARG, "jh_advance needs built-in boundary kinds on every side");
  if (nsteps == 0) return JH_OK;
  if (c->precision == JH_PREC_FP64)
    return advance_impl<double>(c, (const double*)U, (const double*)aux, (double*)U_out,
                                (double*)aux_out, dt, cfl, nsteps, dt_hist);
  return advance_impl<float>(c, (const float*)U, (const float*)aux, (float*)U_out,
                             (float*)aux_out, dt, cfl, nsteps, dt_hist);
}

int jh_rhs(jh_ctx* c, const void* U, const void* aux, double dt, void* out) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux || !out) return fail(c, JH_ERR_ARG, "NULL pointer");
  if (!(dt > 0)) return fail(c, JH_ERR_ARG, "dt must be > 0");
  int rc = take_user_strips(c);
  if (rc) return rc;
  if (c->precision == JH_PREC_FP64)
    return rhs_impl<double>(c, (const double*)U, (const double*)aux, dt, (double*)out);
  return rhs_impl<float>(c, (const float*)U, (const float*)aux, dt, (float*)out);
}

int jh_update_aux(jh_ctx* c, const void* U, const void* aux_in, void* aux_out) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux_in || !aux_out) return fail(c, JH_ERR_ARG, "NULL pointer");
  if (c->halley == JH_HALLEY_GRID)
    reset_grid_slot_kernel<<<1, 64, 0, c->stream>>>(c->sc, c->parity, JH_PROD_UPDATE);
  if (c->precision == JH_PREC_FP64) {
    auto a = make_args<double>(c, 1.0, 0.0);
    update_aux_kernel<double><<<grid_all(c), 256, 0, c->stream>>>(
        a, (const double*)U, (const double*)aux_in, (double*)aux_out);
    halley_fix<double>(c, JH_PROD_UPDATE, c->parity, -1, 0, (double*)aux_out);
  } else {
    auto a = make_args<float>(c, 1.0, 0.0);
    update_aux_kernel<float><<<grid_all(c), 256, 0, c->stream>>>(
        a, (const float*)U, (const float*)aux_in, (float*)aux_out);
    halley_fix<float>(c, JH_PROD_UPDATE, c->parity, -1, 0, (float*)aux_out);
  }
  return check_launch(c, "update_aux_kernel");
}

int jh_reaction_source(jh_ctx* c, const void* U, const void* aux, double dt, void* S) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux || !S) return fail(c, JH_ERR_ARG, "NULL pointer");
  if (!(dt > 0)) return fail(c, JH_ERR_ARG, "dt must be > 0");
  if (c->precision == JH_PREC_FP64) {
    auto a = make_args<double>(c, dt, 0.0);
    chem_kernel<double, CHEM_SOURCE><<<grid_all(c), 256, 0, c->stream>>>(
        a, (double*)U, (double*)aux, (double*)S, nullptr);
  } else {
    auto a = make_args<float>(c, dt, 0.0);
    chem_kernel<float, CHEM_SOURCE><<<grid_all(c), 256, 0, c->stream>>>(
        a, (float*)U, (float*)aux, (float*)S, nullptr);
  }
  return check_launch(c, "chem_kernel");
}

int jh_cfl_dt(jh_ctx* c, const void* U, const void* aux, double cfl, double* dt_dev) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux || !dt_dev) return fail(c, JH_ERR_ARG, "NULL pointer");
  int rc = c->precision == JH_PREC_FP64
               ? prologue<double>(c, (const double*)U, (const double*)aux, 0.0, cfl)
               : prologue<float>(c, (const float*)U, (const float*)aux, 0.0, cfl);
  if (rc) return rc;
  copy_scalars_kernel<<<1, 64, 0, c->stream>>>(c->sc, c->parity, cfl, dt_dev, nullptr);
  return check_launch(c, "copy_scalars_kernel");
}

int jh_lf_alpha(jh_ctx* c, const void* U, const void* aux, double* alpha_dev) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux || !alpha_dev) return fail(c, JH_ERR_ARG, "NULL pointer");
  int rc = take_user_strips(c);
  if (rc) return rc;
  if (c->precision == JH_PREC_FP64) {
    rc = prologue<double>(c, (const double*)U, (const double*)aux, 1.0, 0.0);
    if (rc) return rc;
    auto a = make_args<double>(c, 1.0, 0.0);
    bc_kernel<double><<<bc_blocks(c), kBcThreads, 0, c->stream>>>(a, (const double*)U,
                                                            (const double*)aux, 0);
  } else {
    rc = prologue<float>(c, (const float*)U, (const float*)aux, 1.0, 0.0);
    if (rc) return rc;
    auto a = make_args<float>(c, 1.0, 0.0);
    bc_kernel<float><<<bc_blocks(c), kBcThreads, 0, c->stream>>>(a, (const float*)U,
                                                          (const float*)aux, 0);
  }
  copy_scalars_kernel<<<1, 64, 0, c->stream>>>(c->sc, c->parity, 0.0, nullptr, alpha_dev);
  return check_launch(c, "lf_alpha kernels");
}

namespace {
int vjp_checks(jh_ctx* c, const void* U, const void* aux, double dt, const void* U_bar,
               const void* aux_bar, const void* U_in_bar, const void* aux_in_bar) {
  if (!U || !aux || !U_bar || !aux_bar || !U_in_bar || !aux_in_bar)
    return fail(c, JH_ERR_ARG, "NULL pointer");
  if (c->precision != JH_PREC_FP64) return fail(c, JH_ERR_ARG, "the adjoint is FP64 only");
  if (!(dt > 0)) return fail(c, JH_ERR_ARG, "the adjoint needs a fixed dt > 0");
  if (c->solver != JH_SOLVER_FLUX_SPLITTING || c->split != JH_SPLIT_LF ||
      c->recon != JH_RECON_WENO5_JS)
    return fail(c, JH_ERR_ARG, "the adjoint supports LF flux splitting with WENO5_JS only");
  for (int s = 0; s < 4; ++s)
    if (c->bc[s] == JH_BC_PRESSURE_OUTLET || c->bc[s] == JH_BC_USER ||
        c->bc[s] == JH_BC_INJECTION_INLET)
      return fail(c, JH_ERR_ARG, "the adjoint supports periodic, zero_gradient, slip_wall and "
                                 "no_slip_wall boundaries");
  return JH_OK;
}
}  // namespace

int jh_step_vjp(jh_ctx* c, const void* U, const void* aux, double dt, const void* U_bar,
                const void* aux_bar, void* U_in_bar, void* aux_in_bar, void* U_out,
                void* aux_out) {
  if (!c) return JH_ERR_ARG;
  int rc = vjp_checks(c, U, aux, dt, U_bar, aux_bar, U_in_bar, aux_in_bar);
  if (rc) return rc;
  return step_vjp_impl(c, (const double*)U, (const double*)aux, dt, (const double*)U_bar,
                       (const double*)aux_bar, (double*)U_in_bar, (double*)aux_in_bar,
                       (double*)U_out, (double*)aux_out);
}

int jh_step_vjp_known(jh_ctx* c, const void* U, const void* aux, double dt, const void* U_next,
                      const void* aux_next, const void* U_bar, const void* aux_bar,
                      void* U_in_bar, void* aux_in_bar) {
  if (!c) return JH_ERR_ARG;
  if (!U_next || !aux_next) return fail(c, JH_ERR_ARG, "NULL pointer");
  int rc = vjp_checks(c, U, aux, dt, U_bar, aux_bar, U_in_bar, aux_in_bar);
  if (rc) return rc;
  return step_vjp_impl(c, (const double*)U, (const double*)aux, dt, (const double*)U_bar,
                       (const double*)aux_bar, (double*)U_in_bar, (double*)aux_in_bar, nullptr,
                       nullptr, (const double*)U_next, (const double*)aux_next);
}

int jh_step_res_count(jh_ctx* c, long long* count) {
  if (!c || !count) return JH_ERR_ARG;
  *count = (long long)res_count(c);
  return JH_OK;
}

int jh_step_keep(jh_ctx* c, const void* U, const void* aux, double dt, void* U_next,
                 void* aux_next, void* res) {
  if (!c) return JH_ERR_ARG;
  if (!U_next || !aux_next || !res) return fail(c, JH_ERR_ARG, "NULL pointer");
  // the adjoint's conditions (the cotangent arguments are not used here)
  int rc = vjp_checks(c, U, aux, dt, res, res, res, res);
  if (rc) return rc;
  return step_keep_impl(c, (const double*)U, (const double*)aux, dt, (double*)U_next,
                        (double*)aux_next, (double*)res);
}

int jh_step_vjp_res(jh_ctx* c, const void* U, const void* aux, double dt, const void* res,
                    const void* U_next, const void* aux_next, const void* U_bar,
                    const void* aux_bar, void* U_in_bar, void* aux_in_bar) {
  if (!c) return JH_ERR_ARG;
  if (!res || !U_next || !aux_next) return fail(c, JH_ERR_ARG, "NULL pointer");
  int rc = vjp_checks(c, U, aux, dt, U_bar, aux_bar, U_in_bar, aux_in_bar);
  if (rc) return rc;
  return step_vjp_res_impl(c, (const double*)U, (const double*)aux, dt, (const double*)res,
                           (const double*)U_next, (const double*)aux_next, (const double*)U_bar,
                           (const double*)aux_bar, (double*)U_in_bar, (double*)aux_in_bar);
}

int jh_timing_enable(jh_ctx* c, int every) {
  if (!c) return JH_ERR_ARG;
  if (every < 0) return fail(c, JH_ERR_ARG, "timing period < 0");
  c->timing = every != 0;
  c->timing_every = every;
  c->timed_steps = 0;
  return JH_OK;
}

int jh_timing_read(jh_ctx* c, double* ms, int* counts) {
  if (!c || !ms || !counts) return JH_ERR_ARG;
  for (int k = 0; k < JH_KERNEL_KINDS; ++k) {
    ms[k] = 0.0;
    counts[k] = 0;
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (auto& m : c->marks) {
    float t = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&t, m.a, m.b));
    ms[m.kind] += t;
    counts[m.kind] += 1;
    c->ev_free.push_back(m.a);
    c->ev_free.push_back(m.b);
  }
  c->marks.clear();
  return JH_OK;
}

int jh_ghost_strip(jh_ctx* c, int side, void** ptr, int* count) {
  if (!c || !ptr || !count || side < 0 || side > 3) return JH_ERR_ARG;
  *ptr = c->ghost[side];
  *count = NV * 3 * ((side < 2) ? c->ny : c->nx);
  if (c->bc[side] == JH_BC_USER) c->user_filled[side] = true;   // the caller writes it
  return JH_OK;
}

int jh_set_ghost_strip(jh_ctx* c, int side, const void* src) {
  if (!c) return JH_ERR_ARG;
  if (side < 0 || side > 3 || !src) return fail(c, JH_ERR_ARG, "bad side or NULL strip");
  if (c->bc[side] != JH_BC_USER) return fail(c, JH_ERR_ARG, "side is not JH_BC_USER");
  const size_t bytes = elem_size(c) * NV * 3 * ((side < 2) ? c->ny : c->nx);
  HIP_TRY(c, hipMemcpyAsync(c->ghost[side], src, bytes, hipMemcpyDeviceToDevice, c->stream));
  c->user_filled[side] = true;
  return JH_OK;
}

int jh_step_begin(jh_ctx* c, const void* U, const void* aux, double dt, double cfl) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux) return fail(c, JH_ERR_ARG, "NULL state pointer");
  if (!(dt > 0) && !(cfl > 0)) return fail(c, JH_ERR_ARG, "need dt > 0 or cfl > 0");
  return c->precision == JH_PREC_FP64
             ? prologue<double>(c, (const double*)U, (const double*)aux, dt, cfl)
             : prologue<float>(c, (const float*)U, (const float*)aux, dt, cfl);
}

int jh_step_stage(jh_ctx* c, int stage, const void* Us, const void* As, const void* Un,
                  void* Uo, void* Ao, double dt, double cfl) {
  if (!c) return JH_ERR_ARG;
  if (stage < 0 || stage > 2) return fail(c, JH_ERR_ARG, "stage must be 0, 1 or 2");
  if (!Us || !As || !Un || !Uo || !Ao) return fail(c, JH_ERR_ARG, "NULL state pointer");
  if (!(dt > 0) && !(cfl > 0)) return fail(c, JH_ERR_ARG, "need dt > 0 or cfl > 0");
  if (dt > 0) cfl = 0.0;
  else dt = 0.0;
  int rc = take_user_strips(c);
  if (rc) return rc;
  if (c->precision == JH_PREC_FP64) {
    auto a = make_args<double>(c, dt, cfl);
    bc_kernel<double><<<bc_blocks(c), kBcThreads, 0, c->stream>>>(a, (const double*)Us,
                                                            (const double*)As, stage);
    launch_stage<double, MODE_RK>(c, a, stage, (const double*)Us, (const double*)As,
                                  (const double*)Un, (double*)Uo, (double*)Ao);
    halley_fix<double>(c, stage, c->parity, stage < 2 ? stage + 1 : -1, 0, (double*)Ao);
  } else {
    auto a = make_args<float>(c, dt, cfl);
    bc_kernel<float><<<bc_blocks(c), kBcThreads, 0, c->stream>>>(a, (const float*)Us,
                                                          (const float*)As, stage);
    launch_stage<float, MODE_RK>(c, a, stage, (const float*)Us, (const float*)As,
                                 (const float*)Un, (float*)Uo, (float*)Ao);
    halley_fix<float>(c, stage, c->parity, stage < 2 ? stage + 1 : -1, 0, (float*)Ao);
  }
  return check_launch(c, "step stage kernels");
}

int jh_step_finish(jh_ctx* c, void* U, void* aux, double dt, double cfl, double* dt_dev) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux) return fail(c, JH_ERR_ARG, "NULL state pointer");
  if (!(dt > 0) && !(cfl > 0)) return fail(c, JH_ERR_ARG, "need dt > 0 or cfl > 0");
  if (dt > 0) cfl = 0.0;
  else dt = 0.0;
  if (c->precision == JH_PREC_FP64) {
    auto a = make_args<double>(c, dt, cfl);
    launch_chem_step<double>(c, a, c->stream, (double*)U, (double*)aux, dt_dev, false);
    halley_fix<double>(c, JH_PROD_CHEM, c->parity ^ 1, 0, 1, (double*)aux);
  } else {
    auto a = make_args<float>(c, dt, cfl);
    launch_chem_step<float>(c, a, c->stream, (float*)U, (float*)aux, dt_dev, false);
    halley_fix<float>(c, JH_PROD_CHEM, c->parity ^ 1, 0, 1, (float*)aux);
  }
  const int rc = check_launch(c, "chem_kernel");
  c->parity ^= 1;
  return rc;
}

int jh_health(jh_ctx* c, unsigned int* flags) {
  if (!c || !flags) return JH_ERR_ARG;
  StepScalars h;
  HIP_TRY(c, hipMemcpyAsync(&h, c->sc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *flags = h.flags;
  const unsigned int zero = 0;
  HIP_TRY(c, hipMemcpyAsync(&c->sc->flags, &zero, sizeof(zero), hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return JH_OK;
}

// ------------------------------------------------------------------ batch API

int jh_batch_create(int ncase, const int* nx, const int* ny, const double* dx, const double* dy,
                    const int* bc_kind, const double* bc_param, int precision,
                    const double* mech, int n_mech, void* stream, jh_batch** out) {
  if (!out) return bfail(nullptr, JH_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (ncase < 1) return bfail(nullptr, JH_ERR_ARG, "ncase must be >= 1");
  if (!nx || !ny || !dx || !dy || !bc_kind || !mech)
    return bfail(nullptr, JH_ERR_ARG, "NULL argument");
  if (precision != JH_PREC_FP64 && precision != JH_PREC_MIXED)
    return bfail(nullptr, JH_ERR_ARG, "unknown precision");
  if (n_mech != jh_mech_const_count())
    return bfail(nullptr, JH_ERR_MECH, "mechanism constant count mismatch");
  for (int k = 0; k < ncase; ++k) {
    if (nx[k] < 4 || ny[k] < 4) return bfail(nullptr, JH_ERR_ARG, "nx and ny must be >= 4");
    if ((long long)nx[k] * ny[k] * NU >= (1ll << 31))
      return bfail(nullptr, JH_ERR_ARG, "grid too large for 32-bit plane offsets");
    if (!(dx[k] > 0) || !(dy[k] > 0)) return bfail(nullptr, JH_ERR_ARG, "dx, dy must be > 0");
    for (int s = 0; s < 4; ++s) {
      const int kind = bc_kind[4 * k + s];
      if (kind < JH_BC_PERIODIC || kind > JH_BC_INJECTION_INLET || kind == JH_BC_USER)
        return bfail(nullptr, JH_ERR_ARG, "batched cases take built-in boundary kinds only");
    }
  }
  jh_batch* b = new jh_batch();
  b->ncase = ncase;
  b->precision = precision;
  b->tile = precision == JH_PREC_FP64 ? 3 : 2;   // as jh_ctx_create
  b->user = static_cast<hipStream_t>(stream);
  b->nx.assign(nx, nx + ncase);
  b->ny.assign(ny, ny + ncase);
  b->dx.assign(dx, dx + ncase);
  b->dy.assign(dy, dy + ncase);
  b->bc.assign(bc_kind, bc_kind + 4 * ncase);
  b->mirror = true;
  for (int k : b->bc)
    if (k != JH_BC_PERIODIC && k != JH_BC_ZERO_GRADIENT && k != JH_BC_SLIP_WALL &&
        k != JH_BC_NO_SLIP_WALL)
      b->mirror = false;
  b->bcp.assign(4 * ncase, 0.0);
  if (bc_param) b->bcp.assign(bc_param, bc_param + 4 * ncase);
  b->U.assign(ncase, nullptr);
  b->aux.assign(ncase, nullptr);
  b->inj.assign((size_t)ncase * JH_NS, 0.0);
  for (int k = 0; k < ncase; ++k) b->inj[(size_t)k * JH_NS] = 300.0;
  // block tables: stage tiles per variant, bc lines, reduce, chemistry cells
  const int stride = ncase + 1;
  b->h_off.assign((kTileVariants + 3) * stride, 0);
  for (int v = 0; v < kTileVariants; ++v) b->t_stage[v] = v * stride;
  b->t_bc = kTileVariants * stride;
  b->t_red = (kTileVariants + 1) * stride;
  b->t_chem = (kTileVariants + 2) * stride;
  for (int k = 0; k < ncase; ++k) {
    const long long n = (long long)nx[k] * ny[k];
    for (int v = 0; v < kTileVariants; ++v) {
      const int tx = kTileRows[v];
      b->h_off[b->t_stage[v] + k + 1] =
          b->h_off[b->t_stage[v] + k] + ((nx[k] + tx - 1) / tx) * ((ny[k] + TY - 1) / TY);
    }
    b->h_off[b->t_bc + k + 1] =
        b->h_off[b->t_bc + k] + (2 * nx[k] + 2 * ny[k] + kBcThreads - 1) / kBcThreads;
    const int cb = (int)((n + 255) / 256);
    b->h_off[b->t_red + k + 1] = b->h_off[b->t_red + k] + std::min(cb, 64);
    b->h_off[b->t_chem + k + 1] = b->h_off[b->t_chem + k] + cb;
  }
  const size_t es = precision == JH_PREC_FP64 ? 8 : 4;
  size_t total = 0;
  b->buf_off.resize(ncase);
  for (int k = 0; k < ncase; ++k) {
    b->buf_off[k] = total;
    total += case_bytes(es, nx[k], ny[k]);
  }
  const size_t desc_bytes = precision == JH_PREC_FP64 ? sizeof(CaseDesc<double>) * ncase
                                                      : sizeof(CaseDesc<float>) * ncase;
  hipError_t e = hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking);
  if (!e) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
  if (!e) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
  if (!e) e = hipMalloc(&b->mc, sizeof(MechConst));
  if (!e) e = hipMalloc(&b->sc, sizeof(StepScalars) * ncase);
  if (!e) e = hipMalloc(&b->bufs, total);
  if (!e) e = hipMalloc(&b->desc, desc_bytes);
  if (!e) e = hipMalloc(&b->off, sizeof(int) * b->h_off.size());
  if (!e) e = hipMemcpy(b->mc, mech, sizeof(MechConst), hipMemcpyHostToDevice);
  if (!e) e = hipMemset(b->sc, 0, sizeof(StepScalars) * ncase);
  if (!e) e = hipMemset(b->bufs, 0, total);
  if (!e)
    e = hipMemcpy(b->off, b->h_off.data(), sizeof(int) * b->h_off.size(), hipMemcpyHostToDevice);
  if (e) {
    g_batch_err = std::string("allocation: ") + hipGetErrorString(e);
    jh_batch_destroy(b);
    return JH_ERR_HIP;
  }
  int rc = precision == JH_PREC_FP64 ? setup_grids<double>(b) : setup_grids<float>(b);
  if (!rc) rc = precision == JH_PREC_FP64 ? upload_desc<double>(b) : upload_desc<float>(b);
  if (rc) {
    g_batch_err = b->err;
    jh_batch_destroy(b);
    return rc;
  }
  *out = b;
  return JH_OK;
}

int jh_batch_destroy(jh_batch* b) {
  if (!b) return JH_OK;
  if (b->own) (void)hipStreamSynchronize(b->own);
  drop_graph(b);
  (void)hipFree(b->mc);
  (void)hipFree(b->sc);
  (void)hipFree(b->bufs);
  (void)hipFree(b->desc);
  (void)hipFree(b->off);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->own) (void)hipStreamDestroy(b->own);
  delete b;
  return JH_OK;
}

int jh_batch_bind(jh_batch* b, void* const* U, void* const* aux) {
  if (!b) return JH_ERR_ARG;
  if (!U || !aux) return bfail(b, JH_ERR_ARG, "NULL pointer array");
  for (int k = 0; k < b->ncase; ++k)
    if (!U[k] || !aux[k]) return bfail(b, JH_ERR_ARG, "NULL state pointer");
  b->U.assign(U, U + b->ncase);
  b->aux.assign(aux, aux + b->ncase);
  b->bound = true;
  ++b->gen;
  return b->precision == JH_PREC_FP64 ? upload_desc<double>(b) : upload_desc<float>(b);
}

int jh_batch_set_stream(jh_batch* b, void* stream) {
  if (!b) return JH_ERR_ARG;
  b->user = static_cast<hipStream_t>(stream);
  return JH_OK;
}

int jh_batch_set_tile(jh_batch* b, int variant) {
  if (!b) return JH_ERR_ARG;
  if (variant < 0 || variant >= jh_tile_variants())
    return bfail(b, JH_ERR_ARG, "unknown tile variant");
  b->tile = variant;
  return JH_OK;
}

int jh_batch_get_tile(const jh_batch* b) { return b ? b->tile : -1; }

int jh_batch_set_inlet(jh_batch* b, int k, double T0, const double* Y, int n) {
  if (!b) return JH_ERR_ARG;
  if (k < 0 || k >= b->ncase) return bfail(b, JH_ERR_ARG, "case index out of range");
  if (!Y || n != JH_NS - 1) return bfail(b, JH_ERR_ARG, "Y_inj needs num_species-1 values");
  if (!(T0 > 0)) return bfail(b, JH_ERR_ARG, "T0 must be > 0");
  b->inj[(size_t)k * JH_NS] = T0;
  for (int j = 0; j < n; ++j) b->inj[(size_t)k * JH_NS + 1 + j] = Y[j];
  return b->precision == JH_PREC_FP64 ? upload_desc<double>(b) : upload_desc<float>(b);
}

int jh_batch_advance(jh_batch* b, double dt, double cfl, int nsteps, double* dt_hist,
                     int use_graph) {
  if (!b) return JH_ERR_ARG;
  if (!b->bound) return bfail(b, JH_ERR_ARG, "no state bound (jh_batch_bind)");
  if (nsteps < 1) return bfail(b, JH_ERR_ARG, "nsteps must be >= 1");
  if (!(dt > 0) && !(cfl > 0)) return bfail(b, JH_ERR_ARG, "need dt > 0 or cfl > 0");
  if (dt > 0) cfl = 0.0;
  else dt = 0.0;
  BHIP_TRY(b, hipEventRecord(b->ev_in, b->user));
  BHIP_TRY(b, hipStreamWaitEvent(b->own, b->ev_in, 0));
  if (use_graph) {
    const jh_batch::Key& k = b->key;
    if (!(b->gexec && k.nsteps == nsteps && k.dt == dt && k.cfl == cfl && k.hist == dt_hist &&
          k.tile == b->tile && k.gen == b->gen)) {
      drop_graph(b);
      hipGraph_t g = nullptr;
      BHIP_TRY(b, hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal));
      const int rc = batch_enqueue_any(b, b->own, dt, cfl, nsteps, dt_hist);
      const hipError_t e = hipStreamEndCapture(b->own, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess)
        return bfail(b, JH_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        b->gexec = nullptr;
        return bfail(b, JH_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      }
      b->key.nsteps = nsteps;
      b->key.dt = dt;
      b->key.cfl = cfl;
      b->key.hist = dt_hist;
      b->key.tile = b->tile;
      b->key.gen = b->gen;
    }
    BHIP_TRY(b, hipGraphLaunch(b->gexec, b->own));
  } else {
    const int rc = batch_enqueue_any(b, b->own, dt, cfl, nsteps, dt_hist);
    if (rc) return rc;
  }
  BHIP_TRY(b, hipEventRecord(b->ev_out, b->own));
  BHIP_TRY(b, hipStreamWaitEvent(b->user, b->ev_out, 0));
  return JH_OK;
}

int jh_batch_health(jh_batch* b, unsigned int* flags) {
  if (!b || !flags) return JH_ERR_ARG;
  std::vector<StepScalars> h(b->ncase);
  BHIP_TRY(b, hipStreamSynchronize(b->own));
  BHIP_TRY(b, hipMemcpy(h.data(), b->sc, sizeof(StepScalars) * b->ncase, hipMemcpyDeviceToHost));
  unsigned int f = 0;
  for (int k = 0; k < b->ncase; ++k) {
    f |= h[k].flags;
    h[k].flags = 0;
  }
  BHIP_TRY(b, hipMemcpy(b->sc, h.data(), sizeof(StepScalars) * b->ncase, hipMemcpyHostToDevice));
  *flags = f;
  return JH_OK;
}

const char* jh_batch_last_error(const jh_batch* b) {
  return b ? b->err.c_str() : g_batch_err.c_str();
}

// ------------------------------------------------------------------ AMR blocks API

int jh_blocks_create(int bnx, int bny, double dx, double dy, int precision, const double* mech,
                     int n_mech, void* stream, jh_blocks** out) {
  if (!out) return kfail(nullptr, JH_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (bnx < 1 || bny < 1) return kfail(nullptr, JH_ERR_ARG, "block sizes must be >= 1");
  if (!(dx > 0) || !(dy > 0)) return kfail(nullptr, JH_ERR_ARG, "dx, dy must be > 0");
  if (precision != JH_PREC_FP64 && precision != JH_PREC_MIXED)
    return kfail(nullptr, JH_ERR_ARG, "unknown precision");
  if (!mech) return kfail(nullptr, JH_ERR_ARG, "mech is NULL");
  if (n_mech != jh_mech_const_count())
    return kfail(nullptr, JH_ERR_MECH, "mechanism constant count mismatch");
  jh_blocks* b = new jh_blocks();
  b->nx = bnx;
  b->ny = bny;
  b->dx = dx;
  b->dy = dy;
  b->precision = precision;
  b->stream = static_cast<hipStream_t>(stream);
  hipError_t e = hipMalloc(&b->mc, sizeof(MechConst));
  if (!e) e = hipMemcpy(b->mc, mech, sizeof(MechConst), hipMemcpyHostToDevice);
  // set-up for jh_blocks_stage_noalloc: this device's address of `scratch`
  if (!e) e = scratch_address(&b->scratch);
  if (e) {
    g_blocks_err = std::string("allocation: ") + hipGetErrorString(e);
    jh_blocks_destroy(b);
    return JH_ERR_HIP;
  }
  *out = b;
  return JH_OK;
}

int jh_blocks_fill_ghosts(void* stream, int precision, int n_blk, int bnx, int bny,
                          const void* src, long long sb, long long sv, long long sr,
                          const int* idx, const void* s_u, const void* s_v, void* out) {
  if (!src || !idx || !s_u || !s_v || !out) return kfail(nullptr, JH_ERR_ARG, "NULL pointer");
  if (n_blk < 1 || bnx < 1 || bny < 1 || sb < 1 || sv < 1 || sr < bny)
    return kfail(nullptr, JH_ERR_ARG, "bad block count, size or strides");
  if (precision != JH_PREC_FP64 && precision != JH_PREC_MIXED)
    return kfail(nullptr, JH_ERR_ARG, "unknown precision");
  const long long P = (long long)(bnx + 6) * (bny + 6), m = n_blk * P;
  if (m >= (1ll << 31) || (long long)n_blk * bnx * bny >= (1ll << 31))
    return kfail(nullptr, JH_ERR_ARG, "too many cells for 32-bit indices");
  const dim3 grid((unsigned)((m + 255) / 256), NV);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == JH_PREC_FP64)
    blocks_fill_kernel<double><<<grid, 256, 0, st>>>(
        (const double*)src, sb, sv, sr, bnx, bny, idx, (const double*)s_u, (const double*)s_v,
        (double*)out, n_blk, (int)P);
  else
    blocks_fill_kernel<float><<<grid, 256, 0, st>>>(
        (const float*)src, sb, sv, sr, bnx, bny, idx, (const float*)s_u, (const float*)s_v,
        (float*)out, n_blk, (int)P);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kfail(nullptr, JH_ERR_HIP, std::string("blocks_fill_kernel: ") +
                                                             hipGetErrorString(e));
  return JH_OK;
}

int jh_blocks_destroy(jh_blocks* b) {
  if (!b) return JH_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream); else (void)hipDeviceSynchronize();
  if (b->mc) (void)hipFree(b->mc);
  if (b->sc) (void)hipFree(b->sc);
  delete b;
  return JH_OK;
}

int jh_blocks_set_stream(jh_blocks* b, void* stream) {
  if (!b) return JH_ERR_ARG;
  b->stream = static_cast<hipStream_t>(stream);
  return JH_OK;
}

const char* jh_blocks_last_error(const jh_blocks* b) {
  return b ? b->err.c_str() : g_blocks_err.c_str();
}

int jh_blocks_flux(jh_blocks* b, int n_blk, const void* blk, double dt, void* rhs) {
  if (!b) return JH_ERR_ARG;
  if (!blk || !rhs) return kfail(b, JH_ERR_ARG, "NULL pointer");
  if (blk == rhs) return kfail(b, JH_ERR_ARG, "rhs must not alias the blocks");
  if (!(dt > 0)) return kfail(b, JH_ERR_ARG, "dt must be > 0");
  int rc = blocks_checks(b, n_blk);
  if (rc) return rc;
  if (b->precision == JH_PREC_FP64)
    return blocks_stage_impl<double, MODE_RHS>(b, n_blk, 0, (const double*)blk,
                                               (const double*)blk, (double*)rhs, dt, b->sc);
  return blocks_stage_impl<float, MODE_RHS>(b, n_blk, 0, (const float*)blk, (const float*)blk,
                                            (float*)rhs, dt, b->sc);
}

int jh_blocks_stage(jh_blocks* b, int n_blk, int stage, const void* blk_s, const void* blk_n,
                    double dt, void* blk_out) {
  if (!b) return JH_ERR_ARG;
  int rc = stage_args_check(b, stage, blk_s, blk_n, dt, blk_out);
  if (!rc) rc = blocks_checks(b, n_blk);
  if (rc) return rc;
  if (b->precision == JH_PREC_FP64)
    return blocks_stage_impl<double, MODE_RK>(b, n_blk, stage, (const double*)blk_s,
                                              (const double*)blk_n, (double*)blk_out, dt, b->sc);
  return blocks_stage_impl<float, MODE_RK>(b, n_blk, stage, (const float*)blk_s,
                                           (const float*)blk_n, (float*)blk_out, dt, b->sc);
}

int jh_blocks_stage_noalloc(jh_blocks* b, int n_blk, int stage, const void* blk_s,
                            const void* blk_n, double dt, void* blk_out) {
  if (!b) return JH_ERR_ARG;
  int rc = stage_args_check(b, stage, blk_s, blk_n, dt, blk_out);
  if (!rc && n_blk > kNoallocBlocks)
    rc = kfail(b, JH_ERR_ARG, "n_blk exceeds the no-allocation entry's " +
                                  std::to_string(kNoallocBlocks) + "-block scratch");
  if (!rc) rc = blocks_checks(b, n_blk, false);
  if (rc) return rc;
  if (!b->scratch) return kfail(b, JH_ERR_HIP, "no device address for the scratch array");
  // nothing in the scratch needs clearing: the LF maxima are stored (not
  // folded) by blocks_alpha_kernel before the stage kernel reads them, and
  // the health flags it also holds are not reported by this entry point
  if (b->precision == JH_PREC_FP64)
    return blocks_stage_impl<double, MODE_RK>(b, n_blk, stage, (const double*)blk_s,
                                              (const double*)blk_n, (double*)blk_out, dt,
                                              b->scratch);
  return blocks_stage_impl<float, MODE_RK>(b, n_blk, stage, (const float*)blk_s,
                                           (const float*)blk_n, (float*)blk_out, dt, b->scratch);
}

int jh_blocks_chem(jh_blocks* b, int n_blk, void* blk, double dt) {
  if (!b) return JH_ERR_ARG;
  if (!blk) return kfail(b, JH_ERR_ARG, "NULL pointer");
  if (!(dt > 0)) return kfail(b, JH_ERR_ARG, "dt must be > 0");
  int rc = blocks_checks(b, n_blk);
  if (rc) return rc;
  const int n = b->nx * b->ny;
  const long long bs = (long long)(NU + 2) * n;
  const dim3 grid((n + 255) / 256, n_blk);
  if (b->precision == JH_PREC_FP64)
    blocks_chem_kernel<double><<<grid, 256, 0, b->stream>>>(blocks_args<double>(b, dt, b->sc),
                                                            bs, (double*)blk);
  else
    blocks_chem_kernel<float><<<grid, 256, 0, b->stream>>>(blocks_args<float>(b, dt, b->sc), bs,
                                                           (float*)blk);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kfail(b, JH_ERR_HIP, std::string("block chemistry: ") +
                                                       hipGetErrorString(e));
  return JH_OK;
}

int jh_blocks_health(jh_blocks* b, unsigned int* flags) {
  if (!b || !flags) return JH_ERR_ARG;
  *flags = 0;
  if (!b->cap) return JH_OK;
  std::vector<StepScalars> h(b->cap);
  KHIP_TRY(b, hipStreamSynchronize(b->stream));
  KHIP_TRY(b, hipMemcpy(h.data(), b->sc, sizeof(StepScalars) * b->cap, hipMemcpyDeviceToHost));
  unsigned int f = 0;
  for (auto& x : h) {
    f |= x.flags;
    x.flags = 0;
  }
  KHIP_TRY(b, hipMemcpy(b->sc, h.data(), sizeof(StepScalars) * b->cap, hipMemcpyHostToDevice));
  *flags = f;
  return JH_OK;
}

}  // extern "C"
