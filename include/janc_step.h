/*
 * janc_step.h -- C ABI of the MI355X (gfx950) JANC step library.
 *
 * One context = one configured 2-D single-block solver: grid, boundary kinds,
 * precision form and the mechanism constants.  Every state pointer is a
 * caller-owned device pointer to a C-order (var, nx, ny) array with y
 * contiguous -- the reference's U (12, nx, ny) and aux (2, nx, ny) layout,
 * so torch/JAX contiguous tensors pass through DLPack without a copy.
 * Element type: double for JH_PREC_FP64, float for JH_PREC_MIXED.
 *
 * All calls are asynchronous on the context's HIP stream and return a status
 * code; jh_last_error() gives the message of the last failure.  A context must
 * not be used from two threads at once; distinct contexts are independent.
 * Reference paths below are relative to the luofx23/JANC source tree.
 */
#ifndef JANC_STEP_H
#define JANC_STEP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jh_ctx jh_ctx;

enum jh_status {
  JH_OK = 0,
  JH_ERR_ARG = 1,      /* null pointer, bad size or bad enum */
  JH_ERR_MECH = 2,     /* mechanism constants do not match the compiled mechanism */
  JH_ERR_HIP = 3,      /* a HIP runtime call failed */
};

enum jh_precision { JH_PREC_FP64 = 0, JH_PREC_MIXED = 1 };

/* Boundary kinds, boundary/boundary.py:51-118 and
 * boundary/boundary_conditions_2D/__init__.py:4-22.  JH_BC_USER leaves the
 * side's ghost strip to the caller (a Python callable in the reference,
 * boundary.py:52-57); see jh_ghost_strip(). */
enum jh_bc_kind {
  JH_BC_PERIODIC = 0,
  JH_BC_ZERO_GRADIENT = 1,
  JH_BC_SLIP_WALL = 2,
  JH_BC_NO_SLIP_WALL = 3,
  JH_BC_PRESSURE_OUTLET = 4,
  JH_BC_USER = 5,
  /* Injection inlet: the pressure-feedback injector boundary of the
   * reference's rotating-detonation example (examples/
   * janc_basic_example2.ipynb, bottom_boundary + inj_model, a user callable
   * there), with the manifold total pressure P0 as the side's bc_param and
   * T0 / injected composition from jh_ctx_set_inlet(). */
  JH_BC_INJECTION_INLET = 6
};

/* Sides in the order bc_kind[] / bc_param[] use. */
enum jh_side { JH_SIDE_LEFT = 0, JH_SIDE_RIGHT = 1, JH_SIDE_BOTTOM = 2, JH_SIDE_TOP = 3 };

/* Number of doubles in the mechanism constant block the library was built
 * for, and the digest of the compiled reaction structure.  Replaces the
 * Cantera read of preprocess/load.py:8-135 (the block is produced by
 * janc_amd/mechanism.py). */
int jh_mech_const_count(void);
const char* jh_mech_digest(void);
int jh_num_species(void);

/* Create a context.  Replaces Simulator.__init__'s set_thermo / set_reaction /
 * set_flux_solver / set_boundary (simulation/set_simulator.py:387-396).
 *   bc_kind[4]  : jh_bc_kind per side (left, right, bottom, top)
 *   bc_param[4] : theta['Pb'] for JH_BC_PRESSURE_OUTLET sides, the manifold
 *                 total pressure P0 for JH_BC_INJECTION_INLET sides, else
 *                 ignored
 *   mech        : host array of jh_mech_const_count() doubles
 *   stream      : hipStream_t (NULL = default stream) */
int jh_ctx_create(int nx, int ny, double dx, double dy, const int* bc_kind,
                  const double* bc_param, int precision, const double* mech,
                  int n_mech, void* stream, jh_ctx** out);
int jh_ctx_destroy(jh_ctx* ctx);
/* Per-call boundary parameter (theta['Pb'] of pressure_outlet.py:24-66 is a
 * call argument in the reference). */
int jh_ctx_set_bc_param(jh_ctx* ctx, int side, double value);

/* Manifold total temperature T0 and injected mass fractions Y_inj (the first
 * n = jh_num_species()-1 species) of JH_BC_INJECTION_INLET sides (theta['Yinj']
 * of janc_basic_example2.ipynb; the example's nondimensional T0 = 1). */
int jh_ctx_set_inlet(jh_ctx* ctx, double T0, const double* Y_inj, int n);
int jh_ctx_set_stream(jh_ctx* ctx, void* stream);

/* Flux scheme: the flux_config of set_flux_solver (solver_2D/flux.py:19-40).
 *   solver_type    : 'flux_splitting' (flux.py:56-68) or 'godunov'
 *                    (flux.py:43-54)
 *   split_method   : flux splitting only: 'LF' (flux_splitting/LF.py:4-16)
 *                    or 'SW' (flux_splitting/SW.py:4-38)
 *   reconstruction : interface_reconstruction 'WENO5_JS'
 *                    (reconstruction/WENO5_JS.py:1-117) or 'MUSCL'
 *                    (reconstruction/MUSCL.py:1-65)
 *   riemann_solver : 'HLLC' (riemann_solver/HLLC.py:5-77), godunov only.
 * The reference's other Riemann solver, 'AUSMPW_plus', is not offered: its y
 * flux averages the y faces' gamma with the x faces' (AUSMPW_plus.py:63,67),
 * arrays of shapes (nx, ny+1) and (nx+1, ny) that do not broadcast, so the
 * reference cannot run it on a 2-D grid.  Default: flux splitting, WENO5_JS.
 * jh_step_vjp supports the default scheme only (LF, WENO5_JS). */
enum jh_solver_type { JH_SOLVER_FLUX_SPLITTING = 0, JH_SOLVER_GODUNOV = 1 };
enum jh_split_method { JH_SPLIT_LF = 0, JH_SPLIT_SW = 1 };
enum jh_reconstruction { JH_RECON_WENO5_JS = 0, JH_RECON_MUSCL = 1 };
enum jh_riemann_solver { JH_RIEMANN_HLLC = 0 };
int jh_ctx_set_flux(jh_ctx* ctx, int solver_type, int split_method, int reconstruction,
                    int riemann_solver);

/* Stage-kernel tile variant of the flux-splitting kernel (jh_tile_info gives
 * rows, threads and the waves per SIMD the register budget is sized for).
 * No counterpart in the reference (XLA picks its own fusion); chosen by the
 * tile autotuner (janc_amd/autotune.py).  The Godunov kernel has one tile
 * shape per reconstruction (WENO5_JS: 64 x 3 cells, 256 threads; MUSCL:
 * 64 x 7 cells, 512 threads), the Steger-Warming kernel one (64 x 16 cells,
 * 512 threads). */
int jh_ctx_set_tile(jh_ctx* ctx, int variant);
int jh_ctx_get_tile(const jh_ctx* ctx);

/* Stop rule of the Halley temperature solve after every RK stage, after the
 * chemistry and in jh_update_aux (get_T_nasa7, model/thermo_model.py:
 * 135-153).  JH_HALLEY_GRID is the reference's: every cell makes N* updates,
 * N* = the first iteration at which max|res| over the grid is <= 5e-9
 * (at most 5); each cell's own count and its state after one more update are
 * kept, and a fix-up pass applies N* (one extra light launch per solve; the
 * first call allocates two planes and a byte per cell).  JH_HALLEY_PER_CELL
 * (default) stops each cell on its own residual: equal to the grid-wide stop
 * to rounding in SI units (a converged cell's further updates move T by
 * ~1e-16 relative), ~1e-9 apart with nondimensional energy scales, where the
 * absolute tolerance is a larger fraction of e.  The batch API stops per
 * cell. */
enum jh_halley_stop { JH_HALLEY_PER_CELL = 0, JH_HALLEY_GRID = 1 };
int jh_ctx_set_halley_stop(jh_ctx* ctx, int mode);
int jh_ctx_get_halley_stop(const jh_ctx* ctx);

/* Step tracing, no counterpart in the reference: while set, each step of
 * jh_advance (up to max_steps) writes a record of 22 int64 into records
 * (device): per RK stage, the first ghost-padded index (P = (x+3)(ny+6) +
 * y+3) attaining max|u|+a and the number of padded cells attaining it, the
 * same for max|v|+a (the nanmax of LF.py:10-11, corners excluded); then a
 * hash and the count of the active branches of the chemistry clip
 * (reaction_model.py:191); then, after each of the step's temperature solves
 * (three stages, chemistry), a hash and the count of the cells below the
 * NASA-7 switch temperature (thermo_model.py:101-104; the two fits are not
 * continuous there).  The step is only piecewise smooth in these: reverse
 * mode, and jh_step_vjp, follow the branch taken, so a finite-difference
 * check of the adjoint holds only across perturbations that change none of
 * them.  Adds kernels to every traced step; records = NULL stops tracing. */
int jh_ctx_set_trace(jh_ctx* ctx, long long* records, int max_steps);

/* Chemistry clip branches (reaction_model.py:191), no counterpart in the
 * reference: JH_CLIP_RECORD writes, for each following step of jh_advance
 * (up to max_steps), the branch every cell's clip takes (uint16 per cell:
 * bit k = lower bound for species k, bit 8 + k = upper bound) into masks
 * (device, (max_steps, nx*ny)); JH_CLIP_REPLAY makes the chemistry take
 * those branches instead of comparing.  A replayed rollout from a perturbed
 * state is the branch-wise smooth function the adjoint differentiates (a
 * finite-difference check of jh_step_vjp without the clip's kinks). */
enum jh_clip_mode { JH_CLIP_OFF = 0, JH_CLIP_RECORD = 1, JH_CLIP_REPLAY = 2 };
int jh_ctx_set_clip_branches(jh_ctx* ctx, unsigned short* masks, int max_steps, int mode);
int jh_tile_variants(void);
int jh_tile_info(int variant, int* rows, int* threads, int* waves);
const char* jh_last_error(const jh_ctx* ctx);

/* Advance nsteps steps.  One step is the reference's advance_func
 * (simulation/set_simulator.py:444-447): dt = CFL_2D(U, aux) (:438-443) when
 * dt <= 0, else the fixed dt (:405-408); then advance_one_step (:117-122):
 * TVD_RK3 (time_step/TVD_RK3.py:5-12) of flux_func (:44-47), point-implicit
 * chemistry (model/reaction_model.py:183-203) and update_aux.
 * U_out/aux_out may alias U/aux.  dt_hist (device, nsteps doubles) receives
 * each step's dt; may be NULL. */
int jh_advance(jh_ctx* ctx, const void* U, const void* aux, void* U_out, void* aux_out,
               double dt, double cfl, int nsteps, double* dt_hist);

/* flux_func (set_simulator.py:44-47): rhs = dt * total_flux(BC(U, aux))
 * on the interior, 12 planes. */
int jh_rhs(jh_ctx* ctx, const void* U, const void* aux, double dt, void* rhs_out);

/* update_aux (solver_2D/aux_func.py:21-29): Halley solve of T from e with
 * aux_in's T as the initial guess; writes (gamma, T). aux_out may alias aux_in. */
int jh_update_aux(jh_ctx* ctx, const void* U, const void* aux_in, void* aux_out);

/* reaction_source_terms (model/reaction_model.py:183-203): S (12 planes;
 * planes 0-3 are zero). */
int jh_reaction_source(jh_ctx* ctx, const void* U, const void* aux, double dt, void* S_out);

/* CFL_2D (set_simulator.py:438-443) -> *dt_dev (device double). */
int jh_cfl_dt(jh_ctx* ctx, const void* U, const void* aux, double cfl, double* dt_dev);

/* The LF maxima of LF.py:10-11 over the ghost-padded field of (U, aux):
 * alpha_dev[0] = nanmax(|u|+a), alpha_dev[1] = nanmax(|v|+a) (device). */
int jh_lf_alpha(jh_ctx* ctx, const void* U, const void* aux, double* alpha_dev);

/* Vector-Jacobian product of one fixed-dt step (the reference's reverse-mode
 * differentiation of advance_one_step, set_simulator.py:117-122, with
 * get_T_nasa7's custom VJP, model/thermo_model.py:155-190).  Given the state
 * (U, aux) the step starts from and cotangents (U_bar, aux_bar) of its
 * output, writes the cotangents of U and aux (U_in_bar, aux_in_bar); the
 * step is recomputed, and its output is also written to U_out / aux_out when
 * they are non-NULL.  FP64 contexts with periodic / zero_gradient /
 * slip_wall / no_slip_wall boundaries only.  U_in_bar / aux_in_bar may be
 * the same arrays as U_bar / aux_bar (in place) but must not overlap the
 * state.  A context's first call allocates its workspace and uploads its
 * ghost-fold tables, synchronising the context stream once; later calls
 * neither allocate nor synchronise. */
int jh_step_vjp(jh_ctx* ctx, const void* U, const void* aux, double dt, const void* U_bar,
                const void* aux_bar, void* U_in_bar, void* aux_in_bar, void* U_out,
                void* aux_out);
/* As jh_step_vjp for a step whose output (U_next, aux_next) the caller
 * already holds (a rollout's next state): the RK stages are recomputed, the
 * chemistry's forward pass is not.  No counterpart of its own in the
 * reference (the same reverse-mode derivative). */
int jh_step_vjp_known(jh_ctx* ctx, const void* U, const void* aux, double dt,
                      const void* U_next, const void* aux_next, const void* U_bar,
                      const void* aux_bar, void* U_in_bar, void* aux_in_bar);
/* The step's residuals for reverse mode, as jax.vjp of advance_one_step
 * (simulation/set_simulator.py:117-122) keeps them: jh_step_keep runs one
 * FP64 step (output to U_next / aux_next) and writes into res the three RK
 * stage states and the stages' LF maxima; jh_step_vjp_res is then the step's
 * VJP without recomputing any of it (U, aux: the step's input; U_next,
 * aux_next: its output; cotangent arguments as jh_step_vjp).  res holds
 * jh_step_res_count() doubles (3 x 14 planes of nx x ny, then 6) and is read-only to
 * jh_step_vjp_res; the results equal jh_step_vjp_known's bit for bit.  The
 * same conditions as jh_step_vjp. */
int jh_step_res_count(jh_ctx* ctx, long long* count);
int jh_step_keep(jh_ctx* ctx, const void* U, const void* aux, double dt, void* U_next,
                 void* aux_next, void* res);
int jh_step_vjp_res(jh_ctx* ctx, const void* U, const void* aux, double dt, const void* res,
                    const void* U_next, const void* aux_next, const void* U_bar,
                    const void* aux_bar, void* U_in_bar, void* aux_in_bar);

/* Device pointer and element count of one side's ghost strip (14 planes of
 * 3 x len: vars 0-11 = U, 12-13 = aux; left/right strips are (14, 3, ny) with
 * layer 0 outermost on the left and innermost on the right; bottom/top are
 * (14, 3, nx) likewise).  A JH_BC_USER side must be written here before the
 * stage that reads it (boundary.py:52-57 callables). */
int jh_ghost_strip(jh_ctx* ctx, int side, void** ptr, int* count);
/* Copy a caller-computed strip (device, same layout and element type) into a
 * JH_BC_USER side's ghost strip on the context stream.  Each call that reads
 * ghost cells (jh_rhs, jh_lf_alpha, jh_step_stage) needs every JH_BC_USER
 * side written since the previous such call (by this call, or through the
 * pointer from jh_ghost_strip), else it fails with JH_ERR_ARG. */
int jh_set_ghost_strip(jh_ctx* ctx, int side, const void* strip);

/* One step in pieces, for boundaries computed by the caller between the RK
 * stages (the reference calls its boundary callables inside every stage,
 * boundary.py:127-131 from flux_func, set_simulator.py:44-47):
 *   jh_step_begin  : the step's CFL and stage-1 LF maxima of (U, aux)
 *                    (set_simulator.py:438-443);
 *   jh_step_stage  : ghost fill of the built-in sides + one SSP-RK3 stage
 *                    (TVD_RK3.py:5-12) with update_aux: stage 0 (U,aux)->(U1,A1),
 *                    1 (U1,A1)->(U2,A2), 2 (U2,A2)->(Uo,Ao); Un = the step's
 *                    input U; Uo may alias Un;
 *   jh_step_finish : chemistry + update_aux in place (set_simulator.py:
 *                    117-122), dt of the step to *dt_dev if non-NULL.
 * dt > 0 selects a fixed dt, else the CFL dt with cfl; pass the same
 * (dt, cfl) to all three. */
int jh_step_begin(jh_ctx* ctx, const void* U, const void* aux, double dt, double cfl);
int jh_step_stage(jh_ctx* ctx, int stage, const void* Us, const void* As, const void* Un,
                  void* Uo, void* Ao, double dt, double cfl);
int jh_step_finish(jh_ctx* ctx, void* U, void* aux, double dt, double cfl, double* dt_dev);

/* Kernel classes of jh_timing_read(). */
enum jh_kernel_kind {
  JH_KERNEL_REDUCE = 0,   /* prologue: CFL and stage-1 LF maxima of a new state */
  JH_KERNEL_BC = 1,       /* ghost strips + ghost LF maxima, once per RK stage */
  JH_KERNEL_STAGE = 2,    /* fused WENO5/LF/divergence/RK combine/Halley, per stage */
  JH_KERNEL_CHEM = 3,     /* point-implicit chemistry + final Halley, per step */
  JH_KERNEL_KINDS = 4
};

/* Time launches with HIP events on the context stream (no counterpart in the
 * reference, which has no profiling hooks).  every = 0: off; 1: every
 * launch; k > 1: every launch outside jh_advance, and inside it the launches
 * of every k-th step (counted from this call), which keeps the events' own
 * cost (~5 % of a step when every launch is bracketed) out of the rest. */
int jh_timing_enable(jh_ctx* ctx, int every);
/* Sum of event-measured milliseconds and launch counts per jh_kernel_kind
 * since the last read (ms[4], counts[4]); synchronises the stream. */
int jh_timing_read(jh_ctx* ctx, double* ms, int* counts);

/* Health flags accumulated on the device since the last call (JH_FLAG_*
 * bits); synchronises the stream.  The reference has no run-time error
 * channel (NaNs propagate, the temperature solve and the chemistry solve
 * stop or return silently). */
enum jh_health_flag {
  JH_HEALTH_NONFINITE = 1,        /* a non-finite (gamma, T) or state value */
  JH_HEALTH_NEGATIVE_RHO = 2,     /* rho <= 0 in an updated cell */
  JH_HEALTH_HALLEY_NOCONV = 4,    /* a temperature solve ended at its 5-iteration cap
                                     with |res| > 5e-9 (thermo_model.py:140-152) */
  JH_HEALTH_SINGULAR_PIVOT = 8,   /* a zero or non-finite pivot in a chemistry solve
                                     (reaction_model.py:157) */
  JH_HEALTH_HALLEY_REGRESS = 16,  /* JH_HALLEY_GRID: a converged cell's next update left
                                     the tolerance (the fix-up assumes it does not) */
  JH_HEALTH_NONPOSITIVE_T = 32    /* T <= 0 after a temperature solve */
};
int jh_health(jh_ctx* ctx, unsigned int* flags);

/* ---------------------------------------------------------------- batches
 * Many independent cases of uneven grid sizes advanced together: one launch
 * per kernel kind covers every case (the stage kernel is persistent over the
 * tile list of all cases), and jh_batch_advance can capture its launch
 * sequence in a HIP graph and replay it.  Each case evolves exactly as
 * jh_advance would advance it alone (bit-identical results).  The reference
 * batches same-size blocks with jax.vmap over advance_one_step
 * (simulation/set_simulator.py:37-42, 66-70) and has no uneven-size batch;
 * against it this replaces a Python loop of advance_func calls
 * (set_simulator.py:444-447), one per case. */
typedef struct jh_batch jh_batch;

/* ncase cases; nx, ny, dx, dy have ncase entries; bc_kind / bc_param have
 * 4 per case (JH_SIDE order, no JH_BC_USER; bc_param may be NULL).  All cases
 * share the precision form and the mechanism.  Launches and graph replays
 * run on an internal stream ordered after / before `stream` with events. */
int jh_batch_create(int ncase, const int* nx, const int* ny, const double* dx, const double* dy,
                    const int* bc_kind, const double* bc_param, int precision,
                    const double* mech_consts, int n_consts, void* stream, jh_batch** out);
int jh_batch_destroy(jh_batch* batch);
/* Device state pointers of every case (U[k]: (12, nx, ny), aux[k]:
 * (2, nx, ny)); advanced in place.  Invalidates a captured graph. */
int jh_batch_bind(jh_batch* batch, void* const* U, void* const* aux);
int jh_batch_set_stream(jh_batch* batch, void* stream);
/* Stage-kernel variant for every case (as jh_ctx_set_tile). */
int jh_batch_set_tile(jh_batch* batch, int variant);
/* jh_ctx_set_inlet for case k of the batch. */
int jh_batch_set_inlet(jh_batch* batch, int k, double T0, const double* Y_inj, int n);
int jh_batch_get_tile(const jh_batch* batch);
/* nsteps steps of every case: fixed dt if dt > 0, else each case's own CFL
 * dt (set_simulator.py:438-447).  dt_hist (device, optional) receives
 * (ncase, nsteps) per-case dt.  use_graph != 0 captures the sequence in a
 * HIP graph on first use and replays it while (nsteps, dt, cfl, dt_hist,
 * tile variant, bound pointers) stay the same. */
int jh_batch_advance(jh_batch* batch, double dt, double cfl, int nsteps, double* dt_hist,
                     int use_graph);
/* OR of every case's health flags since the last call; synchronises. */
int jh_batch_health(jh_batch* batch, unsigned int* flags);
const char* jh_batch_last_error(const jh_batch* batch);

/* ---------------------------------------------------------------- AMR blocks
 * The AMR branch of the reference's step (AMR_Simulator,
 * simulation/set_simulator.py:496-532): one refinement level's same-size
 * blocks, with flux_func and update_func vmap-ed over the block axis
 * (set_rhs with is_amr, :37-42), advanced by advance_one_step_amr
 * (time_step/TVD_RK3.py:15-37) and the detailed-chemistry AMR branch of
 * advance_one_step (:107-111).  Blocks are the reference's blk_data layout,
 * (n_blk, 14, PX, PY) = U (12 planes) and aux (gamma, T), C-order.  Between
 * RK stages the reference refills the ghost layers with jaxamr's
 * get_ghost_block_data and keeps the interiors with update_external_boundary
 * (the AMR hierarchy, a third-party library): the caller does both, these
 * entry points do the per-block arithmetic.  Every block has its own
 * Lax-Friedrichs maxima over its padded field, as vmap gives; temperature
 * solves stop per cell (JH_HALLEY_PER_CELL). */
typedef struct jh_blocks jh_blocks;
/* Blocks of bnx x bny interior cells, grid spacing (dx, dy) of the level. */
int jh_blocks_create(int bnx, int bny, double dx, double dy, int precision,
                     const double* mech_consts, int n_consts, void* stream, jh_blocks** out);
int jh_blocks_destroy(jh_blocks* blocks);
/* Ghost refill of n_blk same-size blocks (bnx x bny) tiling one uniform
 * level: the padded blocks out (n_blk, 14, bnx + 6, bny + 6) are gathered
 * from the level's interiors, out[b][v][p] = interior cell idx[b P + p]
 * (flat index block * bnx * bny + x * bny + y) of plane v times its sign
 * table (s_u for plane 1, s_v for plane 2; n_blk P entries, P = (bnx + 6)
 * (bny + 6)).  The source is read through element strides sb (block), sv
 * (plane), sr (row), columns contiguous: interior blocks, or the interior
 * view of padded ones.  Replaces the ghost-layer fill of the reference's AMR
 * hierarchy (jaxamr's get_ghost_block_data, called at time_step/TVD_RK3.py:
 * 16,23,30) for one uniform level with the domain's boundary padding; the
 * tables come from janc_amd.amr.UniformBlockGrid.  Runs on stream (NULL:
 * the default stream); errors in jh_blocks_last_error(NULL). */
int jh_blocks_fill_ghosts(void* stream, int precision, int n_blk, int bnx, int bny,
                          const void* src, long long sb, long long sv, long long sr,
                          const int* idx, const void* s_u, const void* s_v, void* out);
int jh_blocks_set_stream(jh_blocks* blocks, void* stream);
const char* jh_blocks_last_error(const jh_blocks* blocks);
/* flux_func of the AMR branch (set_simulator.py:37-40): blk (n_blk, 14,
 * bnx+6, bny+6) with filled ghost layers -> rhs (n_blk, 12, bnx+6, bny+6) =
 * jnp.pad(dt * total_flux(U, aux, dx, dy), 3) per block (zero ghosts). */
int jh_blocks_flux(jh_blocks* blocks, int n_blk, const void* blk, double dt, void* rhs);
/* One stage of advance_one_step_amr (TVD_RK3.py:21-35) with the AMR
 * flux_func and update_func: blk_s = the stage's input (ghosts filled),
 * blk_n = the step's input U (ghosts filled; read for stages 1 and 2) ->
 * blk_out (n_blk, 14, bnx+6, bny+6): U_s + rhs (stage 0), 3/4 U_n + 1/4
 * (U_s + rhs) (1), 1/3 U_n + 2/3 (U_s + rhs) (2), rhs zero on the ghosts, and
 * (gamma, T) = update_aux(U_out, aux_s) over the padded block.  blk_out must
 * not alias the inputs. */
int jh_blocks_stage(jh_blocks* blocks, int n_blk, int stage, const void* blk_s,
                    const void* blk_n, double dt, void* blk_out);
/* jh_blocks_stage without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  jh_blocks_stage grows the handle's
 * per-block scalars (LF maxima, health flags) with hipMalloc when a call has
 * more blocks than any before; this entry keeps them in a file-scope device
 * array of 2048 blocks instead (each GPU holds it until the process ends) and
 * fails with JH_ERR_ARG for n_blk > 2048.  Same kernels, same results; its
 * blocks' health flags are not reported.  Calls that overlap on one GPU share
 * the array and race. */
int jh_blocks_stage_noalloc(jh_blocks* blocks, int n_blk, int stage, const void* blk_s,
                            const void* blk_n, double dt, void* blk_out);
/* Chemistry of the AMR branch (set_simulator.py:107-111), in place on
 * interior blocks (n_blk, 14, bnx, bny): U += reaction_source_terms(U, aux,
 * dt) (reaction_model.py:183-203), aux = update_aux(U, aux). */
int jh_blocks_chem(jh_blocks* blocks, int n_blk, void* blk, double dt);
/* OR of every block's health flags since the last call; synchronises. */
int jh_blocks_health(jh_blocks* blocks, unsigned int* flags);

#ifdef __cplusplus
}
#endif
#endif /* JANC_STEP_H */
