// C ABI of the gfx950 JANC step library: context management and the launch
// sequence of one SSP-RK3 + point-implicit-chemistry step.  Declarations and
// the reference entry points each call replaces are in include/janc_step.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/janc_step.h"
#include "adjoint.hpp"
#include "stage.hpp"

using namespace jh;

struct jh_ctx {
  int nx = 0, ny = 0;
  double dx = 0, dy = 0;
  int precision = JH_PREC_FP64;
  int bc[4] = {0, 0, 0, 0};
  double bc_p[4] = {0, 0, 0, 0};
  hipStream_t stream = nullptr;
  MechConst* mc = nullptr;        // device
  StepScalars* sc = nullptr;      // device
  void* ghost[4] = {nullptr, nullptr, nullptr, nullptr};
  void* tmpU[2] = {nullptr, nullptr};
  void* tmpA[2] = {nullptr, nullptr};
  int parity = 0;
  int tile = 0;                   // stage-kernel variant (kTileRows / kTileMinBlocks)
  std::string err;
  // kernel timing with HIP events on the context stream (jh_timing_*)
  bool timing = false;
  std::vector<hipEvent_t> ev_free;
  struct Mark { int kind; hipEvent_t a, b; };
  std::vector<Mark> marks;
  // adjoint workspace (allocated on the first jh_step_vjp call)
  double* adj = nullptr;
  size_t adj_len = 0;
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
  return a;
}

int grid_cells(const jh_ctx* c) {
  const int n = c->nx * c->ny;
  int b = (n + 255) / 256;
  return b < 2048 ? b : 2048;
}

// one thread per cell (pointwise kernels without a grid-stride loop)
int grid_all(const jh_ctx* c) { return (c->nx * c->ny + 255) / 256; }

int tile_rows(const jh_ctx* c);

int stage_blocks(const jh_ctx* c) {
  const int tx = tile_rows(c);
  return ((c->nx + tx - 1) / tx) * ((c->ny + TY - 1) / TY);
}

// Stage-kernel variants (jh_ctx_set_tile): rows along x, and the waves per
// SIMD the register budget is sized for.
constexpr int kTileRows[3] = {8, 8, 16};
constexpr int kTileWaves[3] = {3, 4, 4};

template <typename Real, int MODE>
void launch_stage(jh_ctx* c, const StepArgs<Real>& a, int stage, const Real* Us, const Real* As,
                  const Real* Un, Real* Uo, Real* Ao) {
  const int nb = stage_blocks(c);
  switch (c->tile) {
    case 1:
      stage_kernel<Real, MODE, 8, 4><<<nb, Tile<8>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
      break;
    case 2:
      stage_kernel<Real, MODE, 16, 4><<<nb, Tile<16>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
      break;
    default:
      stage_kernel<Real, MODE, 8, 3><<<nb, Tile<8>::NT, 0, c->stream>>>(a, stage, Us, As, Un, Uo, Ao);
  }
}

int tile_rows(const jh_ctx* c) { return kTileRows[c->tile]; }

int bc_blocks(const jh_ctx* c) { return (2 * c->nx + 2 * c->ny + 255) / 256; }

// Prologue for a state handed in by the caller: zero this parity's slots and
// reduce the interior maxima (stage-1 LF alpha, CFL).
template <typename Real>
int prologue(jh_ctx* c, const Real* U, const Real* aux, double dt, double cfl) {
  Timed tm(c, JH_KERNEL_REDUCE);
  reset_kernel<<<1, 64, 0, c->stream>>>(c->sc, c->parity);
  auto a = make_args<Real>(c, dt, cfl);
  reduce_kernel<Real><<<grid_cells(c), 256, 0, c->stream>>>(a, U, aux, c->parity, 1);
  return check_launch(c, "reduce_kernel");
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
  for (int s = 0; s < nsteps; ++s) {
    auto a = make_args<Real>(c, dt, cfl);
    const int bb = bc_blocks(c);
    // stage 1: (Un, An) -> (U1, A1)
    { Timed t(c, JH_KERNEL_BC); bc_kernel<Real><<<bb, 256, 0, c->stream>>>(a, Un, An, 0); }
    { Timed t(c, JH_KERNEL_STAGE);
      launch_stage<Real, MODE_RK>(c, a, 0, Un, An, Un, U1, A1); }
    // stage 2: (U1, A1) -> (U2, A2)
    { Timed t(c, JH_KERNEL_BC); bc_kernel<Real><<<bb, 256, 0, c->stream>>>(a, U1, A1, 1); }
    { Timed t(c, JH_KERNEL_STAGE);
      launch_stage<Real, MODE_RK>(c, a, 1, U1, A1, Un, U2, A2); }
    // stage 3: (U2, A2) -> (Uo, auxo); Uo may alias Un (each thread reads
    // Un at its own cell before writing it)
    { Timed t(c, JH_KERNEL_BC); bc_kernel<Real><<<bb, 256, 0, c->stream>>>(a, U2, A2, 2); }
    { Timed t(c, JH_KERNEL_STAGE);
      launch_stage<Real, MODE_RK>(c, a, 2, U2, A2, Un, Uo, auxo); }
    // chemistry + final update_aux in place, next step's maxima
    { Timed t(c, JH_KERNEL_CHEM);
      chem_kernel<Real, CHEM_STEP><<<grid_all(c), 256, 0, c->stream>>>(
          a, Uo, auxo, nullptr, dt_hist ? dt_hist + s : nullptr); }
    rc = check_launch(c, "step kernels");
    if (rc) return rc;
    c->parity ^= 1;
    Un = Uo;
    An = auxo;
  }
  (void)n;
  return JH_OK;
}

template <typename Real>
int rhs_impl(jh_ctx* c, const Real* U, const Real* aux, double dt, Real* out) {
  int rc = prologue<Real>(c, U, aux, dt, 0.0);
  if (rc) return rc;
  auto a = make_args<Real>(c, dt, 0.0);
  bc_kernel<Real><<<bc_blocks(c), 256, 0, c->stream>>>(a, U, aux, 0);
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

struct AdjWork {
  double *U[5], *A[5];          // forward states of the step: 0 = input, 1-3 RK, 4 = output
  double *Ub, *Uacc, *ab, *rb;  // cotangents
  double *Up, *ap, *Ubp, *abp;  // padded
  double *cx, *cy, *part, *scal;
  unsigned int* cnt;
  int npad, nblk;
};

int adj_work(jh_ctx* c, const double* U0, const double* A0, AdjWork& w) {
  const size_t n = (size_t)c->nx * c->ny;
  const size_t np = (size_t)(c->nx + 6) * (c->ny + 6);
  const size_t nfx = (size_t)(c->nx + 1) * c->ny, nfy = (size_t)c->nx * (c->ny + 1);
  const int nblk = blocks_for(np);
  const size_t len = 4 * (NU + 2) * n + (3 * NU + 2) * n + 2 * (NU + 2) * np +
                     (size_t)NU * 10 * (nfx + nfy) + 2 * nblk + 2 + 2;
  if (c->adj_len < len) {
    if (c->adj) (void)hipFree(c->adj);
    c->adj = nullptr;
    c->adj_len = 0;
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
  w.Up = p; p += NU * np;
  w.ap = p; p += 2 * np;
  w.Ubp = p; p += NU * np;
  w.abp = p; p += 2 * np;
  w.cx = p; p += NU * 10 * nfx;
  w.cy = p; p += NU * 10 * nfy;
  w.part = p; p += 2 * nblk;
  w.scal = p; p += 2;
  w.cnt = reinterpret_cast<unsigned int*>(p);
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

// Flux VJP of one RK stage with input (Us, As) and rhs cotangent rb; the
// results are added into (Ut, At).
void flux_vjp(jh_ctx* c, const AdjArgs& a, AdjWork& w, const double* Us, const double* As,
              const double* rb, const unsigned long long* alpha, double* Ut, double* At) {
  const size_t n = (size_t)c->nx * c->ny;
  const int nfx = (c->nx + 1) * c->ny, nfy = c->nx * (c->ny + 1);
  hipStream_t st = c->stream;
  pad_kernel<<<blocks_for(w.npad), 256, 0, st>>>(a, Us, As, w.Up, w.ap);
  face_vjp_kernel<0><<<blocks_for(nfx), 256, 0, st>>>(a, w.Up, w.ap, rb, alpha, w.cx);
  face_vjp_kernel<1><<<blocks_for(nfy), 256, 0, st>>>(a, w.Up, w.ap, rb, alpha, w.cy);
  cell_vjp_kernel<<<w.nblk, 256, 0, st>>>(a, w.Up, w.ap, w.cx, w.cy, alpha, w.Ubp, w.abp, w.part);
  alpha_reduce_kernel<<<1, 256, 0, st>>>(w.part, w.nblk, w.scal);
  zero_u32_kernel<<<1, 64, 0, st>>>(w.cnt, 2);
  alpha_count_kernel<<<blocks_for(w.npad), 256, 0, st>>>(a, w.Up, w.ap, alpha, w.cnt);
  alpha_apply_kernel<<<blocks_for(w.npad), 256, 0, st>>>(a, w.Up, w.ap, alpha, w.scal, w.cnt,
                                                         w.Ubp, w.abp);
  unpad_kernel<<<blocks_for(n), 256, 0, st>>>(a, w.Ubp, w.abp, Ut, At);
}

int step_vjp_impl(jh_ctx* c, const double* U0, const double* A0, double dt, const double* Ubar,
                  const double* Abar, double* Ub0, double* Ab0, double* Uo, double* Ao) {
  AdjWork w;
  int rc = adj_work(c, U0, A0, w);
  if (rc) return rc;
  const size_t n = (size_t)c->nx * c->ny;
  const int bn = blocks_for(NU * n), b2 = blocks_for(2 * n), bc1 = blocks_for(n);
  hipStream_t st = c->stream;
  // ---- forward recompute, keeping every stage ----
  rc = prologue<double>(c, U0, A0, dt, 0.0);
  if (rc) return rc;
  const int P = c->parity;
  auto a = make_args<double>(c, dt, 0.0);
  const int bb = bc_blocks(c);
  for (int s = 0; s < 3; ++s) {
    bc_kernel<double><<<bb, 256, 0, st>>>(a, w.U[s], w.A[s], s);
    launch_stage<double, MODE_RK>(c, a, s, w.U[s], w.A[s], U0, w.U[s + 1], w.A[s + 1]);
  }
  HIP_TRY(c, hipMemcpyAsync(w.U[4], w.U[3], NU * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(c, hipMemcpyAsync(w.A[4], w.A[3], 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  chem_kernel<double, CHEM_STEP><<<grid_all(c), 256, 0, st>>>(a, w.U[4], w.A[4], nullptr, nullptr);
  rc = check_launch(c, "vjp forward");
  if (rc) return rc;
  c->parity ^= 1;
  if (Uo) HIP_TRY(c, hipMemcpyAsync(Uo, w.U[4], NU * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (Ao) HIP_TRY(c, hipMemcpyAsync(Ao, w.A[4], 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  const AdjArgs aa = adj_args(c, dt);
  const unsigned long long* alpha[3] = {&c->sc->amax[P][0][0], &c->sc->amax[P][1][0],
                                        &c->sc->amax[P][2][0]};
  // ---- reverse sweep ----
  // final update_aux and chemistry: Ub = U4_bar -> U3_bar
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 1.0, Ubar, 0.0, w.Ub);
  getT_vjp_kernel<<<bc1, 256, 0, st>>>(aa, w.U[4], w.A[4], Abar, w.Ub);
  chem_vjp_kernel<<<bc1, 256, 0, st>>>(aa, w.U[3], w.A[3], w.Ub, w.Ub, w.ab);
  getT_vjp_kernel<<<bc1, 256, 0, st>>>(aa, w.U[3], w.A[3], w.ab, w.Ub);
  // stage 3: U3 = 1/3 U + 2/3 (U2 + r(U2))
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 1.0 / 3.0, w.Ub, 0.0, w.Uacc);
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 2.0 / 3.0, w.Ub, 0.0, w.rb);
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 1.0, w.rb, 0.0, w.Ub);
  fill_kernel<<<b2, 256, 0, st>>>(2 * n, 0.0, w.ab);
  flux_vjp(c, aa, w, w.U[2], w.A[2], w.rb, alpha[2], w.Ub, w.ab);
  getT_vjp_kernel<<<bc1, 256, 0, st>>>(aa, w.U[2], w.A[2], w.ab, w.Ub);
  // stage 2: U2 = 3/4 U + 1/4 (U1 + r(U1))
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 3.0 / 4.0, w.Ub, 1.0, w.Uacc);
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 1.0 / 4.0, w.Ub, 0.0, w.rb);
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 1.0, w.rb, 0.0, w.Ub);
  fill_kernel<<<b2, 256, 0, st>>>(2 * n, 0.0, w.ab);
  flux_vjp(c, aa, w, w.U[1], w.A[1], w.rb, alpha[1], w.Ub, w.ab);
  getT_vjp_kernel<<<bc1, 256, 0, st>>>(aa, w.U[1], w.A[1], w.ab, w.Ub);
  // stage 1: U1 = U + r(U)
  axpby_kernel<<<bn, 256, 0, st>>>(NU * n, 1.0, w.Ub, 1.0, w.Uacc);
  fill_kernel<<<b2, 256, 0, st>>>(2 * n, 0.0, w.ab);
  flux_vjp(c, aa, w, w.U[0], w.A[0], w.Ub, alpha[0], w.Uacc, w.ab);
  HIP_TRY(c, hipMemcpyAsync(Ub0, w.Uacc, NU * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(c, hipMemcpyAsync(Ab0, w.ab, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  return check_launch(c, "vjp kernels");
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
    if (bc_kind[s] < JH_BC_PERIODIC || bc_kind[s] > JH_BC_USER)
      return fail(nullptr, JH_ERR_ARG, "unknown boundary kind");
  jh_ctx* c = new jh_ctx();
  c->nx = nx;
  c->ny = ny;
  c->dx = dx;
  c->dy = dy;
  c->precision = precision;
  c->stream = static_cast<hipStream_t>(stream);
  for (int s = 0; s < 4; ++s) {
    c->bc[s] = bc_kind[s];
    c->bc_p[s] = bc_param ? bc_param[s] : 0.0;
  }
  const size_t es = elem_size(c), n = (size_t)nx * ny;
  hipError_t e = hipSuccess;
  e = e ? e : hipMalloc(&c->mc, sizeof(MechConst));
  e = e ? e : hipMalloc(&c->sc, sizeof(StepScalars));
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
  (void)hipFree(c->mc);
  (void)hipFree(c->sc);
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

int jh_ctx_set_tile(jh_ctx* c, int variant) {
  if (!c) return JH_ERR_ARG;
  if (variant < 0 || variant >= jh_tile_variants()) return fail(c, JH_ERR_ARG, "unknown tile variant");
  c->tile = variant;
  return JH_OK;
}

int jh_tile_variants(void) { return 3; }

int jh_tile_info(int variant, int* rows, int* threads, int* waves) {
  if (variant < 0 || variant >= 3 || !rows || !threads || !waves) return JH_ERR_ARG;
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
  for (int s = 0; s < 4; ++s)
    if (c->bc[s] == JH_BC_USER) return fail(c, JH_ERR_ARG, "user boundary strips not filled");
  if (c->precision == JH_PREC_FP64)
    return rhs_impl<double>(c, (const double*)U, (const double*)aux, dt, (double*)out);
  return rhs_impl<float>(c, (const float*)U, (const float*)aux, dt, (float*)out);
}

int jh_update_aux(jh_ctx* c, const void* U, const void* aux_in, void* aux_out) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux_in || !aux_out) return fail(c, JH_ERR_ARG, "NULL pointer");
  if (c->precision == JH_PREC_FP64) {
    auto a = make_args<double>(c, 1.0, 0.0);
    update_aux_kernel<double><<<grid_all(c), 256, 0, c->stream>>>(
        a, (const double*)U, (const double*)aux_in, (double*)aux_out);
  } else {
    auto a = make_args<float>(c, 1.0, 0.0);
    update_aux_kernel<float><<<grid_all(c), 256, 0, c->stream>>>(
        a, (const float*)U, (const float*)aux_in, (float*)aux_out);
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
  int rc;
  if (c->precision == JH_PREC_FP64) {
    rc = prologue<double>(c, (const double*)U, (const double*)aux, 1.0, 0.0);
    if (rc) return rc;
    auto a = make_args<double>(c, 1.0, 0.0);
    bc_kernel<double><<<bc_blocks(c), 256, 0, c->stream>>>(a, (const double*)U,
                                                            (const double*)aux, 0);
  } else {
    rc = prologue<float>(c, (const float*)U, (const float*)aux, 1.0, 0.0);
    if (rc) return rc;
    auto a = make_args<float>(c, 1.0, 0.0);
    bc_kernel<float><<<bc_blocks(c), 256, 0, c->stream>>>(a, (const float*)U,
                                                          (const float*)aux, 0);
  }
  copy_scalars_kernel<<<1, 64, 0, c->stream>>>(c->sc, c->parity, 0.0, nullptr, alpha_dev);
  return check_launch(c, "lf_alpha kernels");
}

int jh_step_vjp(jh_ctx* c, const void* U, const void* aux, double dt, const void* U_bar,
                const void* aux_bar, void* U_in_bar, void* aux_in_bar, void* U_out,
                void* aux_out) {
  if (!c) return JH_ERR_ARG;
  if (!U || !aux || !U_bar || !aux_bar || !U_in_bar || !aux_in_bar)
    return fail(c, JH_ERR_ARG, "NULL pointer");
  if (c->precision != JH_PREC_FP64) return fail(c, JH_ERR_ARG, "the adjoint is FP64 only");
  if (!(dt > 0)) return fail(c, JH_ERR_ARG, "the adjoint needs a fixed dt > 0");
  for (int s = 0; s < 4; ++s)
    if (c->bc[s] == JH_BC_PRESSURE_OUTLET || c->bc[s] == JH_BC_USER)
      return fail(c, JH_ERR_ARG, "the adjoint supports periodic, zero_gradient, slip_wall and "
                                 "no_slip_wall boundaries");
  return step_vjp_impl(c, (const double*)U, (const double*)aux, dt, (const double*)U_bar,
                       (const double*)aux_bar, (double*)U_in_bar, (double*)aux_in_bar,
                       (double*)U_out, (double*)aux_out);
}

int jh_timing_enable(jh_ctx* c, int on) {
  if (!c) return JH_ERR_ARG;
  c->timing = on != 0;
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
  return JH_OK;
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

}  // extern "C"
