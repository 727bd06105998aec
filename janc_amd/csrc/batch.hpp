// Batched launch of many independent cases (uneven grid sizes) of the step.
//
// Each case keeps its own state, stage buffers, ghost strips and reduction
// scalars; a device array of CaseDesc describes them.  One launch per
// kernel kind covers every case: a block (or, for the persistent stage
// kernel, a job) finds its case by binary search in a prefix-sum table of
// per-case block counts, loads that case's arguments and runs the same
// per-tile / per-cell body as the single-case kernels, so every case's result
// is bit-identical to advancing it alone.
#pragma once
#include <hip/hip_runtime.h>

#include "stage.hpp"

namespace jh {

template <typename Real>
struct CaseDesc {
  StepArgs<Real> a;          // geometry, BCs, strips, scalars of the case
  Real* U;                   // state (12 planes), advanced in place
  Real* aux;                 // (gamma, T)
  Real* U1;                  // RK stage buffers
  Real* U2;
  Real* A1;
  Real* A2;
};

// Case owning flat block index b: off[k] <= b < off[k+1] (off has ncase+1
// entries, off[0] = 0).  b is wave-uniform, so this runs on the scalar unit.
__device__ __forceinline__ int find_case(const int* __restrict__ off, int ncase, int b) {
  int lo = 0, hi = ncase;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename Real>
__device__ __forceinline__ StepArgs<Real> case_args(const CaseDesc<Real>& d, int parity,
                                                    double dt_fixed, double cfl) {
  StepArgs<Real> a = d.a;
  a.parity = parity;
  a.dt_fixed = dt_fixed;
  a.cfl = cfl;
  return a;
}

// Zero the parity's reduction slots of every case (thread k -> case k).
template <typename Real>
__global__ void reset_batch_kernel(const CaseDesc<Real>* __restrict__ d, int ncase, int parity) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncase) return;
  StepScalars* sc = d[k].a.sc;
#pragma unroll
  for (int i = 0; i < 6; ++i) (&sc->amax[parity][0][0])[i] = 0ull;
  sc->cmax[parity] = 0ull;
}

template <typename Real>
__global__ __launch_bounds__(256) void reduce_batch_kernel(const CaseDesc<Real>* __restrict__ d,
                                                           const int* __restrict__ off, int ncase,
                                                           int parity, double dt_fixed,
                                                           double cfl) {
  __shared__ double red[4];
  __shared__ unsigned long long redb[4];
  const int k = find_case(off, ncase, blockIdx.x);
  const StepArgs<Real> a = case_args(d[k], parity, dt_fixed, cfl);
  reduce_body<Real>(a, d[k].U, d[k].aux, parity, 1, blockIdx.x - off[k], off[k + 1] - off[k],
                    red, redb);
}

template <typename Real>
__global__ __launch_bounds__(kBcThreads) void bc_batch_kernel(const CaseDesc<Real>* __restrict__ d,
                                                       const int* __restrict__ off, int ncase,
                                                       int stage, int parity, double dt_fixed,
                                                       double cfl) {
  __shared__ double red[4];
  const int k = find_case(off, ncase, blockIdx.x);
  const CaseDesc<Real>& cd = d[k];
  const StepArgs<Real> a = case_args(cd, parity, dt_fixed, cfl);
  const Real* U = stage == 0 ? cd.U : stage == 1 ? cd.U1 : cd.U2;
  const Real* A = stage == 0 ? cd.aux : stage == 1 ? cd.A1 : cd.A2;
  bc_body<Real>(a, U, A, stage, blockIdx.x - off[k], red);
}

// Persistent stage kernel: gridDim.x blocks (a multiple of 8, at most the
// resident capacity) loop over the flat tile list of all cases.  Job j goes
// through xcd_tile so each XCD (blocks are dealt to the 8 XCDs round-robin,
// and j = blockIdx.x + i*gridDim.x keeps j's XCD) walks a contiguous run of
// the list: whole cases or long runs of one case's x-adjacent tiles.
// MIRROR: every case's sides are copy kinds and the halo is read from the
// source cells (no ghost-strip launches).
template <typename Real, int TX, int WPE, bool MIRROR = false>
__global__ __launch_bounds__(Tile<TX>::NT) __attribute__((amdgpu_waves_per_eu(WPE)))
void stage_batch_kernel(const CaseDesc<Real>* __restrict__ d, const int* __restrict__ off,
                        int ncase, int stage, int parity, double dt_fixed, double cfl) {
  __shared__ StageLds<Real, TX> L;
  const int total = off[ncase];
  for (int job = blockIdx.x; job < total; job += gridDim.x) {
    const int t = xcd_tile(job, total);
    const int k = find_case(off, ncase, t);
    const CaseDesc<Real>& cd = d[k];
    const StepArgs<Real> a = case_args(cd, parity, dt_fixed, cfl);
    const Real* Us = stage == 0 ? cd.U : stage == 1 ? cd.U1 : cd.U2;
    const Real* As = stage == 0 ? cd.aux : stage == 1 ? cd.A1 : cd.A2;
    Real* Uo = stage == 0 ? cd.U1 : stage == 1 ? cd.U2 : cd.U;
    Real* Ao = stage == 0 ? cd.A1 : stage == 1 ? cd.A2 : cd.aux;
    stage_body<Real, MODE_RK, TX, RECON_WENO5, SPLIT_LF, Real, false, false, MIRROR>(
        a, stage, Us, As, cd.U, Uo, Ao, t - off[k], L);
    __syncthreads();   // LDS is reused by the next job
  }
}

// Chemistry + final update_aux for every case; dt_hist (optional) is
// (ncase, hist_len) row-major, step s written at column s.
template <typename Real>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void chem_batch_kernel(const CaseDesc<Real>* __restrict__ d,
                                                         const int* __restrict__ off, int ncase,
                                                         int parity, double dt_fixed, double cfl,
                                                         double* dt_hist, int hist_len, int s) {
  __shared__ double red[4];
  __shared__ unsigned long long redb[4];
  const int k = find_case(off, ncase, blockIdx.x);
  const CaseDesc<Real>& cd = d[k];
  const StepArgs<Real> a = case_args(cd, parity, dt_fixed, cfl);
  __shared__ unsigned long long redg[kGridSlotsChem * 4];
  __shared__ MixTables mix;
  __shared__ GibbsTables gt;
  chem_body<Real, CHEM_STEP>(a, cd.U, cd.aux, nullptr,
                             dt_hist ? dt_hist + (size_t)k * hist_len + s : nullptr,
                             blockIdx.x - off[k], red, redb, redg, mix, gt);
}

}  // namespace jh
