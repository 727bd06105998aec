// Shared device/host types of the janc_amd step library.
#pragma once
#include <stdint.h>

#include "mech_dims.h"

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

// Every value of the mechanism, filled from Mechanism.const_array() (Python).
// Field order and sizes must match Mechanism.layout(); jh_mech_const_count()
// reports sizeof(MechConst)/8 so the host can check the match.
struct MechConst {
  double R[JH_NS];                 // 1/M_k: species gas constants (nondim)
  double M[JH_NS];                 // molar masses M/(1000*M0)
  double Tcr;                      // NASA-7 switch temperature (common)
  double cp[2][JH_NS][5];          // [low/high][k] cp per mass, highest power first
  double h[2][JH_NS][6];           // enthalpy per mass
  double hc[2][JH_NA][6];          // h/R per mole (Gibbs)
  double s[2][JH_NA][5];           // s/R without the log term
  double logc[2][JH_NA];           // a1 (log coefficient of s/R)
  double A[JH_NRX];                // pre-exponential (nondim, mol units)
  double B[JH_NRX];                // temperature exponent
  double EaR[JH_NRX];              // Ea/Ru
  double eff[JH_NRX][JH_NS];       // third-body efficiencies (x rho0/M0)
  double effM[JH_NRX][JH_NA];      // eff/M_n (Jacobian of the third-body sum)
  double Pref;                     // 101325/P0
  double T0;                       // nondim T0 (log(T0*T) in the entropy)
};

// Boundary kinds and sides: enum jh_bc_kind / jh_side of the public header.
#include "../../include/janc_step.h"

enum JhSide : int { JH_LEFT = JH_SIDE_LEFT, JH_RIGHT = JH_SIDE_RIGHT, JH_BOTTOM = JH_SIDE_BOTTOM,
                    JH_TOP = JH_SIDE_TOP };

// Number of values N* = max over cells of the Halley update count can take.
constexpr int kNstar = 6;

// Producers of (gamma, T) whose Halley solves are tracked for the grid-wide
// stop: the three RK stage kernels, the chemistry kernel, update_aux alone.
enum : int { JH_PROD_CHEM = 3, JH_PROD_UPDATE = 4, JH_PRODUCERS = 5 };

// Grid-wide Halley stop, one producer launch: N* and, for every candidate
// value of N*, the maxima the producer's epilogue reduces (the LF maxima of
// the next RK stage and, after the chemistry, the CFL maximum), all as IEEE
// bit patterns / integers for a 64-bit atomicMax.  Contiguous so one block
// reduction writes the first 1 + 2 kNstar (stages) or all (chemistry).
struct GridSlots {
  unsigned long long nstar;
  unsigned long long lf[2][kNstar];   // [x/y][N*]
  unsigned long long cfl[kNstar];     // [N*]
};
constexpr int kGridSlotsStage = 1 + 2 * kNstar;
constexpr int kGridSlotsChem = 1 + 3 * kNstar;

// Device-resident step scalars.  Maxima of non-negative doubles are kept as
// their IEEE bit patterns so a 64-bit unsigned atomicMax orders them; a NaN
// (all exponent bits set) compares above every finite value.
struct StepScalars {
  unsigned long long amax[2][3][2];   // [step parity][RK stage][x/y] LF alpha
  unsigned long long cmax[2];         // [step parity] max(sx/dx + sy/dy)
  double dt_last;                     // dt of the last completed step
  unsigned int flags;                 // health flags (JH_FLAG_*)
  unsigned int pad;
  GridSlots gs[2][JH_PRODUCERS];      // [step parity][producer]
};

// Health flags (the public header's jh_health_flag bits).
enum : unsigned int {
  JH_FLAG_NONFINITE = JH_HEALTH_NONFINITE,
  JH_FLAG_NEGATIVE_RHO = JH_HEALTH_NEGATIVE_RHO,
  JH_FLAG_HALLEY_NOCONV = JH_HEALTH_HALLEY_NOCONV,
  JH_FLAG_SINGULAR_PIVOT = JH_HEALTH_SINGULAR_PIVOT,
  JH_FLAG_HALLEY_REGRESS = JH_HEALTH_HALLEY_REGRESS,
  JH_FLAG_NONPOSITIVE_T = JH_HEALTH_NONPOSITIVE_T,
};
