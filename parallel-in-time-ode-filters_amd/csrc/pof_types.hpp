// Plain data shared by the device kernels and the host driver.
#pragma once

#include "pof_problems.hpp"

namespace pof {

// --------------------------------------------------------------- model args
struct ModelArgs {
  const double* F;      // [D*D] preconditioned IWP transition
  const double* LQ;     // [D*D] lower Cholesky factor of the process noise
  const double* N1z;    // [D*D] tria([0 | LQ]): predicted factor from a zero prior
  const double* Pd;     // [D]   Nordsieck scaling (diagonal of P)
  const double* x0m;    // [D]   preconditioned initial mean
  const double* x0L;    // [D*D] preconditioned initial factor
  const double* lin;    // [Ns+1, D] linearisation points (EK1 mode), or null
  const double* Hx;     // [Ns, d, D] explicit observation matrices (affine mode), or null
  const double* cx;     // [Ns, d]    explicit observation offsets (affine mode)
  const double* Rx;     // [Ns, d, d] explicit observation noise factors (affine mode), or null = 0
  int kind;             // ProblemKind (EK1 mode)
  int q;                // IWP order; D = d (q + 1)
  int Ns;               // number of steps (elements)
  int pad;
  double params[kMaxParams];
};

// linearisations of the residual model (pof/observations.py, pof/linearization/unscented.py)
enum LinMethod : int { kLinEK1 = 0, kLinEK0 = 1, kLinUncertain = 2, kLinRegularized = 3, kLinUnscented = 4 };

struct LinParams {
  double lam;     // regularisation weight (linearize_regularized)
  double alpha;   // unscented spread
  double beta;    // unscented prior weight
  double kappa;   // unscented kappa; used when kappa_set
  int kappa_set;  // 0: kappa = 3 + D (unscented.py:47-48)
  int pad;
};

// doubles of per-state work buffer of the uncertain / unscented linearisations
__host__ __device__ constexpr int lin_work_host(int D, int d) {
  return 2 * (2 * D + 1) * d + 3 * D * d + 3 * d * d + 2 * d + 8;
}

// filter element / aggregate arrays (structure of arrays; one entry per element)
struct FiltArrays {
  double* A;    // [n, D, D]
  double* b;    // [n, D]
  double* U;    // [n, D, D]
  double* eta;  // [n, D]
  double* Z;    // [n, D, D]
};
// smoother element / aggregate arrays
struct SmoArrays {
  double* g;    // [n, D]
  double* E;    // [n, D, D]
  double* Dm;   // [n, D, D]
};
// prefix states (b, U) or (g, L)
struct StateArrays {
  double* m;    // [n, D]
  double* L;    // [n, D, D]
};

// --------------------------------------------------------------- batch group
// A group of IVPs that share (D, d, q) and therefore F, LQ and N1z.  IVP j
// owns the entry range [base[j], base[j] + 2 (Ns[j] + 1)) of every array:
// its filter levels (entries base + sum_{t<l} Ns >> t), its smoother levels
// (n_0 = Ns + 1, same rule) and its time points (Ns + 1 entries).
struct GroupArgs {
  const double* F;            // [D*D]
  const double* LQ;           // [D*D]
  const double* N1z;          // [D*D]
  const double* Pd;           // [slots, D]
  const double* x0m;          // [slots, D]
  const double* x0L;          // [slots, D*D]
  const double* params;       // [slots, kMaxParams]
  const int* kind;            // [slots]
  const int* Ns;              // [slots]
  const long long* base;      // [slots] first entry of the IVP's range
  const int* cur;             // [slots] means buffer holding the linearisation points
  const int* wchol;           // [slots] 1: smoother scans also carry the factors
  double* means0;             // [entries, D]
  double* means1;             // [entries, D]
  FiltArrays fel;             // filter elements and up-sweep aggregates
  StateArrays fpre;           // filter prefix states
  SmoArrays sel;              // smoother elements and aggregates
  StateArrays spre;           // smoother prefix states
  double* t_ssq;              // [entries] per-step partial terms
  double* t_nll;
  double* t_obj;
  double* t_viol;
  double* part;               // [chunks, 4] reduction partials
  double* const* out_means;   // [slots] caller-owned outputs (may hold null)
  double* const* out_chols;
  double* const* out_full_means;
  double* const* out_full_chols;
  const double* out_scale;    // [slots] factor scale of the output
  const int* out_pfull;       // [slots] 1: full-state outputs are P (m, L) instead of (m, L)
  const double* Hx;           // affine mode (one IVP at entry 0): [Ns, d, D]
  const double* cx;           // [Ns, d]
  const double* Rx;           // [Ns, d, d] or null (noiseless)
  int q;
  int affine;
  int lq_lower;  // 1: LQ's strict upper part is zero (the smoother-element factorisation skips it)
};

// Work list of one launch: active IVP a (id ivp[a]) owns the work items
// [pre[a], pre[a + 1]); total = pre[m].
struct LaunchMap {
  const int* ivp;
  const int* pre;
  int m;
  int total;
};

}  // namespace pof
