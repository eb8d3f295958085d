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
  int kind;             // ProblemKind (EK1 mode)
  int q;                // IWP order; D = d (q + 1)
  int Ns;               // number of steps (elements)
  int pad;
  double params[kMaxParams];
};

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

}  // namespace pof
