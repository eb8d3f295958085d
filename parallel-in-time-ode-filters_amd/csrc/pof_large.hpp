// Large-state IEKS engine: the same solve as the per-element kernels, for state
// dimensions whose elements do not fit one worker's registers (Lorenz-96 with
// d = 32, q = 3: D = 128).  Every scan level runs as batched primitives over all
// its combines: a blocked Householder LQ (panel kernel plus a one-pass trailing
// update on the FP64 matrix cores), a strided-batched MFMA GEMM for the plain
// products, a triangular-solve kernel, and small element-wise kernels.  The host
// drives JAX's odd/even recursion.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/pof.h"
#include "pof_engine.hpp"

namespace pof {
namespace large {

struct Problem {
  int D, d, q, Ns, kind;
  const double* params;  // [kMaxParams]
  // host arrays of the set-up (row-major)
  const double* F;    // [D, D]
  const double* QL;   // [D, D]
  const double* N1z;  // [D, D] tria([0 | QL])
  const double* P;    // [D] Nordsieck scaling
  const double* x0m;  // [D]
  const double* x0L;  // [D, D]
  const double* m0;   // [Ns + 1, D] initial trajectory (preconditioned)
};

struct Engine;  // per-context state (device, workspace)

Engine* engine_create(int device);
void engine_destroy(Engine* e);

// Whole IEKS solve (pof/solver.py:11-73) on `stream`; `small` may provide the EK1
// linearisation kernel for this (D, d), else the field must have a device
// Jacobian row (Lorenz-96).  Fills out's buffers and info.
int solve(Engine* e, const Problem& p, const pof_opts& opts, pof_out* out, const EngineFns* small, hipStream_t stream,
          std::string& err);

// 1 if the large engine can linearise this field for (D, d) without a small-engine kernel
bool supports(int kind, int D, int d);

}  // namespace large
}  // namespace pof
