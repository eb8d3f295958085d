// Host-side interface between the C-ABI layer and the templated HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "pof_types.hpp"

namespace pof {

constexpr int kMaxLevels = 40;

// Device workspace of one IEKS solve (one IVP) with state dimension D and
// Ns steps.  Scan levels are stored back to back: level l holds
// n_l = floor(n_0 / 2^l) entries starting at entry offset off[l].
struct Workspace {
  int D = 0, d = 0, Ns = 0;
  // filter scan (n_0 = Ns)
  int f_levels = 0;
  size_t f_n[kMaxLevels] = {};
  size_t f_off[kMaxLevels] = {};
  FiltArrays fel{};      // elements and up-sweep aggregates, all levels
  StateArrays fpre{};    // prefix states, all levels
  // smoother scan (n_0 = Ns + 1, stored in reversed time order)
  int s_levels = 0;
  size_t s_n[kMaxLevels] = {};
  size_t s_off[kMaxLevels] = {};
  SmoArrays sel{};
  StateArrays spre{};
  // trajectories
  double* means[2] = {nullptr, nullptr};  // [Ns+1, D] linearisation points / smoothed means
  // per-index partial terms, reduced in a fixed order
  double* t_ssq = nullptr;   // [Ns]
  double* t_nll = nullptr;   // [Ns]
  double* t_obj = nullptr;   // [Ns]
  double* t_viol = nullptr;  // [Ns+1]
  double* scal = nullptr;    // [8] device scalars: nll, obj, ssq, means_converged
  double* part = nullptr;    // [4 * 128] partial sums of the reduction
  // model constants
  double* F = nullptr;
  double* LQ = nullptr;
  double* N1z = nullptr;
  double* Pd = nullptr;
  double* x0m = nullptr;
  double* x0L = nullptr;
  double* Hx = nullptr;      // affine mode only
  double* cx = nullptr;
  std::vector<void*> allocs;
};

// Optional per-kernel timing with HIP events (instrumented runs only).
constexpr int kMaxTimed = 256;
enum KernelClass : int {
  kKFilterLeaf = 0,
  kKFilterUp,
  kKFilterDown,
  kKFilterPost,
  kKSmoothUp,
  kKSmoothDown,
  kKSmoothPost,
  kKReduce,
  kKOutput,
  kNumKernelClasses
};
struct KernelTimer {
  int n = 0;
  hipEvent_t ev[kMaxTimed + 1];
  int cls[kMaxTimed];
  long long work[kMaxTimed];  // elements processed by the launch
  bool ready = false;
};

struct IterOpts {
  int cur = 0;           // means[cur] are the linearisation points; means[1-cur] receive the result
  int affine = 0;        // 1: observations from Hx/cx instead of EK1
  int kind = 0;
  int q = 0;
  double params[kMaxParams] = {};
  KernelTimer* timer = nullptr;  // non-null: record an event after every launch
  int with_chol = 1;             // smoother scan also propagates the factors D / L
};

// Kernel entry points per (D, d) instantiation.
struct EngineFns {
  int D, d;
  // one IEKS iteration: filter scan, smoother elements + S statistics, smoother scan,
  // smoothed means, objective, convergence partials and the reduction into ws.scal
  hipError_t (*iteration)(const Workspace& ws, const IterOpts& o, hipStream_t s);
  // smoother scan with factors over the smoother elements of the last iteration
  // (after means-only iterations): fills the smoothed factors read by output()
  hipError_t (*smooth_chol)(const Workspace& ws, const IterOpts& o, hipStream_t s);
  // project smoothed states: means_out[t] = E0 m_t, chol_out[t] = scale * E0 L_t;
  // full_means / full_chols (optional) receive the preconditioned states.
  hipError_t (*output)(const Workspace& ws, int cur, double scale, double* means_out, double* chol_out,
                       double* full_means, double* full_chols, double* filt_means, double* filt_chols,
                       hipStream_t s);
  // EK1 linearisation at n points: H [n, d, D], c [n, d]
  hipError_t (*linearize)(const Workspace& ws, const IterOpts& o, const double* m, int n, double* H, double* c,
                          hipStream_t s);
};

const EngineFns* find_engine(int D, int d);

}  // namespace pof
