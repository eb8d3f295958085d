// Host-side interface between the C-ABI layer and the templated HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "pof_types.hpp"

namespace pof {

constexpr int kMaxLevels = 40;
constexpr int kChunk = 2048;  // time points per block of the stage-1 reduction

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

// Scan-level widths (work items per launch) from which each scan kernel runs
// in its throughput form (one lane per matrix row, several combines per
// wavefront) instead of its latency form (rows split over several lanes).
// Only state dimensions whose latency form splits rows (D <= 16) have both.
struct Tuning {
  int filter_up = 1 << 30;
  int filter_down = 1 << 30;
  int smooth_up = 1 << 30;
  int smooth_down = 1 << 30;
};

// Launch maps of one active set of a group (host view of a device int buffer):
// the active IVP ids, then one prefix table per launch kind and level.
struct LaunchPlan {
  Tuning tune;
  const int* dev = nullptr;  // device buffer
  int m = 0;                 // active IVPs
  int fl = 0, sl = 0;        // filter / smoother levels (maximum over the active IVPs)
  int leaf_off = 0, leaf_total = 0;    // per IVP: Ns work items
  int one_off = 0, one_total = 0;      // 1 (the first element)
  int leaf1_off = 0, leaf1_total = 0;  // Ns - 1 (the other elements)
  int pts_off = 0, pts_total = 0;      // Ns + 1
  int chunk_off = 0, chunk_total = 0;  // ceil((Ns + 1) / kChunk)
  int fup_off[kMaxLevels] = {}, fup_total[kMaxLevels] = {};  // Ns >> (l + 1)
  int fdn_off[kMaxLevels] = {}, fdn_total[kMaxLevels] = {};  // ((Ns >> l) + 1) / 2
  int sup_off[kMaxLevels] = {}, sup_total[kMaxLevels] = {};  // (Ns + 1) >> (l + 1)
  int sdn_off[kMaxLevels] = {}, sdn_total[kMaxLevels] = {};  // (((Ns + 1) >> l) + 1) / 2
  int fdn_ivps[kMaxLevels] = {}, sdn_ivps[kMaxLevels] = {};  // active IVPs that have level l
  LaunchMap map(int off, int total) const { return LaunchMap{dev, dev + off, m, total}; }
};

// Kernel entry points per (D, d) instantiation, all over one group's active set.
struct EngineFns {
  int D, d;
  // one IEKS iteration for every active IVP: filter scan, smoother elements and
  // S statistics, smoother scan, smoothed means, objective, convergence
  // partials, and the per-IVP reduction into scal[a * 8 + 0..4] (a = active index)
  hipError_t (*iteration)(const GroupArgs& g, const LaunchPlan& p, double* scal, hipStream_t s, KernelTimer* t);
  // smoother scan with factors over the smoother elements of the last iteration
  hipError_t (*smooth_chol)(const GroupArgs& g, const LaunchPlan& p, hipStream_t s, KernelTimer* t);
  // project smoothed states into the caller's outputs (g.out_* tables, g.out_scale)
  hipError_t (*output)(const GroupArgs& g, const LaunchPlan& p, hipStream_t s);
  // filtered states of IVP ivp[0] (x0, then the level-0 filter prefixes)
  hipError_t (*filtered)(const GroupArgs& g, const LaunchPlan& p, double* fm, double* fL, hipStream_t s);
  // sequential filter + smoother iteration (at_pred: the EKF/EKS single pass)
  hipError_t (*seq_iteration)(const GroupArgs& g, const LaunchPlan& p, double* scal, int at_pred, hipStream_t s,
                              KernelTimer* t);
  hipError_t (*seq_smooth_chol)(const GroupArgs& g, const LaunchPlan& p, hipStream_t s, KernelTimer* t);
  // the smoother alone over filtered states in the filter prefix arrays (parallel or sequential)
  hipError_t (*smooth_only)(const GroupArgs& g, const LaunchPlan& p, double* scal, int sequential, hipStream_t s,
                            KernelTimer* t);
  // EK1 linearisation at n points: H [n, d, D], c [n, d]
  hipError_t (*linearize)(const ModelArgs& a, const double* m, int n, double* H, double* c, hipStream_t s);
  // linearisation of the given LinMethod at n states (m [n, D], L [n, D, D]):
  // H [n, dO, D], c [n, dO], R [n, dO, dO]; work: n * lin_work<D, d>() doubles
  hipError_t (*linearize_alt)(const ModelArgs& a, int method, const LinParams& p, const double* m, const double* L,
                              int n, double* H, double* c, double* R, double* work, hipStream_t s);
};

const EngineFns* find_engine(int D, int d);

}  // namespace pof
