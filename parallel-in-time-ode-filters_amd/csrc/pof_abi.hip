// C-ABI of libpof_hip.so: host set-up, device workspace and the IEKS loop.
//
// Host set-up restates pof/convenience.py:13-45 (set_up_solver),
// pof/transitions.py (IWP prior, Nordsieck preconditioner),
// pof/initialization.py:15-22,42-56,75-89 (Taylor-mode x0, constant and
// prior trajectories); the loop restates pof/solver.py:11-73 with the
// convergence test of pof/convergence_criteria.py:4-13.  Per-iteration device
// work is in pof_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pof.h"
#include "pof_engine.hpp"
#include "pof_large.hpp"
#include "pof_problems.hpp"

namespace pof {

// ------------------------------------------------------------ power series
// Truncated Taylor series with kJet coefficients, for Taylor-mode derivatives.
constexpr int kJet = 12;
struct Jet {
  double c[kJet];
};
__host__ __device__ inline Jet jconst(double v) {
  Jet r{};
  r.c[0] = v;
  return r;
}
__host__ __device__ inline Jet operator+(const Jet& a, const Jet& b) {
  Jet r;
  for (int i = 0; i < kJet; ++i) r.c[i] = a.c[i] + b.c[i];
  return r;
}
__host__ __device__ inline Jet operator-(const Jet& a, const Jet& b) {
  Jet r;
  for (int i = 0; i < kJet; ++i) r.c[i] = a.c[i] - b.c[i];
  return r;
}
__host__ __device__ inline Jet operator-(const Jet& a) {
  Jet r;
  for (int i = 0; i < kJet; ++i) r.c[i] = -a.c[i];
  return r;
}
__host__ __device__ inline Jet operator*(const Jet& a, const Jet& b) {
  Jet r{};
  for (int k = 0; k < kJet; ++k) {
    double s = 0.0;
    for (int j = 0; j <= k; ++j) s += a.c[j] * b.c[k - j];
    r.c[k] = s;
  }
  return r;
}
__host__ __device__ inline Jet operator/(const Jet& a, const Jet& b) {
  Jet r{};
  for (int k = 0; k < kJet; ++k) {
    double s = a.c[k];
    for (int j = 0; j < k; ++j) s -= r.c[j] * b.c[k - j];
    r.c[k] = s / b.c[0];
  }
  return r;
}
__host__ __device__ inline Jet operator+(const Jet& a, double b) { Jet r = a; r.c[0] += b; return r; }
__host__ __device__ inline Jet operator+(double b, const Jet& a) { return a + b; }
__host__ __device__ inline Jet operator-(const Jet& a, double b) { Jet r = a; r.c[0] -= b; return r; }
__host__ __device__ inline Jet operator-(double b, const Jet& a) { Jet r = -a; r.c[0] += b; return r; }
__host__ __device__ inline Jet operator*(const Jet& a, double b) {
  Jet r;
  for (int i = 0; i < kJet; ++i) r.c[i] = a.c[i] * b;
  return r;
}
__host__ __device__ inline Jet operator*(double b, const Jet& a) { return a * b; }
__host__ __device__ inline Jet operator/(const Jet& a, double b) {
  Jet r;
  for (int i = 0; i < kJet; ++i) r.c[i] = a.c[i] / b;
  return r;
}
__host__ __device__ inline Jet dsqrt(const Jet& a) {
  Jet s{};
  s.c[0] = sqrt(a.c[0]);
  for (int k = 1; k < kJet; ++k) {
    double acc = a.c[k];
    for (int j = 1; j < k; ++j) acc -= s.c[j] * s.c[k - j];
    s.c[k] = acc / (2.0 * s.c[0]);
  }
  return s;
}

template <int d>
static void field_dispatch_jet(int kind, const double* p, const Jet* y, Jet* f) {
  vector_field<Jet, d>(kind, p, y, f);
}

static bool field_jets(int kind, int d, const double* p, const Jet* y, Jet* f) {
  switch (d) {
    case 1: field_dispatch_jet<1>(kind, p, y, f); return true;
    case 2: field_dispatch_jet<2>(kind, p, y, f); return true;
    case 3: field_dispatch_jet<3>(kind, p, y, f); return true;
    case 4: field_dispatch_jet<4>(kind, p, y, f); return true;
    case 5: field_dispatch_jet<5>(kind, p, y, f); return true;
    case 6: field_dispatch_jet<6>(kind, p, y, f); return true;
    case 8: field_dispatch_jet<8>(kind, p, y, f); return true;
    case 16: field_dispatch_jet<16>(kind, p, y, f); return true;
    case 32: field_dispatch_jet<32>(kind, p, y, f); return true;
    default: return false;
  }
}

// Derivatives y^(k)(t0), k = 0..q: out[k * d + i] (tornadox.init.TaylorMode,
// called at pof/initialization.py:17-19): (k+1) a_{k+1} = [f(y(t))]_k, y^(k) = k! a_k.
static bool taylor_derivatives(int kind, int d, const double* p, const double* y0, int q, double* out) {
  if (q + 1 > kJet) return false;
  std::vector<Jet> y(d), f(d);
  for (int i = 0; i < d; ++i) y[i] = jconst(y0[i]);
  for (int k = 0; k < q; ++k) {
    if (!field_jets(kind, d, p, y.data(), f.data())) return false;
    for (int i = 0; i < d; ++i) y[i].c[k + 1] = f[i].c[k] / (k + 1);
  }
  double fact = 1.0;
  for (int k = 0; k <= q; ++k) {
    if (k > 0) fact *= k;
    for (int i = 0; i < d; ++i) out[k * d + i] = y[i].c[k] * fact;
  }
  return true;
}

static bool field_eval(int kind, int d, const double* p, const double* y, double* f) {
  switch (d) {
#define POF_FE(dd) \
  case dd: vector_field<double, dd>(kind, p, y, f); return true;
    POF_FE(1) POF_FE(2) POF_FE(3) POF_FE(4) POF_FE(5) POF_FE(6) POF_FE(8) POF_FE(16) POF_FE(32)
#undef POF_FE
    default: return false;
  }
}

// ------------------------------------------------------------- host algebra
// Householder LQ with LAPACK's sign convention (the device lq() on the host).
// 1 if the strict upper part of the n x n row-major matrix A is exactly zero
static bool strictly_upper_zero(const double* A, int n) {
  for (int r = 0; r < n; ++r)
    for (int c = r + 1; c < n; ++c)
      if (A[(size_t)r * n + c] != 0.0) return false;
  return true;
}

static void tria_host(const std::vector<double>& M, int nr, int nc, std::vector<double>& L) {
  std::vector<double> a(M);
  const int K = std::min(nr, nc);
  std::vector<double> v(nc);
  for (int i = 0; i < K && i < nc - 1; ++i) {
    double* pr = &a[(size_t)i * nc];
    const double alpha = pr[i];
    double sig = 0.0;
    for (int j = i + 1; j < nc; ++j) sig = std::fma(pr[j], pr[j], sig);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (sig != 0.0) {
      const double nrm = std::sqrt(std::fma(alpha, alpha, sig));
      beta = -copysign(nrm, alpha);  // Fortran SIGN: honours the sign of -0.0
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int j = i + 1; j < nc; ++j) v[j] = pr[j] * scal;
    pr[i] = beta;
    for (int j = i + 1; j < nc; ++j) pr[j] = 0.0;
    for (int r = i + 1; r < nr; ++r) {
      double* rr = &a[(size_t)r * nc];
      double w = rr[i];
      for (int j = i + 1; j < nc; ++j) w = std::fma(rr[j], v[j], w);
      const double tw = tau * w;
      rr[i] -= tw;
      for (int j = i + 1; j < nc; ++j) rr[j] = std::fma(-tw, v[j], rr[j]);
    }
  }
  L.assign((size_t)nr * K, 0.0);
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < K; ++c) L[(size_t)r * K + c] = a[(size_t)r * nc + c];
}

struct Setup {
  int d = 0, q = 0, D = 0, Ns = 0;
  double dt = 0.0;
  std::vector<double> F, LQ, N1z, P, PI, x0m, x0L, A1;
};

// (transitions.py:37-50) A_1d = flip(pascal_lower), L_Q = chol(flip(hilbert)), kron(I_d, .)
static void iwp_1d(int q, std::vector<double>& A, std::vector<double>& L) {
  const int n = q + 1;
  A.assign((size_t)n * n, 0.0);
  std::vector<double> Q((size_t)n * n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      if (j >= i) {
        // C(q - i, j - i)
        double c = 1.0;
        const int N = q - i, k = j - i;
        for (int t = 1; t <= k; ++t) c = c * (N - k + t) / t;
        A[(size_t)i * n + j] = c;
      }
      Q[(size_t)i * n + j] = 1.0 / (double)(2 * q + 1 - i - j);
    }
  }
  L.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double s = Q[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    const double ljj = std::sqrt(s);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double t = Q[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = t / ljj;
    }
  }
}

static std::vector<double> kron_eye(int d, const std::vector<double>& B, int n) {
  const int D = d * n;
  std::vector<double> K((size_t)D * D, 0.0);
  for (int b = 0; b < d; ++b)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) K[(size_t)(b * n + i) * D + (b * n + j)] = B[(size_t)i * n + j];
  return K;
}

// (transitions.py:53-61) P = diag(|dt|^(q-i+1/2) / (q-i)!), PI = diag(|dt|^-(q-i+1/2) (q-i)!)
static void nordsieck(int d, int q, double dt, std::vector<double>& P, std::vector<double>& PI) {
  const int n = q + 1;
  P.resize((size_t)d * n);
  PI.resize((size_t)d * n);
  for (int i = 0; i < n; ++i) {
    const int pw = q - i;
    double fact = 1.0;
    for (int t = 2; t <= pw; ++t) fact *= t;
    const double e = pw + 0.5;
    const double p = std::pow(std::fabs(dt), e) / fact;
    const double pi = std::pow(std::fabs(dt), -e) * fact;
    for (int b = 0; b < d; ++b) {
      P[(size_t)b * n + i] = p;
      PI[(size_t)b * n + i] = pi;
    }
  }
}

static int build_setup(const pof_problem* pr, Setup& s, std::string& err) {
  if (!pr) { err = "null problem"; return POF_E_ARG; }
  if (pr->d < 1 || pr->q < 1 || pr->n_times < 2 || !pr->ts_host || !pr->y0_host) {
    err = "invalid problem: need d >= 1, q >= 1, n_times >= 2, ts and y0";
    return POF_E_ARG;
  }
  const int need = problem_dim(pr->kind);
  if (need < 0) { err = "unknown problem kind " + std::to_string(pr->kind); return POF_E_ARG; }
  if ((need > 0 && need != pr->d) || (need == 0 && pr->d < 4)) {
    err = "problem kind " + std::to_string(pr->kind) + " does not have dimension d=" + std::to_string(pr->d);
    return POF_E_ARG;
  }
  if (pr->n_times - 1 > (int64_t)0x3fffffff) { err = "too many time points"; return POF_E_ARG; }
  s.d = pr->d;
  s.q = pr->q;
  s.D = pr->d * (pr->q + 1);
  s.Ns = (int)(pr->n_times - 1);
  s.dt = pr->ts_host[1] - pr->ts_host[0];
  std::vector<double> A1, L1;
  iwp_1d(s.q, A1, L1);
  s.A1 = A1;
  s.F = kron_eye(s.d, A1, s.q + 1);
  s.LQ = kron_eye(s.d, L1, s.q + 1);
  nordsieck(s.d, s.q, s.dt, s.P, s.PI);
  // N1 of a zero prior: tria([F 0 | LQ])
  std::vector<double> M((size_t)s.D * 2 * s.D, 0.0);
  for (int r = 0; r < s.D; ++r)
    for (int c = 0; c < s.D; ++c) M[(size_t)r * 2 * s.D + s.D + c] = s.LQ[(size_t)r * s.D + c];
  tria_host(M, s.D, 2 * s.D, s.N1z);
  // x0 = PI * TaylorMode (P0 = 0)
  std::vector<double> der((size_t)(s.q + 1) * s.d);
  if (!taylor_derivatives(pr->kind, s.d, pr->params, pr->y0_host, s.q, der.data())) {
    err = "Taylor-mode initialisation unsupported for this (kind, d, q)";
    return POF_E_ARG;
  }
  s.x0m.assign(s.D, 0.0);
  for (int i = 0; i < s.d; ++i)
    for (int k = 0; k <= s.q; ++k) {
      const int j = i * (s.q + 1) + k;
      s.x0m[j] = s.PI[j] * der[(size_t)k * s.d + i];
    }
  s.x0L.assign((size_t)s.D * s.D, 0.0);
  return POF_OK;
}

// get_initial_trajectory (convenience.py:76-92)
// The state every point of a constant initial trajectory holds: [y0, f(y0), 0, ...]
// per dimension, times PI (initialization.py:42-56, convenience.py:84-87).
static int constant_state(const pof_problem* pr, const Setup& s, std::vector<double>& x, std::string& err) {
  const int D = s.D, d = s.d, q1 = s.q + 1;
  std::vector<double> f(d);
  if (!field_eval(pr->kind, d, pr->params, pr->y0_host, f.data())) {
    err = "field unsupported";
    return POF_E_ARG;
  }
  x.assign(D, 0.0);
  for (int i = 0; i < d; ++i) {
    x[(size_t)i * q1] = pr->y0_host[i] * s.PI[(size_t)i * q1];
    x[(size_t)i * q1 + 1] = f[i] * s.PI[(size_t)i * q1 + 1];
  }
  return POF_OK;
}

static int initial_means(const pof_problem* pr, const Setup& s, std::vector<double>& m, std::string& err) {
  const int D = s.D, d = s.d, q1 = s.q + 1;
  const size_t N = (size_t)s.Ns + 1;
  m.assign(N * D, 0.0);
  if (pr->init == POF_INIT_CONSTANT) {
    std::vector<double> x;
    const int rc = constant_state(pr, s, x, err);
    if (rc != POF_OK) return rc;
    for (size_t t = 0; t < N; ++t) std::copy(x.begin(), x.end(), m.begin() + t * D);
    return POF_OK;
  }
  if (pr->init == POF_INIT_PRIOR) {
    // x0 (not preconditioned) predicted with P(h) A PI(h), h = ts[k] (initialization.py:75-89)
    std::vector<double> der((size_t)q1 * d);
    taylor_derivatives(pr->kind, d, pr->params, pr->y0_host, s.q, der.data());
    std::vector<double> m0(D);
    for (int i = 0; i < d; ++i)
      for (int k = 0; k < q1; ++k) m0[(size_t)i * q1 + k] = der[(size_t)k * d + i];
    std::copy(m0.begin(), m0.end(), m.begin());
    // rows t in [lo, hi); each row is independent (two pow calls per derivative order),
    // so long grids are split over host threads with the same arithmetic per row
    auto rows = [&](size_t lo, size_t hi) {
      std::vector<double> Ph, PIh;
      for (size_t t = lo; t < hi; ++t) {
        nordsieck(d, s.q, pr->ts_host[t], Ph, PIh);
        for (int b = 0; b < d; ++b)
          for (int i = 0; i < q1; ++i) {
            double acc = 0.0;
            for (int j = 0; j < q1; ++j) {
              const double Fij = (Ph[(size_t)b * q1 + i] * s.A1[(size_t)i * q1 + j]) * PIh[(size_t)b * q1 + j];
              acc += Fij * m0[(size_t)b * q1 + j];
            }
            m[t * D + (size_t)b * q1 + i] = acc;
          }
      }
    };
    const size_t kRowsPerThread = 8192;
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nt = std::min(hw, (N + kRowsPerThread - 1) / kRowsPerThread);
    if (nt <= 1) {
      rows(1, N);
    } else {
      std::vector<std::thread> pool;
      const size_t chunk = (N - 1 + nt - 1) / nt;
      for (size_t k = 0; k < nt; ++k) {
        const size_t lo = 1 + k * chunk, hi = std::min(N, lo + chunk);
        if (lo < hi) pool.emplace_back(rows, lo, hi);
      }
      for (auto& th : pool) th.join();
    }
    return POF_OK;
  }
  if (pr->init == POF_INIT_COARSE) {
    err = "method=coarse runs a sequential EKS on the device (pof_coarse_means)";
    return POF_E_UNSUPPORTED;
  }
  err = "method=" + std::to_string(pr->init) + " not found";
  return POF_E_INIT;
}

}  // namespace pof

// ---------------------------------------------------------------- context
struct pof_batch;

struct pof_ctx {
  int device = 0;
  std::string err;
  std::mutex err_mu;  // pof_solve_batch runs its groups on several host threads
  // per-kernel-class timing (pof_profile_*)
  bool profiling = false;
  pof::KernelTimer timer;
  double prof_ms[pof::kNumKernelClasses] = {};
  long long prof_launches[pof::kNumKernelClasses] = {};
  long long prof_work[pof::kNumKernelClasses] = {};
  // batches kept for reuse by pof_solve / pof_ieks_step / pof_solve_batch, one per (D, d)
  std::vector<pof_batch*> cache;
  // small batches for the coarse initial trajectory (EKS on 100 points), one per (D, d)
  std::vector<pof_batch*> aux;
  // large-state engine (batched primitives), created on first use
  pof::large::Engine* large = nullptr;
  // replay unchanged iterations from a HIP graph (pof_graphs_enable)
  bool graphs = true;
  // this device's copy of the file-scope scratch array (pof_linearize_model_noalloc), or null with
  // scratch_err saying why its address could not be taken
  double* scratch = nullptr;
  std::string scratch_err;
  // scan-level widths of the throughput kernel forms per state dimension (pof_tuning_set);
  // tune_version counts changes, so that batches rebuild plans and graphs
  std::map<int, pof::Tuning> tuning;
  int tune_version = 0;
};

// A group of IVPs with one (D, d) sharing one device workspace.  IVP slots are
// admitted and retired between iterations; each owns a range of
// 2 (Ns + 1) entries of every array (first fit over the free ranges).
struct pof_batch {
  struct Slot {
    bool used = false;
    int k = 0;  // iterations done
    double nll = 0.0, obj = 0.0, ssq = 0.0;
    bool means_close = false, isnan_ = false;
    int status = POF_STATUS_CONVERGED;
    pof_opts opts{};
    pof_out* out = nullptr;
    long long ticket = 0;
    long long start = 0, len = 0;
    int last_wchol = 0;
    int scale_mode = 0;  // output factor scale: 0 ssq (calibrate) / sqrt(ssq); 1 none; 2 sqrt(ssq) if calibrate
  };

  pof_ctx* ctx = nullptr;
  int D = 0, d = 0, q = 0;
  bool affine = false;
  bool cached = false;
  // 0 parallel IEKS, 1 sequential IEKS, 2 sequential EKS single pass, 3 / 4 parallel /
  // sequential smoother alone over given filtered states (set by the first admission)
  int mode = -1;
  const pof::EngineFns* e = nullptr;
  long long entries = 0;
  int slots = 0;
  hipStream_t st = nullptr;      // stream the batch works on
  hipStream_t own_st = nullptr;  // stream created with the batch
  hipEvent_t ev = nullptr;
  hipEvent_t join_ev = nullptr;  // orders the batch's own stream with a caller's null stream
  std::vector<void*> allocs;
  pof::GroupArgs g{};
  // writable device tables (g holds the same pointers)
  double *F = nullptr, *LQ = nullptr, *N1z = nullptr, *Pd = nullptr, *x0m = nullptr, *x0L = nullptr;
  double* cvec = nullptr;  // [slots, D] the state of a constant initial trajectory
  double *params = nullptr, *scale = nullptr, *scal_dev = nullptr, *scal_host = nullptr;
  int *kind = nullptr, *Ns = nullptr, *cur = nullptr, *wchol = nullptr, *pfull = nullptr;
  long long* base = nullptr;
  double** outp = nullptr;  // [4, slots]: means, chols, full means, full chols
  int* maps_dev = nullptr;
  size_t maps_cap = 0;
  // host mirrors
  std::vector<double> hPd, hx0m, hx0L, hparams, hscale, hcvec;
  std::vector<int> hkind, hNs, hcur, hwchol;  // hcur: 0, uploaded at admission (the device flips it)
  std::vector<int> hpfull;
  std::vector<long long> hbase;
  std::vector<double*> hout;
  std::vector<int> hmaps;
  std::vector<Slot> sl;
  std::map<long long, long long> free_;  // free entry ranges: start -> length
  std::vector<int> dirty_slots;          // per-slot tables to upload before the next launch
  std::vector<std::pair<long long, std::vector<double>>> init_host;  // initial means to upload
  std::vector<std::pair<long long, const double*>> init_dev;         // ... or to copy on the device
  std::vector<long long> init_zero;                                  // ... or to clear (entries, count D each)
  std::vector<long long> init_zero_n;
  std::vector<int> init_const;  // slots whose initial means repeat their cvec row (filled on the device)
  std::vector<int> active;
  pof::LaunchPlan plan;
  bool plan_dirty = true;
  bool launched = false;
  bool need_sync = false;
  bool unsynced = false;  // launches whose host-side copies may still be pending
  long long next_ticket = 1;
  // one iteration's launches captured as a graph, valid while the plan and the
  // mode are unchanged (their kernel arguments are then identical every launch)
  hipGraphExec_t gexec = nullptr;
  int gmode = -1;
  int gstable = 0;  // launches since the plan was last rebuilt
  int tune_version = -1;  // the context's tuning the plan was built with
};

// ------------------------------------------------------ no-allocation scratch
// Device scratch of pof_linearize_model_noalloc: the Nordsieck scaling P and
// the per-state work slices of the linearisation, in a file-scope array that
// every GPU the library runs on holds from the load of the library to the end
// of the process.  Host code never passes `scratch` itself (a host-side
// stand-in): it passes the device address that hipGetSymbolAddress returns for
// the current device, looked up once per device when a context is created.
// All calls on one GPU share the array and it keeps the last call's contents;
// calls that overlap on one GPU race.
constexpr size_t kScratchDoubles = size_t(8) << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

using namespace pof;

int set_err(pof_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) {
    std::lock_guard<std::mutex> lk(c->err_mu);
    c->err = buf;
  }
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return set_err(ctx, POF_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

int n_levels_host(long long n0) {
  int L = 0;
  while (n0 >= 1) {
    ++L;
    if (n0 == 1) break;
    n0 /= 2;
  }
  return L;
}

long long entries_for(long long n_times) { return 2 * n_times; }

// A constant initial trajectory: rows 0 .. n - 1 of m (D each) = v.
__global__ void k_fill_rows(double* m, const double* v, long long n, int D) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n * D) m[e] = v[e % D];
}

// After an iteration: every active IVP's result buffer becomes its next linearisation point.
__global__ void k_flip_cur(int* cur, const int* ivp, int m) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < m) cur[ivp[a]] ^= 1;
}

// after an instrumented run: accumulate event intervals per kernel class
hipError_t collect_timing(pof_ctx* ctx, hipStream_t st) {
  if (!ctx->profiling) return hipSuccess;
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  KernelTimer& t = ctx->timer;
  for (int i = 0; i < t.n; ++i) {
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]);
    if (e != hipSuccess) return e;
    ctx->prof_ms[t.cls[i]] += ms;
    ctx->prof_launches[t.cls[i]] += 1;
    ctx->prof_work[t.cls[i]] += t.work[i];
  }
  t.n = 0;
  return hipSuccess;
}

const EngineFns* engine_or_err(pof_ctx* ctx, int D, int d) {
  const EngineFns* e = find_engine(D, d);
  if (!e) set_err(ctx, POF_E_UNSUPPORTED, "no kernel instantiation for D=%d, d=%d", D, d);
  return e;
}

bool isclose(double a, double b, double rtol, double atol) {
  if (std::isnan(a) || std::isnan(b)) return false;
  if (std::isinf(a) || std::isinf(b)) return a == b;
  return std::fabs(a - b) <= atol + rtol * std::fabs(b);
}

// ------------------------------------------------------------ batch memory
void batch_free(pof_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->scal_host) (void)hipHostFree(b->scal_host);
  if (b->maps_dev) (void)hipFree(b->maps_dev);
  if (b->ev) (void)hipEventDestroy(b->ev);
  if (b->join_ev) (void)hipEventDestroy(b->join_ev);
  if (b->own_st) (void)hipStreamDestroy(b->own_st);
  delete b;
}

template <class T>
hipError_t dalloc(pof_batch* b, size_t n, T** out) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256));
  if (e != hipSuccess) return e;
  b->allocs.push_back(p);
  *out = static_cast<T*>(p);
  return hipSuccess;
}

// Device bytes of a batch of `entries` entries and `slots` IVPs with state dimension D.
size_t batch_bytes(int D, long long entries, int slots) {
  const size_t DD = (size_t)D * D;
  const size_t per_entry = 8 * (7 * DD + 7 * (size_t)D + 4);
  const size_t per_slot = 8 * (2 * D + DD + kMaxParams + 1 + 8 + 4) + 4 * 4 + 8;
  return per_entry * (size_t)entries + per_slot * (size_t)slots + 8 * 4 * ((size_t)entries / kChunk + slots + 1);
}

int batch_new(pof_ctx* ctx, int D, int d, long long entries, int slots, bool affine, pof_batch** out) {
  *out = nullptr;
  const EngineFns* e = engine_or_err(ctx, D, d);
  if (!e) return POF_E_UNSUPPORTED;
  if (entries < 4 || slots < 1) return set_err(ctx, POF_E_ARG, "batch needs entries >= 4 and slots >= 1");
  HIP_TRY(hipSetDevice(ctx->device));
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  const size_t need = batch_bytes(D, entries, slots);
  if (need > free_b)
    return set_err(ctx, POF_E_HIP, "batch needs %zu bytes of device memory, %zu free", need, free_b);
  pof_batch* b = new pof_batch();
  b->ctx = ctx;
  b->D = D;
  b->d = d;
  b->q = D / d - 1;
  b->affine = affine;
  b->e = e;
  b->entries = entries;
  b->slots = slots;
  const size_t DD = (size_t)D * D, n = (size_t)entries, S = (size_t)slots;
  GroupArgs& g = b->g;
  hipError_t he = hipSuccess;
#define POF_A(ptr, count)                       \
  if (he == hipSuccess) he = dalloc(b, (count), &(ptr));
  double *fA, *fb, *fU, *feta, *fZ, *pm, *pL, *sg, *sE, *sD, *qm, *qL, *m0, *m1, *ts, *tn, *to, *tv, *part;
  POF_A(fA, n * DD) POF_A(fb, n * D) POF_A(fU, n * DD) POF_A(feta, n * D) POF_A(fZ, n * DD)
  POF_A(pm, n * D) POF_A(pL, n * DD)
  POF_A(sg, n * D) POF_A(sE, n * DD) POF_A(sD, n * DD) POF_A(qm, n * D) POF_A(qL, n * DD)
  POF_A(m0, n * D) POF_A(m1, n * D) POF_A(ts, n) POF_A(tn, n) POF_A(to, n) POF_A(tv, n)
  POF_A(part, 4 * (n / kChunk + S + 1))
  POF_A(b->F, DD) POF_A(b->LQ, DD) POF_A(b->N1z, DD) POF_A(b->Pd, S * D) POF_A(b->x0m, S * D)
  POF_A(b->x0L, S * DD) POF_A(b->params, S * kMaxParams) POF_A(b->scale, S) POF_A(b->scal_dev, S * 8)
  POF_A(b->kind, S) POF_A(b->Ns, S) POF_A(b->cur, S) POF_A(b->wchol, S) POF_A(b->base, S) POF_A(b->outp, 4 * S)
  POF_A(b->pfull, S) POF_A(b->cvec, S * D)
#undef POF_A
  if (he == hipSuccess) he = hipHostMalloc((void**)&b->scal_host, S * 8 * sizeof(double));
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&b->own_st, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming);
  if (he != hipSuccess) {
    batch_free(b);
    return set_err(ctx, POF_E_HIP, "batch allocation failed: %s", hipGetErrorString(he));
  }
  b->st = b->own_st;
  g.F = b->F;
  g.LQ = b->LQ;
  g.N1z = b->N1z;
  g.Pd = b->Pd;
  g.x0m = b->x0m;
  g.x0L = b->x0L;
  g.params = b->params;
  g.kind = b->kind;
  g.Ns = b->Ns;
  g.base = b->base;
  g.cur = b->cur;
  g.wchol = b->wchol;
  g.means0 = m0;
  g.means1 = m1;
  g.fel = FiltArrays{fA, fb, fU, feta, fZ};
  g.fpre = StateArrays{pm, pL};
  g.sel = SmoArrays{sg, sE, sD};
  g.spre = StateArrays{qm, qL};
  g.t_ssq = ts;
  g.t_nll = tn;
  g.t_obj = to;
  g.t_viol = tv;
  g.part = part;
  g.out_means = b->outp;
  g.out_chols = b->outp + S;
  g.out_full_means = b->outp + 2 * S;
  g.out_full_chols = b->outp + 3 * S;
  g.out_scale = b->scale;
  g.out_pfull = b->pfull;
  g.Hx = nullptr;
  g.cx = nullptr;
  g.q = b->q;
  g.affine = affine ? 1 : 0;
  g.lq_lower = 0;
  b->hPd.assign(S * D, 0.0);
  b->hx0m.assign(S * D, 0.0);
  b->hcvec.assign(S * D, 0.0);
  b->hx0L.assign(S * DD, 0.0);
  b->hparams.assign(S * kMaxParams, 0.0);
  b->hscale.assign(S, 1.0);
  b->hkind.assign(S, 0);
  b->hNs.assign(S, 1);
  b->hcur.assign(S, 0);
  b->hwchol.assign(S, 0);
  b->hpfull.assign(S, 0);
  b->hbase.assign(S, 0);
  b->hout.assign(4 * S, nullptr);
  b->sl.assign(S, pof_batch::Slot{});
  b->free_[0] = entries;
  // the transition model of the group (pof/transitions.py:37-50) and N1 of a zero prior
  if (!affine) {
    std::vector<double> A1, L1;
    iwp_1d(b->q, A1, L1);
    const std::vector<double> Fm = kron_eye(d, A1, b->q + 1), LQm = kron_eye(d, L1, b->q + 1);
    std::vector<double> M(DD * 2, 0.0), N1;
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) M[(size_t)r * 2 * D + D + c] = LQm[(size_t)r * D + c];
    tria_host(M, D, 2 * D, N1);
    he = hipMemcpy(b->F, Fm.data(), DD * 8, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(b->LQ, LQm.data(), DD * 8, hipMemcpyHostToDevice);
    b->g.lq_lower = strictly_upper_zero(LQm.data(), D) ? 1 : 0;
    if (he == hipSuccess) he = hipMemcpy(b->N1z, N1.data(), DD * 8, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
      batch_free(b);
      return set_err(ctx, POF_E_HIP, "batch set-up copy failed: %s", hipGetErrorString(he));
    }
  }
  *out = b;
  return POF_OK;
}

// first-fit range allocator over the batch's entries
bool range_alloc(pof_batch* b, long long len, long long* start) {
  for (auto it = b->free_.begin(); it != b->free_.end(); ++it) {
    if (it->second >= len) {
      *start = it->first;
      const long long rest = it->second - len, s2 = it->first + len;
      b->free_.erase(it);
      if (rest > 0) b->free_[s2] = rest;
      return true;
    }
  }
  return false;
}

void range_free(pof_batch* b, long long start, long long len) {
  auto it = b->free_.emplace(start, len).first;
  auto nx = std::next(it);
  if (nx != b->free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_.erase(nx);
  }
  if (it != b->free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_.erase(it);
    }
  }
}

// ---------------------------------------------------------- launch plans
// Append the launch tables of IVP slots `ids` to h; offsets in p are relative
// to `at` (the tables' position in the device buffer is set after upload).
int plan_tables(pof_batch* b, const std::vector<int>& ids, std::vector<int>& h, LaunchPlan& p) {
  const int at = (int)h.size();
  const int m = (int)ids.size();
  h.insert(h.end(), ids.begin(), ids.end());
  long long maxNs = 1;
  for (int id : ids) maxNs = std::max<long long>(maxNs, b->hNs[id]);
  p = LaunchPlan{};
  p.m = m;
  p.fl = n_levels_host(maxNs);
  p.sl = n_levels_host(maxNs + 1);
  if (p.sl > kMaxLevels) return POF_E_ARG;
  bool overflow = false;
  auto table = [&](auto count, int& off, int& total, int* nivps) {
    off = (int)h.size() - at;
    long long acc = 0;
    int ni = 0;
    h.push_back(0);
    for (int id : ids) {
      const long long c = count((long long)b->hNs[id]);
      acc += c;
      if (c > 0) ++ni;
      h.push_back((int)acc);
    }
    if (acc > 0x7fffffffLL) overflow = true;
    total = (int)acc;
    if (nivps) *nivps = ni;
  };
  table([](long long N) { return N; }, p.leaf_off, p.leaf_total, nullptr);
  table([](long long) { return 1LL; }, p.one_off, p.one_total, nullptr);
  table([](long long N) { return N - 1; }, p.leaf1_off, p.leaf1_total, nullptr);
  table([](long long N) { return N + 1; }, p.pts_off, p.pts_total, nullptr);
  table([](long long N) { return (N + 1 + kChunk - 1) / kChunk; }, p.chunk_off, p.chunk_total, nullptr);
  for (int l = 0; l + 1 < p.fl; ++l)
    table([l](long long N) { return N >> (l + 1); }, p.fup_off[l], p.fup_total[l], nullptr);
  for (int l = 0; l < p.fl; ++l)
    table([l](long long N) { return ((N >> l) + 1) / 2; }, p.fdn_off[l], p.fdn_total[l], &p.fdn_ivps[l]);
  for (int l = 0; l + 1 < p.sl; ++l)
    table([l](long long N) { return (N + 1) >> (l + 1); }, p.sup_off[l], p.sup_total[l], nullptr);
  for (int l = 0; l < p.sl; ++l)
    table([l](long long N) { return (((N + 1) >> l) + 1) / 2; }, p.sdn_off[l], p.sdn_total[l], &p.sdn_ivps[l]);
  return overflow ? POF_E_ARG : POF_OK;
}

// Admit one IVP with its set-up done on the host.  `m_host` (or `m_dev`, or
// neither: zeros) holds the initial means [Ns + 1, D].
int settle(pof_batch* b) {
  pof_ctx* ctx = b->ctx;
  if (b->unsynced) HIP_TRY(hipStreamSynchronize(b->st));
  b->unsynced = false;
  return POF_OK;
}

// m_const (D values), if given, is the state of a constant initial trajectory,
// filled on the device instead of uploading Ns + 1 copies.
int admit_raw(pof_batch* b, int Nsteps, const double* Pd, const double* x0m, const double* x0L, int kind,
              const double* params, std::vector<double>* m_host, const double* m_dev, const pof_opts* opts,
              pof_out* out, int scale_mode, int mode, bool pfull, long long* ticket,
              const double* m_const = nullptr) {
  pof_ctx* ctx = b->ctx;
  if (b->launched) return set_err(ctx, POF_E_ARG, "admit between pof_batch_launch and pof_batch_finish");
  bool busy = false;
  for (const auto& sl : b->sl) busy = busy || sl.used;
  if (busy && b->mode != mode)
    return set_err(ctx, POF_E_ARG, "the batch runs mode %d solves (0 parallel, 1 sequential, 2 EKS), got %d", b->mode,
                   mode);
  // the host tables change below: no earlier copy from them may still be pending
  if (settle(b) != POF_OK) return POF_E_HIP;
  int j = -1;
  for (int i = 0; i < b->slots; ++i)
    if (!b->sl[i].used) {
      j = i;
      break;
    }
  const long long len = entries_for((long long)Nsteps + 1);
  long long start = 0;
  if (j < 0) return set_err(ctx, POF_E_FULL, "no free IVP slot");
  if (!range_alloc(b, len, &start)) return set_err(ctx, POF_E_FULL, "no free range of %lld entries", len);
  const int D = b->D;
  const size_t DD = (size_t)D * D, S = (size_t)b->slots;
  pof_batch::Slot& s = b->sl[j];
  s = pof_batch::Slot{};
  s.used = true;
  s.opts = *opts;
  s.out = out;
  s.ticket = b->next_ticket++;
  s.start = start;
  s.len = len;
  s.scale_mode = scale_mode;
  b->mode = mode;
  std::copy(Pd, Pd + D, &b->hPd[(size_t)j * D]);
  std::copy(x0m, x0m + D, &b->hx0m[(size_t)j * D]);
  std::copy(x0L, x0L + DD, &b->hx0L[(size_t)j * DD]);
  for (int i = 0; i < kMaxParams; ++i) b->hparams[(size_t)j * kMaxParams + i] = params ? params[i] : 0.0;
  b->hkind[j] = kind;
  b->hNs[j] = Nsteps;
  b->hbase[j] = start;
  b->hcur[j] = 0;
  b->hwchol[j] = 0;
  b->hpfull[j] = pfull ? 1 : 0;
  b->hout[0 * S + j] = out->means_dev;
  b->hout[1 * S + j] = out->chols_dev;
  b->hout[2 * S + j] = out->full_means_dev;
  b->hout[3 * S + j] = out->full_chols_dev;
  b->dirty_slots.push_back(j);
  if (m_const) {
    std::copy(m_const, m_const + D, &b->hcvec[(size_t)j * D]);
    b->init_const.push_back(j);
  } else if (m_host) {
    b->init_host.emplace_back(start, std::move(*m_host));
  } else if (m_dev) {
    b->init_dev.emplace_back(start, m_dev);
  } else {
    b->init_zero.push_back(start);
    b->init_zero_n.push_back((long long)Nsteps + 1);
  }
  b->plan_dirty = true;
  if (ticket) *ticket = s.ticket;
  return POF_OK;
}

// Launches of one iteration of b's plan in b's mode, on b->st.
hipError_t launch_iteration(pof_batch* b, KernelTimer* tm) {
  hipError_t e;
  if (b->mode >= 3)
    e = b->e->smooth_only(b->g, b->plan, b->scal_dev, b->mode == 4 ? 1 : 0, b->st, tm);
  else if (b->mode >= 1)
    e = b->e->seq_iteration(b->g, b->plan, b->scal_dev, b->mode == 2 ? 1 : 0, b->st, tm);
  else
    e = b->e->iteration(b->g, b->plan, b->scal_dev, b->st, tm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_flip_cur, dim3((b->plan.m + 255) / 256), dim3(256), 0, b->st, b->cur, b->plan.dev, b->plan.m);
  return hipGetLastError();
}

// The caller has drained b->st (settle): no replay of the graph is in flight.
void drop_graph(pof_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gmode = -1;
}

// Capture one iteration of b's plan and mode into b->gexec.  No capture (eager
// launches) on the null stream or on a stream the caller is capturing already.
hipError_t capture_iteration(pof_batch* b) {
  if (!b->st) return hipSuccess;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(b->st, &cs);
  if (e != hipSuccess || cs != hipStreamCaptureStatusNone) return e;
  e = hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t el = launch_iteration(b, nullptr);
  hipGraph_t gr = nullptr;
  e = hipStreamEndCapture(b->st, &gr);  // ends the capture whatever the launches returned
  if (e == hipSuccess) e = el;
  hipGraphExec_t ge = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
  if (gr) (void)hipGraphDestroy(gr);
  if (e != hipSuccess) return e;
  b->gexec = ge;
  b->gmode = b->mode;
  return hipSuccess;
}

// Iterations launched eagerly after a plan change before the next is captured:
// a plan that changes at every launch (IVPs retiring) is never captured.
constexpr int kGraphAfter = 2;

int batch_launch(pof_batch* b) {
  pof_ctx* ctx = b->ctx;
  if (b->launched) return set_err(ctx, POF_E_ARG, "pof_batch_launch called twice without pof_batch_finish");
  HIP_TRY(hipSetDevice(ctx->device));
  if (b->tune_version != ctx->tune_version) b->plan_dirty = true;  // other kernel forms: new plan and graph
  if (b->plan_dirty) {
    if (settle(b) != POF_OK) return POF_E_HIP;
    drop_graph(b);
    b->gstable = 0;
    b->active.clear();
    for (int i = 0; i < b->slots; ++i)
      if (b->sl[i].used) b->active.push_back(i);
    if (b->active.empty()) {
      b->plan_dirty = false;
      b->plan.m = 0;
      return POF_OK;
    }
    b->hmaps.clear();
    int rc = plan_tables(b, b->active, b->hmaps, b->plan);
    if (rc != POF_OK) return set_err(ctx, rc, "launch tables overflow");
    const auto tu = ctx->tuning.find(b->D);
    if (tu != ctx->tuning.end()) b->plan.tune = tu->second;
    b->tune_version = ctx->tune_version;
    if (b->hmaps.size() > b->maps_cap) {
      HIP_TRY(hipStreamSynchronize(b->st));
      if (b->maps_dev) HIP_TRY(hipFree(b->maps_dev));
      b->maps_dev = nullptr;
      b->maps_cap = 0;
      const size_t cap = std::max<size_t>(b->hmaps.size() * 2, 4096);
      HIP_TRY(hipMalloc((void**)&b->maps_dev, cap * sizeof(int)));
      b->maps_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(b->maps_dev, b->hmaps.data(), b->hmaps.size() * sizeof(int), hipMemcpyHostToDevice,
                           b->st));
    b->plan.dev = b->maps_dev;
    b->plan_dirty = false;
  }
  if (b->plan.m == 0) return POF_OK;
  // newly admitted IVPs: per-slot tables and initial means
  if (!b->dirty_slots.empty() || !b->init_host.empty() || !b->init_dev.empty() || !b->init_zero.empty() ||
      !b->init_const.empty()) {
    const int D = b->D;
    const size_t DD = (size_t)D * D, S = (size_t)b->slots;
    if (!b->dirty_slots.empty()) {
      // one copy per table over the slot range of the new IVPs: the host
      // tables mirror the device ones for every slot (cur through the flips
      // of each launch), so the slots in between are rewritten unchanged
      int lo = b->slots, hi = -1;
      for (int j : b->dirty_slots) {
        lo = std::min(lo, j);
        hi = std::max(hi, j);
      }
      const size_t n = (size_t)(hi - lo + 1), l = (size_t)lo;
      HIP_TRY(hipMemcpyAsync(b->Pd + l * D, &b->hPd[l * D], n * D * 8, hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->x0m + l * D, &b->hx0m[l * D], n * D * 8, hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->x0L + l * DD, &b->hx0L[l * DD], n * DD * 8, hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->params + l * kMaxParams, &b->hparams[l * kMaxParams], n * kMaxParams * 8,
                             hipMemcpyHostToDevice, b->st));
      for (int t = 0; t < 4; ++t)
        HIP_TRY(hipMemcpyAsync(b->outp + t * S + l, &b->hout[t * S + l], n * sizeof(double*), hipMemcpyHostToDevice,
                               b->st));
      HIP_TRY(hipMemcpyAsync(b->cur + l, &b->hcur[l], n * sizeof(int), hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->wchol + l, &b->hwchol[l], n * sizeof(int), hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->pfull + l, &b->hpfull[l], n * sizeof(int), hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->cvec + l * D, &b->hcvec[l * D], n * D * 8, hipMemcpyHostToDevice, b->st));
    }
    b->dirty_slots.clear();
    HIP_TRY(hipMemcpyAsync(b->kind, b->hkind.data(), S * sizeof(int), hipMemcpyHostToDevice, b->st));
    HIP_TRY(hipMemcpyAsync(b->Ns, b->hNs.data(), S * sizeof(int), hipMemcpyHostToDevice, b->st));
    HIP_TRY(hipMemcpyAsync(b->base, b->hbase.data(), S * sizeof(long long), hipMemcpyHostToDevice, b->st));
    for (auto& pr : b->init_host)
      HIP_TRY(hipMemcpyAsync(b->g.means0 + (size_t)pr.first * D, pr.second.data(), pr.second.size() * 8,
                             hipMemcpyHostToDevice, b->st));
    for (auto& pr : b->init_dev) {
      long long n = 0;
      for (int j : b->active)
        if (b->hbase[j] == pr.first) n = (long long)b->hNs[j] + 1;
      HIP_TRY(hipMemcpyAsync(b->g.means0 + (size_t)pr.first * D, pr.second, (size_t)n * D * 8,
                             hipMemcpyDeviceToDevice, b->st));
    }
    for (int j : b->init_const) {
      const long long rows = (long long)b->hNs[j] + 1;
      const long long e = rows * D;
      hipLaunchKernelGGL(k_fill_rows, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, b->st,
                         b->g.means0 + (size_t)b->hbase[j] * D, (const double*)(b->cvec + (size_t)j * D), rows, D);
      HIP_TRY(hipGetLastError());
    }
    for (size_t i = 0; i < b->init_zero.size(); ++i)
      HIP_TRY(hipMemsetAsync(b->g.means0 + (size_t)b->init_zero[i] * D, 0, (size_t)b->init_zero_n[i] * D * 8,
                             b->st));
    // host sources must outlive the copies: wait here (admission only)
    HIP_TRY(hipStreamSynchronize(b->st));
    b->init_host.clear();
    b->init_dev.clear();
    b->init_zero.clear();
    b->init_zero_n.clear();
    b->init_const.clear();
  }
  // factor flag of this iteration: the smoothed factors are needed only from
  // an IVP's last iteration, known in advance only with fixed_iters
  b->need_sync = false;
  bool wchange = false;
  for (int j : b->active) {
    const pof_batch::Slot& s = b->sl[j];
    const bool fixed = s.opts.fixed_iters > 0;
    const int w = (fixed && s.k + 1 == s.opts.fixed_iters) ? 1 : 0;
    wchange = wchange || (w != b->hwchol[j]);
    if (!fixed || s.k + 1 >= s.opts.fixed_iters) b->need_sync = true;
  }
  if (wchange) {
    if (settle(b) != POF_OK) return POF_E_HIP;
    for (int j : b->active) {
      const pof_batch::Slot& s = b->sl[j];
      b->hwchol[j] = (s.opts.fixed_iters > 0 && s.k + 1 == s.opts.fixed_iters) ? 1 : 0;
    }
    HIP_TRY(hipMemcpyAsync(b->wchol, b->hwchol.data(), (size_t)b->slots * sizeof(int), hipMemcpyHostToDevice, b->st));
  }
  KernelTimer* tm = ctx->profiling ? &ctx->timer : nullptr;
  // an instrumented run times every launch from the host: no replay then
  const bool replay = ctx->graphs && !tm;
  if (b->gexec && b->gmode != b->mode) {
    // a replay of the old graph may still be in flight: drain before destroying it
    if (settle(b) != POF_OK) return POF_E_HIP;
    drop_graph(b);
  }
  if (replay && !b->gexec && b->gstable >= kGraphAfter) HIP_TRY(capture_iteration(b));
  if (replay && b->gexec) {
    HIP_TRY(hipGraphLaunch(b->gexec, b->st));
  } else {
    HIP_TRY(launch_iteration(b, tm));
  }
  ++b->gstable;
  for (int j : b->active) b->hcur[j] ^= 1;  // host mirror of the flip
  HIP_TRY(collect_timing(ctx, b->st));
  if (b->need_sync) {
    HIP_TRY(hipMemcpyAsync(b->scal_host, b->scal_dev, (size_t)b->plan.m * 8 * sizeof(double),
                           hipMemcpyDeviceToHost, b->st));
    HIP_TRY(hipEventRecord(b->ev, b->st));
  }
  b->launched = true;
  b->unsynced = true;
  return POF_OK;
}

// Wait for the launched iteration, apply the convergence test of
// pof/solver.py:36-57 per IVP, and retire finished IVPs (factor pass, output
// projection, info).  Writes up to max_done finished tickets; *n_done = count.
int batch_finish(pof_batch* b, int max_done, long long* done, int* n_done) {
  pof_ctx* ctx = b->ctx;
  if (n_done) *n_done = 0;
  if (!b->launched) return POF_OK;
  b->launched = false;
  HIP_TRY(hipSetDevice(ctx->device));
  if (!b->need_sync) {
    for (int j : b->active) b->sl[j].k += 1;  // no IVP can finish: no host round trip
    return POF_OK;
  }
  HIP_TRY(hipEventSynchronize(b->ev));
  b->unsynced = false;
  std::vector<int> fin;
  for (size_t a = 0; a < b->active.size(); ++a) {
    const int j = b->active[a];
    pof_batch::Slot& s = b->sl[j];
    const double* sc = b->scal_host + a * 8;
    s.k += 1;
    s.last_wchol = b->hwchol[j];
    const double obj_old = s.obj;
    s.nll = sc[0];
    s.obj = sc[1];
    s.ssq = sc[2];
    s.means_close = sc[3] != 0.0;
    s.isnan_ = std::isnan(s.obj) || std::isnan(s.nll);
    bool finished = false;
    if (s.opts.fixed_iters > 0) {
      if (s.k >= s.opts.fixed_iters) {
        s.status = POF_STATUS_FIXED;
        finished = true;
      }
    } else {
      const bool obj_conv = isclose(obj_old, s.obj, 1e-6, 1e-9);
      const bool conv = s.isnan_ || obj_conv || s.means_close;
      if (!(s.k <= s.opts.maxiters)) {
        s.status = POF_STATUS_MAXITER;
        finished = true;
      } else if (conv) {
        s.status = s.isnan_ ? POF_STATUS_NAN : POF_STATUS_CONVERGED;
        finished = true;
      }
    }
    if (finished) fin.push_back(j);
  }
  if (fin.empty()) return POF_OK;
  // factor pass for the IVPs whose last iteration carried means only, then the output
  std::vector<int> need;
  for (int j : fin)
    if (!b->sl[j].last_wchol) need.push_back(j);
  const size_t S = (size_t)b->slots;
  for (int j : fin) {
    const pof_batch::Slot& s = b->sl[j];
    b->hwchol[j] = 1;
    // double calibration (step.py:42-44 inside the loop + solver.py:66-68); the
    // linear filter/smoother API returns raw factors; the EKS calibrates once
    // (solver.py:86-89)
    b->hscale[j] = (s.scale_mode == 1)   ? 1.0
                   : (s.scale_mode == 2) ? (s.opts.calibrate ? std::sqrt(s.ssq) : 1.0)
                                         : (s.opts.calibrate ? s.ssq : std::sqrt(s.ssq));
  }
  b->hmaps.clear();
  LaunchPlan pc, po;
  int rc = POF_OK;
  const int start_c = 0;
  if (!need.empty()) rc = plan_tables(b, need, b->hmaps, pc);
  const int start_o = (int)b->hmaps.size();
  if (rc == POF_OK) rc = plan_tables(b, fin, b->hmaps, po);
  if (rc != POF_OK) return set_err(ctx, rc, "launch tables overflow");
  if (b->hmaps.size() > b->maps_cap) {
    HIP_TRY(hipStreamSynchronize(b->st));
    if (b->maps_dev) HIP_TRY(hipFree(b->maps_dev));
    b->maps_dev = nullptr;
    b->maps_cap = 0;
    const size_t cap = std::max<size_t>(b->hmaps.size() * 2, 4096);
    HIP_TRY(hipMalloc((void**)&b->maps_dev, cap * sizeof(int)));
    b->maps_cap = cap;
  }
  HIP_TRY(hipMemcpyAsync(b->maps_dev, b->hmaps.data(), b->hmaps.size() * sizeof(int), hipMemcpyHostToDevice,
                         b->st));
  pc.dev = b->maps_dev + start_c;
  po.dev = b->maps_dev + start_o;
  HIP_TRY(hipMemcpyAsync(b->wchol, b->hwchol.data(), S * sizeof(int), hipMemcpyHostToDevice, b->st));
  HIP_TRY(hipMemcpyAsync(b->scale, b->hscale.data(), S * sizeof(double), hipMemcpyHostToDevice, b->st));
  KernelTimer* tm = ctx->profiling ? &ctx->timer : nullptr;
  if (!need.empty()) {
    if (b->mode == 1 || b->mode == 2 || b->mode == 4) {
      HIP_TRY(b->e->seq_smooth_chol(b->g, pc, b->st, tm));
    } else {
      HIP_TRY(b->e->smooth_chol(b->g, pc, b->st, tm));
    }
    HIP_TRY(collect_timing(ctx, b->st));
  }
  HIP_TRY(b->e->output(b->g, po, b->st));
  HIP_TRY(hipStreamSynchronize(b->st));
  int nd = 0;
  for (int j : fin) {
    pof_batch::Slot& s = b->sl[j];
    if (s.out) {
      s.out->iterations = (double)s.k;
      s.out->nll = s.nll;
      s.out->obj = s.obj;
      s.out->sigma_squared = s.ssq;
      s.out->calibrated = s.opts.calibrate ? 1 : 0;
      s.out->status = s.status;
    }
    if (done && nd < max_done) done[nd] = s.ticket;
    ++nd;
    range_free(b, s.start, s.len);
    s.used = false;
    b->hwchol[j] = 0;
  }
  b->plan_dirty = true;
  if (n_done) *n_done = nd;
  return POF_OK;
}

void batch_abort(pof_batch* b);

// The context's reusable batch for (D, d) with room for `entries` entries and `slots` IVPs.
// A cached batch serves one call at a time and is empty between calls.
int cached_batch(pof_ctx* ctx, int D, int d, long long entries, int slots, pof_batch** out) {
  for (size_t i = 0; i < ctx->cache.size(); ++i) {
    pof_batch* b = ctx->cache[i];
    if (b->D != D || b->d != d) continue;
    if (b->entries >= entries && b->slots >= slots) {
      for (const auto& sl : b->sl)
        if (sl.used) {
          batch_abort(b);
          break;
        }
      *out = b;
      return POF_OK;
    }
    entries = std::max(entries, b->entries);
    slots = std::max(slots, b->slots);
    batch_free(b);
    ctx->cache.erase(ctx->cache.begin() + i);
    break;
  }
  int rc = batch_new(ctx, D, d, entries, slots, false, out);
  if (rc != POF_OK) return rc;
  (*out)->cached = true;
  ctx->cache.push_back(*out);
  return POF_OK;
}

// Run a batch until every admitted IVP has finished.
int run_to_completion(pof_batch* b) {
  for (;;) {
    bool any = false;
    for (int i = 0; i < b->slots; ++i) any = any || b->sl[i].used;
    if (!any) return POF_OK;
    int rc = batch_launch(b);
    if (rc != POF_OK) return rc;
    rc = batch_finish(b, 0, nullptr, nullptr);
    if (rc != POF_OK) return rc;
  }
}

// Make `b` work in the order of the caller's stream for one call.  A non-null
// stream becomes the batch's stream.  The null stream (the default stream, which
// torch uses unless told otherwise) is not captured into graphs and does not
// order the batch's own non-blocking stream by itself, so the batch keeps its
// own stream and is joined to the caller's with events: the work enqueued here
// starts after everything the caller enqueued before the call, and the
// caller's later work starts after the batch's.
struct StreamScope {
  pof_batch* b;
  hipStream_t saved;
  hipStream_t caller;
  hipError_t status = hipSuccess;  // of the entry join; check before enqueuing work
  StreamScope(pof_batch* b_, void* stream) : b(b_), saved(b_->st), caller((hipStream_t)stream) {
    if (caller) {
      b->st = caller;
      return;
    }
    status = hipEventRecord(b->join_ev, nullptr);
    if (status == hipSuccess) status = hipStreamWaitEvent(b->st, b->join_ev, 0);
  }
  ~StreamScope() {
    if (!caller && hipEventRecord(b->join_ev, b->st) == hipSuccess)
      (void)hipStreamWaitEvent(nullptr, b->join_ev, 0);
    b->st = saved;
  }
};

// Drop every IVP of b after a failed call (the cached batches of pof_solve,
// pof_eks_solve, pof_ieks_step and pof_solve_batch serve one call at a time):
// wait for its stream, errors ignored, then free every slot and entry range,
// so that no later call runs a stale IVP or writes through its dangling
// output pointers.  The per-slot device tables are rewritten at the next
// admission of each slot.
void batch_abort(pof_batch* b) {
  (void)hipStreamSynchronize(b->st);
  (void)hipGetLastError();
  drop_graph(b);
  b->gstable = 0;
  b->launched = false;
  b->unsynced = false;
  b->need_sync = false;
  for (auto& s : b->sl) s = pof_batch::Slot{};
  std::fill(b->hwchol.begin(), b->hwchol.end(), 0);
  b->free_.clear();
  b->free_[0] = b->entries;
  b->dirty_slots.clear();
  b->init_host.clear();
  b->init_dev.clear();
  b->init_zero.clear();
  b->init_zero_n.clear();
  b->init_const.clear();
  b->active.clear();
  b->plan.m = 0;
  b->plan_dirty = true;
}

// Single-pass sequential EKS of `prob` on batch `b` (pof/solver.py:76-96):
// one sequential filter (EKF linearised at the predicted means) + smoother
// iteration, factors scaled by sqrt(ssq) when calibrate, outputs E0 (m, L)
// and P (m, L).  Runs to completion.
int eks_run(pof_batch* b, const pof_problem* prob, const Setup& s, int calibrate, pof_out* out) {
  pof_opts o{};
  o.maxiters = 1;
  o.fixed_iters = 1;
  o.calibrate = calibrate;
  o.sequential = 1;
  // the EKF does not read the initial means; zeros feed the convergence partials
  int rc = admit_raw(b, s.Ns, s.P.data(), s.x0m.data(), s.x0L.data(), prob->kind, prob->params, nullptr, nullptr, &o,
                     out, 2, 2, true, nullptr);
  if (rc != POF_OK) return rc;
  rc = run_to_completion(b);
  if (rc != POF_OK) {
    batch_abort(b);
    return rc;
  }
  out->iterations = 1.0;
  return rc;
}

// Coarse initial trajectory (initialization.py:112-129 with convenience.py:80-83):
// EKS on N_coarse = 100 equispaced points of [ts[0], ts[-1]], full states
// P (m, L) looked up at idx = floor(ts[k] / coarse_dt) (a negative index counts
// from the end, out-of-range indices clamp, as a JAX gather), then times PI of
// the fine grid.  Means only: the solve reads nothing else of the initial
// trajectory.  Runs on the context's own small batch, so `prob` may be
// admitted to a batch with IVPs in flight.
int coarse_means(pof_ctx* ctx, const pof_problem* prob, const Setup& fine, std::vector<double>& m) {
  constexpr int kCoarse = 100;
  const int D = fine.D;
  pof_batch* b = nullptr;
  for (pof_batch* x : ctx->aux)
    if (x->D == D && x->d == fine.d) b = x;
  if (!b) {
    int rc = batch_new(ctx, D, fine.d, entries_for(kCoarse), 1, false, &b);
    if (rc != POF_OK) return rc;
    ctx->aux.push_back(b);
  }
  const double t0 = prob->ts_host[0], t1 = prob->ts_host[prob->n_times - 1];
  std::vector<double> cts(kCoarse);
  // jnp.linspace (endpoint): t0 (1 - k/div) + t1 k/div for k < div = N - 1, then t1
  const int div = kCoarse - 1;
  for (int k = 0; k < div; ++k) {
    const double st = (double)k / (double)div;
    const double w0 = t0 * (1.0 - st);  // two rounded products, then the sum (no fused multiply-add)
    const double w1 = t1 * st;
    cts[k] = w0 + w1;
  }
  cts[div] = t1;
  pof_problem cp = *prob;
  cp.init = POF_INIT_CONSTANT;
  cp.n_times = kCoarse;
  cp.ts_host = cts.data();
  Setup cs;
  int rc = build_setup(&cp, cs, ctx->err);
  if (rc != POF_OK) return rc;
  double* full = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMalloc((void**)&full, (size_t)kCoarse * D * sizeof(double)));
  pof_out po{};
  po.full_means_dev = full;
  rc = eks_run(b, &cp, cs, 1, &po);
  std::vector<double> cm((size_t)kCoarse * D);
  hipError_t he = hipSuccess;
  if (rc == POF_OK) he = hipMemcpy(cm.data(), full, cm.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(full);
  if (rc != POF_OK) return rc;
  if (he != hipSuccess) return set_err(ctx, POF_E_HIP, "coarse means copy: %s", hipGetErrorString(he));
  const double cdt = cts[1] - cts[0];
  const size_t N = (size_t)fine.Ns + 1;
  m.assign(N * D, 0.0);
  for (size_t k = 0; k < N; ++k) {
    long long idx = (long long)std::floor(prob->ts_host[k] / cdt);
    if (idx < 0) idx += kCoarse;
    idx = std::min<long long>(std::max<long long>(idx, 0), kCoarse - 1);
    for (int i = 0; i < D; ++i) m[k * D + i] = fine.PI[i] * cm[(size_t)idx * D + i];
  }
  return POF_OK;
}

int admit_problem(pof_batch* b, const pof_problem* prob, const pof_opts* opts, pof_out* out, long long* ticket) {
  pof_ctx* ctx = b->ctx;
  if (opts->products != POF_PRODUCTS_F64)
    return set_err(ctx, POF_E_UNSUPPORTED, "FP32 products exist in the large-state engine only (pof_solve_large)");
  Setup s;
  int rc = build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (s.D != b->D || s.d != b->d)
    return set_err(ctx, POF_E_ARG, "problem (D=%d, d=%d) does not match the batch", s.D, s.d);
  std::vector<double> m0;
  if (prob->init == POF_INIT_CONSTANT) {
    rc = constant_state(prob, s, m0, ctx->err);
    if (rc != POF_OK) return rc;
    return admit_raw(b, s.Ns, s.P.data(), s.x0m.data(), s.x0L.data(), prob->kind, prob->params, nullptr, nullptr,
                     opts, out, 0, opts->sequential ? 1 : 0, false, ticket, m0.data());
  }
  rc = (prob->init == POF_INIT_COARSE) ? coarse_means(ctx, prob, s, m0) : initial_means(prob, s, m0, ctx->err);
  if (rc != POF_OK) return rc;
  return admit_raw(b, s.Ns, s.P.data(), s.x0m.data(), s.x0L.data(), prob->kind, prob->params, &m0, nullptr, opts, out,
                   0, opts->sequential ? 1 : 0, false, ticket);
}

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;

// The device address of `scratch` on `device` (which must be current).
hipError_t scratch_address(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<double*>(p);
  *out = static_cast<double*>(p);
  return hipSuccess;
}

// Doubles of device buffer that pof_linearize_model(_noalloc) uses for n states:
// P (D, padded to 32 doubles), then n work slices for the uncertain / unscented methods.
size_t lin_buffer_doubles(int D, int d, int method, long long n) {
  const bool needs_L = method == POF_LIN_UNCERTAIN || method == POF_LIN_UNSCENTED;
  const size_t head = ((size_t)D + 31) / 32 * 32;
  return head + (needs_L ? (size_t)n * (size_t)lin_work_host(D, d) : 0);
}

}  // namespace

extern "C" {

int pof_ctx_create(int device, pof_ctx** out) {
  if (!out) return POF_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return POF_E_HIP;
  if (device < 0 || device >= n) return POF_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return POF_E_HIP;
  pof_ctx* c = new pof_ctx();
  c->device = device;
  const hipError_t se = scratch_address(device, &c->scratch);
  if (se != hipSuccess) {
    c->scratch = nullptr;
    c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
  }
  *out = c;
  return POF_OK;
}

void pof_ctx_destroy(pof_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (pof_batch* b : ctx->cache) batch_free(b);
  ctx->cache.clear();
  for (pof_batch* b : ctx->aux) batch_free(b);
  ctx->aux.clear();
  pof::large::engine_destroy(ctx->large);
  ctx->large = nullptr;
  if (ctx->timer.ready)
    for (int i = 0; i <= pof::kMaxTimed; ++i) (void)hipEventDestroy(ctx->timer.ev[i]);
  delete ctx;
}

const char* pof_last_error(const pof_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pof_supported(int D, int d) { return pof::find_engine(D, d) != nullptr ? 1 : 0; }

int pof_profile_enable(pof_ctx* ctx, int on) {
  if (!ctx) return POF_E_ARG;
  if (on && !ctx->timer.ready) {
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i <= pof::kMaxTimed; ++i) {
      hipError_t e = hipEventCreate(&ctx->timer.ev[i]);
      if (e != hipSuccess) return set_err(ctx, POF_E_HIP, "hipEventCreate: %s", hipGetErrorString(e));
    }
    ctx->timer.ready = true;
  }
  ctx->profiling = on != 0;
  for (int i = 0; i < pof::kNumKernelClasses; ++i) {
    ctx->prof_ms[i] = 0.0;
    ctx->prof_launches[i] = 0;
    ctx->prof_work[i] = 0;
  }
  return POF_OK;
}

int pof_graphs_enable(pof_ctx* ctx, int on) {
  if (!ctx) return POF_E_ARG;
  ctx->graphs = on != 0;
  return POF_OK;
}

int pof_tuning_set(pof_ctx* ctx, int D, const int32_t* widths) {
  if (!ctx || D < 1 || !widths) return set_err(ctx, POF_E_ARG, "pof_tuning_set: null argument or D < 1");
  for (int i = 0; i < 4; ++i)
    if (widths[i] < 1) return set_err(ctx, POF_E_ARG, "pof_tuning_set: widths must be >= 1");
  pof::Tuning t;
  t.filter_up = widths[0];
  t.filter_down = widths[1];
  t.smooth_up = widths[2];
  t.smooth_down = widths[3];
  ctx->tuning[D] = t;
  ++ctx->tune_version;
  return POF_OK;
}

int pof_tuning_get(const pof_ctx* ctx, int D, int32_t* widths) {
  if (!ctx || !widths) return POF_E_ARG;
  pof::Tuning t;
  const auto it = ctx->tuning.find(D);
  if (it != ctx->tuning.end()) t = it->second;
  widths[0] = t.filter_up;
  widths[1] = t.filter_down;
  widths[2] = t.smooth_up;
  widths[3] = t.smooth_down;
  return POF_OK;
}

int pof_profile_read(const pof_ctx* ctx, int n_classes, double* ms, long long* launches, long long* work) {
  if (!ctx || n_classes < 0) return POF_E_ARG;
  for (int i = 0; i < n_classes && i < pof::kNumKernelClasses; ++i) {
    if (ms) ms[i] = ctx->prof_ms[i];
    if (launches) launches[i] = ctx->prof_launches[i];
    if (work) work[i] = ctx->prof_work[i];
  }
  return pof::kNumKernelClasses;
}

int pof_eval_field(int kind, int d, const double* params, const double* y_host, double* f_host) {
  const int need = pof::problem_dim(kind);
  if (need < 0 || (need > 0 && need != d) || !params || !y_host || !f_host) return POF_E_ARG;
  return pof::field_eval(kind, d, params, y_host, f_host) ? POF_OK : POF_E_ARG;
}

int pof_setup(const pof_problem* prob, double* F_host, double* QL_host, double* P_host, double* PI_host,
              double* x0_mean_host, double* x0_chol_host) {
  pof::Setup s;
  std::string err;
  int rc = pof::build_setup(prob, s, err);
  if (rc != POF_OK) return rc;
  if (F_host) std::copy(s.F.begin(), s.F.end(), F_host);
  if (QL_host) std::copy(s.LQ.begin(), s.LQ.end(), QL_host);
  if (P_host) std::copy(s.P.begin(), s.P.end(), P_host);
  if (PI_host) std::copy(s.PI.begin(), s.PI.end(), PI_host);
  if (x0_mean_host) std::copy(s.x0m.begin(), s.x0m.end(), x0_mean_host);
  if (x0_chol_host) std::copy(s.x0L.begin(), s.x0L.end(), x0_chol_host);
  return POF_OK;
}

int pof_transitions(int d, int q, double dt, double* F_host, double* QL_host, double* P_host, double* PI_host) {
  if (d < 1 || q < 0) return POF_E_ARG;
  std::vector<double> A1, L1, P, PI;
  pof::iwp_1d(q, A1, L1);
  const std::vector<double> F = pof::kron_eye(d, A1, q + 1), LQ = pof::kron_eye(d, L1, q + 1);
  pof::nordsieck(d, q, dt, P, PI);
  if (F_host) std::copy(F.begin(), F.end(), F_host);
  if (QL_host) std::copy(LQ.begin(), LQ.end(), QL_host);
  if (P_host) std::copy(P.begin(), P.end(), P_host);
  if (PI_host) std::copy(PI.begin(), PI.end(), PI_host);
  return POF_OK;
}

int pof_initial_means(const pof_problem* prob, double* means_host) {
  pof::Setup s;
  std::string err;
  int rc = pof::build_setup(prob, s, err);
  if (rc != POF_OK) return rc;
  std::vector<double> m;
  rc = pof::initial_means(prob, s, m, err);
  if (rc != POF_OK) return rc;
  std::copy(m.begin(), m.end(), means_host);
  return POF_OK;
}

int pof_linearize(pof_ctx* ctx, const pof_problem* prob, const double* m_dev, int64_t n, double* H_dev,
                  double* c_dev, void* stream) {
  if (!ctx) return POF_E_ARG;
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (n < 0 || n > 0x3fffffff || (n > 0 && (!m_dev || !H_dev || !c_dev)))
    return set_err(ctx, POF_E_ARG, "bad linearize args");
  const pof::EngineFns* e = engine_or_err(ctx, s.D, s.d);
  if (!e) return POF_E_UNSUPPORTED;
  if (n == 0) return POF_OK;
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipSetDevice(ctx->device));
  double* Pd = nullptr;
  HIP_TRY(hipMalloc((void**)&Pd, s.D * sizeof(double)));
  hipError_t he = hipMemcpyAsync(Pd, s.P.data(), s.D * sizeof(double), hipMemcpyHostToDevice, st);
  pof::ModelArgs a{};
  a.Pd = Pd;
  a.kind = prob->kind;
  a.q = s.q;
  a.Ns = (int)n;
  for (int i = 0; i < pof::kMaxParams; ++i) a.params[i] = prob->params[i];
  if (he == hipSuccess) he = e->linearize(a, m_dev, (int)n, H_dev, c_dev, st);
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  (void)hipFree(Pd);
  if (he != hipSuccess) return set_err(ctx, POF_E_HIP, "linearize failed: %s", hipGetErrorString(he));
  return POF_OK;
}

}  // extern "C"

namespace {
// pof_linearize_model on a device buffer of lin_buffer_doubles(...) doubles:
// `buf` (the no-allocation variant's scratch), or null to allocate one for the call.
int linearize_model(pof_ctx* ctx, const pof_problem* prob, int method, const double* params, const double* m_dev,
                    const double* L_dev, int64_t n, double* H_dev, double* c_dev, double* R_dev, void* stream,
                    double* buf, size_t buf_doubles) {
  if (!ctx) return POF_E_ARG;
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (method < POF_LIN_EK1 || method > POF_LIN_UNSCENTED) return set_err(ctx, POF_E_ARG, "unknown method %d", method);
  const bool needs_L = method == POF_LIN_UNCERTAIN || method == POF_LIN_UNSCENTED;
  if (n < 0 || n > 0x3fffffff || (n > 0 && (!m_dev || !H_dev || !c_dev || !R_dev || (needs_L && !L_dev))))
    return set_err(ctx, POF_E_ARG, "bad linearize_model args");
  pof::LinParams lp{};
  lp.lam = params ? params[0] : 1.0;
  lp.alpha = params ? params[1] : 1.0;
  lp.beta = params ? params[2] : 0.0;
  lp.kappa = params ? params[3] : 0.0;
  lp.kappa_set = (params && !std::isnan(params[3])) ? 1 : 0;
  if (method == POF_LIN_REGULARIZED && !(lp.lam > 0.0)) return set_err(ctx, POF_E_ARG, "lam must be > 0");
  const pof::EngineFns* e = engine_or_err(ctx, s.D, s.d);
  if (!e) return POF_E_UNSUPPORTED;
  const size_t need = lin_buffer_doubles(s.D, s.d, method, n);
  if (buf && need > buf_doubles)
    return set_err(ctx, POF_E_ARG,
                   "n = %lld states need %zu doubles of scratch; the no-allocation scratch holds %zu (n <= %lld here)",
                   (long long)n, need, buf_doubles,
                   needs_L ? (long long)((buf_doubles - lin_buffer_doubles(s.D, s.d, method, 0)) /
                                         (size_t)lin_work_host(s.D, s.d))
                           : 0x3fffffffLL);
  if (n == 0) return POF_OK;
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipSetDevice(ctx->device));
  const bool own = buf == nullptr;
  if (own) HIP_TRY(hipMalloc((void**)&buf, need * sizeof(double)));
  // P, then the per-state work slices (uncertain / unscented); each slice is
  // written before it is read, so nothing needs clearing
  const size_t head = lin_buffer_doubles(s.D, s.d, POF_LIN_EK1, 0);
  hipError_t he = hipMemcpyAsync(buf, s.P.data(), s.D * sizeof(double), hipMemcpyHostToDevice, st);
  pof::ModelArgs a{};
  a.Pd = buf;
  a.kind = prob->kind;
  a.q = s.q;
  a.Ns = (int)n;
  for (int i = 0; i < pof::kMaxParams; ++i) a.params[i] = prob->params[i];
  if (he == hipSuccess)
    he = e->linearize_alt(a, method, lp, m_dev, L_dev, (int)n, H_dev, c_dev, R_dev, needs_L ? buf + head : nullptr,
                          st);
  // the host copy of P must outlive its transfer, and the call is complete on return
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  if (own) (void)hipFree(buf);
  if (he != hipSuccess) return set_err(ctx, POF_E_HIP, "linearize_model failed: %s", hipGetErrorString(he));
  return POF_OK;
}
}  // namespace

extern "C" {

int pof_linearize_model(pof_ctx* ctx, const pof_problem* prob, int method, const double* params,
                        const double* m_dev, const double* L_dev, int64_t n, double* H_dev, double* c_dev,
                        double* R_dev, void* stream) {
  return linearize_model(ctx, prob, method, params, m_dev, L_dev, n, H_dev, c_dev, R_dev, stream, nullptr, 0);
}

int pof_linearize_model_noalloc(pof_ctx* ctx, const pof_problem* prob, int method, const double* params,
                                const double* m_dev, const double* L_dev, int64_t n, double* H_dev, double* c_dev,
                                double* R_dev, void* stream) {
  if (!ctx) return POF_E_ARG;
  if (!ctx->scratch) return set_err(ctx, POF_E_HIP, "no-allocation scratch unavailable: %s", ctx->scratch_err.c_str());
  return linearize_model(ctx, prob, method, params, m_dev, L_dev, n, H_dev, c_dev, R_dev, stream, ctx->scratch,
                         kScratchDoubles);
}

// The whole solve on the large-state engine (batched primitives); `small` is the
// per-worker engine of (D, d) when there is one (its EK1 kernel is reused).
static int solve_large(pof_ctx* ctx, const pof_problem* prob, const pof_opts* opts, pof_out* out, void* stream) {
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (opts->sequential) return set_err(ctx, POF_E_UNSUPPORTED, "the large-state engine runs the parallel IEKS only");
  if (opts->products != POF_PRODUCTS_F64 && opts->products != POF_PRODUCTS_F32)
    return set_err(ctx, POF_E_ARG, "products must be POF_PRODUCTS_F64 or POF_PRODUCTS_F32");
  // FP32 combine products leave rounding of ~1e-7 relative in the filtered means; sigma_squared (the
  // mean square of the whitened one-step residuals, tiny near convergence) is then rounding noise, and
  // with it every calibrated factor
  if (opts->products == POF_PRODUCTS_F32 && opts->calibrate)
    return set_err(ctx, POF_E_UNSUPPORTED,
                   "FP32 combine products do not resolve sigma_squared: use POF_PRODUCTS_F64 to calibrate");
  if (s.Ns > 65534) return set_err(ctx, POF_E_ARG, "the large-state engine takes at most 65534 steps");
  std::vector<double> m0;
  rc = (prob->init == POF_INIT_COARSE) ? POF_E_UNSUPPORTED : pof::initial_means(prob, s, m0, ctx->err);
  if (rc == POF_E_UNSUPPORTED) return set_err(ctx, rc, "init=coarse is not available on the large-state engine");
  if (rc != POF_OK) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if (!ctx->large) {
    ctx->large = pof::large::engine_create(ctx->device);
    if (!ctx->large) return set_err(ctx, POF_E_HIP, "large-state engine creation failed");
  }
  pof::large::Problem lp{};
  lp.D = s.D;
  lp.d = s.d;
  lp.q = s.q;
  lp.Ns = s.Ns;
  lp.kind = prob->kind;
  lp.params = prob->params;
  lp.F = s.F.data();
  lp.QL = s.LQ.data();
  lp.N1z = s.N1z.data();
  lp.P = s.P.data();
  lp.x0m = s.x0m.data();
  lp.x0L = s.x0L.data();
  lp.m0 = m0.data();
  std::string err;
  rc = pof::large::solve(ctx->large, lp, *opts, out, pof::find_engine(s.D, s.d), (hipStream_t)stream, err);
  if (rc != POF_OK) return set_err(ctx, rc, "%s", err.c_str());
  return POF_OK;
}

int pof_solve_large(pof_ctx* ctx, const pof_problem* prob, const pof_opts* opts, pof_out* out, void* stream) {
  if (!ctx || !prob || !opts || !out) return set_err(ctx, POF_E_ARG, "null argument");
  return solve_large(ctx, prob, opts, out, stream);
}

int pof_solve(pof_ctx* ctx, const pof_problem* prob, const pof_opts* opts, pof_out* out, void* stream) {
  if (!ctx || !opts || !out) return set_err(ctx, POF_E_ARG, "null argument");
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (opts->products == POF_PRODUCTS_F32 ||
      (!pof::find_engine(s.D, s.d) && pof::large::supports(prob->kind, s.D, s.d)))
    return solve_large(ctx, prob, opts, out, stream);
  if (!engine_or_err(ctx, s.D, s.d)) return POF_E_UNSUPPORTED;
  pof_batch* b = nullptr;
  rc = cached_batch(ctx, s.D, s.d, entries_for(prob->n_times), 1, &b);
  if (rc != POF_OK) return rc;
  StreamScope scope(b, stream);
  HIP_TRY(scope.status);
  rc = admit_problem(b, prob, opts, out, nullptr);
  if (rc == POF_OK) rc = run_to_completion(b);
  if (rc != POF_OK) batch_abort(b);
  return rc;
}

int pof_coarse_means(pof_ctx* ctx, const pof_problem* prob, double* means_host) {
  if (!ctx || !prob || !means_host) return set_err(ctx, POF_E_ARG, "null argument");
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (!engine_or_err(ctx, s.D, s.d)) return POF_E_UNSUPPORTED;
  std::vector<double> m;
  rc = coarse_means(ctx, prob, s, m);
  if (rc != POF_OK) return rc;
  std::copy(m.begin(), m.end(), means_host);
  return POF_OK;
}

int pof_eks_solve(pof_ctx* ctx, const pof_problem* prob, int calibrate, pof_out* out, void* stream) {
  if (!ctx || !prob || !out) return set_err(ctx, POF_E_ARG, "null argument");
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (!engine_or_err(ctx, s.D, s.d)) return POF_E_UNSUPPORTED;
  pof_batch* b = nullptr;
  rc = cached_batch(ctx, s.D, s.d, entries_for(prob->n_times), 1, &b);
  if (rc != POF_OK) return rc;
  StreamScope scope(b, stream);
  HIP_TRY(scope.status);
  return eks_run(b, prob, s, calibrate, out);
}

int pof_ieks_step(pof_ctx* ctx, const pof_problem* prob, const double* states_mean_dev, int sequential,
                  double* out_mean_dev, double* out_chol_dev, double* nll, double* obj, double* ssq, void* stream) {
  if (!ctx || !states_mean_dev) return set_err(ctx, POF_E_ARG, "null argument");
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (!engine_or_err(ctx, s.D, s.d)) return POF_E_UNSUPPORTED;
  pof_batch* b = nullptr;
  rc = cached_batch(ctx, s.D, s.d, entries_for(prob->n_times), 1, &b);
  if (rc != POF_OK) return rc;
  StreamScope scope(b, stream);
  HIP_TRY(scope.status);
  pof_opts o{};
  o.maxiters = 1;
  o.fixed_iters = 1;
  o.calibrate = 0;  // factors scaled by sqrt(ssq) once (step.py:42-44)
  pof_out po{};
  po.full_means_dev = out_mean_dev;
  po.full_chols_dev = out_chol_dev;
  o.sequential = sequential ? 1 : 0;
  rc = admit_raw(b, s.Ns, s.P.data(), s.x0m.data(), s.x0L.data(), prob->kind, prob->params, nullptr,
                 states_mean_dev, &o, &po, 0, o.sequential, false, nullptr);
  if (rc == POF_OK) rc = run_to_completion(b);
  if (rc != POF_OK) {
    batch_abort(b);
    return rc;
  }
  if (nll) *nll = po.nll;
  if (obj) *obj = po.obj;
  if (ssq) *ssq = po.sigma_squared;
  return POF_OK;
}

namespace {
// A caller's transition model (F, QL) for an affine batch: F, QL, N1 = tria([0 | QL]),
// and whether QL is lower triangular (its strict upper part exactly zero), which
// lets the smoother-element factorisation skip QL's upper columns.
hipError_t upload_transition(pof_batch* b, const double* F_host, const double* QL_host) {
  const int D = b->D;
  const size_t DD = (size_t)D * D;
  std::vector<double> M(DD * 2, 0.0), N1z;
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) M[(size_t)r * 2 * D + D + c] = QL_host[(size_t)r * D + c];
  pof::tria_host(M, D, 2 * D, N1z);
  b->g.lq_lower = pof::strictly_upper_zero(QL_host, D) ? 1 : 0;
  hipError_t e = hipMemcpy(b->F, F_host, DD * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->LQ, QL_host, DD * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->N1z, N1z.data(), DD * 8, hipMemcpyHostToDevice);
  return e;
}
}  // namespace

int pof_smoothing(pof_ctx* ctx, int D, int64_t Ns, const double* F_host, const double* QL_host,
                  const double* filt_mean_dev, const double* filt_chol_dev, int sequential, double* smooth_mean_dev,
                  double* smooth_chol_dev, double* obj, void* stream) {
  if (!ctx) return POF_E_ARG;
  if (D < 1 || Ns < 1 || Ns > 0x3fffffff || !F_host || !QL_host || !filt_mean_dev || !filt_chol_dev)
    return set_err(ctx, POF_E_ARG, "invalid smoothing arguments");
  // the smoother does not depend on the observation dimension: any kernel table entry of this D
  const pof::EngineFns* e = nullptr;
  int d = 0;
  for (int dd = 1; dd <= 4 && !e; ++dd)
    if (D % dd == 0 && (e = pof::find_engine(D, dd)) != nullptr) d = dd;
  if (!e) return set_err(ctx, POF_E_UNSUPPORTED, "no kernels for state dimension %d", D);
  HIP_TRY(hipSetDevice(ctx->device));
  pof_batch* b = nullptr;
  int rc = batch_new(ctx, D, d, entries_for(Ns + 1), 1, true, &b);
  if (rc != POF_OK) return rc;
  struct Guard {
    pof_batch* b;
    ~Guard() { batch_free(b); }
  } guard{b};
  StreamScope scope(b, stream);
  HIP_TRY(scope.status);
  const size_t DD = (size_t)D * D;
  HIP_TRY(upload_transition(b, F_host, QL_host));
  // state 0 enters as the prior of the slot, states 1..Ns as the level-0 filter prefixes
  std::vector<double> m0(D), L0(DD);
  HIP_TRY(hipMemcpy(m0.data(), filt_mean_dev, D * 8, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(L0.data(), filt_chol_dev, DD * 8, hipMemcpyDeviceToHost));
  std::vector<double> P(D, 1.0);
  pof_opts o{};
  o.maxiters = 1;
  o.fixed_iters = 1;
  pof_out po{};
  po.full_means_dev = smooth_mean_dev;
  po.full_chols_dev = smooth_chol_dev;
  rc = admit_raw(b, (int)Ns, P.data(), m0.data(), L0.data(), 0, nullptr, nullptr, nullptr, &o, &po, 1,
                 sequential ? 4 : 3, false, nullptr);
  if (rc != POF_OK) return rc;
  const long long base = b->sl[0].start;
  HIP_TRY(hipMemcpyAsync(b->g.fpre.m + (size_t)base * D, filt_mean_dev + D, (size_t)Ns * D * 8,
                         hipMemcpyDeviceToDevice, b->st));
  HIP_TRY(hipMemcpyAsync(b->g.fpre.L + (size_t)base * DD, filt_chol_dev + DD, (size_t)Ns * DD * 8,
                         hipMemcpyDeviceToDevice, b->st));
  rc = run_to_completion(b);
  if (rc != POF_OK) return rc;
  if (obj) *obj = po.obj;
  return POF_OK;
}

int pof_ekf(pof_ctx* ctx, const pof_problem* prob, const double* x0_mean_host, const double* x0_chol_host,
            const double* F_host, const double* QL_host, double* filt_mean_dev, double* filt_chol_dev, double* nll,
            double* ssq, void* stream) {
  if (!ctx || !prob || !x0_mean_host || !x0_chol_host || !F_host || !QL_host)
    return set_err(ctx, POF_E_ARG, "null argument");
  pof::Setup s;
  int rc = pof::build_setup(prob, s, ctx->err);
  if (rc != POF_OK) return rc;
  if (!engine_or_err(ctx, s.D, s.d)) return POF_E_UNSUPPORTED;
  HIP_TRY(hipSetDevice(ctx->device));
  pof_batch* b = nullptr;
  rc = batch_new(ctx, s.D, s.d, entries_for(prob->n_times), 1, false, &b);
  if (rc != POF_OK) return rc;
  struct Guard {
    pof_batch* b;
    ~Guard() { batch_free(b); }
  } guard{b};
  StreamScope scope(b, stream);
  HIP_TRY(scope.status);
  const int D = s.D;
  const size_t DD = (size_t)D * D;
  HIP_TRY(upload_transition(b, F_host, QL_host));
  pof_opts o{};
  o.maxiters = 1;
  o.fixed_iters = 1;
  o.sequential = 1;
  pof_out po{};
  rc = admit_raw(b, s.Ns, s.P.data(), x0_mean_host, x0_chol_host, prob->kind, prob->params, nullptr, nullptr, &o,
                 &po, 1, 2, false, nullptr);
  if (rc != POF_OK) return rc;
  rc = batch_launch(b);
  if (rc != POF_OK) return rc;
  HIP_TRY(hipEventSynchronize(b->ev));
  if (filt_mean_dev || filt_chol_dev) HIP_TRY(b->e->filtered(b->g, b->plan, filt_mean_dev, filt_chol_dev, b->st));
  rc = batch_finish(b, 0, nullptr, nullptr);
  if (rc != POF_OK) return rc;
  if (nll) *nll = po.nll;
  if (ssq) *ssq = po.sigma_squared;
  return POF_OK;
}

int pof_linear_filtsmooth(pof_ctx* ctx, int D, int d, int64_t Ns, const double* x0_mean_host,
                          const double* x0_chol_host, const double* F_host, const double* QL_host,
                          const double* H_dev, const double* c_dev, const double* R_dev, int sequential,
                          double* smooth_mean_dev,
                          double* smooth_chol_dev, double* filt_mean_dev, double* filt_chol_dev, double* nll,
                          double* obj, double* ssq, void* stream) {
  if (!ctx) return POF_E_ARG;
  if (D < 1 || d < 1 || Ns < 1 || Ns > 0x3fffffff || !x0_mean_host || !x0_chol_host || !F_host || !QL_host ||
      !H_dev || !c_dev)
    return set_err(ctx, POF_E_ARG, "invalid linear_filtsmooth arguments");
  if (!engine_or_err(ctx, D, d)) return POF_E_UNSUPPORTED;
  pof_batch* b = nullptr;
  int rc = batch_new(ctx, D, d, entries_for(Ns + 1), 1, true, &b);
  if (rc != POF_OK) return rc;
  struct Guard {
    pof_batch* b;
    ~Guard() { batch_free(b); }
  } guard{b};
  StreamScope scope(b, stream);
  HIP_TRY(scope.status);
  // the caller's model: F, QL, N1 = tria([0 | QL]), x0; no preconditioning (P = 1)
  const size_t DD = (size_t)D * D;
  HIP_TRY(upload_transition(b, F_host, QL_host));
  b->g.Hx = H_dev;
  b->g.cx = c_dev;
  b->g.Rx = R_dev;
  std::vector<double> P(D, 1.0);
  pof_opts o{};
  o.maxiters = 1;
  o.fixed_iters = 1;
  o.calibrate = 0;
  pof_out po{};
  po.full_means_dev = smooth_mean_dev;
  po.full_chols_dev = smooth_chol_dev;
  // the means buffer only feeds the convergence partials here; it starts at zero
  o.sequential = sequential ? 1 : 0;
  rc = admit_raw(b, (int)Ns, P.data(), x0_mean_host, x0_chol_host, 0, nullptr, nullptr, nullptr, &o, &po, 1,
                 o.sequential, false, nullptr);
  if (rc != POF_OK) return rc;
  rc = batch_launch(b);
  if (rc != POF_OK) return rc;
  // filtered states are read before the finish pass reuses nothing of the filter
  HIP_TRY(hipEventSynchronize(b->ev));
  if (filt_mean_dev || filt_chol_dev) HIP_TRY(b->e->filtered(b->g, b->plan, filt_mean_dev, filt_chol_dev, b->st));
  rc = batch_finish(b, 0, nullptr, nullptr);
  if (rc != POF_OK) return rc;
  if (nll) *nll = po.nll;
  if (obj) *obj = po.obj;
  if (ssq) *ssq = po.sigma_squared;
  return POF_OK;
}

// ------------------------------------------------------------------ batches
int64_t pof_batch_entries_for(int64_t n_times) { return entries_for(n_times); }

int64_t pof_batch_bytes(int D, int64_t entries, int32_t slots) { return (int64_t)batch_bytes(D, entries, slots); }

int pof_batch_create(pof_ctx* ctx, int D, int d, int64_t entries, int32_t slots, pof_batch** out) {
  if (!ctx || !out) return POF_E_ARG;
  return batch_new(ctx, D, d, entries, slots, false, out);
}

void pof_batch_destroy(pof_batch* b) { batch_free(b); }

int pof_batch_admit(pof_batch* b, const pof_problem* prob, const pof_opts* opts, pof_out* out, int64_t* ticket) {
  if (!b || !prob || !opts || !out) return POF_E_ARG;
  long long t = 0;
  int rc = admit_problem(b, prob, opts, out, &t);
  if (rc == POF_OK && ticket) *ticket = t;
  return rc;
}

int pof_batch_launch(pof_batch* b) { return b ? batch_launch(b) : POF_E_ARG; }

int pof_batch_ready(const pof_batch* b) {
  if (!b) return POF_E_ARG;
  if (!b->launched || !b->need_sync) return 1;
  const hipError_t q = hipEventQuery(b->ev);
  if (q == hipErrorNotReady) return 0;
  if (q != hipSuccess) return set_err(b->ctx, POF_E_HIP, "hipEventQuery: %s", hipGetErrorString(q));
  return 1;
}

int pof_batch_finish(pof_batch* b, int32_t max_done, int64_t* done, int32_t* n_done) {
  if (!b) return POF_E_ARG;
  std::vector<long long> tmp(std::max(max_done, 0));
  int nd = 0;
  int rc = batch_finish(b, max_done, tmp.data(), &nd);
  for (int i = 0; i < nd && i < max_done; ++i)
    if (done) done[i] = tmp[i];
  if (n_done) *n_done = nd;
  return rc;
}

int pof_batch_active(const pof_batch* b) {
  if (!b) return 0;
  int n = 0;
  for (const auto& s : b->sl) n += s.used ? 1 : 0;
  return n;
}

int64_t pof_batch_free_entries(const pof_batch* b) {
  if (!b) return 0;
  long long best = 0;
  for (const auto& kv : b->free_) best = std::max(best, kv.second);
  return best;
}

int pof_solve_batch(pof_ctx* ctx, const pof_problem* probs, const pof_opts* opts, pof_out* outs, int32_t n) {
  if (!ctx || (n > 0 && (!probs || !opts || !outs)) || n < 0) return set_err(ctx, POF_E_ARG, "null argument");
  // group by (D, d): one batch (and stream) per group
  struct Grp {
    int D, d;
    long long entries = 0;
    int count = 0;
    pof_batch* b = nullptr;
  };
  std::vector<Grp> groups;
  std::vector<int> gid(n);
  for (int i = 0; i < n; ++i) {
    const pof_problem& p = probs[i];
    if (p.d < 1 || p.q < 1 || p.n_times < 2) return set_err(ctx, POF_E_ARG, "problem %d: invalid sizes", i);
    const int D = p.d * (p.q + 1);
    if (!engine_or_err(ctx, D, p.d)) return POF_E_UNSUPPORTED;
    int k = -1;
    for (size_t g = 0; g < groups.size(); ++g)
      if (groups[g].D == D && groups[g].d == p.d) k = (int)g;
    if (k < 0) {
      groups.push_back(Grp{D, p.d});
      k = (int)groups.size() - 1;
    }
    groups[k].entries += entries_for(p.n_times);
    groups[k].count += 1;
    gid[i] = k;
  }
  for (auto& g : groups) {
    int rc = cached_batch(ctx, g.D, g.d, g.entries, g.count, &g.b);
    if (rc != POF_OK) return rc;
  }
  // on any failure after the first admission, every group's batch is emptied
  // (a later call must not run this call's IVPs)
  auto fail = [&](int rc) {
    for (auto& g : groups)
      if (g.b) batch_abort(g.b);
    return rc;
  };
  for (int i = 0; i < n; ++i) {
    int rc = admit_problem(groups[gid[i]].b, &probs[i], &opts[i], &outs[i], nullptr);
    if (rc != POF_OK) return fail(rc);
  }
  // iterate every group on its own stream, each at its own pace: a group's next
  // iteration is enqueued as soon as its previous one has been finished, so a
  // group with short iterations never waits for one with long ones.  (One host
  // thread per group, so that one group's blocking waits would not delay another's
  // launches, gained 2-3% on a 2,048-IVP batch but crashed once inside the HIP
  // runtime from a worker thread under rocprofv3's kernel trace; one host thread
  // drives every group.)
  for (auto& g : groups) {
    if (pof_batch_active(g.b) == 0) continue;
    int rc = batch_launch(g.b);
    if (rc != POF_OK) return fail(rc);
  }
  for (;;) {
    bool any = false;
    for (auto& g : groups) {
      if (!g.b->launched) continue;
      any = true;
      if (g.b->need_sync) {
        const hipError_t q = hipEventQuery(g.b->ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return fail(set_err(ctx, POF_E_HIP, "hipEventQuery: %s", hipGetErrorString(q)));
      }
      int rc = batch_finish(g.b, 0, nullptr, nullptr);
      if (rc != POF_OK) return fail(rc);
      if (pof_batch_active(g.b) > 0) {
        rc = batch_launch(g.b);
        if (rc != POF_OK) return fail(rc);
      }
    }
    if (!any) break;
  }
  return POF_OK;
}

}  // extern "C"
