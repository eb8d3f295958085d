/* C ABI of libpof_hip.so: the MI355X (gfx950) parallel-in-time IEKS ODE filter.
 *
 * The reference (nathanaelbosch/parallel-in-time-ode-filters) is pure Python
 * over JAX and has no FFI of its own; its drop-in surface is the Python
 * function pof.solver.solve (pof/solver.py:11-73).  Each entry point below is
 * the native counterpart of one reference function on that path; the Python
 * package `pof` in this repository binds them with ctypes (see
 * INTEGRATION.md) and keeps the reference's names and argument meanings.
 *
 * Conventions: plain pointers and sizes only.  Arrays named *_dev are device
 * pointers (row-major float64), arrays named *_host are host pointers.
 * Every function returns POF_OK (0) or a negative POF_E_* status; the message
 * of the last failure on a context is pof_last_error(ctx).  Work is enqueued
 * on `stream` (a hipStream_t; NULL = the default stream) and is complete when
 * the call returns (scalar results are read back).  A context is not
 * thread-safe: use one per host thread.
 */
#ifndef POF_H_
#define POF_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POF_OK 0
#define POF_E_ARG -1        /* invalid argument (shape, kind, size) */
#define POF_E_UNSUPPORTED -2 /* no kernel instantiation for (D, d) */
#define POF_E_HIP -3        /* HIP runtime error */
#define POF_E_INIT -4       /* unknown initial-trajectory method */
#define POF_E_FULL -5       /* batch has no free IVP slot or entry range right now */

/* ODE vector fields available on the device (pof/ivp.py of the reference,
 * plus fhn_classic, lorenz63, lorenz96). */
#define POF_LOGISTIC 1
#define POF_LOTKAVOLTERRA 2
#define POF_VANDERPOL 3
#define POF_FITZHUGHNAGUMO 4
#define POF_FHN_CLASSIC 5
#define POF_ROBER 6
#define POF_RIGID_BODY 7
#define POF_SEIR 8
#define POF_THREEBODY 9
#define POF_HENONHEILES 10
#define POF_LORENZ63 11
#define POF_LORENZ96 12

/* initial trajectory (pof/convenience.py:76-92) */
#define POF_INIT_PRIOR 0
#define POF_INIT_CONSTANT 1
#define POF_INIT_COARSE 2   /* sequential EKS on 100 points, nearest-left lookup (initialization.py:112-129) */

#define POF_MAX_PARAMS 8

/* solve status */
#define POF_STATUS_CONVERGED 0 /* crit() fired (obj or means) */
#define POF_STATUS_MAXITER 1   /* k > maxiters */
#define POF_STATUS_NAN 2       /* obj or nll is NaN */
#define POF_STATUS_FIXED 3     /* ran exactly fixed_iters iterations */

typedef struct pof_ctx pof_ctx;
typedef struct pof_batch pof_batch;

/* One IVP y' = f(y) on the grid ts (pof/solver.py:11-13 arguments f, y0, ts, order, init). */
typedef struct {
  int32_t kind;                    /* POF_LOGISTIC ... */
  int32_t d;                       /* ODE dimension (len(y0)) */
  int32_t q;                       /* order = number of derivatives of the IWP prior, >= 1 */
  int32_t init;                    /* POF_INIT_PRIOR / POF_INIT_CONSTANT / POF_INIT_COARSE */
  int64_t n_times;                 /* len(ts) = Ns + 1 >= 2 */
  const double* ts_host;           /* [n_times]; uniform grid assumed, dt = ts[1] - ts[0] */
  const double* y0_host;           /* [d] */
  double params[POF_MAX_PARAMS];   /* field parameters (see pof/ivp.py factories) */
} pof_problem;

typedef struct {
  int32_t maxiters;    /* the loop runs while k <= maxiters (reference semantics) */
  int32_t fixed_iters; /* > 0: run exactly this many iterations, no convergence test */
  int32_t calibrate;   /* 1: returned factors scaled by sqrt(ssq) twice (solver.py:66-68 + step.py:42-44) */
  int32_t sequential;  /* 1: sequential filter + smoother per iteration (solve(sequential=True), step.py:39-40) */
  int32_t products;    /* POF_PRODUCTS_F64 (default) or POF_PRODUCTS_F32 (large-state engine only) */
} pof_opts;

/* Precision of the scan combines' matrix products.  No counterpart in the
 * reference (JAX runs one dtype); F32: the large-state engine forms the filter
 * and smoother combines' matrix-matrix products with FP32 MFMA GEMMs (FP32
 * accumulate), while every triangularisation, triangular solve, element and
 * observation statistic stays FP64.  Experimental: F32 needs calibrate = 0
 * (POF_E_UNSUPPORTED otherwise; sigma_squared is not resolved in FP32) and its
 * means miss the FP64 ones by about 1.3e-6 relative at D = 128, above the
 * 1e-6 bar; the Python API refuses it. */
#define POF_PRODUCTS_F64 0
#define POF_PRODUCTS_F32 1

typedef struct {
  /* caller-owned device outputs (NULL = not wanted) */
  double* means_dev;       /* [n_times, d]     E0 m_t                    */
  double* chols_dev;       /* [n_times, d, D]  E0 L_t (calibrated)        */
  double* full_means_dev;  /* [n_times, D]     preconditioned states      */
  double* full_chols_dev;  /* [n_times, D, D]  preconditioned factors     */
  /* info (pof/solver.py:58-64) */
  double iterations;
  double nll;
  double obj;
  double sigma_squared;
  int32_t calibrated;
  int32_t status;          /* POF_STATUS_* */
} pof_out;

/* Context: device binding and reusable device workspace. */
int pof_ctx_create(int device, pof_ctx** out);
void pof_ctx_destroy(pof_ctx* ctx);
const char* pof_last_error(const pof_ctx* ctx);

/* 1 if a kernel instantiation exists for state dimension D = d (q + 1) and ODE dimension d. */
int pof_supported(int D, int d);

/* Per-kernel timing with HIP events around every launch of the solve loop
 * (instrumented runs only; no reference counterpart).  Classes, in order:
 * filter_leaf, filter_up, filter_down, filter_post, smooth_up, smooth_down,
 * smooth_post, reduce.  pof_profile_enable(ctx, 1) resets the counters;
 * pof_profile_read fills up to n_classes entries of milliseconds, launch
 * counts and elements processed, and returns the number of classes. */
int pof_profile_enable(pof_ctx* ctx, int on);
int pof_profile_read(const pof_ctx* ctx, int n_classes, double* ms, long long* launches, long long* work);

/* Replay of IEKS iterations from a HIP graph (on by default; no reference
 * counterpart: the reference's loop is an XLA while_loop, pof/solver.py:36-57).
 * Once a batch's set of IVPs has been unchanged for two launches, one
 * iteration's kernel launches are captured into a graph and each later
 * iteration is one graph launch, until an IVP is admitted or retires.  Results
 * are bitwise the same either way; off (on = 0) launches every kernel from the
 * host.  Instrumented runs (pof_profile_enable) always launch from the host. */
int pof_graphs_enable(pof_ctx* ctx, int on);

/* Kernel forms of the scan levels (no reference counterpart; the reference's
 * scan is jax.lax.associative_scan, parallel_filtsmooth/filter.py:35 and
 * smoother.py:15-17).  Each scan kernel of state dimension D <= 16 exists in a
 * latency form (a matrix row split over 2 or 4 lanes, one combine per
 * wavefront) and a throughput form (one lane per row, several combines per
 * wavefront); a level with at least widths[k] work items (combines, summed
 * over the batch's IVPs) runs in the throughput form.  widths = {filter
 * up-sweep, filter down-sweep, smoother up-sweep, smoother down-sweep}, each
 * >= 1 (1: always the throughput form; a width above every level: never).
 * Results are bitwise the same in either form.  Set per context and state
 * dimension, typically from the autotuner's cache (pof/autotune.py); the
 * built-in default is 8192 for each. */
int pof_tuning_set(pof_ctx* ctx, int D, const int32_t* widths);
int pof_tuning_get(const pof_ctx* ctx, int D, int32_t* widths);

/* Full parallel IEKS solve.  Replaces pof.solver.solve (pof/solver.py:11-73). */
int pof_solve(pof_ctx* ctx, const pof_problem* prob, const pof_opts* opts, pof_out* out, void* stream);

/* The same solve on the large-state engine: every scan level as batched
 * primitives of this library (a blocked Householder LQ whose trailing updates
 * run on the FP64 matrix cores, a strided-batched FP64 MFMA GEMM and a
 * triangular-solve kernel), for state dimensions beyond the per-worker
 * kernels (Lorenz-96 with d = 32, q = 3: D = 128).  pof_solve uses it by itself
 * for such problems; it runs any (D, d) whose field it can linearise (every
 * pof_supported pair, and Lorenz-96 for D <= 128).  Parallel IEKS only;
 * init "prior" or "constant". */
int pof_solve_large(pof_ctx* ctx, const pof_problem* prob, const pof_opts* opts, pof_out* out, void* stream);

/* Solver set-up on the host.  Replaces pof.convenience.set_up_solver (pof/convenience.py:13-45):
 * F, QL [D, D] (transitions.py:37-50), P, PI diagonals [D] (transitions.py:53-68),
 * x0 mean [D] and factor [D, D] (preconditioned Taylor-mode init, initialization.py:15-22,
 * convenience.py:30-31).  Any output may be NULL. */
int pof_setup(const pof_problem* prob, double* F_host, double* QL_host, double* P_host, double* PI_host,
              double* x0_mean_host, double* x0_chol_host);

/* IWP prior of order q in d dimensions.  Replaces pof.transitions.preconditioned_discretize
 * and nordsieck_preconditioner (transitions.py:37-68): F, QL [D, D]; P, PI diagonals [D]. */
int pof_transitions(int d, int q, double dt, double* F_host, double* QL_host, double* P_host, double* PI_host);

/* Initial trajectory means [n_times, D] on the host.  Replaces
 * pof.convenience.get_initial_trajectory (convenience.py:76-92) for "constant"
 * (PI-scaled) and "prior" (not PI-scaled, steps = ts[1:]); "coarse" needs the
 * GPU (pof_coarse_means) and returns POF_E_UNSUPPORTED here. */
int pof_initial_means(const pof_problem* prob, double* means_host);

/* "coarse" initial trajectory means [n_times, D] on the host: sequential EKS
 * on 100 equispaced points on the GPU, looked up at floor(ts / coarse_dt),
 * times PI.  Replaces pof.initialization.coarse_ekf_init (initialization.py:112-129)
 * with convenience.py:80-83. */
int pof_coarse_means(pof_ctx* ctx, const pof_problem* prob, double* means_host);

/* Single-pass sequential EKS.  Replaces pof.solver.sequential_eks_solve
 * (pof/solver.py:76-96): sequential square-root EKF (linearised at the
 * predicted means) + RTS smoother; factors scaled by sqrt(ssq) when
 * calibrate.  out->means_dev / chols_dev receive E0 (m, L); full_means_dev /
 * full_chols_dev (return_full_states) receive P (m, L).  info: nll is the
 * summed log-likelihood, as in the reference. */
int pof_eks_solve(pof_ctx* ctx, const pof_problem* prob, int calibrate, pof_out* out, void* stream);

/* Vector field f(y) at one point on the host (pof/ivp.py f(None, y)). */
int pof_eval_field(int kind, int d, const double* params, const double* y_host, double* f_host);

/* EK1 linearisation of r(x) = E1 x - f(E0 x) at n device points m [n, D].
 * Replaces pof.step.linearize_at_previous_states / pof.observations.linearize
 * (step.py:12-22, observations.py:35-40): H [n, d, D], c [n, d], cholR = 0. */
int pof_linearize(pof_ctx* ctx, const pof_problem* prob, const double* m_dev, int64_t n, double* H_dev,
                  double* c_dev, void* stream);

/* Linearisations of the residual r(x) = E1 x - f(E0 x) (method codes) */
#define POF_LIN_EK1 0          /* observations.py:35-40: H = E1 - J E0, cholR = 0 */
#define POF_LIN_EK0 1          /* observations.py:52-61: H = E1 of the IWP without the preconditioner, cholR = 0 */
#define POF_LIN_UNCERTAIN 2    /* observations.py:43-49: EK1 with cholR = tria(H L) */
#define POF_LIN_REGULARIZED 3  /* observations.py:64-89: H = [J; I], b = [r(m); -m], cholR = blockdiag(0, I/sqrt(lam)) */
#define POF_LIN_UNSCENTED 4    /* linearization/unscented.py:19-70 */

/* Linearise at n device states (m [n, D], L [n, D, D]; L is read by UNCERTAIN and
 * UNSCENTED only).  Replaces pof.observations.{linearize, linearize_ek0,
 * uncertain_linearize, linearize_regularized} and
 * pof.linearization.unscented.linearize_unscented, one state each, vectorised.
 * params (may be NULL): {lam, alpha, beta, kappa}; kappa NaN = 3 + D (defaults
 * lam 1, alpha 1, beta 0).  Outputs H [n, dO, D], c [n, dO], R [n, dO, dO] with
 * dO = d, or d + D for REGULARIZED. */
int pof_linearize_model(pof_ctx* ctx, const pof_problem* prob, int method, const double* params,
                        const double* m_dev, const double* L_dev, int64_t n, double* H_dev, double* c_dev,
                        double* R_dev, void* stream);

/* pof_linearize_model with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as pof_linearize_model, but its device
 * buffer (P, then a work slice per state for UNCERTAIN / UNSCENTED: 362
 * doubles per state at D = 15, d = 3) is a library-owned 64 MiB array that
 * each GPU holds until the process ends, instead of a hipMalloc per call.
 * POF_E_ARG when n needs more than that array (n <= 23,000 or so at D = 15 for
 * UNCERTAIN / UNSCENTED; EK1, EK0 and REGULARIZED need no per-state scratch).
 * Every call on one GPU shares the array: calls that overlap on one GPU
 * (several host threads or streams at once) race. */
int pof_linearize_model_noalloc(pof_ctx* ctx, const pof_problem* prob, int method, const double* params,
                                const double* m_dev, const double* L_dev, int64_t n, double* H_dev, double* c_dev,
                                double* R_dev, void* stream);

/* One IEKS iteration.  Replaces pof.step.ieks_step (step.py:33-45): linearise at
 * states_dev [Ns+1, D], parallel (sequential = 0) or sequential (1) filter +
 * smoother, chol *= sqrt(ssq).  Outputs full states [Ns+1, D] / [Ns+1, D, D]
 * and nll, obj, ssq (nll is +sum log-likelihood on the sequential branch, as
 * in the reference). */
int pof_ieks_step(pof_ctx* ctx, const pof_problem* prob, const double* states_mean_dev, int sequential,
                  double* out_mean_dev, double* out_chol_dev, double* nll, double* obj, double* ssq, void* stream);

/* Noiseless linear filter + smoother with a constant transition (F, QL) and
 * observations y_k = H_k x_{k+1} + c_k = 0, k < Ns.  Replaces
 * pof.parallel_filtsmooth.linear_filtsmooth (parallel_filtsmooth/__init__.py:5-10,
 * filter.py:18-47, smoother.py:8-22) and, with sequential = 1,
 * pof.sequential_filtsmooth.linear_filtsmooth (sequential_filtsmooth/__init__.py:13-16).
 * x0 mean [D] / factor [D, D], F/QL [D, D] on the host; H [Ns, d, D], c [Ns, d],
 * the observation noise factors R [Ns, d, d] (NULL = noiseless) and outputs on
 * the device.  filt_* (may be NULL) receive the filtered states [Ns+1, ...]. */
int pof_linear_filtsmooth(pof_ctx* ctx, int D, int d, int64_t Ns, const double* x0_mean_host,
                          const double* x0_chol_host, const double* F_host, const double* QL_host,
                          const double* H_dev, const double* c_dev, const double* R_dev, int sequential,
                          double* smooth_mean_dev,
                          double* smooth_chol_dev, double* filt_mean_dev, double* filt_chol_dev, double* nll,
                          double* obj, double* ssq, void* stream);

/* RTS smoother alone over given filtered states [Ns+1, D] / [Ns+1, D, D] (device),
 * constant transition (F, QL) on the host.  Replaces
 * pof.parallel_filtsmooth.smoothing (smoother.py:8-22; sequential = 0) and
 * pof.sequential_filtsmooth.smoothing (sequential_filtsmooth/smoother.py:8-28;
 * sequential = 1).  Outputs the smoothed states and obj (utils.py:97-101). */
int pof_smoothing(pof_ctx* ctx, int D, int64_t Ns, const double* F_host, const double* QL_host,
                  const double* filt_mean_dev, const double* filt_chol_dev, int sequential, double* smooth_mean_dev,
                  double* smooth_chol_dev, double* obj, void* stream);

/* Square-root EKF of the problem's residual model, linearised at each predicted
 * mean, with the caller's prior x0 and constant transition (F, QL) (host), for
 * n_times - 1 steps.  Replaces pof.sequential_filtsmooth.extended_kalman_filter
 * (sequential_filtsmooth/filter.py:9-30): filtered states [Ns+1, ...] on the
 * device (may be NULL), nll = summed log-likelihood, ssq. */
int pof_ekf(pof_ctx* ctx, const pof_problem* prob, const double* x0_mean_host, const double* x0_chol_host,
            const double* F_host, const double* QL_host, double* filt_mean_dev, double* filt_chol_dev, double* nll,
            double* ssq, void* stream);

/* ------------------------------------------------------------------ batches
 * Many IVPs of one state dimension D = d (q + 1) solved together: every scan
 * level of every IVP in the group runs in one launch, each IVP keeps its own
 * iterate state in HBM and its own convergence test, and IVPs can be admitted
 * and retired between iterations (continuous batching).  No counterpart in
 * the reference, which solves one IVP per call; each IVP's arithmetic is the
 * same as pof_solve's (results are identical alone and in a batch). */

/* Entries of batch workspace one IVP with n_times time points occupies (2 n_times). */
int64_t pof_batch_entries_for(int64_t n_times);
/* Device bytes of a batch with `entries` entries and `slots` IVP slots. */
int64_t pof_batch_bytes(int D, int64_t entries, int32_t slots);
/* Create a batch (allocates all of its device memory and its own stream). */
int pof_batch_create(pof_ctx* ctx, int D, int d, int64_t entries, int32_t slots, pof_batch** out);
void pof_batch_destroy(pof_batch* batch);
/* Admit one IVP (solve options as in pof_solve; `out` must stay valid until
 * the IVP finishes).  POF_E_FULL when no slot or entry range is free now. */
int pof_batch_admit(pof_batch* batch, const pof_problem* prob, const pof_opts* opts, pof_out* out,
                    int64_t* ticket);
/* Enqueue one IEKS iteration for every admitted, unfinished IVP (asynchronous). */
int pof_batch_launch(pof_batch* batch);
/* 1 if pof_batch_finish would not wait (the launched iteration has completed, or
 * nothing is launched), 0 while it runs. */
int pof_batch_ready(const pof_batch* batch);
/* Wait for the launched iteration, apply each IVP's convergence test
 * (pof/solver.py:36-57) and retire finished IVPs: factor pass, output
 * projection into their pof_out buffers, info fields.  Writes up to max_done
 * tickets of finished IVPs to done; *n_done = number finished. */
int pof_batch_finish(pof_batch* batch, int32_t max_done, int64_t* done, int32_t* n_done);
/* Admitted, unfinished IVPs. */
int pof_batch_active(const pof_batch* batch);
/* Largest free entry range (an IVP with n_times points fits if 2 n_times <= this). */
int64_t pof_batch_free_entries(const pof_batch* batch);
/* Solve n independent IVPs (any mix of (D, d)) to completion: one batch per
 * (D, d) group, groups iterated concurrently on their own streams. */
int pof_solve_batch(pof_ctx* ctx, const pof_problem* probs, const pof_opts* opts, pof_out* outs, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* POF_H_ */
