// HIP kernels of one parallel IEKS iteration for gfx950 (templates; the
// instantiations per ODE dimension live in pof_inst_d*.hip).
//
// Filter: JAX's odd/even associative scan (pof/parallel_filtsmooth/filter.py:35)
// reproduced level by level -- up-sweep level l pairs (2i, 2i+1), down-sweep
// fills the even positions -- so every factor is built by the same sequence
// of Householder LQs as the reference.  Smoother: the same scan structure
// on the elements stored in reversed time order (smoother.py:15-17,
// reverse=True).  All arithmetic is float64.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "pof_core.hpp"
#include "pof_engine.hpp"

namespace pof {

namespace {

inline ModelArgs model_args(const Workspace& ws, const IterOpts& o) {
  ModelArgs a{};
  a.F = ws.F;
  a.LQ = ws.LQ;
  a.N1z = ws.N1z;
  a.Pd = ws.Pd;
  a.x0m = ws.x0m;
  a.x0L = ws.x0L;
  a.lin = o.affine ? nullptr : ws.means[o.cur];
  a.Hx = ws.Hx;
  a.cx = ws.cx;
  a.kind = o.kind;
  a.q = o.q;
  a.Ns = ws.Ns;
  for (int i = 0; i < kMaxParams; ++i) a.params[i] = o.params[i];
  return a;
}

inline FiltArrays at(const FiltArrays& x, size_t off, int D) {
  const size_t DD = (size_t)D * D;
  return FiltArrays{x.A + off * DD, x.b + off * D, x.U + off * DD, x.eta + off * D, x.Z + off * DD};
}
inline SmoArrays at(const SmoArrays& x, size_t off, int D) {
  const size_t DD = (size_t)D * D;
  return SmoArrays{x.g + off * D, x.E + off * DD, x.Dm + off * DD};
}
inline StateArrays at(const StateArrays& x, size_t off, int D) {
  return StateArrays{x.m + off * D, x.L + off * (size_t)D * D};
}

}  // namespace

// ------------------------------------------------------------------ kernels
// Leaf elements: worker k builds filter element k (k = 0 carries x0).
template <int D, int d>
__global__ __launch_bounds__((Geo<D, kLdsLeaf>::BLOCK)) void k_filter_leaf(ModelArgs a, FiltArrays out, int n) {
  extern __shared__ double smem[];
  Ctx<D> cx = make_ctx<D, kLdsLeaf>(a, smem);
  const int k = cx.wid;
  if (k >= n) return;
  double A[D], U[D], Z[D];
  double b, eta;
  filter_element<D, d>(cx, a, k, k == 0, A, b, U, eta, Z);
  const int r = cx.lane;
  if (r < D) {
    const size_t o = (size_t)k * D + r;
    store_row(out.A + o * D, A);
    store_row(out.U + o * D, U);
    store_row(out.Z + o * D, Z);
    out.b[o] = b;
    out.eta[o] = eta;
  }
}

// Up-sweep: out[i] = in[2i] o in[2i+1]
template <int D>
__global__ __launch_bounds__((Geo<D, kLdsFull>::BLOCK)) void k_filter_up(ModelArgs a, FiltArrays in, FiltArrays out, int n_out) {
  extern __shared__ double smem[];
  Ctx<D> cx = make_ctx<D, kLdsFull>(a, smem);
  const int i = cx.wid;
  if (i >= n_out) return;
  StateArrays none{};
  filter_combine<D, false>(cx, in, none, 2 * (size_t)i, in, 2 * (size_t)i + 1, out, none, (size_t)i);
}

// Down-sweep of level l: positions 2i and 2i+1 of the level-l prefix states.
//   p[0] = state(e[0]); p[2i+1] = up[i]; p[2i] = up[i-1] o e[2i]  (i >= 1)
// With top != 0 (a level of one element) only p[0] = state(e[0]) is written.
template <int D>
__global__ __launch_bounds__((Geo<D, kLdsState>::BLOCK)) void k_filter_down(ModelArgs a, FiltArrays el, StateArrays up,
                                                               StateArrays p, int n) {
  extern __shared__ double smem[];
  Ctx<D> cx = make_ctx<D, kLdsState>(a, smem);
  const int i = cx.wid;
  const int j0 = 2 * i;
  if (j0 >= n) return;
  const int r = cx.lane;
  if (j0 == 0) {
    if (r < D) {
      p.m[r] = el.b[r];
#pragma unroll
      for (int c = 0; c < D; ++c) p.L[r * D + c] = el.U[r * D + c];
    }
  } else {
    FiltArrays none{};
    filter_combine<D, true>(cx, none, up, (size_t)i - 1, el, (size_t)j0, none, p, (size_t)j0);
  }
  const int j1 = j0 + 1;
  if (j1 < n && r < D) {
    p.m[(size_t)j1 * D + r] = up.m[(size_t)i * D + r];
#pragma unroll
    for (int c = 0; c < D; ++c) p.L[((size_t)j1 * D + r) * D + c] = up.L[((size_t)i * D + r) * D + c];
  }
}

// Per filtered state s = 0..Ns: smoother element (smoother.py:37-50) stored at
// reversed index Ns - s (terminal element (m, 0, L) for s = Ns), and for
// s < Ns the predicted observation statistics of filter.py:84-114:
//   S = tria([H_s tria([F L_s | LQ]) | 0]), yhat = H_s F m_s + c_s,
//   ssq_s = |S^-T yhat|^2 (utils.py:110-112), nll_s = -loglik(yhat, S) (utils.py:22-30).
template <int D, int d>
__global__ __launch_bounds__((Geo<D, kLdsPost>::BLOCK)) void k_filter_post(ModelArgs a, StateArrays pre0, SmoArrays sel,
                                                               double* t_ssq, double* t_nll) {
  extern __shared__ double smem[];
  Ctx<D> cx = make_ctx<D, kLdsPost>(a, smem);
  constexpr int LD = D + 1;
  const int s = cx.wid;
  const int Ns = a.Ns;
  if (s > Ns) return;
  const int r = cx.lane;
  const bool act = r < D;
  const double* mp = (s == 0) ? a.x0m : pre0.m + (size_t)(s - 1) * D;
  const double* Lp = (s == 0) ? a.x0L : pre0.L + (size_t)(s - 1) * D * D;
  const size_t ro = (size_t)(Ns - s);
  if (s == Ns) {
    if (act) {
      sel.g[ro * D + r] = mp[r];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        sel.E[(ro * D + r) * D + c] = 0.0;
        sel.Dm[(ro * D + r) * D + c] = Lp[r * D + c];
      }
    }
    return;
  }
  double* TL = cx.tile(0);
  double* TP11 = cx.tile(1);
  double* TE = cx.tile(2);
  double* TD = cx.tile(3);
  double* Vm = cx.vec(0);
  double* Vfm = cx.vec(1);
  // global reads first: the filtered state and the linearisation point of obs model s
  double H[d][D];
  double cv[d];
  observation<D, d>(a, s, H, cv);
  {
    const int rr = act ? r : 0;
    double lr[D];
    load_row(lr, Lp + rr * D);
    const double mr = mp[rr];
    if (act) {
#pragma unroll
      for (int c = 0; c < D; ++c) TL[r * LD + c] = lr[c];
      Vm[r] = mr;
    }
  }
  wsync();
  // Phi_ = [[F L, LQ], [L, 0]]
  double ph[2 * D];
  double fm = 0.0;
  if (r < D) {
    double fr[D];
    load_row(fr, cx.sF + r * LD);
    double t[D];
    row_times_tile<D, D, LD>(fr, TL, t);
#pragma unroll
    for (int k = 0; k < D; ++k) fm = fma(fr[k], Vm[k], fm);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      ph[c] = t[c];
      ph[D + c] = cx.sLQ[r * LD + c];
    }
    Vfm[r] = fm;
  } else if (r < 2 * D) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      ph[c] = TL[(r - D) * LD + c];
      ph[D + c] = 0.0;
    }
  } else {
    zero(ph);
  }
  lq<2 * D, 2 * D>(ph, r, cx.bc());
  if (r < D) {
#pragma unroll
    for (int c = 0; c < D; ++c) TP11[r * LD + c] = ph[c];
  }
  wsync();
  // E row = Phi21 row Phi11^-1 ; D = Phi22 ; moved to lane r - D
  if (r >= D && r < 2 * D) {
    double e[D];
#pragma unroll
    for (int c = 0; c < D; ++c) e[c] = ph[c];
    rsolve_lower<D, LD>(e, TP11);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      TE[(r - D) * LD + c] = e[c];
      TD[(r - D) * LD + c] = ph[D + c];
    }
  }
  wsync();
  // observation statistics with obs model s (the smoother element stays in
  // LDS until the end so that no store is outstanding during the LQ below)
  double sr[D + d];
  {
    double h[D];
    select_row<d, D>(H, r, h);
    double t[D];
    row_times_tile<D, D, LD>(h, TP11, t);
#pragma unroll
    for (int c = 0; c < D; ++c) sr[c] = (r < d) ? t[c] : 0.0;
#pragma unroll
    for (int c = 0; c < d; ++c) sr[D + c] = 0.0;
  }
  lq<d, D + d>(sr, r, cx.bc());
  double* TS = cx.tile(4);
  if (r < d) {
#pragma unroll
    for (int c = 0; c < d; ++c) TS[r * LD + c] = sr[c];
  }
  wsync();
  if (r == 0) {
    double yh[d];
#pragma unroll
    for (int i = 0; i < d; ++i) {
      double acc = cv[i];
#pragma unroll
      for (int j = 0; j < D; ++j) acc = fma(H[i][j], Vfm[j], acc);
      yh[i] = acc;
    }
    double wz[d], yz[d];
#pragma unroll
    for (int i = 0; i < d; ++i) {
      wz[i] = yh[i];
      yz[i] = yh[i];
    }
    lsolve_lowerT<d, LD>(wz, TS);  // S^T w = yhat
    lsolve_lower<d, LD>(yz, TS);   // S y = yhat
    double ss = 0.0, nn = 0.0, ld = 0.0;
#pragma unroll
    for (int i = 0; i < d; ++i) {
      ss = fma(wz[i], wz[i], ss);
      nn = fma(yz[i], yz[i], nn);
      ld += log(fabs(TS[i * LD + i]));
    }
    t_ssq[s] = ss;
    // -loglik = 0.5 |S^-1 y|^2 + sum log|S_ii| + d log(2 pi) / 2
    t_nll[s] = 0.5 * nn + (ld + d * log(2.0 * M_PI) / 2.0);
  }
  if (act) {
    double e[D];
    load_row(e, TE + r * LD);
    double gg = Vm[r];
#pragma unroll
    for (int k = 0; k < D; ++k) gg = fma(-e[k], Vfm[k], gg);
    sel.g[ro * D + r] = gg;
    store_row(sel.E + (ro * D + r) * D, e);
#pragma unroll
    for (int c = 0; c < D; ++c) sel.Dm[(ro * D + r) * D + c] = TD[r * LD + c];
  }
  wsync();
}

// Smoother up-sweep
template <int D>
__global__ __launch_bounds__((Geo<D, kLdsSmooth>::BLOCK)) void k_smooth_up(ModelArgs a, SmoArrays in, SmoArrays out, int n_out,
                                                             int with_chol) {
  extern __shared__ double smem[];
  Ctx<D> cx = make_ctx<D, kLdsSmooth>(a, smem);
  const int i = cx.wid;
  if (i >= n_out) return;
  StateArrays none{};
  smoother_combine<D, false>(cx, in, none, 2 * (size_t)i, in, 2 * (size_t)i + 1, out, none, (size_t)i,
                             with_chol != 0);
}

template <int D>
__global__ __launch_bounds__((Geo<D, kLdsSmooth>::BLOCK)) void k_smooth_down(ModelArgs a, SmoArrays el, StateArrays up,
                                                               StateArrays p, int n, int with_chol) {
  extern __shared__ double smem[];
  Ctx<D> cx = make_ctx<D, kLdsSmooth>(a, smem);
  const int i = cx.wid;
  const int j0 = 2 * i;
  if (j0 >= n) return;
  const int r = cx.lane;
  if (j0 == 0) {
    if (r < D) {
      p.m[r] = el.g[r];
      if (with_chol) {
#pragma unroll
        for (int c = 0; c < D; ++c) p.L[r * D + c] = el.Dm[r * D + c];
      }
    }
  } else {
    SmoArrays none{};
    smoother_combine<D, true>(cx, none, up, (size_t)i - 1, el, (size_t)j0, none, p, (size_t)j0, with_chol != 0);
  }
  const int j1 = j0 + 1;
  if (j1 < n && r < D) {
    p.m[(size_t)j1 * D + r] = up.m[(size_t)i * D + r];
    if (with_chol) {
#pragma unroll
      for (int c = 0; c < D; ++c) p.L[((size_t)j1 * D + r) * D + c] = up.L[((size_t)i * D + r) * D + c];
    }
  }
}

// Thread t: smoothed mean m_t (reversed prefix Ns - t) -> means_new[t];
// obj_t = |LQ^-1 (m_t - F m_{t+1})|^2 (utils.py:97-101 with the argument order
// of smoother.py:20); viol_t = 1 if any |m_old - m_new| > 1e-8 + 1e-13 |m_new|
// (convergence_criteria.py:10, jnp.isclose).
template <int D>
__global__ void k_smooth_post(const double* F, const double* LQ, StateArrays spre0, const double* mold, double* mnew,
                              double* t_obj, double* t_viol, int Ns) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > Ns) return;
  const double* m = spre0.m + (size_t)(Ns - t) * D;
  double mt[D];
  load_row(mt, m);
  store_row(mnew + (size_t)t * D, mt);
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double o = mold[(size_t)t * D + i];
    const bool close = fabs(o - mt[i]) <= 1e-8 + 1e-13 * fabs(mt[i]);
    v = close ? v : 1.0;
  }
  t_viol[t] = v;
  if (t < Ns) {
    const double* mn = spre0.m + (size_t)(Ns - t - 1) * D;
    double rr[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double acc = mt[i];
#pragma unroll
      for (int j = 0; j < D; ++j) acc = fma(-F[i * D + j], mn[j], acc);
      rr[i] = acc;
    }
    // forward substitution with LQ (lower)
    double o = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = rr[i];
#pragma unroll
      for (int j = 0; j < i; ++j) s = fma(-LQ[i * D + j], rr[j], s);
      rr[i] = s / LQ[i * D + i];
      o = fma(rr[i], rr[i], o);
    }
    t_obj[t] = o;
  }
}

// Fixed-order reduction of the per-index terms, stage 1: block b sums the
// contiguous segment [b * seg, (b + 1) * seg) into part[b * 4 + 0..3]
// (ssq, nll, obj: sums; violations: max).
constexpr int kRedBlocks = 128;
static __global__ void k_reduce_part(const double* t_ssq, const double* t_nll, const double* t_obj,
                                     const double* t_viol, int Ns, double* part) {
  __shared__ double sh[4][256];
  const int tid = threadIdx.x;
  const int seg = (Ns + 1 + kRedBlocks - 1) / kRedBlocks;
  const int lo = blockIdx.x * seg;
  const int hi = min(lo + seg, Ns + 1);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int i = lo + tid; i < hi; i += 256) {
    if (i < Ns) {
      s0 += t_ssq[i];
      s1 += t_nll[i];
      s2 += t_obj[i];
    }
    s3 = fmax(s3, t_viol[i]);
  }
  sh[0][tid] = s0;
  sh[1][tid] = s1;
  sh[2][tid] = s2;
  sh[3][tid] = s3;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      sh[0][tid] += sh[0][tid + w];
      sh[1][tid] += sh[1][tid + w];
      sh[2][tid] += sh[2][tid + w];
      sh[3][tid] = fmax(sh[3][tid], sh[3][tid + w]);
    }
    __syncthreads();
  }
  if (tid < 4) part[blockIdx.x * 4 + tid] = sh[tid][0];
}

// Stage 2 (one block of kRedBlocks threads): scal[0] = nll, scal[1] = obj,
// scal[2] = ssq, scal[3] = 1 if all means are close, scal[4] = NaN flag.
static __global__ void k_reduce_final(const double* part, int Ns, int d, double* scal) {
  __shared__ double sh[4][kRedBlocks];
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) sh[k][tid] = part[tid * 4 + k];
  __syncthreads();
  for (int w = kRedBlocks / 2; w > 0; w >>= 1) {
    if (tid < w) {
      sh[0][tid] += sh[0][tid + w];
      sh[1][tid] += sh[1][tid + w];
      sh[2][tid] += sh[2][tid + w];
      sh[3][tid] = fmax(sh[3][tid], sh[3][tid + w]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    scal[0] = sh[1][0];
    scal[1] = sh[2][0];
    scal[2] = sh[0][0] / Ns / d;
    scal[3] = (sh[3][0] == 0.0) ? 1.0 : 0.0;
    scal[4] = (isnan(scal[0]) || isnan(scal[1])) ? 1.0 : 0.0;
  }
}

// Output projection: means_out[t][i] = P_ii0 m_t[i0], chol_out[t][i][:] = scale * P_ii0 L_t[i0][:]
// with i0 = i (q+1) (E0 = E0_raw P, pof/convenience.py:26-27; solver.py:71).
template <int D, int d>
__global__ void k_output(const double* Pd, StateArrays spre0, int Ns, double scale, double* means_out,
                         double* chol_out, double* full_means, double* full_chols) {
  constexpr int q1 = D / d;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > Ns) return;
  const double* m = spre0.m + (size_t)(Ns - t) * D;
  const double* L = spre0.L + (size_t)(Ns - t) * D * D;
#pragma unroll
  for (int i = 0; i < d; ++i) {
    const double p = Pd[i * q1];
    if (means_out) means_out[(size_t)t * d + i] = p * m[i * q1];
    if (chol_out) {
#pragma unroll
      for (int c = 0; c < D; ++c) chol_out[((size_t)t * d + i) * D + c] = (scale * p) * L[i * q1 * D + c];
    }
  }
  if (full_means) {
#pragma unroll
    for (int i = 0; i < D; ++i) full_means[(size_t)t * D + i] = m[i];
  }
  if (full_chols) {
#pragma unroll
    for (int i = 0; i < D * D; ++i) full_chols[(size_t)t * D * D + i] = scale * L[i];
  }
}

// Filtered states (state s = x0 for s = 0, else level-0 filter prefix s-1).
template <int D>
__global__ void k_filtered_out(const double* x0m, const double* x0L, StateArrays fpre0, int Ns, double* fm,
                               double* fL) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > Ns) return;
  const double* m = (s == 0) ? x0m : fpre0.m + (size_t)(s - 1) * D;
  const double* L = (s == 0) ? x0L : fpre0.L + (size_t)(s - 1) * D * D;
  if (fm) {
#pragma unroll
    for (int i = 0; i < D; ++i) fm[(size_t)s * D + i] = m[i];
  }
  if (fL) {
#pragma unroll
    for (int i = 0; i < D * D; ++i) fL[(size_t)s * D * D + i] = L[i];
  }
}

// EK1 at n points (thread per point): H [n, d, D], c [n, d]
template <int D, int d>
__global__ void k_linearize(ModelArgs a, const double* m, int n, double* H, double* c) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  ModelArgs b = a;
  b.lin = m - D;  // observation(k) reads lin[k + 1]
  b.Hx = nullptr;
  double Hk[d][D];
  double ck[d];
  observation<D, d>(b, k, Hk, ck);
#pragma unroll
  for (int i = 0; i < d; ++i) {
#pragma unroll
    for (int j = 0; j < D; ++j) H[((size_t)k * d + i) * D + j] = Hk[i][j];
    c[(size_t)k * d + i] = ck[i];
  }
}

// ------------------------------------------------------------ launch logic
template <int D, int K>
inline unsigned grid_workers(size_t nworkers) {
  return (unsigned)((nworkers + Geo<D, K>::WPB - 1) / Geo<D, K>::WPB);
}
#define POF_LAUNCH(kernel, K, nworkers, ...)                                                        \
  hipLaunchKernelGGL(kernel, dim3((grid_workers<D, K>(nworkers))), dim3((Geo<D, K>::BLOCK)),        \
                     (Geo<D, K>::lds_bytes()), __VA_ARGS__)

namespace {
// With a timer, event i+1 is recorded after launch i (event 0 before the first).
struct Rec {
  KernelTimer* t;
  hipStream_t st;
  Rec(KernelTimer* t_, hipStream_t s) : t(t_), st(s) {
    if (t) {
      t->n = 0;
      (void)hipEventRecord(t->ev[0], st);
    }
  }
  void operator()(int cls, long long work) {
    if (!t || t->n >= kMaxTimed) return;
    t->cls[t->n] = cls;
    t->work[t->n] = work;
    (void)hipEventRecord(t->ev[t->n + 1], st);
    ++t->n;
  }
};
}  // namespace

template <int D>
void smooth_scan(const Workspace& ws, const ModelArgs& a, int wc, hipStream_t st, Rec& rec) {
  for (int l = 0; l + 1 < ws.s_levels; ++l) {
    const int nout = (int)ws.s_n[l + 1];
    POF_LAUNCH((k_smooth_up<D>), kLdsSmooth, nout, st, a,
                       at(ws.sel, ws.s_off[l], D), at(ws.sel, ws.s_off[l + 1], D), nout, wc);
    rec(kKSmoothUp, nout);
  }
  for (int l = ws.s_levels - 1; l >= 0; --l) {
    const int n = (int)ws.s_n[l];
    StateArrays up = (l + 1 < ws.s_levels) ? at(ws.spre, ws.s_off[l + 1], D) : StateArrays{};
    POF_LAUNCH((k_smooth_down<D>), kLdsSmooth, (n + 1) / 2, st, a,
                       at(ws.sel, ws.s_off[l], D), up, at(ws.spre, ws.s_off[l], D), n, wc);
    rec(kKSmoothDown, (n + 1) / 2 - 1);
  }
}

template <int D, int d>
hipError_t run_smooth_chol(const Workspace& ws, const IterOpts& o, hipStream_t st) {
  const ModelArgs a = model_args(ws, o);
  Rec rec(o.timer, st);
  smooth_scan<D>(ws, a, 1, st, rec);
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_iteration(const Workspace& ws, const IterOpts& o, hipStream_t st) {
  const ModelArgs a = model_args(ws, o);
  const int Ns = ws.Ns;
  Rec rec(o.timer, st);
  // ---- filter: leaf elements, up-sweep, down-sweep
  {
    FiltArrays l0 = at(ws.fel, ws.f_off[0], D);
    POF_LAUNCH((k_filter_leaf<D, d>), kLdsLeaf, Ns, st, a, l0, Ns);
    rec(kKFilterLeaf, Ns);
    for (int l = 0; l + 1 < ws.f_levels; ++l) {
      const int nout = (int)ws.f_n[l + 1];
      POF_LAUNCH((k_filter_up<D>), kLdsFull, nout, st, a,
                         at(ws.fel, ws.f_off[l], D), at(ws.fel, ws.f_off[l + 1], D), nout);
      rec(kKFilterUp, nout);
    }
    for (int l = ws.f_levels - 1; l >= 0; --l) {
      const int n = (int)ws.f_n[l];
      StateArrays up = (l + 1 < ws.f_levels) ? at(ws.fpre, ws.f_off[l + 1], D) : StateArrays{};
      POF_LAUNCH((k_filter_down<D>), kLdsState, (n + 1) / 2, st, a,
                         at(ws.fel, ws.f_off[l], D), up, at(ws.fpre, ws.f_off[l], D), n);
      rec(kKFilterDown, (n + 1) / 2 - 1);
    }
  }
  // ---- smoother elements + observation statistics
  {
    SmoArrays s0 = at(ws.sel, ws.s_off[0], D);
    POF_LAUNCH((k_filter_post<D, d>), kLdsPost, Ns + 1, st, a,
                       at(ws.fpre, ws.f_off[0], D), s0, ws.t_ssq, ws.t_nll);
    rec(kKFilterPost, Ns);
  }
  // ---- smoother scan (reversed order)
  smooth_scan<D>(ws, a, o.with_chol, st, rec);
  // ---- smoothed means, objective, convergence partials, reduction
  {
    const int nt = Ns + 1;
    hipLaunchKernelGGL((k_smooth_post<D>), dim3((nt + 127) / 128), dim3(128), 0, st, ws.F, ws.LQ,
                       at(ws.spre, ws.s_off[0], D), ws.means[o.cur], ws.means[1 - o.cur], ws.t_obj, ws.t_viol, Ns);
    rec(kKSmoothPost, Ns + 1);
    hipLaunchKernelGGL(k_reduce_part, dim3(kRedBlocks), dim3(256), 0, st, ws.t_ssq, ws.t_nll, ws.t_obj,
                       ws.t_viol, Ns, ws.part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedBlocks), 0, st, ws.part, Ns, d, ws.scal);
    rec(kKReduce, Ns);
  }
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_output(const Workspace& ws, int cur, double scale, double* means_out, double* chol_out,
                      double* full_means, double* full_chols, double* filt_means, double* filt_chols,
                      hipStream_t st) {
  (void)cur;
  const int nt = ws.Ns + 1;
  hipLaunchKernelGGL((k_output<D, d>), dim3((nt + 127) / 128), dim3(128), 0, st, ws.Pd,
                     at(ws.spre, ws.s_off[0], D), ws.Ns, scale, means_out, chol_out, full_means, full_chols);
  if (filt_means || filt_chols) {
    hipLaunchKernelGGL((k_filtered_out<D>), dim3((nt + 127) / 128), dim3(128), 0, st, ws.x0m, ws.x0L,
                       at(ws.fpre, ws.f_off[0], D), ws.Ns, filt_means, filt_chols);
  }
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_linearize(const Workspace& ws, const IterOpts& o, const double* m, int n, double* H, double* c,
                         hipStream_t st) {
  const ModelArgs a = model_args(ws, o);
  hipLaunchKernelGGL((k_linearize<D, d>), dim3((n + 127) / 128), dim3(128), 0, st, a, m, n, H, c);
  return hipGetLastError();
}

}  // namespace pof
