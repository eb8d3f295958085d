// HIP kernels of one parallel IEKS iteration for gfx950 (templates; the
// instantiations per ODE dimension live in pof_inst_d*.hip).
//
// Every launch serves all active IVPs of one group (same D, d, q): work item w
// of a launch belongs to the active IVP a with pre[a] <= w < pre[a + 1]
// (LaunchMap), and every IVP reads and writes only its own entry range, so a
// result does not depend on which other IVPs share the launch.
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
#include "pof_linearization.hpp"

namespace pof {

namespace {

__host__ __device__ inline FiltArrays at(const FiltArrays& x, size_t off, int D) {
  const size_t DD = (size_t)D * D;
  return FiltArrays{x.A + off * DD, x.b + off * D, x.U + off * DD, x.eta + off * D, x.Z + off * DD};
}
__host__ __device__ inline SmoArrays at(const SmoArrays& x, size_t off, int D) {
  const size_t DD = (size_t)D * D;
  return SmoArrays{x.g + off * D, x.E + off * DD, x.Dm + off * DD};
}
__host__ __device__ inline StateArrays at(const StateArrays& x, size_t off, int D) {
  return StateArrays{x.m + off * D, x.L + off * (size_t)D * D};
}

}  // namespace

// --------------------------------------------------------------- work lists
// Active IVP index of work item w: the largest a with pre[a] <= w.
__device__ __forceinline__ int find_seg(const int* pre, int m, int w) {
  int lo = 0, hi = m - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// IVP id j and local index i of work item w (false past the end).
__device__ __forceinline__ bool locate(const LaunchMap& lm, int w, int& j, int& i) {
  if (w >= lm.total) return false;
  const int a = find_seg(lm.pre, lm.m, w);
  j = lm.ivp[a];
  i = w - lm.pre[a];
  return true;
}

// Entry offset of scan level l inside an IVP's range (level t holds n0 >> t entries).
__device__ __forceinline__ long long lev_off(long long n0, int l) {
  long long o = 0;
  for (int t = 0; t < l; ++t) o += n0 >> t;
  return o;
}
// Number of scan levels of n0 >= 1 elements (n0, n0 / 2, ..., 1).
__device__ __forceinline__ int n_levels(long long n0) { return 64 - __clzll(n0); }

// Model view of IVP j of the group.
template <int D>
__device__ __forceinline__ ModelArgs view(const GroupArgs& g, int j) {
  ModelArgs a;
  const long long base = g.base[j];
  a.F = g.F;
  a.LQ = g.LQ;
  a.N1z = g.N1z;
  a.Pd = g.Pd + (size_t)j * D;
  a.x0m = g.x0m + (size_t)j * D;
  a.x0L = g.x0L + (size_t)j * D * D;
  a.lin = g.affine ? nullptr : ((g.cur[j] ? g.means1 : g.means0) + (size_t)base * D);
  a.Hx = g.Hx;
  a.cx = g.cx;
  a.Rx = g.Rx;
  a.kind = g.kind[j];
  a.q = g.q;
  a.Ns = g.Ns[j];
  a.pad = 0;
#pragma unroll
  for (int i = 0; i < kMaxParams; ++i) a.params[i] = g.params[(size_t)j * kMaxParams + i];
  return a;
}

// ------------------------------------------------------------------ kernels
// Launch bounds of the element kernels on the IEKS path (EK1), where LDS
// allows more waves per SIMD than their natural register allocation: the
// bound trades a small spill for occupancy.  Measured to pay at D = 12 only
// (leaf 3 -> 4, post 2 -> 3 waves, +0.9% on Lorenz-63 Ns = 65536); the same
// bounds on leaf<15>, post<10> and post<15> ran slower.
template <int D, bool EK1>
constexpr int leaf_min_waves(bool first) {
  return (first || !EK1) ? 1 : (D == 12) ? 4 : 1;
}
template <int D, bool EK1>
constexpr int post_min_waves() {
  return !EK1 ? 1 : (D == 12) ? 3 : 1;
}

// Leaf elements.  FIRST: element 0 of each IVP (it carries the prior x0;
// one work item per IVP); otherwise work item i builds element i + 1.
template <int D, int d, bool EK1, bool FIRST>
__global__ __launch_bounds__((Geo<D, FIRST ? kLdsLeaf1 : kLdsLeaf>::BLOCK), (leaf_min_waves<D, EK1>(FIRST))) void
k_filter_leaf(GroupArgs g, LaunchMap lm) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, FIRST ? kLdsLeaf1 : kLdsLeaf>(g.F, g.LQ, g.N1z, smem);
  int j, k;
  if (!locate(lm, cx.wid, j, k)) return;
  k = FIRST ? 0 : k + 1;
  const ModelArgs a = view<D>(g, j);
  const FiltArrays out = at(g.fel, g.base[j], D);
  constexpr int S = decltype(cx)::S, NL = Lay<D, S>::NL;
  double A[NL], U[NL], Z[NL];
  double b, eta;
  filter_element<D, d, EK1, FIRST>(cx, a, k, A, b, U, eta, Z);
  const int r = cx.row;
  if (r < D) {
    const size_t o = (size_t)k * D + r;
    store_dist<S, D>(out.A + o * D, A, cx.sub);
    store_dist<S, D>(out.U + o * D, U, cx.sub);
    store_dist<S, D>(out.Z + o * D, Z, cx.sub);
    if (cx.sub == 0) {
      out.b[o] = b;
      out.eta[o] = eta;
    }
  }
}

// Waves per SIMD the up- and down-sweep combines are compiled for: 5 at D = 10,
// the natural allocation elsewhere.  A 4-wave budget (128 VGPRs) for the D = 12
// up-sweep or the D = 15 down-sweep spills (24 and 44 bytes per lane) and is no
// faster than their natural 3 waves.
template <int D>
constexpr int up_min_waves() {
  return (D == 10) ? 5 : 1;
}
template <int D>
constexpr int down_min_waves() {
  return (D == 10) ? 5 : 1;
}

// Waves per SIMD of a scan kernel of kind K: the tuned bound of the
// latency-oriented kinds; at least 2 for the throughput kinds, whose natural
// allocation can exceed half the register file (D = 15: 256 VGPRs + 4 AGPRs,
// one wave per SIMD).
template <int D, int K>
constexpr int scan_min_waves() {
  return (K == kLdsFull) ? up_min_waves<D>() : (K == kLdsState) ? down_min_waves<D>() : is_throughput_kind(K) ? 2 : 1;
}

// Up-sweep: out[i] = in[2i] o in[2i+1]  (K: kLdsFull, or kLdsFullT on wide levels)
template <int D, int K>
__global__ __launch_bounds__((Geo<D, K>::BLOCK), (scan_min_waves<D, K>())) void k_filter_up(GroupArgs g, LaunchMap lm,
                                                                                            int l) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, K>(g.F, g.LQ, g.N1z, smem);
  int j, i;
  if (!locate(lm, cx.wid, j, i)) return;
  const long long N = g.Ns[j], base = g.base[j];
  const FiltArrays in = at(g.fel, base + lev_off(N, l), D);
  const FiltArrays out = at(g.fel, base + lev_off(N, l + 1), D);
  StateArrays none{};
  filter_combine<D, false>(cx, in, none, 2 * (size_t)i, in, 2 * (size_t)i + 1, out, none, (size_t)i);
}

// Down-sweep of level l: positions 2i and 2i+1 of the level-l prefix states.
//   p[0] = state(e[0]); p[2i+1] = up[i]; p[2i] = up[i-1] o e[2i]  (i >= 1)
// With top != 0 (a level of one element) only p[0] = state(e[0]) is written.
template <int D, int K>
__global__ __launch_bounds__((Geo<D, K>::BLOCK), (scan_min_waves<D, K>())) void k_filter_down(GroupArgs g,
                                                                                            LaunchMap lm, int l) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, K>(g.F, g.LQ, g.N1z, smem);
  int j, i;
  if (!locate(lm, cx.wid, j, i)) return;
  const long long N = g.Ns[j], base = g.base[j];
  const int n = (int)(N >> l);
  const FiltArrays el = at(g.fel, base + lev_off(N, l), D);
  const StateArrays p = at(g.fpre, base + lev_off(N, l), D);
  const StateArrays up = (l + 1 < n_levels(N)) ? at(g.fpre, base + lev_off(N, l + 1), D) : StateArrays{};
  constexpr int S = decltype(cx)::S, NL = Lay<D, S>::NL;
  const int j0 = 2 * i;
  const int r = cx.row, sb = cx.sub;
  if (j0 == 0) {
    if (r < D) {
      double u[NL];
      load_dist<S, D>(u, el.U + r * D, sb);
      store_dist<S, D>(p.L + r * D, u, sb);
      if (sb == 0) p.m[r] = el.b[r];
    }
  } else {
    FiltArrays none{};
    filter_combine<D, true>(cx, none, up, (size_t)i - 1, el, (size_t)j0, none, p, (size_t)j0);
  }
  const int j1 = j0 + 1;
  if (j1 < n && r < D) {
    double u[NL];
    load_dist<S, D>(u, up.L + ((size_t)i * D + r) * D, sb);
    store_dist<S, D>(p.L + ((size_t)j1 * D + r) * D, u, sb);
    if (sb == 0) p.m[(size_t)j1 * D + r] = up.m[(size_t)i * D + r];
  }
}

// Per filtered state s = 0..Ns: smoother element (smoother.py:37-50) stored at
// reversed index Ns - s (terminal element (m, 0, L) for s = Ns), and for
// s < Ns the predicted observation statistics of filter.py:84-114:
//   S = tria([H_s tria([F L_s | LQ]) | 0]), yhat = H_s F m_s + c_s,
//   ssq_s = |S^-T yhat|^2 (utils.py:110-112), nll_s = -loglik(yhat, S) (utils.py:22-30).
// The smoother factor D = Phi22 comes from the last D pivots of the Phi
// factorisation and is formed only when the IVP's smoothed factors are wanted
// this iteration (wchol).  mode kPostDm recomputes just D, for the factor pass
// of an IVP whose last iteration carried means only; kPostNoObs forms the
// smoother elements without observation statistics (smoother on its own).
enum PostMode : int { kPostFull = 0, kPostDm = 1, kPostNoObs = 2 };
template <int D, int d, bool EK1>
__global__ __launch_bounds__((Geo<D, kLdsPost>::BLOCK), (post_min_waves<D, EK1>())) void k_filter_post(GroupArgs g,
                                                                                                      LaunchMap lm,
                                                                                                      int mode) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, kLdsPost>(g.F, g.LQ, g.N1z, smem);
  using L = Lay<D, decltype(cx)::S>;
  constexpr int S = L::S, LD = L::LD, NL = L::NL;
  constexpr int NLd = padded<d, S>() / S;
  int j, s;
  if (!locate(lm, cx.wid, j, s)) return;
  const ModelArgs a = view<D>(g, j);
  const long long base = g.base[j];
  const StateArrays pre0 = at(g.fpre, base, D);
  const SmoArrays sel = at(g.sel, base, D);
  double* t_ssq = g.t_ssq + base;
  double* t_nll = g.t_nll + base;
  const int Ns = a.Ns;
  const int r = cx.row, sb = cx.sub;
  const bool act = r < D;
  const int rr = act ? r : 0;
  const double* mp = (s == 0) ? a.x0m : pre0.m + (size_t)(s - 1) * D;
  const double* Lp = (s == 0) ? a.x0L : pre0.L + (size_t)(s - 1) * D * D;
  const size_t ro = (size_t)(Ns - s);
  const bool dm_only = mode == kPostDm;
  const bool wc = dm_only || g.wchol[j] != 0;
  if (s == Ns) {
    if (act) {
      double lr[NL], zr[NL];
      load_dist<S, D>(lr, Lp + r * D, sb);
      zero(zr);
      store_dist<S, D>(sel.Dm + (ro * D + r) * D, lr, sb);
      store_dist<S, D>(sel.E + (ro * D + r) * D, zr, sb);
      if (sb == 0) sel.g[ro * D + r] = mp[r];
    }
    return;
  }
  // two tiles: L (dead once Phi is formed) then E; Phi11 (dead once S's rows
  // are formed) then S.  D = Phi22 is stored from the lanes that hold it.
  double* TL = cx.tile(0);
  double* TP11 = cx.tile(1);
  double* TE = TL;
  double* TS = TP11;
  double* Vm = cx.vec(0);
  double* Vfm = cx.vec(1);
  // global reads first: the filtered state and the linearisation point of obs model s
  Obs<D, d, EK1> ob;
  if (mode == kPostFull) ob.load(a, s);
  {
    double lr[NL];
    load_dist<S, D>(lr, Lp + rr * D, sb);
    const double mr = mp[rr];
    if (act) {
#pragma unroll
      for (int t = 0; t < NL; ++t) TL[r * LD + t * S + sb] = lr[t];
      Vm[r] = mr;
    }
  }
  wsync();
  // Phi_ = [[F L, LQ], [L, 0]]: rows 0..2D-1, column blocks (D, D)
  double ph[2 * NL];
  {
    double fr[D];
    load_row(fr, cx.sF + rr * LD);
    double t1[NL];
    row_times_tile<S, D, D, LD>(fr, TL, sb, t1);
    double fm = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) fm = fma(fr[k], Vm[k], fm);
    if (act) Vfm[r] = fm;
    double lq_r[NL], lr[NL];
    load_dist<S, D>(lq_r, cx.sLQ + rr * LD, sb);
    const int r2 = (r >= D && r < 2 * D) ? r - D : 0;
    load_dist<S, D>(lr, TL + r2 * LD, sb);
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      ph[t] = act ? t1[t] : ((r < 2 * D) ? lr[t] : 0.0);
      ph[NL + t] = act ? lq_r[t] : 0.0;
    }
  }
  // second block [LQ; 0]: banded when LQ is lower triangular (the IWP prior's
  // Cholesky factor; a caller's model is checked on the host)
  if (g.lq_lower) {
    lq<S, 2 * D, D, D, 1, 0, D, false, 0, L::W>(ph, r, sb, cx.bc());
  } else {
    lq<S, 2 * D, D, D, 1, 0, D, false, 1 << 20, L::W>(ph, r, sb, cx.bc());
  }
  if (wc) {
    lq<S, 2 * D, D, D, 1, D, 1 << 20, true, 1 << 20, L::W>(ph, r, sb, cx.bc());  // Phi22
  } else {
    lq<S, 2 * D, D, D, 1, D, D, true, 1 << 20, L::W>(ph, r, sb, cx.bc());  // no pivot: clear the strict upper part
  }
  if (dm_only) {
    if (r >= D && r < 2 * D) {
      double dd[NL];
#pragma unroll
      for (int t = 0; t < NL; ++t) dd[t] = ph[NL + t];
      store_dist<S, D>(sel.Dm + (ro * D + (r - D)) * D, dd, sb);
    }
    return;
  }
  if (act) {
#pragma unroll
    for (int t = 0; t < NL; ++t) TP11[r * LD + t * S + sb] = ph[t];
  }
  wsync();
  // E row = Phi21 row Phi11^-1 (moved to row r - D through LDS); D = Phi22,
  // stored when the smoothed factors are wanted
  {
    double e[NL];
#pragma unroll
    for (int t = 0; t < NL; ++t) e[t] = ph[t];
    rsolve_lower_s<S, D, LD>(e, TP11, sb);
    if (r >= D && r < 2 * D) {
#pragma unroll
      for (int t = 0; t < NL; ++t) TE[(r - D) * LD + t * S + sb] = e[t];
      if (wc) {
        double dd[NL];
#pragma unroll
        for (int t = 0; t < NL; ++t) dd[t] = ph[NL + t];
        store_dist<S, D>(sel.Dm + (ro * D + (r - D)) * D, dd, sb);
      }
    }
  }
  wsync();
  // observation statistics with obs model s (the smoother element stays in
  // LDS until the end so that no store is outstanding during the LQ below)
  if (mode == kPostNoObs) {
    if (r == 0 && sb == 0) {
      t_ssq[s] = 0.0;
      t_nll[s] = 0.0;
    }
  } else {
  double sr[NL + NLd];
  {
    double t1[NL];
    ob.template row_times_tile<S, D, LD>(r, TP11, sb, t1);
#pragma unroll
    for (int t = 0; t < NL; ++t) sr[t] = (r < d) ? t1[t] : 0.0;
#pragma unroll
    for (int t = 0; t < NLd; ++t) sr[NL + t] = (r < d) ? ob.r_at(r, t * S + sb) : 0.0;
  }
  lq<S, d, D, d, 1, 0, 1 << 20, true, 1 << 20, L::W>(sr, r, sb, cx.bc());
  if (r < d) store_dist<S, d>(TS + r * LD, sr, sb);
  wsync();
  if (r == 0 && sb == 0) {
    double yh[d];
#pragma unroll
    for (int i = 0; i < d; ++i) yh[i] = ob.c[i] + ob.dot(i, Vfm, 1);
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
  }
  if (act) {
    double e[NL];
    load_dist<S, D>(e, TE + r * LD, sb);
    double gg = Vm[r];
#pragma unroll
    for (int k = 0; k < D; ++k) gg = fma(-TE[r * LD + k], Vfm[k], gg);
    if (sb == 0) sel.g[ro * D + r] = gg;
    store_dist<S, D>(sel.E + (ro * D + r) * D, e, sb);
  }
  wsync();
}

// Smoother up-sweep
template <int D, int K>
__global__ __launch_bounds__((Geo<D, K>::BLOCK)) void k_smooth_up(GroupArgs g, LaunchMap lm, int l) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, K>(g.F, g.LQ, g.N1z, smem);
  int j, i;
  if (!locate(lm, cx.wid, j, i)) return;
  const long long n0 = (long long)g.Ns[j] + 1, base = g.base[j];
  const SmoArrays in = at(g.sel, base + lev_off(n0, l), D);
  const SmoArrays out = at(g.sel, base + lev_off(n0, l + 1), D);
  const int with_chol = g.wchol[j];
  StateArrays none{};
  smoother_combine<D, false>(cx, in, none, 2 * (size_t)i, in, 2 * (size_t)i + 1, out, none, (size_t)i,
                             with_chol != 0);
}

template <int D, int K>
__global__ __launch_bounds__((Geo<D, K>::BLOCK)) void k_smooth_down(GroupArgs g, LaunchMap lm, int l) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, K>(g.F, g.LQ, g.N1z, smem);
  int j, i;
  if (!locate(lm, cx.wid, j, i)) return;
  const long long n0 = (long long)g.Ns[j] + 1, base = g.base[j];
  const int n = (int)(n0 >> l);
  const SmoArrays el = at(g.sel, base + lev_off(n0, l), D);
  const StateArrays p = at(g.spre, base + lev_off(n0, l), D);
  const StateArrays up = (l + 1 < n_levels(n0)) ? at(g.spre, base + lev_off(n0, l + 1), D) : StateArrays{};
  const int with_chol = g.wchol[j];
  constexpr int S = decltype(cx)::S, NL = Lay<D, S>::NL;
  const int j0 = 2 * i;
  const int r = cx.row, sb = cx.sub;
  if (j0 == 0) {
    if (r < D) {
      if (sb == 0) p.m[r] = el.g[r];
      if (with_chol) {
        double u[NL];
        load_dist<S, D>(u, el.Dm + r * D, sb);
        store_dist<S, D>(p.L + r * D, u, sb);
      }
    }
  } else {
    SmoArrays none{};
    smoother_combine<D, true>(cx, none, up, (size_t)i - 1, el, (size_t)j0, none, p, (size_t)j0, with_chol != 0);
  }
  const int j1 = j0 + 1;
  if (j1 < n && r < D) {
    if (sb == 0) p.m[(size_t)j1 * D + r] = up.m[(size_t)i * D + r];
    if (with_chol) {
      double u[NL];
      load_dist<S, D>(u, up.L + ((size_t)i * D + r) * D, sb);
      store_dist<S, D>(p.L + ((size_t)j1 * D + r) * D, u, sb);
    }
  }
}

// Thread t: smoothed mean m_t (reversed prefix Ns - t) -> means_new[t];
// obj_t = |LQ^-1 (m_t - F m_{t+1})|^2 (utils.py:97-101 with the argument order
// of smoother.py:20); viol_t = 1 if any |m_old - m_new| > 1e-8 + 1e-13 |m_new|
// (convergence_criteria.py:10, jnp.isclose).
template <int D>
__global__ void k_smooth_post(GroupArgs g, LaunchMap lm) {
  int j, t;
  if (!locate(lm, blockIdx.x * blockDim.x + threadIdx.x, j, t)) return;
  const int Ns = g.Ns[j];
  const long long base = g.base[j];
  const StateArrays spre0 = at(g.spre, base, D);
  const double* mold = (g.cur[j] ? g.means1 : g.means0) + (size_t)base * D;
  double* mnew = (g.cur[j] ? g.means0 : g.means1) + (size_t)base * D;
  double* t_obj = g.t_obj + base;
  double* t_viol = g.t_viol + base;
  const double* F = g.F;
  const double* LQ = g.LQ;
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

// Fixed-order reduction of the per-step terms of each IVP, stage 1: block b
// (chunk c of IVP j) sums the time points [c kChunk, (c + 1) kChunk) into
// part[b * 4 + 0..3] (ssq, nll, obj: sums over t < Ns; violations: max over t <= Ns).
static __global__ void k_reduce_part(GroupArgs g, LaunchMap lm) {
  __shared__ double sh[4][256];
  const int tid = threadIdx.x;
  int j, c;
  if (!locate(lm, blockIdx.x, j, c)) return;
  const int Ns = g.Ns[j];
  const long long base = g.base[j];
  const int lo = c * kChunk;
  const int hi = min(lo + kChunk, Ns + 1);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int i = lo + tid; i < hi; i += 256) {
    if (i < Ns) {
      s0 += g.t_ssq[base + i];
      s1 += g.t_nll[base + i];
      s2 += g.t_obj[base + i];
    }
    s3 = fmax(s3, g.t_viol[base + i]);
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
  if (tid < 4) g.part[(size_t)blockIdx.x * 4 + tid] = sh[tid][0];
}

// Stage 2, block a per active IVP: its chunks in order.  scal[a * 8 + k]:
// 0 nll, 1 obj, 2 ssq (sum / (Ns d)), 3 = 1 if all means are close, 4 = NaN flag.
static __global__ void k_reduce_final(GroupArgs g, LaunchMap lm, int d, double* scal) {
  const int a = blockIdx.x;
  if (a >= lm.m || threadIdx.x != 0) return;
  const int Ns = g.Ns[lm.ivp[a]];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int b = lm.pre[a]; b < lm.pre[a + 1]; ++b) {
    s0 += g.part[(size_t)b * 4 + 0];
    s1 += g.part[(size_t)b * 4 + 1];
    s2 += g.part[(size_t)b * 4 + 2];
    s3 = fmax(s3, g.part[(size_t)b * 4 + 3]);
  }
  double* o = scal + (size_t)a * 8;
  o[0] = s1;
  o[1] = s2;
  o[2] = s0 / Ns / d;
  o[3] = (s3 == 0.0) ? 1.0 : 0.0;
  o[4] = (isnan(s1) || isnan(s2)) ? 1.0 : 0.0;
}

// Output projection: means_out[t][i] = P_ii0 m_t[i0], chol_out[t][i][:] = scale * P_ii0 L_t[i0][:]
// with i0 = i (q+1) (E0 = E0_raw P, pof/convenience.py:26-27; solver.py:71).
template <int D, int d>
__global__ void k_output(GroupArgs g, LaunchMap lm) {
  constexpr int q1 = D / d;
  int j, t;
  if (!locate(lm, blockIdx.x * blockDim.x + threadIdx.x, j, t)) return;
  const int Ns = g.Ns[j];
  const StateArrays spre0 = at(g.spre, g.base[j], D);
  const double* Pd = g.Pd + (size_t)j * D;
  const double scale = g.out_scale[j];
  double* means_out = g.out_means[j];
  double* chol_out = g.out_chols[j];
  double* full_means = g.out_full_means[j];
  double* full_chols = g.out_full_chols[j];
  const bool pfull = g.out_pfull[j] != 0;
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
    for (int i = 0; i < D; ++i) full_means[(size_t)t * D + i] = pfull ? Pd[i] * m[i] : m[i];
  }
  if (full_chols) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const double f = pfull ? scale * Pd[i] : scale;
#pragma unroll
      for (int c = 0; c < D; ++c) full_chols[((size_t)t * D + i) * D + c] = f * L[i * D + c];
    }
  }
}

// Filtered states (state s = x0 for s = 0, else level-0 filter prefix s-1).
template <int D>
__global__ void k_filtered_out(GroupArgs g, LaunchMap lm, double* fm, double* fL) {
  int j, s;
  if (!locate(lm, blockIdx.x * blockDim.x + threadIdx.x, j, s)) return;
  const double* x0m = g.x0m + (size_t)j * D;
  const double* x0L = g.x0L + (size_t)j * D * D;
  const StateArrays fpre0 = at(g.fpre, g.base[j], D);
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
  Obs<D, d, true> ob;
  ob.load(b, k);
  double Hk[d][D];
  ob.dense(Hk);
  const double (&ck)[d] = ob.c;
#pragma unroll
  for (int i = 0; i < d; ++i) {
#pragma unroll
    for (int j = 0; j < D; ++j) H[((size_t)k * d + i) * D + j] = Hk[i][j];
    c[(size_t)k * d + i] = ck[i];
  }
}

// --------------------------------------------------------------- sequential
// Sequential square-root Kalman filter, one IVP per worker, one lane per row
// (pof/sequential_filtsmooth/filter.py:9-92): for k = 0..Ns-1, predict with
// (F, QL) (_sqrt_predict, 59-66), linearise (EK1 at lin[k+1]; at the
// predicted mean when AT_PRED, the EKF of sequential_eks_solve, filter.py:9-30;
// or the explicit affine model), update (_sqrt_update, 69-92).  Filtered state
// k+1 goes to level 0 of the filter prefix arrays (entry k), the step's
// |I^-T y|^2 and log-likelihood to t_ssq[k] and t_nll[k].
template <int D, int d, bool EK1, bool AT_PRED>
__global__ __launch_bounds__((Geo<D, kLdsSeq>::BLOCK)) void k_seq_filter(GroupArgs g, LaunchMap lm) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, kLdsSeq>(g.F, g.LQ, g.N1z, smem);
  constexpr int LD = Lay<D, 1>::LD;
  int j, i0;
  if (!locate(lm, cx.wid, j, i0)) return;
  const ModelArgs a = view<D>(g, j);
  const long long base = g.base[j];
  const StateArrays fp = at(g.fpre, base, D);
  double* t_ssq = g.t_ssq + base;
  double* t_nll = g.t_nll + base;
  const int Ns = a.Ns;
  const int r = cx.row;
  const bool act = r < D;
  const int rr = act ? r : 0;
  double* TL = cx.tile(0);  // L_k, then L_{k+1}
  double* TP = cx.tile(1);  // predicted factor
  double* TI = cx.tile(2);  // I = S[:d, :d]
  double* TG = cx.tile(3);  // G = S[d:, :d]
  double* Vm = cx.vec(0);   // m_k
  double* Vp = cx.vec(1);   // predicted mean
  if (act) {
#pragma unroll
    for (int c = 0; c < D; ++c) TL[r * LD + c] = a.x0L[r * D + c];
    Vm[r] = a.x0m[r];
  }
  wsync();
  for (int k = 0; k < Ns; ++k) {
    // predict: mp = F m, Lp = tria([F L | QL])
    {
      double fr[D];
      load_row(fr, cx.sF + rr * LD);
      double fl[D];
      row_times_tile<1, D, D, LD>(fr, TL, 0, fl);
      double mp = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) mp = fma(fr[c], Vm[c], mp);
      double row[2 * D];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        row[c] = act ? fl[c] : 0.0;
        row[D + c] = act ? cx.sLQ[rr * LD + c] : 0.0;
      }
      lq<1, D, D, D>(row, r, 0, cx.bc());
      if (act) {
#pragma unroll
        for (int c = 0; c < D; ++c) TP[r * LD + c] = row[c];
        Vp[r] = mp;
      }
    }
    wsync();
    Obs<D, d, EK1> ob;
    if constexpr (AT_PRED) {
      ob.load_at(a, Vp);
    } else {
      ob.load(a, k);
    }
    // update: S = tria([[H Lp, R], [Lp, 0]]) (R = 0 for EK1); column blocks (D, d)
    double psi[D + d];
    {
      double hn[D];
      ob.template row_times_tile<1, D, LD>(r, TP, 0, hn);
      const int rn = (r >= d && r < d + D) ? r - d : 0;
#pragma unroll
      for (int c = 0; c < D; ++c) psi[c] = (r < d) ? hn[c] : ((r < d + D) ? TP[rn * LD + c] : 0.0);
#pragma unroll
      for (int c = 0; c < d; ++c) psi[D + c] = (r < d) ? ob.r_at(r, c) : 0.0;
    }
    lq<1, d + D, D, d>(psi, r, 0, cx.bc());
    if (r < d) {
#pragma unroll
      for (int c = 0; c < d; ++c) TI[r * LD + c] = psi[c];
    } else if (r < d + D) {
#pragma unroll
      for (int c = 0; c < d; ++c) TG[(r - d) * LD + c] = psi[c];
#pragma unroll
      for (int c = d; c < D + d; ++c) TL[(r - d) * LD + (c - d)] = psi[c];
    }
    wsync();
    // y = 0 - (H mp + c); m_{k+1} = mp + G I^-1 y
    double z[d], w[d];
#pragma unroll
    for (int i = 0; i < d; ++i) {
      z[i] = -(ob.c[i] + ob.dot(i, Vp, 1));
      w[i] = z[i];
    }
    lsolve_lower<d, LD>(z, TI);   // I z = y
    lsolve_lowerT<d, LD>(w, TI);  // I^T w = y (whiten)
    double mn = Vp[rr];
#pragma unroll
    for (int i = 0; i < d; ++i) mn = fma(TG[rr * LD + i], z[i], mn);
    wsync();
    if (act) {
      Vm[r] = mn;
      fp.m[(size_t)k * D + r] = mn;
#pragma unroll
      for (int c = 0; c < D; ++c) fp.L[((size_t)k * D + r) * D + c] = TL[r * LD + c];
    }
    if (r == 0) {
      double ss = 0.0, nn = 0.0, ld = 0.0;
#pragma unroll
      for (int i = 0; i < d; ++i) {
        ss = fma(w[i], w[i], ss);
        nn = fma(z[i], z[i], nn);
        ld += log(fabs(TI[i * LD + i]));
      }
      t_ssq[k] = ss;
      // mvn_loglikelihood (utils.py:22-30): -|I^-1 y|^2 / 2 - sum log|I_ii| - d log(2 pi) / 2
      t_nll[k] = -0.5 * nn - (ld + d * log(2.0 * M_PI) / 2.0);
    }
    wsync();
  }
}

// Sequential square-root RTS smoother, one IVP per worker
// (pof/sequential_filtsmooth/smoother.py:8-48, _sqrt_smooth 31-48), backward
// over the filtered states of k_seq_filter.  Smoothed state t goes to level 0
// of the smoother prefix arrays at entry Ns - t (the parallel smoother's
// layout, read by k_smooth_post and k_output); factors only when wchol is set.
template <int D>
__global__ __launch_bounds__((Geo<D, kLdsSeq>::BLOCK)) void k_seq_smooth(GroupArgs g, LaunchMap lm) {
  extern __shared__ double smem[];
  auto cx = make_ctx<D, kLdsSeq>(g.F, g.LQ, g.N1z, smem);
  constexpr int LD = Lay<D, 1>::LD;
  int j, i0;
  if (!locate(lm, cx.wid, j, i0)) return;
  const long long base = g.base[j];
  const int Ns = g.Ns[j];
  const bool wc = g.wchol[j] != 0;
  const StateArrays fp = at(g.fpre, base, D);
  const StateArrays sp = at(g.spre, base, D);
  const double* x0m = g.x0m + (size_t)j * D;
  const double* x0L = g.x0L + (size_t)j * D * D;
  const int r = cx.row;
  const bool act = r < D;
  const int rr = act ? r : 0;
  double* TLs = cx.tile(0);  // smoothed factor of state k+1
  double* TP11 = cx.tile(1);
  double* TE = cx.tile(2);   // gain
  double* TD = cx.tile(3);   // Phi22
  double* TF = cx.tile(4);   // filtered factor of state k
  double* Vs = cx.vec(0);    // smoothed mean of state k+1
  double* Vf = cx.vec(1);    // filtered mean of state k
  double* Vfm = cx.vec(2);   // F m_f
  // state Ns: smoothed = filtered
  {
    const double* m = (Ns == 0) ? x0m : fp.m + (size_t)(Ns - 1) * D;
    const double* L = (Ns == 0) ? x0L : fp.L + (size_t)(Ns - 1) * D * D;
    if (act) {
      const double mv = m[r];
      Vs[r] = mv;
      sp.m[r] = mv;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double v = L[r * D + c];
        TLs[r * LD + c] = v;
        if (wc) sp.L[r * D + c] = v;
      }
    }
  }
  wsync();
  for (int k = Ns - 1; k >= 0; --k) {
    const double* mf = (k == 0) ? x0m : fp.m + (size_t)(k - 1) * D;
    const double* Lf = (k == 0) ? x0L : fp.L + (size_t)(k - 1) * D * D;
    if (act) {
      Vf[r] = mf[r];
#pragma unroll
      for (int c = 0; c < D; ++c) TF[r * LD + c] = Lf[r * D + c];
    }
    wsync();
    // Phi = tria([[F Lf, QL], [Lf, 0]])
    double ph[2 * D];
    {
      double fr[D];
      load_row(fr, cx.sF + rr * LD);
      double t1[D];
      row_times_tile<1, D, D, LD>(fr, TF, 0, t1);
      double fm = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) fm = fma(fr[c], Vf[c], fm);
      if (act) Vfm[r] = fm;
      const int r2 = (r >= D && r < 2 * D) ? r - D : 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        ph[c] = act ? t1[c] : ((r < 2 * D) ? TF[r2 * LD + c] : 0.0);
        ph[D + c] = act ? cx.sLQ[rr * LD + c] : 0.0;
      }
    }
    lq<1, 2 * D, D, D>(ph, r, 0, cx.bc());
    if (act) {
#pragma unroll
      for (int c = 0; c < D; ++c) TP11[r * LD + c] = ph[c];
    }
    wsync();
    // gain = Phi21 Phi11^-1, Phi22 (rows D..2D-1 move to r - D)
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
    // mean = mf + gain (ms - F mf); chol = tria([Phi22 | gain Ls])
    double mean = Vf[rr];
#pragma unroll
    for (int c = 0; c < D; ++c) mean = fma(TE[rr * LD + c], Vs[c] - Vfm[c], mean);
    double cr[2 * D];
    if (wc) {
      double ge[D];
      load_row(ge, TE + rr * LD);
      double t1[D];
      row_times_tile<1, D, D, LD>(ge, TLs, 0, t1);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        cr[c] = act ? TD[rr * LD + c] : 0.0;
        cr[D + c] = act ? t1[c] : 0.0;
      }
      lq<1, D, D, D>(cr, r, 0, cx.bc());
    }
    wsync();
    if (act) {
      Vs[r] = mean;
      sp.m[(size_t)(Ns - k) * D + r] = mean;
      if (wc) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          TLs[r * LD + c] = cr[c];
          sp.L[((size_t)(Ns - k) * D + r) * D + c] = cr[c];
        }
      }
    }
    wsync();
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

// Launch one scan level: the throughput kind KT when the level has at least
// `thr` work items (and the kinds differ at this D), else the latency kind KL.
#define POF_LAUNCH_LEVEL(kernel, KL, KT, thr, nw, ...)                                        \
  do {                                                                                          \
    if constexpr (has_throughput_kinds<D>()) {                                                  \
      if ((nw) >= (thr)) {                                                                      \
        POF_LAUNCH((kernel<D, KT>), KT, nw, __VA_ARGS__);                                       \
        break;                                                                                  \
      }                                                                                         \
    }                                                                                           \
    POF_LAUNCH((kernel<D, KL>), KL, nw, __VA_ARGS__);                                           \
  } while (0)

template <int D>
void smooth_scan(const GroupArgs& g, const LaunchPlan& p, hipStream_t st, Rec& rec) {
  for (int l = 0; l + 1 < p.sl; ++l) {
    const int nw = p.sup_total[l];
    if (nw == 0) continue;
    POF_LAUNCH_LEVEL(k_smooth_up, kLdsSmooth, kLdsSmoothT, p.tune.smooth_up, nw, st, g, p.map(p.sup_off[l], nw), l);
    rec(kKSmoothUp, nw);
  }
  for (int l = p.sl - 1; l >= 0; --l) {
    const int nw = p.sdn_total[l];
    if (nw == 0) continue;
    POF_LAUNCH_LEVEL(k_smooth_down, kLdsSmooth, kLdsSmoothT, p.tune.smooth_down, nw, st, g, p.map(p.sdn_off[l], nw),
                     l);
    rec(kKSmoothDown, nw - p.sdn_ivps[l]);
  }
}

template <int D, int d>
hipError_t run_smooth_chol(const GroupArgs& g, const LaunchPlan& p, hipStream_t st, KernelTimer* t) {
  Rec rec(t, st);
  // the smoother elements' D of the last (means-only) iteration, then the factor scan
  if (g.affine) {
    POF_LAUNCH((k_filter_post<D, d, false>), kLdsPost, p.pts_total, st, g, p.map(p.pts_off, p.pts_total), kPostDm);
  } else {
    POF_LAUNCH((k_filter_post<D, d, true>), kLdsPost, p.pts_total, st, g, p.map(p.pts_off, p.pts_total), kPostDm);
  }
  rec(kKFilterPost, p.leaf_total);
  smooth_scan<D>(g, p, st, rec);
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_iteration(const GroupArgs& g, const LaunchPlan& p, double* scal, hipStream_t st, KernelTimer* t) {
  Rec rec(t, st);
  // ---- filter: leaf elements, up-sweep, down-sweep
  if (g.affine) {
    POF_LAUNCH((k_filter_leaf<D, d, false, true>), kLdsLeaf1, p.one_total, st, g, p.map(p.one_off, p.one_total));
    if (p.leaf1_total > 0)
      POF_LAUNCH((k_filter_leaf<D, d, false, false>), kLdsLeaf, p.leaf1_total, st, g,
                 p.map(p.leaf1_off, p.leaf1_total));
  } else {
    POF_LAUNCH((k_filter_leaf<D, d, true, true>), kLdsLeaf1, p.one_total, st, g, p.map(p.one_off, p.one_total));
    if (p.leaf1_total > 0)
      POF_LAUNCH((k_filter_leaf<D, d, true, false>), kLdsLeaf, p.leaf1_total, st, g,
                 p.map(p.leaf1_off, p.leaf1_total));
  }
  rec(kKFilterLeaf, p.leaf_total);
  for (int l = 0; l + 1 < p.fl; ++l) {
    const int nw = p.fup_total[l];
    if (nw == 0) continue;
    POF_LAUNCH_LEVEL(k_filter_up, kLdsFull, kLdsFullT, p.tune.filter_up, nw, st, g, p.map(p.fup_off[l], nw), l);
    rec(kKFilterUp, nw);
  }
  for (int l = p.fl - 1; l >= 0; --l) {
    const int nw = p.fdn_total[l];
    if (nw == 0) continue;
    POF_LAUNCH_LEVEL(k_filter_down, kLdsState, kLdsStateT, p.tune.filter_down, nw, st, g, p.map(p.fdn_off[l], nw),
                     l);
    rec(kKFilterDown, nw - p.fdn_ivps[l]);
  }
  // ---- smoother elements + observation statistics
  if (g.affine) {
    POF_LAUNCH((k_filter_post<D, d, false>), kLdsPost, p.pts_total, st, g, p.map(p.pts_off, p.pts_total), kPostFull);
  } else {
    POF_LAUNCH((k_filter_post<D, d, true>), kLdsPost, p.pts_total, st, g, p.map(p.pts_off, p.pts_total), kPostFull);
  }
  rec(kKFilterPost, p.leaf_total);
  // ---- smoother scan (reversed order)
  smooth_scan<D>(g, p, st, rec);
  // ---- smoothed means, objective, convergence partials, per-IVP reduction
  hipLaunchKernelGGL((k_smooth_post<D>), dim3((p.pts_total + 127) / 128), dim3(128), 0, st, g,
                     p.map(p.pts_off, p.pts_total));
  rec(kKSmoothPost, p.pts_total);
  hipLaunchKernelGGL(k_reduce_part, dim3(p.chunk_total), dim3(256), 0, st, g, p.map(p.chunk_off, p.chunk_total));
  hipLaunchKernelGGL(k_reduce_final, dim3(p.m), dim3(64), 0, st, g, p.map(p.chunk_off, p.chunk_total), d, scal);
  rec(kKReduce, p.leaf_total);
  return hipGetLastError();
}

// One sequential IEKS iteration (pof/step.py:33-45 with sequential=True) or,
// with at_pred, the single pass of sequential_eks_solve (pof/solver.py:76-96):
// sequential filter and smoother, then the smoothed means, objective,
// convergence partials and per-IVP reduction of run_iteration.
template <int D, int d>
hipError_t run_seq_iteration(const GroupArgs& g, const LaunchPlan& p, double* scal, int at_pred, hipStream_t st,
                             KernelTimer* t) {
  Rec rec(t, st);
  const LaunchMap one = p.map(p.one_off, p.one_total);
  if (g.affine) {
    POF_LAUNCH((k_seq_filter<D, d, false, false>), kLdsSeq, p.one_total, st, g, one);
  } else if (at_pred) {
    POF_LAUNCH((k_seq_filter<D, d, true, true>), kLdsSeq, p.one_total, st, g, one);
  } else {
    POF_LAUNCH((k_seq_filter<D, d, true, false>), kLdsSeq, p.one_total, st, g, one);
  }
  rec(kKFilterLeaf, p.leaf_total);
  POF_LAUNCH((k_seq_smooth<D>), kLdsSeq, p.one_total, st, g, one);
  rec(kKSmoothUp, p.leaf_total);
  hipLaunchKernelGGL((k_smooth_post<D>), dim3((p.pts_total + 127) / 128), dim3(128), 0, st, g,
                     p.map(p.pts_off, p.pts_total));
  rec(kKSmoothPost, p.pts_total);
  hipLaunchKernelGGL(k_reduce_part, dim3(p.chunk_total), dim3(256), 0, st, g, p.map(p.chunk_off, p.chunk_total));
  hipLaunchKernelGGL(k_reduce_final, dim3(p.m), dim3(64), 0, st, g, p.map(p.chunk_off, p.chunk_total), d, scal);
  rec(kKReduce, p.leaf_total);
  return hipGetLastError();
}

// The smoother alone over filtered states already in the filter prefix arrays
// (parallel: smoother elements, reverse scan; sequential: the RTS recursion),
// then the smoothed means, objective and reduction of run_iteration.
template <int D, int d>
hipError_t run_smooth_only(const GroupArgs& g, const LaunchPlan& p, double* scal, int sequential, hipStream_t st,
                           KernelTimer* t) {
  Rec rec(t, st);
  if (sequential) {
    POF_LAUNCH((k_seq_smooth<D>), kLdsSeq, p.one_total, st, g, p.map(p.one_off, p.one_total));
    rec(kKSmoothUp, p.leaf_total);
  } else {
    POF_LAUNCH((k_filter_post<D, d, false>), kLdsPost, p.pts_total, st, g, p.map(p.pts_off, p.pts_total),
               kPostNoObs);
    rec(kKFilterPost, p.leaf_total);
    smooth_scan<D>(g, p, st, rec);
  }
  hipLaunchKernelGGL((k_smooth_post<D>), dim3((p.pts_total + 127) / 128), dim3(128), 0, st, g,
                     p.map(p.pts_off, p.pts_total));
  rec(kKSmoothPost, p.pts_total);
  hipLaunchKernelGGL(k_reduce_part, dim3(p.chunk_total), dim3(256), 0, st, g, p.map(p.chunk_off, p.chunk_total));
  hipLaunchKernelGGL(k_reduce_final, dim3(p.m), dim3(64), 0, st, g, p.map(p.chunk_off, p.chunk_total), d, scal);
  rec(kKReduce, p.leaf_total);
  return hipGetLastError();
}

// Smoothed factors of the last sequential iteration (after means-only passes).
template <int D, int d>
hipError_t run_seq_smooth_chol(const GroupArgs& g, const LaunchPlan& p, hipStream_t st, KernelTimer* t) {
  Rec rec(t, st);
  POF_LAUNCH((k_seq_smooth<D>), kLdsSeq, p.one_total, st, g, p.map(p.one_off, p.one_total));
  rec(kKSmoothUp, p.leaf_total);
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_output(const GroupArgs& g, const LaunchPlan& p, hipStream_t st) {
  hipLaunchKernelGGL((k_output<D, d>), dim3((p.pts_total + 127) / 128), dim3(128), 0, st, g,
                     p.map(p.pts_off, p.pts_total));
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_filtered(const GroupArgs& g, const LaunchPlan& p, double* fm, double* fL, hipStream_t st) {
  hipLaunchKernelGGL((k_filtered_out<D>), dim3((p.pts_total + 127) / 128), dim3(128), 0, st, g,
                     p.map(p.pts_off, p.pts_total), fm, fL);
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_linearize(const ModelArgs& a, const double* m, int n, double* H, double* c, hipStream_t st) {
  hipLaunchKernelGGL((k_linearize<D, d>), dim3((n + 127) / 128), dim3(128), 0, st, a, m, n, H, c);
  return hipGetLastError();
}

template <int D, int d>
hipError_t run_linearize_alt(const ModelArgs& a, int method, const LinParams& p, const double* m, const double* L,
                             int n, double* H, double* c, double* R, double* work, hipStream_t st) {
  hipLaunchKernelGGL((k_linearize_alt<D, d>), dim3((n + 63) / 64), dim3(64), 0, st, a, method, p, m, L, n, H, c, R,
                     work);
  return hipGetLastError();
}

}  // namespace pof
