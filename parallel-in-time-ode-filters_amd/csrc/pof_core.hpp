// Device building blocks of the square-root parallel IEKS on CDNA (gfx950).
//
// Execution model: one "worker" of W lanes (32 or 64) owns one scan element.
// Row r of every matrix of the element is spread over S adjacent lanes
// (lanes r S .. r S + S - 1; S = 4 for D <= 8, 2 for D <= 16, else 1, in the
// scan kernels; S = 1 in the per-element kernels): the lane
// with sub-index s holds the columns c = t S + s, t = 0, 1, ... of the row
// in registers ("split rows").  Reductions along a row and broadcasts from
// one column owner to the others are DPP quad permutations; rows that other
// lanes need are staged through the worker's slice of LDS.  A worker never
// synchronises with other workers, only with its own wavefront.
//
// Rows of D (or d) columns are padded with zeros to a multiple of S.  A
// matrix built from two column blocks (e.g. [M | U2]) keeps each block
// padded; the Householder LQ maps pivot i to its padded column, so the
// padding changes no reflector.
//
// Arithmetic follows the reference (pof/parallel_filtsmooth/filter.py:50-142,
// smoother.py:37-63, utils.py:33-41) in float64.  tria() is a Householder LQ
// with LAPACK dgeqrf/dlarfg's sign convention (R_ii = -sign(alpha) * norm,
// tau = 0 when the row tail is zero), because the reference's
// sigma_squared (whiten() solves L^T r = m, pof/utils.py:110-112) depends on
// the column signs of the factors it produces.
#pragma once

#include <hip/hip_runtime.h>

#include "pof_problems.hpp"
#include "pof_types.hpp"

// No implicit contraction of a * b + c into fused multiply-adds in these
// building blocks: HIP compiles with -ffp-contract=fast by default, and which
// products the compiler then fuses depends on the shape of the surrounding
// code.  Every fused multiply-add here is an explicit fma(), so two code shapes
// that perform the same operations in the same order (the latency and
// throughput forms of a scan kernel) round identically.
#pragma clang fp contract(off)

namespace pof {

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  asm volatile("" ::: "memory");
}

// ------------------------------------------------------------------ layout
template <int n, int S>
constexpr int padded() {
  return (n + S - 1) / S * S;
}

// Widest row split for state dimension D: 2D rows of S lanes fit in 64 lanes.
template <int D>
constexpr int split_of() {
  return (8 * D <= 64) ? 4 : (4 * D <= 64) ? 2 : 1;
}

// Largest observation (ODE) dimension d <= 4 with d | D: the widest leaf
// element Z = [Zr | 0] (nonzero in its first d columns) an engine of state
// dimension D combines.
constexpr int max_obs_dim(int D) {
  return (D % 4 == 0) ? 4 : (D % 3 == 0) ? 3 : (D % 2 == 0) ? 2 : 1;
}

// Lanes per worker and the padded row length for state dimension D with
// rows split over S lanes.  W_ > 0 overrides the worker width (a power of two
// dividing 64, at least 2 D S).
template <int D_, int S_, int W_ = 0>
struct Lay {
  static constexpr int D = D_;
  static constexpr int S = S_;
  static constexpr int W = (W_ > 0) ? W_ : (2 * D_ * S <= 32) ? 32 : 64;
  static_assert(W >= 2 * D_ * S || W == 64, "a worker holds the 2 D rows of a combine");
  static constexpr int DP = padded<D_, S>();
  static constexpr int LD = DP + 1;  // LDS row stride (odd: no bank conflicts down a column)
  static constexpr int NL = DP / S;  // registers of a D-column row
};

// LDS layout kinds: how many tiles / broadcast vectors a worker of each
// kernel needs and whether the block stages the model constants F, LQ, N1z.
// The *T kinds are the throughput forms of the three scan combines: the same
// LDS needs, but one lane per matrix row and the narrowest worker that holds
// the 2D rows (several combines per wavefront), for scan levels wide enough
// to fill the chip, where the per-pivot work that every lane of a worker
// repeats, not the latency of one combine, sets the time.
enum LdsKind : int { kLdsFull = 0, kLdsState, kLdsSmooth, kLdsLeaf, kLdsPost, kLdsSeq, kLdsLeaf1, kLdsFullT,
                     kLdsStateT, kLdsSmoothT };

// the latency-oriented kind of a throughput kind (same LDS needs)
constexpr int base_kind(int K) {
  return (K == kLdsFullT) ? kLdsFull : (K == kLdsStateT) ? kLdsState : (K == kLdsSmoothT) ? kLdsSmooth : K;
}
constexpr bool is_throughput_kind(int K) { return K == kLdsFullT || K == kLdsStateT || K == kLdsSmoothT; }

// Widest observation dimension of the per-worker engines.
constexpr int kMaxObs = 4;

template <int K_>
struct LdsNeeds {
  static constexpr int K = base_kind(K_);
  static constexpr int NT = (K == kLdsFull) ? 5
                            : (K == kLdsState) ? 4
                            : (K == kLdsSeq) ? 5
                            : (K == kLdsPost) ? 2
                            : (K == kLdsLeaf) ? 1
                                              : 2;  // kLdsSmooth, kLdsLeaf1
  static constexpr int NV = (K == kLdsFull || K == kLdsState || K == kLdsSeq) ? ((K == kLdsFull) ? 8 : 4)
                            : (K == kLdsPost) ? 2
                                              : 1;
  static constexpr bool CONSTS = (K == kLdsLeaf || K == kLdsLeaf1 || K == kLdsPost || K == kLdsSeq);
  // the element kernels' small blocks after the tiles: Psi11 (at most kMaxObs
  // rows of stride LD) and the gain K (D rows of stride kMaxObs)
  static constexpr bool SMALL = (K == kLdsLeaf || K == kLdsLeaf1);
};

// Row split per kernel kind: the scan combines use the widest split (their
// work is per column); the per-element kernels, dominated by work every lane
// of a row repeats, keep one lane per row and two workers per wavefront.
template <int D, int K>
constexpr int kind_split() {
  return (K == kLdsLeaf || K == kLdsLeaf1 || K == kLdsPost || K == kLdsSeq || is_throughput_kind(K)) ? 1
                                                                                                     : split_of<D>();
}

// Worker width: the throughput kinds take the narrowest power of two (at least
// 8 lanes) that holds the 2D rows; the others the default of Lay.
constexpr int pow2_at_least(int n, int lo) { return (lo >= n) ? lo : pow2_at_least(n, 2 * lo); }
template <int D, int K>
constexpr int kind_width() {
  return is_throughput_kind(K) ? pow2_at_least(2 * D, 8) : 0;
}

// State dimensions whose scan combines form their D x D tile products on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64, one 16 x 16 tile): whole-wavefront workers (rows over two lanes,
// D > 8) that stage operands in one more LDS tile.  D = 9 .. 11 keep their throughput forms,
// which the autotuner picks on wide levels and whose results must stay bitwise equal to the
// latency form's (the matrix core sums in an order of its own).  At D = 15 the extra tile costs
// one of twelve resident up-sweep workers per CU (LDS), and the products save 14% of the
// up-sweep's instructions: the batch's D = 15 group gained (6.26 -> 6.16 s for the batch).
template <int D>
constexpr bool mfma_products() {
  return D == 12 || D == 15 || D == 16;
}

// 1 when the throughput kinds differ from the latency-oriented ones (rows split over several lanes)
template <int D>
constexpr bool has_throughput_kinds() {
  return split_of<D>() > 1 && !mfma_products<D>();
}

template <int D_, int K = kLdsFull>
struct Geo : Lay<D_, kind_split<D_, K>(), kind_width<D_, K>()> {
  using L = Lay<D_, kind_split<D_, K>(), kind_width<D_, K>()>;
  static constexpr int TILE = D_ * L::LD;
  // tiles per worker; the combines with matrix-core products stage operands and results in one more
  static constexpr int NT =
      LdsNeeds<K>::NT + ((mfma_products<D_>() && (base_kind(K) == kLdsFull || base_kind(K) == kLdsState ||
                                                   base_kind(K) == kLdsSmooth))
                             ? 1
                             : 0);
  static constexpr int NV = LdsNeeds<K>::NV;        // broadcast vectors per worker (2D entries each)
  // pivot-row broadcast of two matrices of 2 DP columns; the throughput kinds broadcast rows in the
  // padded layout of the latency kinds' split (whose reduction order they reproduce)
  static constexpr int BC = 4 * (is_throughput_kind(K) ? padded<D_, split_of<D_>()>() : L::DP) + 2;
  static constexpr int SMALL = LdsNeeds<K>::SMALL ? kMaxObs * L::LD + kMaxObs * D_ : 0;
  static constexpr int WORKER = BC + NV * 2 * D_ + NT * TILE + SMALL;
  static constexpr int CONSTS = LdsNeeds<K>::CONSTS ? 3 * TILE : 0;  // F, LQ, N1z
  static constexpr size_t bytes_for(int wpb) { return sizeof(double) * (CONSTS + wpb * WORKER); }
  // workers per block: as many as fit in 64 KiB of LDS, at most 256 threads
  // (throughput kinds: the largest power of two up to 256 threads that fits)
  static constexpr int wpb_fit(int w) { return (w <= 1 || bytes_for(w) <= 65536) ? w : wpb_fit(w / 2); }
  // resident workers per CU allowed by its 160 KiB of LDS with w workers per block (counted up to
  // 16, four waves per SIMD, beyond which these kernels' registers limit them)
  static constexpr int lds_workers(int w) {
    return ((163840 / (int)bytes_for(w)) * w < 16) ? (163840 / (int)bytes_for(w)) * w : 16;
  }
  // the kinds with the extra tile of the matrix-core products: the block size (4, 2 or 1 whole-
  // wavefront workers) that keeps the most workers resident by LDS
  static constexpr bool EXTRA = NT > LdsNeeds<K>::NT;
  static constexpr int wpb_lds() {
    return (lds_workers(4) >= lds_workers(2) && lds_workers(4) >= lds_workers(1)) ? 4
           : (lds_workers(2) >= lds_workers(1))                                  ? 2
                                                                                 : 1;
  }
  static constexpr int WPB = is_throughput_kind(K)                   ? wpb_fit(256 / L::W)
                             : EXTRA                                 ? wpb_lds()
                             : (L::W == 32 && bytes_for(8) <= 65536) ? 8
                             : (bytes_for(256 / L::W) <= 65536)      ? 256 / L::W
                             : (bytes_for(2) <= 65536)               ? 2
                                                                     : 1;
  static constexpr int BLOCK = L::W * WPB;
  static constexpr size_t lds_bytes() { return bytes_for(WPB); }
  static_assert(bytes_for(WPB) <= 65536, "state dimension too large for one worker per block");
};

template <int D, int S_>
struct Ctx {
  static constexpr int S = S_;
  static constexpr int LD = Lay<D, S_>::LD;
  static constexpr int TILE = D * LD;
  int lane;           // lane inside the worker
  int row;            // matrix row of this lane
  int sub;            // column class of this lane inside its row
  int wid;            // global worker index
  double* w;          // worker LDS slice (starts with the Householder broadcast rows)
  double* vbase;      // broadcast vectors
  double* tbase;      // tiles
  const double* sF;   // block LDS constants, row stride LD (kernels that stage them)
  const double* sLQ;
  const double* sN1;
  __device__ double* tile(int i) const { return tbase + i * TILE; }
  __device__ double* vec(int i) const { return vbase + i * 2 * D; }
  __device__ double* bc() const { return w; }
};

// Opaque copy of a value: keeps the optimiser from correlating the many lane
// tests of one function (which otherwise duplicates whole regions per branch
// and multiplies register pressure).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ size_t opaque_sz(size_t v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Stage F, LQ, N1z into block LDS (kinds that need them; padding columns are
// zero) and carve out this thread's worker slice.  Must be called by every
// thread of the block before any early exit.
template <int D, int K>
__device__ __forceinline__ Ctx<D, Geo<D, K>::S> make_ctx(const double* F, const double* LQ, const double* N1z,
                                                          double* smem) {
  using G = Geo<D, K>;
  if constexpr (G::CONSTS > 0) {
    for (int t = threadIdx.x; t < D * G::DP; t += G::BLOCK) {
      const int r = t / G::DP, c = t % G::DP;
      const bool in = c < D;
      const int cc = in ? c : 0;
      smem[r * G::LD + c] = in ? F[r * D + cc] : 0.0;
      smem[G::TILE + r * G::LD + c] = in ? LQ[r * D + cc] : 0.0;
      smem[2 * G::TILE + r * G::LD + c] = in ? N1z[r * D + cc] : 0.0;
    }
    __syncthreads();
  }
  Ctx<D, G::S> cx;
  cx.lane = threadIdx.x % G::W;
  cx.row = opaque(cx.lane / G::S);
  cx.sub = opaque(cx.lane % G::S);
  const int wl = threadIdx.x / G::W;
  cx.wid = blockIdx.x * G::WPB + wl;
  cx.w = smem + G::CONSTS + wl * G::WORKER;
  cx.vbase = cx.w + G::BC;
  cx.tbase = cx.vbase + G::NV * 2 * D;
  cx.sF = smem;
  cx.sLQ = smem + G::TILE;
  cx.sN1 = smem + 2 * G::TILE;
  return cx;
}

// ----------------------------------------------------------------- helpers
// Pin accumulators at a point of the instruction stream.  Within one basic
// block the instruction selector is free to emit every load of an unrolled
// product first and the arithmetic last, which keeps each loaded value in a
// register pair until the end (hundreds of registers, then spills).  Passing
// the accumulators through an empty volatile asm after each step forces the
// arithmetic of that step to be emitted before the loads of the next one.
template <int N>
__device__ __forceinline__ void pin(double (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) asm volatile("" : "+v"(x[i])::"memory");
}
__device__ __forceinline__ void pin1(double& x) { asm volatile("" : "+v"(x)::"memory"); }

template <int N>
__device__ __forceinline__ void zero(double (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = 0.0;
}

template <int N>
__device__ __forceinline__ void load_row(double (&x)[N], const double* p) {
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = p[i];
}

template <int N>
__device__ __forceinline__ void store_row(double* p, const double (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) p[i] = x[i];
}

// 1/x to about one ulp: the hardware estimate refined by two Newton steps
// (x is a nonzero norm-sized value here; no denormal or overflow handling).
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// sqrt(x) and 1/sqrt(x) for a normal x > 0, to about one ulp: the hardware
// reciprocal square root estimate refined by Goldschmidt's iteration
// (g -> sqrt(x), h -> 1 / (2 sqrt(x))) and a final Newton correction of g.
__device__ __forceinline__ void sqrt_rsqrt(double x, double& sq, double& rsq) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  sq = fma(d, h, g);
  rsq = 2.0 * h;
}

// ------------------------------------------------------- row communication
// Quad permutation of a double (two 32-bit DPP moves).
template <int CTRL>
__device__ __forceinline__ double dppd(double v) {
  const long long x = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the S lanes of a row; every lane of the row gets the same bits.
template <int S>
__device__ __forceinline__ double rowsum(double v) {
  if constexpr (S >= 2) v = v + dppd<0xB1>(v);  // quad_perm [1,0,3,2]
  if constexpr (S >= 4) v = v + dppd<0x4E>(v);  // quad_perm [2,3,0,1]
  return v;
}

// The value held by sub-lane K of this lane's row.
template <int S, int K>
__device__ __forceinline__ double from_sub(double v) {
  if constexpr (S == 1) {
    return v;
  } else if constexpr (S == 2) {
    return dppd<(K == 0) ? 0xA0 : 0xF5>(v);  // quad_perm [K,K,2+K,2+K]
  } else {
    return dppd<(K & 3) * 0x55>(v);  // quad_perm [K,K,K,K]
  }
}
// k is a compile-time constant after unrolling
template <int S>
__device__ __forceinline__ double from_sub_k(double v, int k) {
  if (S == 1 || k == 0) return from_sub<S, 0>(v);
  if (k == 1) return from_sub<S, 1>(v);
  if (k == 2) return from_sub<S, 2>(v);
  return from_sub<S, 3>(v);
}

// The value of v in wavefront lane L (a compile-time lane): two scalar reads.
__device__ __forceinline__ double readlane_d(double v, int L) {
  const long long x = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)x, L);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), L);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// x[t] = p[t S + sub] for columns < n, 0 for the padding columns.
template <int S, int n, int NX>
__device__ __forceinline__ void load_dist(double (&x)[NX], const double* p, int sub) {
#pragma unroll
  for (int t = 0; t < NX; ++t) {
    const int c = t * S + sub;
    if ((t + 1) * S <= n) {
      x[t] = p[c];
    } else if (t * S < n) {
      const bool in = c < n;
      const double v = p[in ? c : 0];
      x[t] = in ? v : 0.0;
    } else {
      x[t] = 0.0;
    }
  }
}

// p[t S + sub] = x[t] for columns < n.
template <int S, int n, int NX>
__device__ __forceinline__ void store_dist(double* p, const double (&x)[NX], int sub) {
#pragma unroll
  for (int t = 0; t < NX; ++t) {
    const int c = t * S + sub;
    if ((t + 1) * S <= n) {
      p[c] = x[t];
    } else if (t * S < n) {
      if (c < n) p[c] = x[t];
    }
  }
}

// Replicated copy f[0..n) of a row held in split form.
template <int S, int n, int NX>
__device__ __forceinline__ void gather(const double (&x)[NX], double (&f)[n]) {
#pragma unroll
  for (int c = 0; c < n; ++c) f[c] = from_sub_k<S>(x[c / S], c % S);
}

// The lane's columns t S + sub of a replicated row f[0..n) (0 past n).
template <int S, int n, int NX>
__device__ __forceinline__ void own(const double (&f)[n], int sub, double (&x)[NX]) {
#pragma unroll
  for (int t = 0; t < NX; ++t) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int c = t * S + k;
      if (c < n) v = (sub == k) ? f[c] : v;
    }
    x[t] = v;
  }
}

// ------------------------------------------------ reductions in a split's order
// A row reduction of the split form (S lanes per row) sums each lane's columns
// t S + sub in a chain and adds the lanes' partial sums in the order of the DPP
// butterfly of rowsum<S>.  The throughput forms keep a row in one lane (S = 1)
// but reduce "as if split over V lanes" (V = the latency form's split), so the
// two forms give bitwise equal results: V partial sums, combined like rowsum<V>.
template <int V>
__device__ __forceinline__ double vcombine(const double (&p)[V]) {
  if constexpr (V == 1) {
    return p[0];
  } else if constexpr (V == 2) {
    return p[0] + p[1];
  } else {
    static_assert(V == 4, "splits of 1, 2 or 4 lanes");
    return (p[0] + p[1]) + (p[2] + p[3]);
  }
}

// compact column of virtual column cv of a row laid out as two blocks (B1, B2),
// each padded to a multiple of V (P1 = padded<B1, V>), or -1 for a padding column
template <int B1, int B2, int P1>
__host__ __device__ constexpr int vcol(int cv) {
  return (cv < P1) ? ((cv < B1) ? cv : -1) : ((cv - P1 < B2) ? B1 + (cv - P1) : -1);
}

// sum_c x[c] v[c] over the row's n columns (x split, v replicated in LDS);
// every lane of the row gets the sum.  V > S (S = 1 only): the order of a split over V lanes.
template <int S, int n, int V = S, int NX>
__device__ __forceinline__ double row_dot(const double (&x)[NX], const double* v, int sub) {
  if constexpr (V != S) {
    static_assert(S == 1, "a reduction order of V lanes is reproduced on one lane per row");
    constexpr int NXV = padded<n, V>() / V;
    double part[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < NXV; ++t) {
        const int c = t * V + u;
        const double xv = (c < n) ? x[(c < n) ? c : 0] : 0.0;
        if ((t + 1) * V <= n) {
          acc = fma(xv, v[c], acc);
        } else if (t * V < n) {
          const double vv = v[c < n ? c : 0];
          acc = fma(xv, (c < n) ? vv : 0.0, acc);
        }
      }
      part[u] = acc;
    }
    return vcombine<V>(part);
  }
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < NX; ++t) {
    const int c = t * S + sub;
    if ((t + 1) * S <= n) {
      acc = fma(x[t], v[c], acc);
    } else if (t * S < n) {
      const double vv = v[c < n ? c : 0];
      acc = fma(x[t], (c < n) ? vv : 0.0, acc);
    }
  }
  return rowsum<S>(acc);
}

// sum_{k < n} p[k * stride] v[k] (a row (stride 1) or a column (stride LD) of
// an LDS tile times an LDS vector), split over the row's lanes and reduced
// with DPP; every lane of the row gets the sum.
template <int S, int n, int stride, int V = S>
__device__ __forceinline__ double lds_dot(const double* p, const double* v, int sub) {
  if constexpr (V != S) {
    static_assert(S == 1, "a reduction order of V lanes is reproduced on one lane per row");
    double part[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < (n + V - 1) / V; ++t) {
        const int k = t * V + u;
        if ((t + 1) * V <= n) {
          acc = fma(p[k * stride], v[k], acc);
        } else {
          const int kk = (k < n) ? k : 0;
          const double x = p[kk * stride] * v[kk];
          acc += (k < n) ? x : 0.0;
        }
      }
      part[u] = acc;
    }
    return vcombine<V>(part);
  }
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < (n + S - 1) / S; ++t) {
    const int k = t * S + sub;
    if ((t + 1) * S <= n) {
      acc = fma(p[k * stride], v[k], acc);
    } else {
      const int kk = (k < n) ? k : 0;
      const double x = p[kk * stride] * v[kk];
      acc += (k < n) ? x : 0.0;
    }
  }
  return rowsum<S>(acc);
}

// sum_t x[t] y[t] over a row held in registers (both split the same way; padding
// registers zero), reduced along the row; V > S as in row_dot.
template <int S, int n, int V = S, int NX>
__device__ __forceinline__ double reg_dot(const double (&x)[NX], const double (&y)[NX]) {
  if constexpr (V != S) {
    static_assert(S == 1, "a reduction order of V lanes is reproduced on one lane per row");
    constexpr int NXV = padded<n, V>() / V;
    double part[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < NXV; ++t) {
        const int c = t * V + u;
        const double xv = (c < n) ? x[(c < n) ? c : 0] : 0.0;
        const double yv = (c < n) ? y[(c < n) ? c : 0] : 0.0;
        acc = fma(xv, yv, acc);
      }
      part[u] = acc;
    }
    return vcombine<V>(part);
  }
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < NX; ++t) acc = fma(x[t], y[t], acc);
  return rowsum<S>(acc);
}

// ----------------------------------------------------------------- products
// out[t] = sum_{k < K} x[k] T[k][t S + sub] over columns < n (T: stride LD)
template <int S, int K, int n, int LD, int NO, int NXF>
__device__ __forceinline__ void row_times_tile(const double (&x)[NXF], const double* T, int sub, double (&out)[NO]) {
  zero(out);
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int t = 0; t < NO; ++t) {
      const int c = t * S + sub;
      if ((t + 1) * S <= n) {
        out[t] = fma(x[k], T[k * LD + c], out[t]);
      } else if (t * S < n) {
        const double v = T[k * LD + (c < n ? c : 0)];
        out[t] = fma(x[k], (c < n) ? v : 0.0, out[t]);
      }
    }
    pin(out);
  }
}

// out[t] = sum_{k < K} T1[k][r] T2[k][t S + sub] over columns < n
template <int S, int K, int n, int LD, int NO>
__device__ __forceinline__ void tile_t_times_tile(int r, const double* T1, const double* T2, int sub,
                                                  double (&out)[NO]) {
  zero(out);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double a = T1[k * LD + r];
#pragma unroll
    for (int t = 0; t < NO; ++t) {
      const int c = t * S + sub;
      if ((t + 1) * S <= n) {
        out[t] = fma(a, T2[k * LD + c], out[t]);
      } else if (t * S < n) {
        const double v = T2[k * LD + (c < n ? c : 0)];
        out[t] = fma(a, (c < n) ? v : 0.0, out[t]);
      }
    }
    pin(out);
  }
}

// Out = op(A) op(B) for D x D tiles in LDS (row stride LD) on the FP64 matrix
// cores, D <= 16: one v_mfma_f64_16x16x4f64 per four summation indices, every
// lane of the wavefront (the worker) taking part.  Lane l supplies A[l % 16][k]
// and B[k][l % 16] for k = 4 step + l / 16, and its accumulator value v is
// element (l / 16 + 4 v, l % 16); rows and columns past D are zero and not
// written.  TA: A is stored transposed (A[i][k] = Ap[k LD + i]); TB likewise.
// Out may be an operand's tile: the operands' reads feed the products, so they
// have completed before the first write.  The caller syncs the worker before
// reading Out.
typedef double f64x4_t __attribute__((ext_vector_type(4)));
template <int D, int LD, bool TA, bool TB>
__device__ __forceinline__ void tile_mfma(const double* Ap, const double* Bp, double* Out) {
  static_assert(D <= 16, "one 16 x 16 tile");
  const int lane = threadIdx.x & 63;
  const int i = lane & 15, kq = lane >> 4;
  f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int st = 0; st < (D + 3) / 4; ++st) {
    const int k = 4 * st + kq;
    const bool ok = k < D && i < D;
    const int kc = ok ? k : 0, ic = ok ? i : 0;
    const double a = TA ? Ap[kc * LD + ic] : Ap[ic * LD + kc];
    const double b = TB ? Bp[ic * LD + kc] : Bp[kc * LD + ic];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ok ? a : 0.0, ok ? b : 0.0, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = kq + 4 * v;
    if (row < D && i < D) Out[row * LD + i] = acc[v];
  }
}

// The row (split) x of each lane's row into tile T (rows < n of the worker's lanes that hold them).
template <int S, int n, int LD, int NX>
__device__ __forceinline__ void stage_row(double* T, int r, int sub, const double (&x)[NX]) {
  if (r < n) {
#pragma unroll
    for (int t = 0; t < NX; ++t) T[r * LD + t * S + sub] = x[t];  // padded columns < LD hold zeros
  }
}

// ---------------------------------------------------------- Householder LQ
// LQ of NG independent NR x (B1 + B2) matrices whose columns come in two
// blocks (B1, B2), each padded to a multiple of S.  Row r of matrix g sits in
// lane row g NR + r, split over its S lanes.  On return that row holds row r
// of the lower-triangular factor L (L L^T = M M^T, compact columns > r zero,
// padding columns zero): the reflectors of LAPACK dgeqrf on M^T (dlarfg:
// beta = -sign(alpha) norm, tau = (beta - alpha) / beta, v = [1, x / (alpha - beta)],
// tau = 0 for a zero tail, no reflector for a length-1 tail).
//
// Per pivot step the pivot row broadcasts its raw tail through LDS; every
// lane forms partial sums over its own columns, reduces them along its row
// with DPP and derives (beta, tau, scale) itself, so no lane sits in a
// divergent pivot section.
//
// The pivots [I0, I1) are applied; FINISH clears the strict upper part at the
// end (so one factorisation can be run in pieces, with other work between).
//
// W2 < B2 declares a banded second block: row r's entries there are zero
// beyond compact column max(r, W2) (W2 = 0: lower triangular).  Each pivot row
// i < B1 then keeps that band (the reflectors before it span columns <= i +
// the band), so the reflector of pivot i is zero beyond column max(i, W2) of
// the second block and those columns are left out of its sums and updates.
// They would only add exact zeros: the factor is the same.
//
// WV is the worker's width when the caller knows it (0: unknown).  In a worker
// that is a whole wavefront the pivot row's squared tail is read from the
// pivot row's own dot product (its lanes multiply their row by a copy of it,
// in the order of the chains below: the same bits) instead of a second chain
// in every lane.
template <int S, int NR, int B1, int B2, int NG = 1, int I0 = 0, int I1 = 1 << 20, bool FINISH = true,
          int W2 = 1 << 20, int WV = 0>
__device__ __forceinline__ void lq(double (&a)[(padded<B1, S>() + padded<B2, S>()) / S], int row, int sub,
                                   double* bc) {
  static_assert(NG == 1 || NG == 2, "one or two matrices");
  constexpr int P1 = padded<B1, S>();
  constexpr int NP = P1 + padded<B2, S>();
  constexpr int NX = NP / S;
  constexpr int NC = B1 + B2;
  constexpr int K = (NR < NC) ? NR : NC;
  // a worker that is a whole wavefront (its width WV = 64 declared by the caller, or rows over
  // more than 32 lanes): lane row S + sub of the worker is that lane of the wavefront
  constexpr bool kFullWave = WV == 64 || NG * NR * S > 32;
  const int g = (NG == 2 && row >= NR) ? 1 : 0;
  const int li = row - g * NR;
  const bool valid = row < NG * NR;
  double* brow = bc + g * NP;  // broadcast row of this lane's matrix
#pragma unroll
  for (int i = I0; i < ((I1 < K) ? I1 : K); ++i) {
    if (i == NC - 1) break;  // a length-1 tail: H = I
    const int pc = (i < B1) ? i : P1 + (i - B1);  // padded column of pivot i
    const int t0 = pc / S, k0 = pc % S;
    // end (exclusive) of the register columns the reflector can touch
    const int band = (i > W2) ? i : W2;
    const int te = (i < B1 && W2 < B2 && P1 + band < NP) ? (P1 + band) / S + 1 : NX;
    if (valid && li == i) {
#pragma unroll
      for (int t = t0; t < te; ++t) brow[t * S + sub] = a[t];
    }
    wsync();
    const double alpha = brow[pc];
    double p[NX];
#pragma unroll
    for (int t = t0; t < te; ++t) p[t] = brow[t * S + sub];
    // the pivot's own group of S columns: only columns after pc are tail
    const bool tl = sub > k0;
    const double p0 = tl ? p[t0] : 0.0;
    // (CH independent chains each: one lane sums NX - t0 columns)
    constexpr int CH = (NX > 16) ? 4 : 2;
    double dt[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) dt[c] = 0.0;
    dt[0] = a[t0] * p0;
#pragma unroll
    for (int t = t0 + 1; t < te; ++t) dt[(t - t0) % CH] = fma(a[t], p[t], dt[(t - t0) % CH]);
    const double dts = (CH == 4) ? (dt[0] + dt[1]) + (dt[2] + dt[3]) : dt[0] + dt[1];
    const double ai = from_sub_k<S>(a[t0], k0);  // this row's entry in the pivot column
    const double dot = rowsum<S>(dts);
    double sig;
    if constexpr (kFullWave) {
      // the pivot row's own dot is its squared tail (p is a copy of its tail), summed in the
      // order below: read it from the row's first lane instead of a chain in every lane
      const double s0 = readlane_d(dot, i * S);
      if constexpr (NG == 2) {
        const double s1 = readlane_d(dot, (NR + i) * S);
        sig = g ? s1 : s0;
      } else {
        sig = s0;
      }
    } else {
      double sq[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) sq[c] = 0.0;
      sq[0] = p0 * p0;
#pragma unroll
      for (int t = t0 + 1; t < te; ++t) sq[(t - t0) % CH] = fma(p[t], p[t], sq[(t - t0) % CH]);
      const double sqs = (CH == 4) ? (sq[0] + sq[1]) + (sq[2] + sq[3]) : sq[0] + sq[1];
      sig = rowsum<S>(sqs);
    }
    double nrm, rnrm;
    sqrt_rsqrt(fma(alpha, alpha, sig), nrm, rnrm);
    const double bc_ = -copysign(nrm, alpha);  // Fortran SIGN: honours the sign of -0.0
    const double den = alpha - bc_;
    const double sc_ = rcp(den);
    const double tc_ = den * copysign(rnrm, alpha);  // (beta - alpha) / beta
    const bool nz = sig != 0.0;
    const double beta = nz ? bc_ : alpha;
    const double scal = nz ? sc_ : 0.0;
    const double tau = nz ? tc_ : 0.0;
    const bool owner = sub == k0;
    if (valid && li == i && owner) a[t0] = beta;
    if (valid && li > i) {
      const double tw = tau * fma(scal, dot, ai);
      const double tws = tw * scal;
      a[t0] = owner ? a[t0] - tw : fma(-tws, p0, a[t0]);
#pragma unroll
      for (int t = t0 + 1; t < te; ++t) a[t] = fma(-tws, p[t], a[t]);
    }
  }
  if constexpr (FINISH) {
    // strict upper part: compact column index of each padded column
#pragma unroll
    for (int t = 0; t < NX; ++t) {
      const int c = t * S + sub;
      const int j = (c < P1) ? c : B1 + (c - P1);
      a[t] = (valid && j > li) ? 0.0 : a[t];
    }
  }
  wsync();
}

// lq() of a split over V lanes reproduced on one lane per row (the throughput
// forms): the same reflectors, partial sums, chains and combination order as
// lq<V, ...> on V lanes, so the factor is bitwise the same.  The lane holds the
// row's compact columns; virtual column cv = t V + u (u = the virtual sub-lane)
// maps to compact column vcol(cv), padding columns are zero.
template <int V, int NR, int B1, int B2, int NG = 1, int I0 = 0, int I1 = 1 << 20, bool FINISH = true,
          int W2 = 1 << 20>
__device__ __forceinline__ void lq_virtual(double (&a)[B1 + B2], int row, double* bc) {
  static_assert(NG == 1 || NG == 2, "one or two matrices");
  constexpr int P1 = padded<B1, V>();
  constexpr int NP = P1 + padded<B2, V>();
  constexpr int NX = NP / V;
  constexpr int NC = B1 + B2;
  constexpr int K = (NR < NC) ? NR : NC;
  const int g = (NG == 2 && row >= NR) ? 1 : 0;
  const int li = row - g * NR;
  const bool valid = row < NG * NR;
  double* brow = bc + g * NP;
#pragma unroll
  for (int i = I0; i < ((I1 < K) ? I1 : K); ++i) {
    if (i == NC - 1) break;  // a length-1 tail: H = I
    const int pc = (i < B1) ? i : P1 + (i - B1);
    const int t0 = pc / V, k0 = pc % V;
    const int band = (i > W2) ? i : W2;
    const int te = (i < B1 && W2 < B2 && P1 + band < NP) ? (P1 + band) / V + 1 : NX;
    if (valid && li == i) {
#pragma unroll
      for (int cv = t0 * V; cv < te * V; ++cv) {
        const int j = vcol<B1, B2, P1>(cv);
        brow[cv] = (j >= 0) ? a[(j >= 0) ? j : 0] : 0.0;
      }
    }
    wsync();
    const double alpha = brow[pc];
    double pv[NP];
#pragma unroll
    for (int cv = t0 * V; cv < te * V; ++cv) pv[cv] = brow[cv];
    constexpr int CH = (NX > 16) ? 4 : 2;
    double sqp[V], dtp[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int c0 = t0 * V + u;
      const int j0 = vcol<B1, B2, P1>(c0);
      const double p0 = (u > k0) ? pv[c0] : 0.0;
      const double a0 = (j0 >= 0) ? a[(j0 >= 0) ? j0 : 0] : 0.0;
      double sq[CH], dt[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) sq[c] = dt[c] = 0.0;
      sq[0] = p0 * p0;
      dt[0] = a0 * p0;
#pragma unroll
      for (int t = t0 + 1; t < te; ++t) {
        const int cv = t * V + u;
        const int j = vcol<B1, B2, P1>(cv);
        const double av = (j >= 0) ? a[(j >= 0) ? j : 0] : 0.0;
        sq[(t - t0) % CH] = fma(pv[cv], pv[cv], sq[(t - t0) % CH]);
        dt[(t - t0) % CH] = fma(av, pv[cv], dt[(t - t0) % CH]);
      }
      if constexpr (CH == 4) {
        sqp[u] = (sq[0] + sq[1]) + (sq[2] + sq[3]);
        dtp[u] = (dt[0] + dt[1]) + (dt[2] + dt[3]);
      } else {
        sqp[u] = sq[0] + sq[1];
        dtp[u] = dt[0] + dt[1];
      }
    }
    const int jp = vcol<B1, B2, P1>(pc);  // the pivot column is never padding
    const double ai = a[jp];
    const double sig = vcombine<V>(sqp);
    const double dot = vcombine<V>(dtp);
    double nrm, rnrm;
    sqrt_rsqrt(fma(alpha, alpha, sig), nrm, rnrm);
    const double bc_ = -copysign(nrm, alpha);
    const double den = alpha - bc_;
    const double sc_ = rcp(den);
    const double tc_ = den * copysign(rnrm, alpha);
    const bool nz = sig != 0.0;
    const double beta = nz ? bc_ : alpha;
    const double scal = nz ? sc_ : 0.0;
    const double tau = nz ? tc_ : 0.0;
    if (valid && li == i) a[jp] = beta;
    if (valid && li > i) {
      const double tw = tau * fma(scal, dot, ai);
      const double tws = tw * scal;
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const int c0 = t0 * V + u;
        const int j0 = vcol<B1, B2, P1>(c0);
        if (j0 >= 0) {
          const double p0 = (u > k0) ? pv[c0] : 0.0;
          a[j0] = (u == k0) ? a[j0] - tw : fma(-tws, p0, a[j0]);
        }
      }
#pragma unroll
      for (int t = t0 + 1; t < te; ++t) {
#pragma unroll
        for (int u = 0; u < V; ++u) {
          const int cv = t * V + u;
          const int j = vcol<B1, B2, P1>(cv);
          if (j >= 0) a[j] = fma(-tws, pv[cv], a[j]);
        }
      }
    }
  }
  if constexpr (FINISH) {
#pragma unroll
    for (int j = 0; j < NC; ++j) a[j] = (valid && j > li) ? 0.0 : a[j];
  }
  wsync();
}

// lq() with the reduction order of a split over V lanes: lq<S> itself when the
// row is split over V = S lanes, lq_virtual<V> on one lane per row.
template <int S, int V, int NR, int B1, int B2, int NG = 1, int I0 = 0, int I1 = 1 << 20, bool FINISH = true,
          int W2 = 1 << 20, int WV = 0>
__device__ __forceinline__ void lqv(double (&a)[(padded<B1, S>() + padded<B2, S>()) / S], int row, int sub,
                                    double* bc) {
  if constexpr (S == V) {
    lq<S, NR, B1, B2, NG, I0, I1, FINISH, W2, WV>(a, row, sub, bc);
  } else {
    static_assert(S == 1, "a reduction order of V lanes is reproduced on one lane per row");
    lq_virtual<V, NR, B1, B2, NG, I0, I1, FINISH, W2>(a, row, bc);
  }
}

// --------------------------------------------------- triangular solves (split)
// y L^T = v (forward substitution; = L y = v), L lower n x n in LDS (stride
// LD), y split over the row's lanes, holding v on entry.
template <int S, int n, int LD, int NY>
__device__ __forceinline__ void rsolve_lowerT_s(double (&y)[NY], const double* L, int sub) {
  double rd[NY];
#pragma unroll
  for (int t = 0; t < NY; ++t) {
    const int c = t * S + sub;
    const int cc = (c < n) ? c : 0;
    rd[t] = rcp(L[cc * LD + cc]);
  }
#pragma unroll
  for (int j = 0; j < n; ++j) {
    const int t = j / S, k = j % S;
    const double yj = from_sub_k<S>(y[t] * rd[t], k);
    if (sub == k) y[t] = yj;
#pragma unroll
    for (int tt = t; tt < NY; ++tt) {
      const int i = tt * S + sub;
      const bool upd = (tt > t || sub > k) && i < n;
      const double l = L[(i < n ? i : 0) * LD + j];
      y[tt] = upd ? fma(-yj, l, y[tt]) : y[tt];
    }
    pin1(y[t]);
  }
}

// y L = v (backward substitution; = L^T y = v), L lower n x n in LDS.
template <int S, int n, int LD, int NY>
__device__ __forceinline__ void rsolve_lower_s(double (&y)[NY], const double* L, int sub) {
  double rd[NY];
#pragma unroll
  for (int t = 0; t < NY; ++t) {
    const int c = t * S + sub;
    const int cc = (c < n) ? c : 0;
    rd[t] = rcp(L[cc * LD + cc]);
  }
#pragma unroll
  for (int j = n - 1; j >= 0; --j) {
    const int t = j / S, k = j % S;
    const double yj = from_sub_k<S>(y[t] * rd[t], k);
    if (sub == k) y[t] = yj;
#pragma unroll
    for (int tt = 0; tt <= t; ++tt) {
      const int i = tt * S + sub;
      const bool upd = tt < t || sub < k;
      const double l = L[j * LD + i];
      y[tt] = upd ? fma(-yj, l, y[tt]) : y[tt];
    }
    pin1(y[t]);
  }
}

// ---------------------------------------------- triangular solves (replicated)
// Small systems (n <= d) solved by every lane on a full copy.
// y L = v with L lower triangular (n x n in LDS, stride LD); y holds v on entry.
template <int n, int LD, int NY>
__device__ __forceinline__ void rsolve_lower(double (&y)[NY], const double* L) {
#pragma unroll
  for (int j = n - 1; j >= 0; --j) {
    double s0 = y[j], s1 = 0.0;
#pragma unroll
    for (int l = j + 1; l < n; ++l) {
      if ((l - j) & 1) s0 = fma(-y[l], L[l * LD + j], s0);
      else s1 = fma(-y[l], L[l * LD + j], s1);
    }
    y[j] = (s0 + s1) / L[j * LD + j];
    pin1(y[j]);
  }
}

// y L^T = v with L lower triangular (forward substitution); y holds v on entry.
template <int n, int LD, int NY>
__device__ __forceinline__ void rsolve_lowerT(double (&y)[NY], const double* L) {
#pragma unroll
  for (int j = 0; j < n; ++j) {
    double s0 = y[j], s1 = 0.0;
#pragma unroll
    for (int l = 0; l < j; ++l) {
      if (l & 1) s1 = fma(-y[l], L[j * LD + l], s1);
      else s0 = fma(-y[l], L[j * LD + l], s0);
    }
    y[j] = (s0 + s1) / L[j * LD + j];
    pin1(y[j]);
  }
}

// x = L^-1 v (forward) with L lower; x holds v on entry.
template <int n, int LD, int NY>
__device__ __forceinline__ void lsolve_lower(double (&x)[NY], const double* L) {
  rsolve_lowerT<n, LD, NY>(x, L);  // L x = v  <=>  x^T L^T = v^T
}

// x = L^-T v (backward) with L lower; x holds v on entry.
template <int n, int LD, int NY>
__device__ __forceinline__ void lsolve_lowerT(double (&x)[NY], const double* L) {
  rsolve_lower<n, LD, NY>(x, L);   // L^T x = v  <=>  x^T L = v^T
}

// ---------------------------------------------------------- observation model
// Affine observation of element k (noiseless): y = H x + c.
//
// EK1 (pof/step.py:12-22, pof/observations.py:35-40) linearises at lin[k+1]:
// H = E1 - J E0, c = J E0 m - f(E0 m).  In the IWP state layout row i of H has
// hc[i][l] = -J[i][l] Pd[l q1] at the columns l q1 (l < d) and hd[i] =
// Pd[i q1 + 1] at column i q1 + 1, zeros elsewhere; the products below use
// only those d + 1 entries (skipping exact zeros leaves finite results
// unchanged).  The affine variant reads an explicit dense H.  Every lane
// holds the whole (small) model.
template <int D, int d, bool EK1>
struct Obs;

template <int D, int d>
struct Obs<D, d, true> {
  static constexpr int q1 = D / d;
  double hc[d][d];
  double hd[d];
  double c[d];

  __device__ __forceinline__ void load(const ModelArgs& a, int k) { load_at(a, a.lin + (size_t)(k + 1) * D); }
  // linearise at the (preconditioned) state mean m
  __device__ __forceinline__ void load_at(const ModelArgs& a, const double* m) {
    double y[d];
#pragma unroll
    for (int i = 0; i < d; ++i) y[i] = a.Pd[i * q1] * m[i * q1];
    double f[d];
    double J[d][d];
    field_and_jacobian<d>(a.kind, a.params, y, f, J);
#pragma unroll
    for (int i = 0; i < d; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < d; ++l) {
        hc[i][l] = -J[i][l] * a.Pd[l * q1];
        acc = fma(J[i][l], y[l], acc);
      }
      hd[i] = a.Pd[i * q1 + 1];
      c[i] = acc - f[i];
    }
  }
  // the noise factor R of EK1 is zero (observations.py:39)
  static __device__ __forceinline__ constexpr double r_at(int, int) { return 0.0; }
  // sum_j H[i][j] v[j st]
  __device__ __forceinline__ double dot(int i, const double* v, int st) const {
    double acc = 0.0;
#pragma unroll
    for (int l = 0; l < d; ++l) acc = fma(hc[i][l], v[l * q1 * st], acc);
    return fma(hd[i], v[(i * q1 + 1) * st], acc);
  }
  // row r (< d, any lane value) of H times the tile T: out[t] = sum_j H[r][j] T[j][t S + sub]
  template <int S, int n, int LD, int NO>
  __device__ __forceinline__ void row_times_tile(int r, const double* T, int sub, double (&out)[NO]) const {
    double hr[d + 1];
#pragma unroll
    for (int l = 0; l < d; ++l) hr[l] = hc[0][l];
    hr[d] = hd[0];
#pragma unroll
    for (int i = 1; i < d; ++i) {
#pragma unroll
      for (int l = 0; l < d; ++l) hr[l] = (r == i) ? hc[i][l] : hr[l];
      hr[d] = (r == i) ? hd[i] : hr[d];
    }
    const int rd = ((r < d) ? r : 0) * q1 + 1;
    zero(out);
#pragma unroll
    for (int l = 0; l <= d; ++l) {
      const double* Tr = T + ((l < d) ? l * q1 : rd) * LD;
#pragma unroll
      for (int t = 0; t < NO; ++t) {
        const int cc = t * S + sub;
        if ((t + 1) * S <= n) {
          out[t] = fma(hr[l], Tr[cc], out[t]);
        } else if (t * S < n) {
          const double v = Tr[(cc < n) ? cc : 0];
          out[t] = fma(hr[l], (cc < n) ? v : 0.0, out[t]);
        }
      }
      pin(out);
    }
  }
  // (K H)[j] = sum_i K[i] H[i][j] for the row vector K (length d), as the
  // coefficients of the columns l q1 (kh[l]) and i q1 + 1 (kh[d + i])
  __device__ __forceinline__ void left_mul(const double (&K)[d], double (&kh)[2 * d]) const {
#pragma unroll
    for (int l = 0; l < d; ++l) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < d; ++i) acc = fma(K[i], hc[i][l], acc);
      kh[l] = acc;
    }
#pragma unroll
    for (int i = 0; i < d; ++i) kh[d + i] = K[i] * hd[i];
  }
  // out[t] -= sum_j (K H)[j] T[j][t S + sub]
  template <int S, int n, int LD, int NO>
  __device__ __forceinline__ void sub_kh_tile(const double (&kh)[2 * d], const double* T, int sub,
                                              double (&out)[NO]) const {
#pragma unroll
    for (int m = 0; m < 2 * d; ++m) {
      const double* Tr = T + ((m < d) ? m * q1 : (m - d) * q1 + 1) * LD;
#pragma unroll
      for (int t = 0; t < NO; ++t) {
        const int cc = t * S + sub;
        if ((t + 1) * S <= n) {
          out[t] = fma(-kh[m], Tr[cc], out[t]);
        } else if (t * S < n) {
          const double v = Tr[(cc < n) ? cc : 0];
          out[t] = fma(-kh[m], (cc < n) ? v : 0.0, out[t]);
        }
      }
      pin(out);
    }
  }
  __device__ __forceinline__ void dense(double (&H)[d][D]) const {
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int j = 0; j < D; ++j) H[i][j] = 0.0;
#pragma unroll
      for (int l = 0; l < d; ++l) H[i][l * q1] = hc[i][l];
      H[i][i * q1 + 1] += hd[i];
    }
  }
};

template <int D, int d>
struct Obs<D, d, false> {
  double H[d][D];
  double c[d];
  double R[d][d];

  __device__ __forceinline__ void load(const ModelArgs& a, int k) {
    const double* Hk = a.Hx + (size_t)k * d * D;
    const double* Rk = a.Rx ? a.Rx + (size_t)k * d * d : nullptr;
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int j = 0; j < D; ++j) H[i][j] = Hk[i * D + j];
      c[i] = a.cx[(size_t)k * d + i];
#pragma unroll
      for (int j = 0; j < d; ++j) R[i][j] = Rk ? Rk[i * d + j] : 0.0;
    }
  }
  // R[r][col] for a lane's row r and column col (0 outside d x d)
  __device__ __forceinline__ double r_at(int r, int col) const {
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int j = 0; j < d; ++j) v = (r == i && col == j) ? R[i][j] : v;
    }
    return v;
  }
  __device__ __forceinline__ double dot(int i, const double* v, int st) const {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) acc = fma(H[i][j], v[j * st], acc);
    return acc;
  }
  template <int S, int n, int LD, int NO>
  __device__ __forceinline__ void row_times_tile(int r, const double* T, int sub, double (&out)[NO]) const {
    double h[D];
#pragma unroll
    for (int j = 0; j < D; ++j) h[j] = H[0][j];
#pragma unroll
    for (int i = 1; i < d; ++i) {
#pragma unroll
      for (int j = 0; j < D; ++j) h[j] = (r == i) ? H[i][j] : h[j];
    }
    pof::row_times_tile<S, D, n, LD>(h, T, sub, out);
  }
  __device__ __forceinline__ void left_mul(const double (&K)[d], double (&kh)[D]) const {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < d; ++i) acc = fma(K[i], H[i][j], acc);
      kh[j] = acc;
    }
  }
  template <int S, int n, int LD, int NO>
  __device__ __forceinline__ void sub_kh_tile(const double (&kh)[D], const double* T, int sub,
                                              double (&out)[NO]) const {
#pragma unroll
    for (int j = 0; j < D; ++j) {
#pragma unroll
      for (int t = 0; t < NO; ++t) {
        const int cc = t * S + sub;
        if ((t + 1) * S <= n) {
          out[t] = fma(-kh[j], T[j * LD + cc], out[t]);
        } else if (t * S < n) {
          const double v = T[j * LD + ((cc < n) ? cc : 0)];
          out[t] = fma(-kh[j], (cc < n) ? v : 0.0, out[t]);
        }
      }
      pin(out);
    }
  }
};

// length of the (K H) coefficient vector of Obs::left_mul
template <int D, int d, bool EK1>
constexpr int kh_len() {
  return EK1 ? 2 * d : D;
}

// ----------------------------------------------------------- filter element
// _get_element (pof/parallel_filtsmooth/filter.py:50-81) for element k with
// prior (x0m, x0L) when first, else the zero prior.  Row r < D of A, U, Z is
// returned split over the row's lanes; b and eta (entry r) in every lane of row r.
template <int D, int d, bool EK1, bool first, int SS>
__device__ void filter_element(const Ctx<D, SS>& cx, const ModelArgs& a, int k,
                               double (&A)[Lay<D, SS>::NL], double& b, double (&U)[Lay<D, SS>::NL], double& eta,
                               double (&Z)[Lay<D, SS>::NL]) {
  using L = Lay<D, SS>;
  constexpr int S = L::S, LD = L::LD, NL = L::NL;
  constexpr int NLd = padded<d, S>() / S;
  const int r = cx.row, s = cx.sub;
  const int rr = (r < D) ? r : 0;
  Obs<D, d, EK1> ob;
  ob.load(a, k);

  // LDS (kLdsLeaf / kLdsLeaf1 layouts): tile 0 U rows; tile 1 N1 (first
  // element only); then Psi11 (d rows, stride LD) and K (D rows, stride kMaxObs)
  static_assert(d <= kMaxObs, "observation dimension beyond the element layout");
  double* T2 = cx.tile(0);
  double* T0 = cx.tile(1);
  double* T1 = cx.tile(first ? 2 : 1);
  double* T3 = T1 + kMaxObs * LD;
  double* V0 = cx.vec(0);
  const double* N1 = cx.sN1;
  double m1r = 0.0;
  if constexpr (first) {
    // m1 = F m0, N1 = tria([F L0 | LQ])
    double row[2 * NL];
    {
      double fr[D];
      load_row(fr, cx.sF + rr * LD);
      double fl[NL];
      zero(fl);
#pragma unroll
      for (int k2 = 0; k2 < D; ++k2) {
#pragma unroll
        for (int t = 0; t < NL; ++t) {
          const int c = t * S + s;
          const double v = a.x0L[k2 * D + ((c < D) ? c : 0)];
          fl[t] = fma(fr[k2], (c < D) ? v : 0.0, fl[t]);
        }
        m1r = fma(fr[k2], a.x0m[k2], m1r);
        pin(fl);
      }
      double lq_r[NL];
      load_dist<S, D>(lq_r, cx.sLQ + rr * LD, s);
#pragma unroll
      for (int t = 0; t < NL; ++t) {
        row[t] = (r < D) ? fl[t] : 0.0;
        row[NL + t] = (r < D) ? lq_r[t] : 0.0;
      }
    }
    lq<S, D, D, D, 1, 0, 1 << 20, true, 1 << 20, L::W>(row, r, s, cx.bc());
    if (r < D) {
#pragma unroll
      for (int t = 0; t < NL; ++t) T0[r * LD + t * S + s] = row[t];
      V0[r] = m1r;
    }
    wsync();
    N1 = T0;
  }

  // Psi_ = [[H N1, R], [N1, 0]]  (R = 0 for EK1): rows 0..d-1 and d..d+D-1, column blocks (D, d)
  double psi[NL + NLd];
  {
    double hn[NL];
    ob.template row_times_tile<S, D, LD>(r, N1, s, hn);
    double nr[NL];
    const int rn = (r >= d && r < d + D) ? r - d : 0;
    load_dist<S, D>(nr, N1 + rn * LD, s);
#pragma unroll
    for (int t = 0; t < NL; ++t) psi[t] = (r < d) ? hn[t] : ((r < d + D) ? nr[t] : 0.0);
#pragma unroll
    for (int t = 0; t < NLd; ++t) psi[NL + t] = (r < d) ? ob.r_at(r, t * S + s) : 0.0;
  }
  lq<S, d + D, D, d, 1, 0, 1 << 20, true, 1 << 20, L::W>(psi, r, s, cx.bc());

  // Psi11 -> T1 (compact columns < d lie in the first block)
  if (r < d) store_dist<S, d>(T1 + r * LD, psi, s);
  wsync();
  // K row = Psi21 row Psi11^-1 -> T3, U row = Psi22 -> T2 (rows d..d+D-1 move to r - d)
  {
    double kk[d];
    gather<S, d>(psi, kk);
    if (r >= d && r < d + D) {
      rsolve_lower<d, LD>(kk, T1);
      if (s == 0) {
#pragma unroll
        for (int c = 0; c < d; ++c) T3[(r - d) * kMaxObs + c] = kk[c];
      }
      // compact columns d .. D + d - 1 of the row are U's columns 0 .. D - 1
      constexpr int P1 = L::DP;
#pragma unroll
      for (int t = 0; t < NL + NLd; ++t) {
        const int c = t * S + s;
        const int j = (c < P1) ? c : D + (c - P1);
        const bool in = (c < P1) ? (c < D) : (c - P1 < d);
        if (in && j >= d) T2[(r - d) * LD + (j - d)] = psi[t];
      }
    }
  }
  wsync();

  // hm1 = H m1 + c (every lane); m1 = 0 without a prior
  double hm1[d];
#pragma unroll
  for (int i = 0; i < d; ++i) hm1[i] = first ? ob.c[i] + ob.dot(i, V0, 1) : ob.c[i];
  // w = Psi11^-1 c
  double wv[d];
#pragma unroll
  for (int i = 0; i < d; ++i) wv[i] = ob.c[i];
  lsolve_lower<d, LD>(wv, T1);

  {
    double kr[d];
#pragma unroll
    for (int i = 0; i < d; ++i) kr[i] = T3[rr * kMaxObs + i];
    load_dist<S, D>(U, T2 + rr * LD, s);
    // A = F - (K H) F ; b = m1 - K (H m1 + c)
    double kh[kh_len<D, d, EK1>()];
    ob.left_mul(kr, kh);
    load_dist<S, D>(A, cx.sF + rr * LD, s);
    ob.template sub_kh_tile<S, D, LD>(kh, cx.sF, s, A);
    double bb = m1r;
#pragma unroll
    for (int i = 0; i < d; ++i) bb = fma(-kr[i], hm1[i], bb);
    b = bb;
  }
  // Z row = (Psi11^-1 (H F)[:, r])^T ; eta = -Z Psi11^-1 c
  double z[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    z[i] = ob.dot(i, cx.sF + rr, LD);
    pin1(z[i]);
  }
  lsolve_lower<d, LD>(z, T1);
  {
    double e = 0.0;
#pragma unroll
    for (int i = 0; i < d; ++i) e = fma(-z[i], wv[i], e);
    eta = e;
  }
  if constexpr (D > d) {
    // Z = [Zr | 0]
    own<S, d>(z, s, Z);
  } else {
    // nx <= ny: Z = tria(Zr) with Zr of shape D x d (D == d here)
    double zd[NLd];
    own<S, d>(z, s, zd);
    lq<S, D, d, 0, 1, 0, 1 << 20, true, 1 << 20, L::W>(zd, r, s, cx.bc());
#pragma unroll
    for (int t = 0; t < NL; ++t) Z[t] = zd[t];
  }
  wsync();
}

// ---------------------------------------------------------- filter operators
// Full combine (pof/parallel_filtsmooth/filter.py:117-142): e = e1 o e2, e1
// earlier.  Reads e1 at index i1 of `x1`, e2 at index i2 of `x2`, writes index
// o of `y`.  If `state_only`, e1 is a prefix state (b1, U1) read from
// (s1.m, s1.L) and only (b, U) of the result are produced, into (so.m, so.L).
template <int D, bool state_only, int SS>
__device__ void filter_combine(const Ctx<D, SS>& cx, const FiltArrays& x1, const StateArrays& s1, size_t i1,
                               const FiltArrays& x2, size_t i2, const FiltArrays& y, const StateArrays& so,
                               size_t o) {
  using L = Lay<D, SS>;
  constexpr int S = L::S, LD = L::LD, NL = L::NL;
  constexpr int VS = split_of<D>();  // reduction order: that of the latency form's split
  const int r = cx.row, s = cx.sub;
  const bool act = r < D;
  const int rr = act ? r : 0;  // clamped row for branch-free loads
  // the full combine factors U (rows 0..D-1) and Z (rows D..2D-1) together
  const bool zl = !state_only && r >= D && r < 2 * D;
  const int rz = zl ? r - D : 0;
  double* TU1 = cx.tile(0);
  double* TZ2 = cx.tile(1);
  double* TX11 = cx.tile(2);
  double* TX21 = cx.tile(3);
  // tile 4 and vectors 4-7 exist only in the full-combine LDS layout; Xi22
  // takes the tile of Z2, whose last use precedes the Xi factorisation
  double* TX22 = state_only ? nullptr : TZ2;
  double* TA1 = state_only ? nullptr : cx.tile(4);
  double* Veta2 = cx.vec(0);
  double* Vb1 = cx.vec(1);
  double* Vu = cx.vec(2);
  double* Vs = cx.vec(3);

  const double* U1p = state_only ? s1.L : x1.U;
  const double* b1p = state_only ? s1.m : x1.b;
  // every global read of the combine is issued here, in one batch, so that
  // the memory latency is paid once rather than once per phase
  const double b1r = b1p[i1 * D + rr];
  const double eta2r = x2.eta[i2 * D + rr];
  double a2[NL], u2[NL], z1[NL];
  double b2r, eta1r = 0.0;
  {
    double u1[NL], z2[NL], a1[NL];
    load_dist<S, D>(u1, U1p + (i1 * D + rr) * D, s);
    load_dist<S, D>(z2, x2.Z + (i2 * D + rr) * D, s);
    if constexpr (!state_only) {
      load_dist<S, D>(a1, x1.A + (i1 * D + rr) * D, s);
    }
    if (act) {
#pragma unroll
      for (int t = 0; t < NL; ++t) {
        TU1[r * LD + t * S + s] = u1[t];
        TZ2[r * LD + t * S + s] = z2[t];
        if constexpr (!state_only) TA1[r * LD + t * S + s] = a1[t];
      }
      Vb1[r] = b1r;
      Veta2[r] = eta2r;
    }
  }
  wsync();
  // the terms of eta that need Z2: s2 = eta2 - Z2 Z2^T b1 and U1^T s2 (the
  // rest of eta follows the factorisation)
  if constexpr (!state_only) {
    double* Vt = cx.vec(4);
    double* Vs2 = cx.vec(5);
    double* Vu2 = cx.vec(6);
    {
      const double t = lds_dot<S, D, LD, VS>(TZ2 + rr, Vb1, s);
      if (act) Vt[r] = t;
    }
    wsync();
    const double s2 = eta2r - lds_dot<S, D, 1, VS>(TZ2 + rr * LD, Vt, s);
    if (act) Vs2[r] = s2;
    wsync();
    {
      const double uu = lds_dot<S, D, LD, VS>(TU1 + rr, Vs2, s);
      if (act) Vu2[r] = uu;
    }
  }
  // Xi_ = [[U1^T Z2, I], [Z2, 0]]: rows 0..2D-1, column blocks (D, D)
  // the combines of mfma_products<D>() (latency form) form their tile products on the matrix
  // cores, staging row operands and results in one more LDS tile
  constexpr bool MP = mfma_products<D>() && S == split_of<D>();
  double* TS = MP ? cx.tile(state_only ? 4 : 5) : nullptr;
  double xr[2 * NL];
  {
    double t1[NL];
    if constexpr (MP) {
      tile_mfma<D, LD, true, false>(TU1, TZ2, TX11);  // U1^T Z2 (TX11 is free until Xi11 is stored)
      wsync();
      load_dist<S, D>(t1, TX11 + rr * LD, s);
    } else {
      tile_t_times_tile<S, D, D, LD>(rr, TU1, TZ2, s, t1);
    }
    double zr[NL];
    const int r2 = (r >= D && r < 2 * D) ? r - D : 0;
    load_dist<S, D>(zr, TZ2 + r2 * LD, s);
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      xr[t] = act ? t1[t] : ((r < 2 * D) ? zr[t] : 0.0);
      xr[NL + t] = (act && t * S + s == r) ? 1.0 : 0.0;
    }
  }
  // the rows of A2, U2 (and Z1) are needed only after this factorisation: their
  // loads are issued half way through it, where fewer registers are live and
  // the remaining pivot steps hide the memory latency
  if constexpr (state_only) {
    // Xi11 and Xi21 are final after the first D pivots; the last D pivots only
    // form Xi22, which enters Z alone
    lqv<S, VS, 2 * D, D, D, 1, 0, D / 2, false, 0, L::W>(xr, r, s, cx.bc());
    load_dist<S, D>(a2, x2.A + (i2 * D + rr) * D, s);
    load_dist<S, D>(u2, x2.U + (i2 * D + rr) * D, s);
    b2r = x2.b[i2 * D + rr];
    zero(z1);
    lqv<S, VS, 2 * D, D, D, 1, D / 2, D, true, 0, L::W>(xr, r, s, cx.bc());
  } else {
    lqv<S, VS, 2 * D, D, D, 1, 0, D, false, 0, L::W>(xr, r, s, cx.bc());
    load_dist<S, D>(a2, x2.A + (i2 * D + rr) * D, s);
    load_dist<S, D>(u2, x2.U + (i2 * D + rr) * D, s);
    load_dist<S, D>(z1, x1.Z + (i1 * D + rz) * D, s);  // Z1 rows feed the rows D..2D-1
    b2r = x2.b[i2 * D + rr];
    eta1r = x1.eta[i1 * D + rr];
    lqv<S, VS, 2 * D, D, D, 1, D, 1 << 20, true, 1 << 20, L::W>(xr, r, s, cx.bc());
  }
  if (act) {
#pragma unroll
    for (int t = 0; t < NL; ++t) TX11[r * LD + t * S + s] = xr[t];
  } else if (r < 2 * D) {
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      TX21[(r - D) * LD + t * S + s] = xr[t];
      if constexpr (!state_only) TX22[(r - D) * LD + t * S + s] = xr[NL + t];  // full layout only
    }
  }
  wsync();

  // MT = A2 U1 Xi11^-T ; W = A2 - MT Xi21^T
  double mt[NL], wr[NL];
  if constexpr (MP) {
    stage_row<S, D, LD>(TS, r, s, a2);
    wsync();
    tile_mfma<D, LD, false, false>(TS, TU1, TS);  // A2 U1
    wsync();
    load_dist<S, D>(mt, TS + rr * LD, s);
  } else {
    double a2f[D];
    gather<S, D>(a2, a2f);
    row_times_tile<S, D, D, LD>(a2f, TU1, s, mt);
  }
  rsolve_lowerT_s<S, D, LD>(mt, TX11, s);
  if constexpr (MP) {
    wsync();  // every row has read A2 U1
    stage_row<S, D, LD>(TS, r, s, mt);
    wsync();
    tile_mfma<D, LD, false, true>(TS, TX21, TS);  // MT Xi21^T
    wsync();
    double p[NL];
    load_dist<S, D>(p, TS + rr * LD, s);
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      wr[t] = (t * S + s < D) ? a2[t] - p[t] : 0.0;
      pin1(wr[t]);
    }
  } else {
    double mtf[D];
    gather<S, D>(mt, mtf);
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int c = t * S + s;
      const int cc = (c < D) ? c : 0;
      double acc = a2[t];
#pragma unroll
      for (int j = 0; j < D; ++j) acc = fma(-mtf[j], TX21[cc * LD + j], acc);
      wr[t] = (c < D) ? acc : 0.0;
      pin1(wr[t]);
    }
  }

  // b = W (b1 + U1 U1^T eta2) + b2   (all global stores are issued at the end:
  // a store in flight would make the next wait on a load also wait for it)
  double bb;
  {
    const double u = lds_dot<S, D, LD, VS>(TU1 + rr, Veta2, s);
    if (act) Vu[r] = u;
    wsync();
    const double sv = Vb1[rr] + lds_dot<S, D, 1, VS>(TU1 + rr * LD, Vu, s);
    if (act) Vs[r] = sv;
    wsync();
    bb = b2r + row_dot<S, D, VS>(wr, Vs, s);
    pin1(bb);  // formed here: sunk to its store, it holds b2 and partial sums through the last factorisation
  }
  // eta = A1^T (I - Xi21 Xi11^-1 U1^T)(eta2 - Z2 Z2^T b1) + eta1 and A = W A1
  // (A's rows go to the LDS of U1, unused from here on, so that they are not
  // held in registers through the factorisation below)
  double eo = 0.0;
  if constexpr (!state_only) {
    const double* Vs2 = cx.vec(5);
    const double* Vu2 = cx.vec(6);
    double* Vs3 = cx.vec(7);
    const double s2 = Vs2[rr];
    double w2[NL];
    load_dist<S, D>(w2, Vu2, s);
    rsolve_lowerT_s<S, D, LD>(w2, TX11, s);  // Xi11^-1 (U1^T s2)
    {
      double tx[NL];
      load_dist<S, D>(tx, TX21 + rr * LD, s);
      const double v2 = reg_dot<S, D, VS>(tx, w2);
      if (act) Vs3[r] = s2 - v2;
    }
    wsync();
    eo = eta1r + lds_dot<S, D, LD, VS>(TA1 + rr, Vs3, s);
    pin1(eo);
    if constexpr (MP) {
      wsync();  // every row has read MT Xi21^T
      stage_row<S, D, LD>(TS, r, s, wr);
      wsync();
      tile_mfma<D, LD, false, false>(TS, TA1, TU1);  // A = W A1 into U1's tile: U1's last reads were above
      wsync();
    } else {
      double wf[D];
      gather<S, D>(wr, wf);
      double ar[NL];
      row_times_tile<S, D, D, LD>(wf, TA1, s, ar);
      wsync();  // every row has read U1 for the last time
      if (act) {
#pragma unroll
        for (int t = 0; t < NL; ++t) TU1[r * LD + t * S + s] = ar[t];
      }
    }
  }
  // U = tria([MT | U2]) in rows 0..D-1 and, for the full combine,
  // Z = tria([A1^T Xi22 | Z1]) in rows D..2D-1, as one paired LQ
  double jr[2 * NL];
  {
    double tz[NL];
    if constexpr (!state_only && MP) {
      tile_mfma<D, LD, true, false>(TA1, TX22, TS);  // A1^T Xi22
      wsync();
      load_dist<S, D>(tz, TS + rz * LD, s);
    } else if constexpr (!state_only) {
      tile_t_times_tile<S, D, D, LD>(rz, TA1, TX22, s, tz);
    } else {
      zero(tz);
    }
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      jr[t] = act ? mt[t] : (zl ? tz[t] : 0.0);
      jr[NL + t] = act ? u2[t] : (zl ? z1[t] : 0.0);
    }
  }
  // bands: U2 is lower triangular; Z1 too, except a leaf element's [Zr | 0]
  // (nonzero in its first d columns only)
  lqv<S, VS, D, D, D, state_only ? 1 : 2, 0, 1 << 20, true, state_only ? 0 : max_obs_dim(D) - 1, L::W>(jr, r, s,
                                                                                                        cx.bc());
  // output offsets formed here, not hoisted above the factorisation (where
  // they would hold registers through it)
  o = opaque_sz(o);
  if (act) {
    if (s == 0) (state_only ? so.m : y.b)[o * D + r] = bb;
    store_dist<S, D>((state_only ? so.L : y.U) + (o * D + r) * D, jr, s);
    if constexpr (!state_only) {
      double ar[NL];
      load_dist<S, D>(ar, TU1 + r * LD, s);
      store_dist<S, D>(y.A + (o * D + r) * D, ar, s);
      if (s == 0) y.eta[o * D + r] = eo;
    }
  } else if (zl) {
    store_dist<S, D>(y.Z + (o * D + rz) * D, jr, s);
  }
  wsync();
}

// ------------------------------------------------------- smoother operators
// (pof/parallel_filtsmooth/smoother.py:53-63) under reverse=True: e1 is the
// aggregate later in time.  full: (g, E, D) = (E2 g1 + g2, E2 E1,
// tria([E2 D1 | D2])).  state_only: (m, L) = (E2 m1 + g2, tria([E2 L1 | D2])).
// with_chol = false skips every D/L term (means-only iterations).
template <int D, bool state_only, int SS>
__device__ void smoother_combine(const Ctx<D, SS>& cx, const SmoArrays& x1, const StateArrays& s1, size_t i1,
                                 const SmoArrays& x2, size_t i2, const SmoArrays& y, const StateArrays& so,
                                 size_t o, bool with_chol) {
  using L = Lay<D, SS>;
  constexpr int S = L::S, LD = L::LD, NL = L::NL;
  constexpr int VS = split_of<D>();  // reduction order: that of the latency form's split
  const int r = cx.row, s = cx.sub;
  const bool act = r < D;
  const int rr = act ? r : 0;
  double* TE1 = cx.tile(0);
  double* TD1 = cx.tile(1);
  double* Vg1 = cx.vec(0);
  const double* g1p = state_only ? s1.m : x1.g;
  const double* D1p = state_only ? s1.L : x1.Dm;
  // all global reads first (one memory latency)
  double e2[NL], d2[NL];
  const double g2r = x2.g[i2 * D + rr];
  load_dist<S, D>(e2, x2.E + (i2 * D + rr) * D, s);
  if (with_chol) load_dist<S, D>(d2, x2.Dm + (i2 * D + rr) * D, s);
  {
    const double g1 = g1p[i1 * D + rr];
    double e1[NL], d1[NL];
    if (!state_only) load_dist<S, D>(e1, x1.E + (i1 * D + rr) * D, s);
    if (with_chol) load_dist<S, D>(d1, D1p + (i1 * D + rr) * D, s);
    if (act) {
      Vg1[r] = g1;
      if (!state_only) {
#pragma unroll
        for (int t = 0; t < NL; ++t) TE1[r * LD + t * S + s] = e1[t];
      }
      if (with_chol) {
#pragma unroll
        for (int t = 0; t < NL; ++t) TD1[r * LD + t * S + s] = d1[t];
      }
    }
  }
  wsync();
  // results stay in registers; every global store is issued at the end so
  // that no wait on outstanding memory operations lands inside the LQ below
  const double gg = g2r + row_dot<S, D, VS>(e2, Vg1, s);
  // E2 E1 and E2 D1: on the matrix cores for mfma_products<D>() (latency form), E2's rows staged in
  // the extra tile and each product written over its right operand's tile
  constexpr bool MP = mfma_products<D>() && S == split_of<D>();
  double er[state_only ? 1 : NL];
  double t1[NL];
  if constexpr (MP) {
    if (!state_only || with_chol) {
      double* TS = cx.tile(2);
      stage_row<S, D, LD>(TS, r, s, e2);
      wsync();
      if constexpr (!state_only) tile_mfma<D, LD, false, false>(TS, TE1, TE1);
      if (with_chol) tile_mfma<D, LD, false, false>(TS, TD1, TD1);
      wsync();
      if constexpr (!state_only) load_dist<S, D>(er, TE1 + rr * LD, s);
      if (with_chol) load_dist<S, D>(t1, TD1 + rr * LD, s);
    }
  } else {
    double e2f[D];
    gather<S, D>(e2, e2f);
    if constexpr (!state_only) row_times_tile<S, D, D, LD>(e2f, TE1, s, er);
    if (with_chol) row_times_tile<S, D, D, LD>(e2f, TD1, s, t1);
  }
  double dr[2 * NL];
  if (with_chol) {
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      dr[t] = t1[t];
      dr[NL + t] = d2[t];
    }
    lqv<S, VS, D, D, D, 1, 0, 1 << 20, true, 1 << 20, L::W>(dr, r, s, cx.bc());
  }
  if (act) {
    if (s == 0) (state_only ? so.m : y.g)[o * D + r] = gg;
    if constexpr (!state_only) store_dist<S, D>(y.E + (o * D + r) * D, er, s);
    if (with_chol) store_dist<S, D>((state_only ? so.L : y.Dm) + (o * D + r) * D, dr, s);
  }
  wsync();
}

}  // namespace pof

#pragma clang fp contract(fast)
