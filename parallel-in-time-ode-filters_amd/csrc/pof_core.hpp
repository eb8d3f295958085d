// Device building blocks of the square-root parallel IEKS on CDNA (gfx950).
//
// Execution model: one "worker" of W lanes (W = 32 for D <= 16, 64 for
// D <= 32; two or one workers per 64-wide wavefront) owns one scan element.
// Lane r of a worker holds row r of every matrix of the element in registers
// (row-per-lane layout), so the row-local parts of products, triangular
// solves and Householder updates need no communication.  Rows that other
// lanes need are staged through the worker's slice of LDS; a worker never
// synchronises with other workers, only with its own wavefront.
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

namespace pof {

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  asm volatile("" ::: "memory");
}

// ------------------------------------------------------------------ layout
// LDS layout kinds: how many tiles / broadcast vectors a worker of each
// kernel needs and whether the block stages the model constants F, LQ, N1z.
enum LdsKind : int { kLdsFull = 0, kLdsState, kLdsSmooth, kLdsLeaf, kLdsPost };

template <int K>
struct LdsNeeds {
  static constexpr int NT = (K == kLdsFull) ? 6 : (K == kLdsState || K == kLdsLeaf) ? 4 : (K == kLdsPost) ? 5 : 2;
  static constexpr int NV = (K == kLdsFull) ? 8 : (K == kLdsState) ? 4 : (K == kLdsPost) ? 2 : 1;
  static constexpr bool CONSTS = (K == kLdsLeaf || K == kLdsPost);
};

template <int D_, int K = kLdsFull>
struct Geo {
  static constexpr int D = D_;
  static constexpr int W = (2 * D_ <= 32) ? 32 : 64;      // lanes per worker
  static constexpr int LD = D_ + 1;                      // padded LDS row stride
  static constexpr int TILE = D_ * LD;
  static constexpr int NT = LdsNeeds<K>::NT;             // tiles per worker
  static constexpr int NV = LdsNeeds<K>::NV;             // broadcast vectors per worker
  static constexpr int BC = 4 * D_ + 2;                  // Householder broadcast rows (two matrices)
  static constexpr int WORKER = BC + NV * 2 * D_ + NT * TILE;
  static constexpr int CONSTS = LdsNeeds<K>::CONSTS ? 3 * TILE : 0;  // F, LQ, N1z
  static constexpr size_t bytes_for(int wpb) { return sizeof(double) * (CONSTS + wpb * WORKER); }
  // workers per block: as many as fit in 64 KiB of LDS, at most 256 threads
  static constexpr int WPB = (W == 32 && bytes_for(8) <= 65536)   ? 8
                             : (W == 32 && bytes_for(4) <= 65536) ? 4
                             : (bytes_for(2) <= 65536)            ? 2
                                                                  : 1;
  static constexpr int BLOCK = W * WPB;
  static constexpr size_t lds_bytes() { return bytes_for(WPB); }
  static_assert(bytes_for(WPB) <= 65536, "state dimension too large for one worker per block");
};

template <int D>
struct Ctx {
  static constexpr int LD = D + 1;
  static constexpr int TILE = D * LD;
  int lane;           // lane inside the worker
  int wid;            // global worker index
  double* w;          // worker LDS slice (starts with the Householder broadcast row)
  double* vbase;      // broadcast vectors
  double* tbase;      // tiles
  const double* sF;   // block LDS constants, row stride LD (kernels that stage them)
  const double* sLQ;
  const double* sN1;
  __device__ double* tile(int i) const { return tbase + i * TILE; }
  __device__ double* vec(int i) const { return vbase + i * 2 * D; }
  __device__ double* bc() const { return w; }
};

// Stage F, LQ, N1z into block LDS (kinds that need them) and carve out this
// thread's worker slice.  Must be called by every thread of the block before
// any early exit.
template <int D, int K>
__device__ __forceinline__ Ctx<D> make_ctx(const double* F, const double* LQ, const double* N1z, double* smem) {
  using G = Geo<D, K>;
  if constexpr (G::CONSTS > 0) {
    for (int t = threadIdx.x; t < D * D; t += G::BLOCK) {
      const int r = t / D, c = t % D;
      smem[r * G::LD + c] = F[t];
      smem[G::TILE + r * G::LD + c] = LQ[t];
      smem[2 * G::TILE + r * G::LD + c] = N1z[t];
    }
    __syncthreads();
  }
  Ctx<D> cx;
  cx.lane = threadIdx.x % G::W;
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
// Opaque copy of a value: keeps the optimiser from correlating the many lane
// tests of one function (which otherwise duplicates whole regions per branch
// and multiplies register pressure).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

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

// out[c] = sum_k x[k] * T[k][c]   (row of X @ T), T with row stride LD
template <int K, int NC, int LD, int NX>
__device__ __forceinline__ void row_times_tile(const double (&x)[NX], const double* T, double (&out)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = fma(x[k], T[k * LD + c], out[c]);
    pin(out);
  }
}

// out[c] = sum_k T1[k][r] * T2[k][c]   (row r of T1^T @ T2)
template <int K, int NC, int LD>
__device__ __forceinline__ void tile_t_times_tile(int r, const double* T1, const double* T2, double (&out)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t = T1[k * LD + r];
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = fma(t, T2[k * LD + c], out[c]);
    pin(out);
  }
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

// Householder LQ of NG independent NR x NC matrices; row r of matrix g sits
// in lane g * NR + r.  On return that lane holds row r of the lower-
// triangular factor L of its matrix (columns > r are zero), L L^T = M M^T:
// the reflectors of LAPACK dgeqrf on M^T (dlarfg: beta = -sign(alpha) *
// norm, tau = (beta - alpha) / beta, v = [1, x / (alpha - beta)], tau = 0
// when the tail x is zero, no reflector for a length-1 tail).
//
// Per pivot step the pivot lane broadcasts its raw row through LDS and every
// lane derives (beta, tau, scale) itself, so there is no divergent pivot
// section; the dot products with the unscaled tail overlap the square root
// and reciprocals.  Finished rows are not cleared per step; the strict upper
// part is zeroed once at the end.
template <int NR, int NC, int NG = 1>
__device__ __forceinline__ void lq(double (&a)[NC], int lane_in, double* bc) {
  static_assert(NG == 1 || NG == 2, "one or two matrices");
  constexpr int K = (NR < NC) ? NR : NC;
  // opaque copy of the lane index: stops the optimiser from threading the
  // per-pivot lane tests into specialised copies of the surrounding code
  int lane = lane_in;
  asm volatile("" : "+v"(lane));
  const int g = (NG == 2 && lane >= NR) ? 1 : 0;
  const int li = lane - g * NR;
  const bool valid = lane < NG * NR;
  double* row = bc + g * NC;  // broadcast row of this lane's matrix
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i == NC - 1) break;  // a length-1 tail: H = I
    if (valid && li == i) {
#pragma unroll
      for (int j = i; j < NC; ++j) row[j] = a[j];
    }
    wsync();
    double p[NC];
#pragma unroll
    for (int j = i; j < NC; ++j) p[j] = row[j];
    const double alpha = p[i];
    // sum of squares of the pivot tail and this row's dot with it, in four
    // independent chains each
    double sq[4] = {0.0, 0.0, 0.0, 0.0};
    double dt[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = i + 1; j < NC; ++j) {
      sq[(j - i - 1) & 3] = fma(p[j], p[j], sq[(j - i - 1) & 3]);
      dt[(j - i - 1) & 3] = fma(a[j], p[j], dt[(j - i - 1) & 3]);
    }
    const double sig = (sq[0] + sq[1]) + (sq[2] + sq[3]);
    const double dot = (dt[0] + dt[1]) + (dt[2] + dt[3]);
    const double nrm = sqrt(fma(alpha, alpha, sig));
    const double bc_ = -copysign(nrm, alpha);  // Fortran SIGN: honours the sign of -0.0
    const double den = alpha - bc_;
    const double sc_ = rcp(den);
    const double tc_ = -den * rcp(bc_);
    const bool nz = sig != 0.0;
    const double beta = nz ? bc_ : alpha;
    const double scal = nz ? sc_ : 0.0;
    const double tau = nz ? tc_ : 0.0;
    if (valid && li == i) a[i] = beta;
    if (valid && li > i) {
      const double tw = tau * fma(scal, dot, a[i]);
      a[i] -= tw;
      const double tws = tw * scal;
#pragma unroll
      for (int j = i + 1; j < NC; ++j) a[j] = fma(-tws, p[j], a[j]);
    }
  }
#pragma unroll
  for (int j = 1; j < NC; ++j) a[j] = (valid && j > li) ? 0.0 : a[j];
  wsync();
}

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
// Affine observation of element k (noiseless): y = H x + c, either from EK1 at
// lin[k+1] (pof/step.py:12-22, pof/observations.py:35-40: H = E1 - J E0,
// c = J E0 m - f(E0 m)) or from explicit arrays.  Every lane computes the
// whole (small) H and c.
template <int D, int d>
__device__ __forceinline__ void observation(const ModelArgs& a, int k, double (&H)[d][D], double (&c)[d]) {
  if (a.lin != nullptr) {
    constexpr int q1 = D / d;
    const double* m = a.lin + (size_t)(k + 1) * D;
    double y[d];
#pragma unroll
    for (int i = 0; i < d; ++i) y[i] = a.Pd[i * q1] * m[i * q1];
    double f[d];
    double J[d][d];
    field_and_jacobian<d>(a.kind, a.params, y, f, J);
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int j = 0; j < D; ++j) H[i][j] = 0.0;
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < d; ++l) {
        H[i][l * q1] = -J[i][l] * a.Pd[l * q1];
        acc = fma(J[i][l], y[l], acc);
      }
      H[i][i * q1 + 1] += a.Pd[i * q1 + 1];
      c[i] = acc - f[i];
    }
  } else {
    const double* Hk = a.Hx + (size_t)k * d * D;
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int j = 0; j < D; ++j) H[i][j] = Hk[i * D + j];
      c[i] = a.cx[(size_t)k * d + i];
    }
  }
}

// select row `lane` (< d) of a register matrix without dynamic indexing
template <int d, int D>
__device__ __forceinline__ void select_row(const double (&H)[d][D], int lane, double (&h)[D]) {
#pragma unroll
  for (int j = 0; j < D; ++j) h[j] = H[0][j];
#pragma unroll
  for (int i = 1; i < d; ++i) {
#pragma unroll
    for (int j = 0; j < D; ++j) h[j] = (lane == i) ? H[i][j] : h[j];
  }
}

// ----------------------------------------------------------- filter element
// _get_element (pof/parallel_filtsmooth/filter.py:50-81) for element k with
// prior (x0m, x0L) when first, else the zero prior.  Lane r < D returns row r.
template <int D, int d>
__device__ void filter_element(const Ctx<D>& cx, const ModelArgs& a, int k, bool first, double (&A)[D],
                               double& b, double (&U)[D], double& eta, double (&Z)[D]) {
  constexpr int LD = D + 1;
  const int r = opaque(cx.lane);
  double H[d][D];
  double cv[d];
  observation<D, d>(a, k, H, cv);

  double* T0 = cx.tile(0);
  double* T1 = cx.tile(1);
  double* T2 = cx.tile(2);
  double* T3 = cx.tile(3);
  double* V0 = cx.vec(0);
  const double* N1 = cx.sN1;
  double m1r = 0.0;
  if (first) {
    // m1 = F m0, N1 = tria([F L0 | LQ])
    double row[2 * D];
    if (r < D) {
      double fr[D];
      load_row(fr, cx.sF + r * LD);
      double fl[D];
      zero(fl);
#pragma unroll
      for (int k2 = 0; k2 < D; ++k2) {
#pragma unroll
        for (int c = 0; c < D; ++c) fl[c] = fma(fr[k2], a.x0L[k2 * D + c], fl[c]);
        m1r = fma(fr[k2], a.x0m[k2], m1r);
        pin(fl);
      }
#pragma unroll
      for (int c = 0; c < D; ++c) {
        row[c] = fl[c];
        row[D + c] = cx.sLQ[r * LD + c];
      }
    } else {
      zero(row);
    }
    lq<D, 2 * D>(row, r, cx.bc());
    if (r < D) {
#pragma unroll
      for (int c = 0; c < D; ++c) T0[r * LD + c] = row[c];
      V0[r] = m1r;
    }
    wsync();
    N1 = T0;
  }

  // Psi_ = [[H N1, R], [N1, 0]]  (R = 0), rows in lanes 0..d+D-1
  double psi[D + d];
  {
    double h[D];
    select_row<d, D>(H, r, h);
    if (r < d) {
      double hn[D];
      row_times_tile<D, D, LD>(h, N1, hn);
#pragma unroll
      for (int c = 0; c < D; ++c) psi[c] = hn[c];
    } else if (r < d + D) {
#pragma unroll
      for (int c = 0; c < D; ++c) psi[c] = N1[(r - d) * LD + c];
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) psi[c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < d; ++c) psi[D + c] = 0.0;
  }
  lq<d + D, D + d>(psi, r, cx.bc());

  // Psi11 -> T1
  if (r < d) {
#pragma unroll
    for (int c = 0; c < d; ++c) T1[r * LD + c] = psi[c];
  }
  wsync();
  // K row = Psi21 row Psi11^-1 (lanes d..d+D-1), then move K and U rows to lane r - d
  if (r >= d && r < d + D) {
    double kk[d];
#pragma unroll
    for (int c = 0; c < d; ++c) kk[c] = psi[c];
    rsolve_lower<d, LD>(kk, T1);
#pragma unroll
    for (int c = 0; c < d; ++c) T3[(r - d) * LD + c] = kk[c];
#pragma unroll
    for (int c = 0; c < D; ++c) T2[(r - d) * LD + c] = psi[d + c];
  }
  wsync();

  // HF = H F (every lane), hm1 = H m1 + c
  double HF[d][D];
  double hm1[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
#pragma unroll
    for (int c = 0; c < D; ++c) HF[i][c] = 0.0;
    hm1[i] = cv[i];
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    double fj[D];
    load_row(fj, cx.sF + j * LD);
    const double m1j = first ? V0[j] : 0.0;
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int c = 0; c < D; ++c) HF[i][c] = fma(H[i][j], fj[c], HF[i][c]);
      hm1[i] = fma(H[i][j], m1j, hm1[i]);
      pin(HF[i]);
    }
  }
  pin(hm1);
  // w = Psi11^-1 c
  double wv[d];
#pragma unroll
  for (int i = 0; i < d; ++i) wv[i] = cv[i];
  lsolve_lower<d, LD>(wv, T1);

  const int rr = (r < D) ? r : 0;
  {
    double kr[d];
#pragma unroll
    for (int i = 0; i < d; ++i) kr[i] = T3[rr * LD + i];
    load_row(U, T2 + rr * LD);
    // A = F - K H F ; b = m1 - K (H m1 + c)
    double fr[D];
    load_row(fr, cx.sF + rr * LD);
    double bb = m1r;
#pragma unroll
    for (int c = 0; c < D; ++c) A[c] = fr[c];
#pragma unroll
    for (int i = 0; i < d; ++i) {
#pragma unroll
      for (int c = 0; c < D; ++c) A[c] = fma(-kr[i], HF[i][c], A[c]);
      bb = fma(-kr[i], hm1[i], bb);
    }
    b = bb;
  }
  // Z row = (Psi11^-1 HF[:, r])^T ; eta = -Z Psi11^-1 c
  double z[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) acc = fma(H[i][j], cx.sF[j * LD + rr], acc);
    z[i] = acc;
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
#pragma unroll
    for (int c = 0; c < D; ++c) Z[c] = (c < d) ? z[(c < d) ? c : 0] : 0.0;
  } else {
    // nx <= ny: Z = tria(Z) with Z of shape D x d
    lq<D, d>(z, r, cx.bc());
#pragma unroll
    for (int c = 0; c < D; ++c) Z[c] = z[c];
  }
  wsync();
}

// ---------------------------------------------------------- filter operators
// Full combine (pof/parallel_filtsmooth/filter.py:117-142): e = e1 o e2, e1
// earlier.  Reads e1 at index i1 of `x1`, e2 at index i2 of `x2`, writes index
// o of `y`.  If `state_only`, e1 is a prefix state (b1, U1) read from
// (s1.m, s1.L) and only (b, U) of the result are produced, into (so.m, so.L).
template <int D, bool state_only>
__device__ void filter_combine(const Ctx<D>& cx, const FiltArrays& x1, const StateArrays& s1, size_t i1,
                               const FiltArrays& x2, size_t i2, const FiltArrays& y, const StateArrays& so,
                               size_t o) {
  constexpr int LD = D + 1;
  const int r = opaque(cx.lane);
  const bool act = r < D;
  const int rr = act ? r : 0;  // clamped row for branch-free loads
  // the full combine factors U (lanes 0..D-1) and Z (lanes D..2D-1) together
  const bool zl = !state_only && r >= D && r < 2 * D;
  const int rz = zl ? r - D : 0;
  double* TU1 = cx.tile(0);
  double* TZ2 = cx.tile(1);
  double* TX11 = cx.tile(2);
  double* TX21 = cx.tile(3);
  // tiles 4-5 and vectors 4-7 exist only in the full-combine LDS layout
  double* TX22 = state_only ? nullptr : cx.tile(4);
  double* TA1 = state_only ? nullptr : cx.tile(5);
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
  const double b2r = x2.b[i2 * D + rr];
  double a2[D], u2[D], z1[D];
  double eta1r = 0.0;
  {
    double u1[D], z2[D], a1[D];
    load_row(u1, U1p + (i1 * D + rr) * D);
    load_row(z2, x2.Z + (i2 * D + rr) * D);
    load_row(a2, x2.A + (i2 * D + rr) * D);
    load_row(u2, x2.U + (i2 * D + rr) * D);
    if constexpr (!state_only) {
      load_row(a1, x1.A + (i1 * D + rr) * D);
      load_row(z1, x1.Z + (i1 * D + rz) * D);  // Z1 rows feed the lanes D..2D-1
      eta1r = x1.eta[i1 * D + rr];
    }
    if (act) {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        TU1[r * LD + c] = u1[c];
        TZ2[r * LD + c] = z2[c];
        if constexpr (!state_only) TA1[r * LD + c] = a1[c];
      }
      Vb1[r] = b1r;
      Veta2[r] = eta2r;
    }
  }
  wsync();
  // Xi_ = [[U1^T Z2, I], [Z2, 0]]
  double xr[2 * D];
  if (r < D) {
    double t[D];
    tile_t_times_tile<D, D, LD>(r, TU1, TZ2, t);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      xr[c] = t[c];
      xr[D + c] = (c == r) ? 1.0 : 0.0;
    }
  } else if (r < 2 * D) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      xr[c] = TZ2[(r - D) * LD + c];
      xr[D + c] = 0.0;
    }
  } else {
    zero(xr);
  }
  lq<2 * D, 2 * D>(xr, r, cx.bc());
  if (r < D) {
#pragma unroll
    for (int c = 0; c < D; ++c) TX11[r * LD + c] = xr[c];
  } else if (r < 2 * D) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      TX21[(r - D) * LD + c] = xr[c];
      if constexpr (!state_only) TX22[(r - D) * LD + c] = xr[D + c];  // tile 4 exists in the full layout only
    }
  }
  wsync();

  // MT = A2 U1 Xi11^-T ; W = A2 - MT Xi21^T
  double mt[D], wr[D];
  row_times_tile<D, D, LD>(a2, TU1, mt);
  rsolve_lowerT<D, LD>(mt, TX11);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double acc = a2[c];
#pragma unroll
    for (int j = 0; j < D; ++j) acc = fma(-mt[j], TX21[c * LD + j], acc);
    wr[c] = acc;
    pin1(wr[c]);
  }

  // b = W (b1 + U1 U1^T eta2) + b2   (all global stores are issued at the end:
  // a store in flight would make the next wait on a load also wait for it)
  double bb;
  {
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) u = fma(TU1[k * LD + rr], Veta2[k], u);
    if (act) Vu[r] = u;
    wsync();
    double sv = b1r;
#pragma unroll
    for (int j = 0; j < D; ++j) sv = fma(TU1[rr * LD + j], Vu[j], sv);
    if (act) Vs[r] = sv;
    wsync();
    bb = b2r;
#pragma unroll
    for (int k = 0; k < D; ++k) bb = fma(wr[k], Vs[k], bb);
  }
  // A = W A1 and eta = A1^T (I - Xi21 Xi11^-1 U1^T)(eta2 - Z2 Z2^T b1) + eta1
  double ar[state_only ? 1 : D];
  double eo = 0.0;
  if constexpr (!state_only) {
    double* Vt = cx.vec(4);
    double* Vs2 = cx.vec(5);
    double* Vu2 = cx.vec(6);
    double* Vs3 = cx.vec(7);
    row_times_tile<D, D, LD>(wr, TA1, ar);
    {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) t = fma(TZ2[k * LD + rr], Vb1[k], t);
      if (act) Vt[r] = t;
    }
    wsync();
    double s2 = eta2r;
#pragma unroll
    for (int j = 0; j < D; ++j) s2 = fma(-TZ2[rr * LD + j], Vt[j], s2);
    if (act) Vs2[r] = s2;
    wsync();
    {
      double uu = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) uu = fma(TU1[k * LD + rr], Vs2[k], uu);
      if (act) Vu2[r] = uu;
    }
    wsync();
    double w2[D];
    load_row(w2, Vu2);
    lsolve_lower<D, LD>(w2, TX11);
    {
      double v2 = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) v2 = fma(TX21[rr * LD + j], w2[j], v2);
      if (act) Vs3[r] = s2 - v2;
    }
    wsync();
    eo = eta1r;
#pragma unroll
    for (int k = 0; k < D; ++k) eo = fma(TA1[k * LD + rr], Vs3[k], eo);
  }
  // U = tria([MT | U2]) in lanes 0..D-1 and, for the full combine,
  // Z = tria([A1^T Xi22 | Z1]) in lanes D..2D-1, as one paired LQ
  double jr[2 * D];
  if (act) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      jr[c] = mt[c];
      jr[D + c] = u2[c];
    }
  } else if (zl) {
    double t[D];
    tile_t_times_tile<D, D, LD>(rz, TA1, TX22, t);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      jr[c] = t[c];
      jr[D + c] = z1[c];
    }
  } else {
    zero(jr);
  }
  lq<D, 2 * D, state_only ? 1 : 2>(jr, r, cx.bc());
  const double(&lo)[D] = *reinterpret_cast<const double(*)[D]>(jr);
  if (act) {
    (state_only ? so.m : y.b)[o * D + r] = bb;
    store_row((state_only ? so.L : y.U) + (o * D + r) * D, lo);
    if constexpr (!state_only) {
      store_row(y.A + (o * D + r) * D, ar);
      y.eta[o * D + r] = eo;
    }
  } else if (zl) {
    store_row(y.Z + (o * D + rz) * D, lo);
  }
  wsync();
}

// ------------------------------------------------------- smoother operators
// (pof/parallel_filtsmooth/smoother.py:53-63) under reverse=True: e1 is the
// aggregate later in time.  full: (g, E, D) = (E2 g1 + g2, E2 E1,
// tria([E2 D1 | D2])).  state_only: (m, L) = (E2 m1 + g2, tria([E2 L1 | D2])).
// with_chol = false skips every D/L term (means-only iterations).
template <int D, bool state_only>
__device__ void smoother_combine(const Ctx<D>& cx, const SmoArrays& x1, const StateArrays& s1, size_t i1,
                                 const SmoArrays& x2, size_t i2, const SmoArrays& y, const StateArrays& so,
                                 size_t o, bool with_chol) {
  constexpr int LD = D + 1;
  const int r = opaque(cx.lane);
  const bool act = r < D;
  const int rr = act ? r : 0;
  double* TE1 = cx.tile(0);
  double* TD1 = cx.tile(1);
  double* Vg1 = cx.vec(0);
  const double* g1p = state_only ? s1.m : x1.g;
  const double* D1p = state_only ? s1.L : x1.Dm;
  // all global reads first (one memory latency)
  double e2[D], d2[D];
  const double g2r = x2.g[i2 * D + rr];
  load_row(e2, x2.E + (i2 * D + rr) * D);
  if (with_chol) load_row(d2, x2.Dm + (i2 * D + rr) * D);
  {
    const double g1 = g1p[i1 * D + rr];
    double e1[D], d1[D];
    if (!state_only) load_row(e1, x1.E + (i1 * D + rr) * D);
    if (with_chol) load_row(d1, D1p + (i1 * D + rr) * D);
    if (act) {
      Vg1[r] = g1;
      if (!state_only) {
#pragma unroll
        for (int c = 0; c < D; ++c) TE1[r * LD + c] = e1[c];
      }
      if (with_chol) {
#pragma unroll
        for (int c = 0; c < D; ++c) TD1[r * LD + c] = d1[c];
      }
    }
  }
  wsync();
  // results stay in registers; every global store is issued at the end so
  // that no wait on outstanding memory operations lands inside the LQ below
  double gg = g2r;
#pragma unroll
  for (int k = 0; k < D; ++k) gg = fma(e2[k], Vg1[k], gg);
  double er[state_only ? 1 : D];
  if constexpr (!state_only) row_times_tile<D, D, LD>(e2, TE1, er);
  double dr[2 * D];
  if (with_chol) {
    {
      double t[D];
      row_times_tile<D, D, LD>(e2, TD1, t);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        dr[c] = t[c];
        dr[D + c] = d2[c];
      }
    }
    lq<D, 2 * D>(dr, r, cx.bc());
  }
  if (act) {
    (state_only ? so.m : y.g)[o * D + r] = gg;
    if constexpr (!state_only) store_row(y.E + (o * D + r) * D, er);
    if (with_chol) {
      double* Do = state_only ? so.L : y.Dm;
#pragma unroll
      for (int c = 0; c < D; ++c) Do[(o * D + r) * D + c] = dr[c];
    }
  }
  wsync();
}

}  // namespace pof
