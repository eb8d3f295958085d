// Large-state IEKS engine (state dimension beyond the per-worker kernels,
// e.g. Lorenz-96 d = 32, q = 3: D = 128).
//
// Each combine of a scan level is a sequence of batched primitives over every
// combine of that level: the blocked Householder LQ (a panel kernel, one wave
// per matrix, and a one-pass trailing update on the matrix cores; LAPACK
// dlarfg's reflector convention, the same as the per-worker lq()), a
// strided-batched GEMM kernel on the matrix cores (v_mfma_f64_16x16x4f64, or
// v_mfma_f32_16x16x4f32 for the FP32-product option), blocked triangular solves
// on the matrix cores, fused chains of matrix-vector products, and small copy /
// fill / reduction kernels.  The host drives JAX's odd/even associative_scan
// recursion (pof/parallel_filtsmooth filter.py:35, smoother.py:15-17) over
// strided views, so every factor is built by the same sequence of operations as
// in the reference.
//
// Matrices are row-major float64.
#include "pof_large.hpp"

#include "pof_problems.hpp"
#include "pof_wave.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <initializer_list>
#include <type_traits>
#include <utility>
#include <vector>

namespace pof {
namespace large {

namespace {

// ------------------------------------------------------------------ views
// A strided batch of row-major matrices (or vectors, ld = 1): item i at p + i * s.
struct V {
  double* p;
  long long s;  // doubles between items (0: one matrix shared by every item)
  int ld;       // doubles between rows
  V at(int r, int c) const { return V{p + (size_t)r * ld + c, s, ld}; }
  V item(long long i) const { return V{p + i * s, s, ld}; }
  V every(int k) const { return V{p, s * k, ld}; }
};

// ------------------------------------------------------------------ kernels
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kLqThreads = 256;
constexpr int kLqMaxN = 512;

// In-place Householder LQ of the m x n matrix of item blockIdx.x: afterwards its
// first min(m, n) columns hold tria(A) = qr(A^T).R^T (utils.py:33-41) with
// LAPACK's reflectors (beta = -sign(alpha) ||x||, tau = (beta - alpha) / beta,
// v = x / (alpha - beta), no reflector for a zero tail).  Pivot row i's tail is
// reduced by the whole block; every later row is updated by one wave.
__global__ __launch_bounds__(kLqThreads) void k_lq(double* A, long long stride, int m, int n, int ld) {
  __shared__ double v[kLqMaxN];
  __shared__ double red[kLqThreads / 64];
  double* a = A + (long long)blockIdx.x * stride;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int NW = kLqThreads / 64;
  const int K = m < n ? m : n;
  for (int i = 0; i < K && i < n - 1; ++i) {
    double* pr = a + (size_t)i * ld;
    double s = 0.0;
    for (int j = i + 1 + tid; j < n; j += kLqThreads) s = fma(pr[j], pr[j], s);
    s = wave_sum(s);
    if (lane == 0) red[wv] = s;
    __syncthreads();
    const double sig = (red[0] + red[1]) + (red[2] + red[3]);
    const double alpha = pr[i];
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (sig != 0.0) {
      const double nrm = sqrt(fma(alpha, alpha, sig));
      beta = -copysign(nrm, alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int j = i + 1 + tid; j < n; j += kLqThreads) v[j] = pr[j] * scal;
    __syncthreads();
    if (tau != 0.0) {
      for (int r = i + 1 + wv; r < m; r += NW) {
        double* rr = a + (size_t)r * ld;
        double w = 0.0;
        for (int j = i + 1 + lane; j < n; j += 64) w = fma(rr[j], v[j], w);
        w = wave_sum(w) + rr[i];
        const double tw = tau * w;
        for (int j = i + 1 + lane; j < n; j += 64) rr[j] = fma(-tw, v[j], rr[j]);
        if (lane == 0) rr[i] -= tw;
      }
    }
    if (tid == 0) pr[i] = beta;
    for (int j = i + 1 + tid; j < n; j += kLqThreads) pr[j] = 0.0;
    __syncthreads();
  }
}

constexpr int kNB = 16;  // rows per panel of the blocked LQ

// f(integral_constant<int, K>) for K = 0, 1, ... in order: a loop whose index
// is a compile-time constant in every step
template <typename F, int... K>
__device__ __forceinline__ void unroll_steps(F& f, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>()), ...);
}

// One panel of the blocked LQ (LAPACK dgelqf: dgelq2 on the panel, dlarft
// forward / rowwise): rows i0 .. i0 + np - 1 of item blockIdx.x, columns
// i0 .. n - 1, factored by one wave with the panel in registers (lane holds
// columns lane + 64 q) and the reflectors of k_lq.  Writes the panel's rows of
// L back, the reflectors V (np x (n - i0), leading 1, zeros before it) and the
// triangular factor T (np x np) with H_1 ... H_np = I - V^T T V.  Pivot column
// k < kNB < 64 lives in slot 0 of lane k, so only slot 0 needs the test c > k;
// columns c >= n - i0 are loaded as zeros and stay zero (every update adds a
// multiple of the pivot row, zero there too), so they need no test.  Rows
// np .. kNB - 1 are zeros too: their updates leave them so and they are never
// written, so the steps need no row test either.
//
// Step k sums 16 - k values over the wave at once (the squared tail of pivot
// row k and the products of rows k + 1 .. kNB - 1 with it) in registers: a
// transposed wave sum leaves each total in a group of lanes, and scalar reads
// of one lane broadcast it, so a step has no LDS round trip and no barrier.
typedef double lq_panel_d4 __attribute__((ext_vector_type(4)));
// Three waves per SIMD at four column slots: the Gram matrix accumulates in
// ordinary VGPRs then, and the panel fits in 156 registers.
template <int NQ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_lq_panel(
    double* A, long long sA, int n, int ld, int i0, int np, double* Vb, long long sV, int ldV, double* Tb,
    long long sT, int ldT) {
  __shared__ double G[kNB][kNB + 1];
  __shared__ double taus[kNB];
  const int lane = threadIdx.x;
  double* a = A + (long long)blockIdx.x * sA + (size_t)i0 * ld + i0;
  double* Vo = Vb + (long long)blockIdx.x * sV;
  const int nc = n - i0;
  double P[kNB][NQ];
#pragma unroll
  for (int r = 0; r < kNB; ++r) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int c = lane + 64 * q;
      P[r][q] = (r < np && c < nc) ? a[(size_t)r * ld + c] : 0.0;
    }
  }
  // Householder step k (a compile-time constant); NV, a power of two >= kNB - k,
  // values per wave sum
  auto step = [&](auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    constexpr int NV = k < 8 ? 16 : (k < 12 ? 8 : (k < 14 ? 4 : (k < 15 ? 2 : 1)));
    constexpr int G1 = 64 / NV;  // lanes per total after the wave sum
    if (k >= np) return;
    // value 0: the squared tail of pivot row k; value j >= 1: row k + j . tail of row k.
    // Slot 0 holds a tail column only in lanes beyond k: masked once, here.
    const double pk0 = (lane > k) ? P[k][0] : 0.0;
    double x[NV];
    {
      double sq = pk0 * pk0;
#pragma unroll
      for (int q = 1; q < NQ; ++q) sq = fma(P[k][q], P[k][q], sq);
      x[0] = sq;
    }
#pragma unroll
    for (int j = 1; j < NV; ++j) {
      const int r = k + j;
      double w = 0.0;
      if (r < kNB) {
        w = P[r][0] * pk0;
#pragma unroll
        for (int q = 1; q < NQ; ++q) w = fma(P[r][q], P[k][q], w);
      }
      x[j] = w;
    }
    wave::sum_scatter<NV>(x, lane);
    const double sig = wave::lane_value(x[0], 0);
    const double alpha = wave::lane_value(P[k][0], k);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (sig != 0.0) {
      const double nrm = sqrt(fma(alpha, alpha, sig));
      beta = -copysign(nrm, alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    // the reflector in slot 0: 1 in the pivot lane, the scaled tail beyond it, 0 before
    const double v0 = (lane > k) ? P[k][0] * scal : ((lane == k) ? 1.0 : 0.0);
    P[k][0] = (lane > k) ? v0 : P[k][0];
#pragma unroll
    for (int q = 1; q < NQ; ++q) P[k][q] *= scal;
    // row k is final: L row k back to A, V row k out (slots q >= 1 hold columns
    // 64 q + lane > k: zero in L, tail in V)
    if (lane < nc) {
      a[(size_t)k * ld + lane] = (lane < k) ? P[k][0] : ((lane == k) ? beta : 0.0);
      Vo[(size_t)k * ldV + lane] = (lane < k) ? 0.0 : ((lane == k) ? 1.0 : P[k][0]);
    }
#pragma unroll
    for (int q = 1; q < NQ; ++q) {
      const int c = lane + 64 * q;
      if (c < nc) {
        a[(size_t)k * ld + c] = 0.0;
        Vo[(size_t)k * ldV + c] = P[k][q];
      }
    }
    if (lane == 0) taus[k] = tau;
    if (tau != 0.0) {
#pragma unroll
      for (int r = k + 1; r < kNB; ++r) {
        // row r . v with v = [1, scal * tail]; the pivot column of row r sits in lane k
        const double tr = wave::lane_value(x[0], (r - k) * G1);
        const double ck = wave::lane_value(P[r][0], k);
        const double tw = tau * fma(scal, tr, ck);
        P[r][0] = fma(-tw, v0, P[r][0]);
#pragma unroll
        for (int q = 1; q < NQ; ++q) P[r][q] = fma(-tw, P[k][q], P[r][q]);
      }
    }
  };
  unroll_steps(step, std::make_integer_sequence<int, kNB>());
  // T (dlarft, forward / rowwise): G = V V^T on the matrix cores, with V read back from the
  // rows just stored (L2; the panel's registers are free by then), 64 columns per batch of
  // loads and four accumulation chains (a Gram matrix sums over every column, so the order
  // of the columns is free); then lane j forms row j of T:
  // T[j][k] = -tau_k sum_{l = j}^{k - 1} T[j][l] G[l][k], T[k][k] = tau_k
  __threadfence_block();
  __syncthreads();
  const int gj = lane & 15, gg = lane >> 4;
  const double* vr = Vo + (size_t)gj * ldV;
  lq_panel_d4 ga[4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (64 * q < nc) {
      double x[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int c = 64 * q + 16 * (e >> 2) + 4 * gg + (e & 3);
        x[e] = (gj < np && c < nc) ? vr[c] : 0.0;
      }
#pragma unroll
      for (int e = 0; e < 16; ++e) ga[e & 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[e], x[e], ga[e & 3], 0, 0, 0);
    }
  }
  // the FP64 accumulator holds element (l / 16 + 4 v, l % 16) in value v of lane l
#pragma unroll
  for (int v = 0; v < 4; ++v) G[gg + 4 * v][gj] = (ga[0][v] + ga[1][v]) + (ga[2][v] + ga[3][v]);
  __syncthreads();
  if (lane < np) {
    double trow[kNB];
#pragma unroll
    for (int k = 0; k < kNB; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < k; ++l) acc = fma(trow[l], -taus[k] * G[l][k], acc);  // trow[l] = 0 for l < lane
      trow[k] = (k < np) ? ((lane < k) ? acc : ((lane == k) ? taus[k] : 0.0)) : 0.0;
    }
    double* T = Tb + (long long)blockIdx.x * sT + (size_t)lane * ldT;
#pragma unroll
    for (int k = 0; k < kNB; ++k)
      if (k < np) T[k] = trow[k];
  }
}

typedef double lq_d4 __attribute__((ext_vector_type(4)));

// Trailing update of one panel of the blocked LQ, in one pass over the rows
// below the panel: A_t <- A_t (I - V^T T V) for A_t = rows i0 + np .. i0 + np
// + mr - 1 and columns i0 .. i0 + nc - 1 of an item, with V (np x nc) and T
// (np x np) from k_lq_panel.  Per strip of 16 rows, on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64): W^T = V A_s^T, W2^T = T^T W^T, A_s^T -= V^T W2^T.
// The strip stays in registers between the three products, so A_t is read
// and written once per panel.
//
// Register layout of a strip: tile t holds columns 16 t .. 16 t + 15, lane
// j + 16 g value v holds (row j, column 16 t + 4 g + v).  These registers are
// at once the B operand of K-step v of W^T = V A_s^T (summation index 4 v + g
// standing for column 4 g + v) and the accumulator of A_s^T's tile when its
// row i stands for column 4 (i % 4) + i / 4 (the FP64 accumulator holds
// element (l / 16 + 4 v, l % 16) in value v of lane l).  W^T's accumulator
// value v is likewise the B operand of K-step v of the next product.
//
// Four waves per block, one strip each; V (zero-padded to 16 rows and 16 NT
// columns) and T are staged in LDS.  Blocks are numbered so that the blocks of
// one item run on one XCD (work-groups go round-robin over the 8 XCDs) and
// share V through its L2.
constexpr int kLqUpdWaves = 4;
template <int NT, bool VEC>
__global__ __launch_bounds__(64 * kLqUpdWaves, 2) void k_lq_update(double* A, long long sA, int ld, int i0, int np,
                                                                    int mr, int nc, const double* Vb, long long sV,
                                                                    int ldV, const double* Tb, long long sT, int ldT,
                                                                    int cnt, int sbk) {
  constexpr int LDV = 16 * NT + 1;
  __shared__ double Vs[16 * LDV];
  __shared__ double Ts[16 * 17];
  const int L = blockIdx.x;
  const int rr = L / 8;
  const long long item = (long long)(rr / sbk) * 8 + (L % 8);
  const int sb = rr % sbk;
  if (item >= cnt) return;  // uniform over the block
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* vsrc = Vb + item * sV;
  for (int e = tid; e < 16 * 16 * NT; e += 64 * kLqUpdWaves) {
    const int p = e / (16 * NT), c = e % (16 * NT);
    Vs[p * LDV + c] = (p < np && c < nc) ? vsrc[(size_t)p * ldV + c] : 0.0;
  }
  {
    const int p = tid / 16, c = tid % 16;
    Ts[p * 17 + c] = (p < np && c < np) ? Tb[item * sT + (size_t)p * ldT + c] : 0.0;
  }
  __syncthreads();
  const int r0 = 16 * (sb * kLqUpdWaves + wave);
  if (r0 >= mr) return;  // no barrier follows
  const int j = lane & 15, g = lane >> 4;
  const int ntr = (nc + 15) / 16;
  const int nfull = nc / 16;  // tiles without a column beyond nc
  const bool rowok = r0 + j < mr;
  // a row past the trailing block reads the block's last row: every product below keeps
  // the strip's rows apart (row j of the strip only meets column j of W^T), and it is not stored
  double* arow = A + item * sA + (size_t)(i0 + np + min(r0 + j, mr - 1)) * ld + i0 + 4 * g;
  lq_d4 X[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < nfull) {
      if constexpr (VEC) {
        const double2 lo = *(const double2*)(arow + 16 * t);
        const double2 hi = *(const double2*)(arow + 16 * t + 2);
        X[t] = lq_d4{lo.x, lo.y, hi.x, hi.y};
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) X[t][v] = arow[16 * t + v];
      }
    } else if (t < ntr) {
#pragma unroll
      for (int v = 0; v < 4; ++v) X[t][v] = (16 * t + 4 * g + v < nc) ? arow[16 * t + v] : 0.0;
    } else {
      X[t] = lq_d4{0.0, 0.0, 0.0, 0.0};
    }
  }
  // W^T (rows: panel rows p, columns: strip rows j) = V A_s^T
  lq_d4 wt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < ntr) {
#pragma unroll
      for (int s = 0; s < 4; ++s)
        wt = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[j * LDV + 16 * t + 4 * g + s], X[t][s], wt, 0, 0, 0);
    }
  }
  // W2^T = T^T W^T
  lq_d4 w2t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s) w2t = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[(4 * s + g) * 17 + j], wt[s], w2t, 0, 0, 0);
  // A_s^T -= V^T W2^T: accumulator row i of tile t stands for column 16 t + 4 (i % 4) + i / 4
  const int ci = 4 * (j & 3) + (j >> 2);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < ntr) {
#pragma unroll
      for (int s = 0; s < 4; ++s)
        X[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-Vs[(4 * s + g) * LDV + 16 * t + ci], w2t[s], X[t], 0, 0, 0);
    }
  }
  if (rowok) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < nfull) {
        if constexpr (VEC) {
          *(double2*)(arow + 16 * t) = double2{X[t][0], X[t][1]};
          *(double2*)(arow + 16 * t + 2) = double2{X[t][2], X[t][3]};
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v) arow[16 * t + v] = X[t][v];
        }
      } else if (t < ntr) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (16 * t + 4 * g + v < nc) arow[16 * t + v] = X[t][v];
      }
    }
  }
}

constexpr int kTrsmMaxN = 128;  // order of the triangular matrix at most (D <= 128)
constexpr int kTrsmWaves = 8;

// op(L) X = B in place for item blockIdx.x: L lower n x n (n <= 128), B n x m
// (m <= 128), op = L (forward substitution) or L^T (TRANS, backward), on the FP64
// matrix cores (v_mfma_f64_16x16x4f64) in blocks of 16 rows:
//   X_K = inv(L_KK) (B_K - sum_J op(L)_KJ X_J)
// over the earlier blocks J (later ones for L^T).  The lower block triangle of L
// (36 blocks of 16 x 16, zero beyond n) is staged in LDS in one coalesced pass;
// the diagonal blocks are then inverted in place, one thread per column of an
// inverse (forward substitution; rows beyond n act as rows of the identity).
// Wave w takes columns 16 w .. 16 w + 15 of B: its blocks are loaded once (from
// B itself, or from a separate source R: B = R, or R^T when RT) and solved in registers, in the accumulator layout (lane j + 16 g value
// v holds row 4 v + g, column j of a 16 x 16 block), which is also the layout
// of the B operand of the products with later blocks.
typedef double trsm_d4 __attribute__((ext_vector_type(4)));
constexpr int kTrsmNB = kTrsmMaxN / 16;
constexpr int kTrsmBS = 16 * 17;  // one padded 16 x 16 block
__host__ __device__ constexpr int trsm_tri(int K, int J) { return K * (K + 1) / 2 + J; }
// Four waves per SIMD (128 VGPRs): two blocks per CU, 2 x 78 KiB of LDS.
template <bool TRANS>
__global__ __launch_bounds__(64 * kTrsmWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_trsm(
    const double* L, long long sL, int ldL, double* B, long long sB, int ldB, int n, int m, const double* R,
    long long sR, int ldR, int RT) {
  constexpr int NB = kTrsmNB;
  __shared__ double Lt[trsm_tri(NB, 0) * kTrsmBS];
  const double* l = L + (long long)blockIdx.x * sL;
  double* b = B + (long long)blockIdx.x * sB;
  const double* rb = R + (long long)blockIdx.x * sR;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int nb = (n + 15) >> 4;
  const int c = 16 * wave + li;
  const bool cok = c < m;
  for (int e = tid; e < trsm_tri(nb, 0) * 256; e += 64 * kTrsmWaves) {
    const int bi = e >> 8, i = (e >> 4) & 15, j = e & 15;
    int K = 0;
    while (trsm_tri(K + 1, 0) <= bi) ++K;
    const int J = bi - trsm_tri(K, 0);
    const int r = 16 * K + i, cc = 16 * J + j;
    const bool ok = r < n && cc < n;
    Lt[bi * kTrsmBS + i * 17 + j] = ok ? l[(size_t)(ok ? r : 0) * ldL + (ok ? cc : 0)] : 0.0;
  }
  __syncthreads();
  {
    double y[16];  // column j of inv(L_KK)
    const int K = tid >> 4, j = tid & 15;
    double* Lk = Lt + trsm_tri(K < nb ? K : 0, K < nb ? K : 0) * kTrsmBS;
    if (K < nb) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        double acc = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int t = 0; t < i; ++t) acc = fma(-Lk[i * 17 + t], y[t], acc);
        y[i] = (i < j) ? 0.0 : acc / ((16 * K + i < n) ? Lk[i * 17 + i] : 1.0);
      }
    }
    __syncthreads();
    if (K < nb) {
#pragma unroll
      for (int i = 0; i < 16; ++i) Lk[i * 17 + j] = y[i];
    }
  }
  __syncthreads();
  if (16 * wave >= m) return;  // uniform over the wave; no barrier follows
  // this wave's blocks of B (two blocks per CU: the other block's work covers the latency)
  trsm_d4 X[NB];
#pragma unroll
  for (int K = 0; K < NB; ++K) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * K + 4 * v + lg;
      const bool ok = K < nb && r < n && cok;
      const int rc = ok ? r : 0, cc = ok ? c : 0;
      X[K][v] = ok ? (RT ? rb[(size_t)cc * ldR + rc] : rb[(size_t)rc * ldR + cc]) : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int K = TRANS ? NB - 1 - q : q;
    if (K >= nb) continue;  // uniform
    // two accumulation chains, so that consecutive products do not wait on each other
    trsm_d4 acc0 = X[K], acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int J = 0; J < NB; ++J) {
      if (TRANS ? (J > K && J < nb) : (J < K)) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          // A = op(L)_KJ: row li, column 4 s + lg of the block
          const double a = TRANS ? Lt[trsm_tri(J, K) * kTrsmBS + (4 * s + lg) * 17 + li]
                                 : Lt[trsm_tri(K, J) * kTrsmBS + li * 17 + 4 * s + lg];
          if (J & 1)
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a, X[J][s], acc1, 0, 0, 0);
          else
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a, X[J][s], acc0, 0, 0, 0);
        }
      }
    }
    const trsm_d4 rhs = acc0 + acc1;
    trsm_d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      // inv(L_KK) (or its transpose): row li, column 4 s + lg
      const double a = TRANS ? Lt[trsm_tri(K, K) * kTrsmBS + (4 * s + lg) * 17 + li]
                             : Lt[trsm_tri(K, K) * kTrsmBS + li * 17 + 4 * s + lg];
      x = __builtin_amdgcn_mfma_f64_16x16x4f64(a, rhs[s], x, 0, 0, 0);
    }
    X[K] = x;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * K + 4 * v + lg;
      if (r < n && cok) b[(size_t)r * ldB + c] = x[v];
    }
  }
}

// dst item b (rows x cols) = alpha * src item b' (r, c), or src (c, r) when trans;
// b' = cnt - 1 - b when rev.  grid (ceil(rows * cols / 256), cnt)
__global__ void k_copy(double* dst, long long ds, int dld, const double* src, long long ss, int sld, int rows,
                       int cols, int cnt, int trans, int rev, double alpha) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= cnt || e >= rows * cols) return;
  const int r = e / cols, c = e % cols;
  const long long bs = rev ? (long long)(cnt - 1 - b) : (long long)b;
  const double x = trans ? src[bs * ss + (size_t)c * sld + r] : src[bs * ss + (size_t)r * sld + c];
  dst[(long long)b * ds + (size_t)r * dld + c] = alpha * x;
}

// Transposed copy through a 32 x 32 LDS tile: dst item b (rows x cols) =
// alpha * src item b' (c, r), both sides read and written along rows.
// grid (ceil(cols / 32), ceil(rows / 32), cnt), 256 threads (32 x 8).
__global__ __launch_bounds__(256) void k_copy_t(double* dst, long long ds, int dld, const double* src, long long ss,
                                                int sld, int rows, int cols, int cnt, int rev, double alpha) {
  __shared__ double t[32][33];
  const int b = blockIdx.z;
  if (b >= cnt) return;
  const long long bs = rev ? (long long)(cnt - 1 - b) : (long long)b;
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // t[k][x] = src (c0 + k, r0 + x) = dst (r0 + x, c0 + k)
  for (int k = ty; k < 32; k += 8) {
    const int sr = c0 + k, sc = r0 + tx;
    t[k][tx] = (sr < cols && sc < rows) ? src[bs * ss + (size_t)sr * sld + sc] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int dr = r0 + k, dc = c0 + tx;
    if (dr < rows && dc < cols) dst[(long long)b * ds + (size_t)dr * dld + dc] = alpha * t[tx][k];
  }
}

// dst item (rows x cols) = dv on the diagonal, 0 elsewhere
__global__ void k_fill(double* dst, long long ds, int dld, int rows, int cols, int cnt, double dv) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= cnt || e >= rows * cols) return;
  const int r = e / cols, c = e % cols;
  dst[(long long)b * ds + (size_t)r * dld + c] = (r == c) ? dv : 0.0;
}

// Up to kMaxCopyOps block copies (or identity fills, src = nullptr: dst = fill I) in one
// launch: op blockIdx.z of item blockIdx.y.  grid (ceil(max rows x cols / 256), cnt, ops)
struct CopyOp {
  double* dst;
  long long ds;
  int dld;
  const double* src;
  long long ss;
  int sld;
  int rows, cols;
  double fill;
};
constexpr int kMaxCopyOps = 5;
struct CopyOps {
  CopyOp op[kMaxCopyOps];
};
__global__ void k_copy_multi(CopyOps a) {
  const int z = blockIdx.z;
  // the op by compile-time indices (uniform selects; no dynamic index into the arguments)
  CopyOp o = a.op[0];
#pragma unroll
  for (int q = 1; q < kMaxCopyOps; ++q)
    if (z == q) o = a.op[q];
  const long long b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= o.rows * o.cols) return;
  const int r = e / o.cols, c = e % o.cols;
  const double x = o.src ? o.src[b * o.ss + (size_t)r * o.sld + c] : ((r == c) ? o.fill : 0.0);
  o.dst[b * o.ds + (size_t)r * o.dld + c] = x;
}

// EK1 of the Lorenz-96 residual at n states (rows of H as in the per-worker
// kernels: -J[i][l] P[l q1] at column l q1, + P[i q1 + 1] at column i q1 + 1;
// c = J y - f(y)).  One thread per (state, row).
__global__ void k_lin_l96(const double* m, const double* P, double Fo, int n, int d, int q1, double* H, double* c) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * d) return;
  const long long k = e / d;
  const int i = (int)(e % d);
  const int D = d * q1;
  const double* mk = m + k * D;
  const int ip1 = (i + 1) % d, im1 = (i + d - 1) % d, im2 = (i + d - 2) % d;
  const double yi = P[i * q1] * mk[i * q1];
  const double yp1 = P[ip1 * q1] * mk[ip1 * q1];
  const double ym1 = P[im1 * q1] * mk[im1 * q1];
  const double ym2 = P[im2 * q1] * mk[im2 * q1];
  const double f = (yp1 - ym2) * ym1 - yi + Fo;
  // J[i][l] for the four columns, summed in ascending l for c
  int cols[4] = {ip1, im1, im2, i};
  double jv[4] = {ym1, yp1 - ym2, -ym1, -1.0};
  double ys[4] = {yp1, ym1, ym2, yi};
  for (int a = 0; a < 4; ++a)
    for (int b2 = a + 1; b2 < 4; ++b2)
      if (cols[b2] < cols[a]) {
        int t = cols[a]; cols[a] = cols[b2]; cols[b2] = t;
        double u = jv[a]; jv[a] = jv[b2]; jv[b2] = u;
        u = ys[a]; ys[a] = ys[b2]; ys[b2] = u;
      }
  double acc = 0.0;
  for (int a = 0; a < 4; ++a) acc = fma(jv[a], ys[a], acc);
  double* Hr = H + (k * d + i) * D;
  for (int j = 0; j < D; ++j) Hr[j] = 0.0;
  for (int a = 0; a < 4; ++a) Hr[cols[a] * q1] = -jv[a] * P[cols[a] * q1];
  Hr[i * q1 + 1] += P[i * q1 + 1];
  c[k * d + i] = acc - f;
}

// Observation terms of state s from S (d x d lower, leading dimension ldS) and
// yhat: ssq_s = |S^-T y|^2 (utils.py:110-112), nll_s = -loglik (utils.py:22-30).
__global__ void k_obs_terms(const double* S, long long sS, int ldS, const double* y, long long sy, int d, int n,
                            double* ssq, double* nll, double* tmp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double* Ss = S + (long long)s * sS;
  const double* ys = y + (long long)s * sy;
  double* w = tmp + (long long)s * 2 * d;
  double* z = w + d;
  double ld = 0.0;
  for (int i = 0; i < d; ++i) {
    double v = ys[i];
    for (int t = 0; t < i; ++t) v = fma(-Ss[(size_t)i * ldS + t], z[t], v);
    z[i] = v / Ss[(size_t)i * ldS + i];
    ld += log(fabs(Ss[(size_t)i * ldS + i]));
  }
  for (int i = d - 1; i >= 0; --i) {
    double v = ys[i];
    for (int t = i + 1; t < d; ++t) v = fma(-Ss[(size_t)t * ldS + i], w[t], v);
    w[i] = v / Ss[(size_t)i * ldS + i];
  }
  double ss = 0.0, nn = 0.0;
  for (int i = 0; i < d; ++i) {
    ss = fma(w[i], w[i], ss);
    nn = fma(z[i], z[i], nn);
  }
  ssq[s] = ss;
  nll[s] = 0.5 * nn + (ld + d * log(2.0 * M_PI) / 2.0);
}

// out[s] = |x_s|^2 (x: n vectors of length D, stride D)
__global__ void k_sqnorm(const double* x, int D, int n, double* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double a = 0.0;
  for (int i = 0; i < D; ++i) a = fma(x[(size_t)s * D + i], x[(size_t)s * D + i], a);
  out[s] = a;
}

// means-close test of the convergence rule (convergence_criteria.py:10-12):
// out[t] = 1 if any |old - new| > 1e-8 + 1e-13 |new| (or NaN) in state t
__global__ void k_viol(const double* mnew, const double* mold, int D, int n, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double v = 0.0;
  for (int i = 0; i < D; ++i) {
    const double o = mold[(size_t)t * D + i], x = mnew[(size_t)t * D + i];
    v = (fabs(o - x) <= 1e-8 + 1e-13 * fabs(x)) ? v : 1.0;
  }
  out[t] = v;
}

// Fixed-order reductions of n terms: out[0] = sum a, out[1] = sum b, out[2] = sum c, out[3] = max v.
__global__ __launch_bounds__(256) void k_reduce4(const double* a, const double* b, const double* c, int nabc,
                                                 const double* v, int nv, double* out) {
  __shared__ double sh[4][256];
  const int t = threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int i = t; i < nabc; i += 256) {
    s0 += a[i];
    s1 += b[i];
    s2 += c[i];
  }
  for (int i = t; i < nv; i += 256) s3 = fmax(s3, v[i]);
  sh[0][t] = s0;
  sh[1][t] = s1;
  sh[2][t] = s2;
  sh[3][t] = s3;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] += sh[0][t + w];
      sh[1][t] += sh[1][t + w];
      sh[2][t] += sh[2][t + w];
      sh[3][t] = fmax(sh[3][t], sh[3][t + w]);
    }
    __syncthreads();
  }
  if (t < 4) out[t] = sh[t][0];
}

// Output projection (solver.py:66-71): means_out[t][i] = P_i0 m_t[i0],
// chol_out[t][i][:] = scale P_i0 L_t[i0][:], i0 = i q1; full outputs m_t and scale L_t.
__global__ void k_project(const double* m, const double* L, const double* P, int D, int d, int n, double scale,
                          double* means, double* chols, double* fmeans, double* fchols) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q1 = D / d;
  if (e >= (long long)n * D) return;
  const long long t = e / D;
  const int r = (int)(e % D);
  if (fmeans) fmeans[t * D + r] = m[t * D + r];
  if (fchols)
    for (int c = 0; c < D; ++c) fchols[(t * D + r) * D + c] = scale * L[(t * D + r) * D + c];
  if (r % q1 == 0) {
    const int i = r / q1;
    if (means) means[t * d + i] = P[r] * m[t * D + r];
    if (chols)
      for (int c = 0; c < D; ++c) chols[(t * d + i) * D + c] = (scale * P[r]) * L[(t * D + r) * D + c];
  }
}

// ------------------------------------------------------------------ GEMM
// C = alpha op(A) op(B) + beta C for item blockIdx.x of a strided batch
// (row-major, op = transpose when tA / tB; a stride of 0 shares a matrix over
// the batch; beta = 0 does not read C).  A workgroup of four waves computes a
// BM x BN tile of C (blockIdx.y, blockIdx.z); the waves form a WM x WN grid,
// each computing (BM / WM) x (BN / WN) as 16 x 16 tiles on the matrix cores.
// K advances in chunks of 16 staged through LDS (zero-padded at the edges),
// loaded along whichever index is contiguous in memory; the next chunk is
// loaded into registers while the current one is multiplied.  F32: the
// operands are rounded to FP32 as they enter LDS and multiplied with FP32
// accumulate (v_mfma_f32_16x16x4f32); the epilogue is FP64.  Lane layout of the
// 16 x 16 x 4 MFMAs: A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k;
// value v of lane l holds result element (4 (l / 16) + v, l % 16) for FP32 but
// (l / 16 + 4 v, l % 16) for FP64 (both checked on gfx950 with an in-bounds
// probe of the builtins).  Three blocks per CU: the 128 x 64 FP64 tile then
// keeps its accumulators in 153 VGPRs (2 blocks with 156 VGPRs + 64 AGPRs
// otherwise).
typedef double gemm_d4 __attribute__((ext_vector_type(4)));
typedef float gemm_f4 __attribute__((ext_vector_type(4)));

constexpr int kGemmKc = 16;

template <bool F32, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(256, 3) void k_gemm(int tA, int tB, int M, int N, int K, double alpha, const double* A,
                                              long long sA, int lda, const double* B, long long sB, int ldb,
                                              double beta, double* C, long long sC, int ldc) {
  static_assert(WM * WN == 4, "four waves");
  using T = typename std::conditional<F32, float, double>::type;
  using V4 = typename std::conditional<F32, gemm_f4, gemm_d4>::type;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;  // MFMA tiles per wave
  constexpr int LA = BM * kGemmKc / 256, LB = BN * kGemmKc / 256;  // elements each thread loads per chunk
  static_assert(LA >= 1 && LB >= 1 && TM >= 1 && TN >= 1, "tile shape");
  __shared__ T As[BM][kGemmKc + 1];
  __shared__ T Bs[kGemmKc][BN + 1];
  const long long z = blockIdx.x;
  const double* Az = A + z * sA;
  const double* Bz = B + z * sB;
  double* Cz = C + z * sC;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.z * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  // element e (0 .. BM * 16 - 1) of the A chunk -> (i, k): along i when A is stored K x M (tA)
  auto a_ik = [&](int e, int& i, int& k) {
    if (tA) {
      i = e % BM;
      k = e / BM;
    } else {
      k = e % kGemmKc;
      i = e / kGemmKc;
    }
  };
  auto b_kj = [&](int e, int& kb, int& j) {
    if (tB) {
      kb = e % kGemmKc;
      j = e / kGemmKc;
    } else {
      j = e % BN;
      kb = e / BN;
    }
  };
  double ra[LA], rb[LB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      int i, k;
      a_ik(tid + 256 * r, i, k);
      const int gi = m0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? (tA ? Az[(size_t)gk * lda + gi] : Az[(size_t)gi * lda + gk]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      int kb, j;
      b_kj(tid + 256 * r, kb, j);
      const int gj = n0 + j, gk = k0 + kb;
      rb[r] = (gj < N && gk < K) ? (tB ? Bz[(size_t)gj * ldb + gk] : Bz[(size_t)gk * ldb + gj]) : 0.0;
    }
  };
  V4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = V4{0, 0, 0, 0};
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += kGemmKc) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      int i, k;
      a_ik(tid + 256 * r, i, k);
      As[i][k] = (T)ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      int kb, j;
      b_kj(tid + 256 * r, kb, j);
      Bs[kb][j] = (T)rb[r];
    }
    __syncthreads();
    if (k0 + kGemmKc < K) fetch(k0 + kGemmKc);  // in flight during the products below
#pragma unroll
    for (int kk = 0; kk < kGemmKc; kk += 4) {
      T af[TM], bf[TN];
#pragma unroll
      for (int a = 0; a < TM; ++a) af[a] = As[(BM / WM) * wr + 16 * a + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int b = 0; b < TN; ++b) bf[b] = Bs[kk + (lane >> 4)][(BN / WN) * wc + 16 * b + (lane & 15)];
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          if constexpr (F32) {
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
          } else {
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
          }
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi =
            m0 + (BM / WM) * wr + 16 * a + (F32 ? 4 * (lane >> 4) + v : (lane >> 4) + 4 * v);
        const int gj = n0 + (BN / WN) * wc + 16 * b + (lane & 15);
        if (gi < M && gj < N) {
          double* c = Cz + (size_t)gi * ldc + gj;
          const double prod = alpha * (double)acc[a][b][v];
          *c = (beta == 0.0) ? prod : fma(beta, *c, prod);
        }
      }
}

// The N = 1 case (matrix-vector, plus the scalar-product rows): one thread per row of C,
// FP64 (vector ALU): c_i = alpha sum_k op(A)[i][k] b_k + beta c_i.
__global__ void k_gemv(int tA, int tB, int M, int K, double alpha, const double* A, long long sA, int lda,
                       const double* B, long long sB, int ldb, double beta, double* C, long long sC, int ldc) {
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const long long z = blockIdx.x;
  const double* Az = A + z * sA;
  const double* Bz = B + z * sB;
  double acc = 0.0;
  for (int k = 0; k < K; ++k) {
    const double a = tA ? Az[(size_t)k * lda + i] : Az[(size_t)i * lda + k];
    const double b = tB ? Bz[k] : Bz[(size_t)k * ldb];
    acc = fma(a, b, acc);
  }
  double* c = C + z * sC + (size_t)i * ldc;
  const double prod = alpha * acc;
  *c = (beta == 0.0) ? prod : fma(beta, *c, prod);
}

// The N = 1 case with A stored row-major (not transposed): one wave per row of
// C, lanes along the row (coalesced), then a wave reduction.
constexpr int kGemvRowWaves = 4;
__global__ __launch_bounds__(64 * kGemvRowWaves) void k_gemv_rows(int tB, int M, int K, double alpha, const double* A,
                                                                  long long sA, int lda, const double* B,
                                                                  long long sB, int ldb, double beta, double* C,
                                                                  long long sC, int ldc) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.y * kGemvRowWaves + (threadIdx.x >> 6);
  if (i >= M) return;  // uniform per wave; no barrier follows
  const long long z = blockIdx.x;
  const double* ar = A + z * sA + (size_t)i * lda;
  const double* Bz = B + z * sB;
  double acc = 0.0;
  for (int k = lane; k < K; k += 64) acc = fma(ar[k], tB ? Bz[k] : Bz[(size_t)k * ldb], acc);
  acc = wave_sum(acc);
  if (lane == 0) {
    double* c = C + z * sC + (size_t)i * ldc;
    const double prod = alpha * acc;
    *c = (beta == 0.0) ? prod : fma(beta, *c, prod);
  }
}

// A chain of five matrix-vector products over one item (blockIdx.x), D <= 128:
//   t1 = op1(M1) x;  t2 = y + s2 op2(M2) t1;  t3 = op3(M3) t2;  t2 -= op4(M4) t3;
//   out = z + op5(M5) t2
// (op = M or M^T), the vectors in LDS: one launch for what the combine's b and eta
// would take seven (five products, two vector copies).  M x: a wave takes 16 rows at
// once, lanes along the rows (coalesced), and a transposed wave sum finishes them;
// M^T x: a thread per column over half the rows (coalesced along the rows), the two
// halves added in LDS.
struct ChainArgs {
  const double* m[5];
  long long ms[5];
  int mld[5];
  int mt[5];
  const double *x, *y, *z;
  long long xs, ys, zs;
  double s2;
  double* out;
  long long outs;
  int D;
};
constexpr int kChainMaxD = 128;
template <typename Put>
__device__ __forceinline__ void chain_mv(const double* M, int ld, int trans, const double* xs, double* red, int D,
                                         Put put) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (!trans) {
    for (int r0 = 16 * wave; r0 < D; r0 += 64) {  // uniform per wave
      double acc[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int r = r0 + j;
        double sum = 0.0;
        if (r < D) {
          const double* mr = M + (size_t)r * ld;
          if (lane < D) sum = mr[lane] * xs[lane];
          if (lane + 64 < D) sum = fma(mr[lane + 64], xs[lane + 64], sum);
        }
        acc[j] = sum;
      }
      wave::sum_scatter<16>(acc, lane);
      const int r = r0 + (lane >> 2);
      if ((lane & 3) == 0 && r < D) put(r, acc[0]);
    }
  } else {
    const int c = tid & 127, h = tid >> 7;
    const int half = (D + 1) / 2;
    const int rb = h * half, re = (h == 0) ? half : D;
    double sum = 0.0;
    if (c < D) {
#pragma unroll 8
      for (int r = rb; r < re; ++r) sum = fma(M[(size_t)r * ld + c], xs[r], sum);
    }
    red[tid] = sum;
    __syncthreads();
    if (tid < D) put(tid, red[tid] + red[tid + 128]);
  }
  __syncthreads();
}
__global__ __launch_bounds__(256) void k_vec_chain(ChainArgs a) {
  __shared__ double v0[kChainMaxD], v1[kChainMaxD], v2[kChainMaxD], v3[kChainMaxD];
  __shared__ double red[256];
  const long long it = blockIdx.x;
  const int tid = threadIdx.x, D = a.D;
  const double* M[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) M[q] = a.m[q] + it * a.ms[q];
  if (tid < D) v0[tid] = a.x[it * a.xs + tid];
  __syncthreads();
  const double* y = a.y + it * a.ys;
  const double* z = a.z + it * a.zs;
  double* out = a.out + it * a.outs;
  const double s2 = a.s2;
  chain_mv(M[0], a.mld[0], a.mt[0], v0, red, D, [&](int i, double t) { v1[i] = t; });
  chain_mv(M[1], a.mld[1], a.mt[1], v1, red, D, [&](int i, double t) { v2[i] = fma(s2, t, y[i]); });
  chain_mv(M[2], a.mld[2], a.mt[2], v2, red, D, [&](int i, double t) { v3[i] = t; });
  chain_mv(M[3], a.mld[3], a.mt[3], v3, red, D, [&](int i, double t) { v2[i] -= t; });
  chain_mv(M[4], a.mld[4], a.mt[4], v2, red, D, [&](int i, double t) { out[i] = z[i] + t; });
}

// ------------------------------------------------------------- host helpers
struct Arena;
double* arena_take(Arena* ar, size_t n);
size_t arena_mark(Arena* ar);
void arena_release(Arena* ar, size_t mark);

struct Ops {
  hipStream_t st;
  Arena* ar;
  int err = 0;
  std::string msg;
  bool f32 = false;          // gemm() forms matrix-matrix products as FP32 GEMMs (set inside the combines)
  bool f32_combines = false;  // the solve asked for FP32 combine products
  int zw2 = 1 << 20;  // band of the filter elements' Z blocks (d - 1: leaf Z = [Zr | 0]; none by default)

  void fail(const char* what, int code) {
    if (!err) {
      err = code;
      char b[160];
      std::snprintf(b, sizeof b, "%s failed (%d)", what, code);
      msg = b;
    }
  }
  // row-major C (M x N) = alpha op(A) op(B) + beta C: FP64 on the matrix cores, or with f32 the
  // product op(A) op(B) of matrix-matrix shapes with FP32 operands and FP32 accumulate, scaled and
  // added to C in FP64
  void gemm(bool tA, bool tB, int M, int N, int K, double alpha, V A, V B, double beta, V C, int cnt) {
    gemm_impl(f32 && M > 1 && N > 1 && K > 1, tA, tB, M, N, K, alpha, A, B, beta, C, cnt);
  }
  void gemm64(bool tA, bool tB, int M, int N, int K, double alpha, V A, V B, double beta, V C, int cnt) {
    gemm_impl(false, tA, tB, M, N, K, alpha, A, B, beta, C, cnt);
  }
  // out = z + op5(M5) (t2 - op4(M4) op3(M3) t2), t2 = y + s2 op2(M2) op1(M1) x, per item (k_vec_chain)
  void vec_chain(int D, V m1, bool t1, V m2, bool t2, V m3, bool t3, V m4, bool t4, V m5, bool t5, V x, V y,
                 double s2, V z, V out, int cnt) {
    if (err || cnt <= 0) return;
    if (D > kChainMaxD) {
      fail("vector chain: D beyond 128", -1);
      return;
    }
    ChainArgs a{};
    const V mm[5] = {m1, m2, m3, m4, m5};
    const bool tt[5] = {t1, t2, t3, t4, t5};
    for (int q = 0; q < 5; ++q) {
      a.m[q] = mm[q].p;
      a.ms[q] = mm[q].s;
      a.mld[q] = mm[q].ld;
      a.mt[q] = tt[q] ? 1 : 0;
    }
    a.x = x.p;
    a.xs = x.s;
    a.y = y.p;
    a.ys = y.s;
    a.z = z.p;
    a.zs = z.s;
    a.s2 = s2;
    a.out = out.p;
    a.outs = out.s;
    a.D = D;
    hipLaunchKernelGGL(k_vec_chain, dim3(cnt), dim3(256), 0, st, a);
    check("k_vec_chain");
  }
  void gemm_impl(bool use32, bool tA, bool tB, int M, int N, int K, double alpha, V A, V B, double beta, V C,
                 int cnt) {
    if (err || cnt <= 0 || M <= 0 || N <= 0) return;
    if (N == 1) {
      if (tA) {
        hipLaunchKernelGGL(k_gemv, dim3(cnt, (M + 127) / 128), dim3(128), 0, st, 1, tB ? 1 : 0, M, K, alpha, A.p,
                           A.s, A.ld, B.p, B.s, B.ld, beta, C.p, C.s, C.ld);
      } else {
        hipLaunchKernelGGL(k_gemv_rows, dim3(cnt, (M + kGemvRowWaves - 1) / kGemvRowWaves), dim3(64 * kGemvRowWaves),
                           0, st, tB ? 1 : 0, M, K, alpha, A.p, A.s, A.ld, B.p, B.s, B.ld, beta, C.p, C.s, C.ld);
      }
      check("k_gemv");
      return;
    }
    // tile shape by the product's shape: 128 x 16 for the skinny panel products (N <= 16),
    // 128 x 64 for the large ones, 64 x 64 otherwise
    if (N <= 16) {
      launch_gemm<128, 16, 4, 1>(use32, tA, tB, M, N, K, alpha, A, B, beta, C, cnt);
    } else if (M >= 128 && N >= 64) {
      launch_gemm<128, 64, 2, 2>(use32, tA, tB, M, N, K, alpha, A, B, beta, C, cnt);
    } else {
      launch_gemm<64, 64, 2, 2>(use32, tA, tB, M, N, K, alpha, A, B, beta, C, cnt);
    }
    check("k_gemm");
  }
  template <int BM, int BN, int WM, int WN>
  void launch_gemm(bool use32, bool tA, bool tB, int M, int N, int K, double alpha, V A, V B, double beta, V C,
                   int cnt) {
    const dim3 grid(cnt, (M + BM - 1) / BM, (N + BN - 1) / BN);
    if (use32) {
      hipLaunchKernelGGL((k_gemm<true, BM, BN, WM, WN>), grid, dim3(256), 0, st, tA ? 1 : 0, tB ? 1 : 0, M, N, K,
                         alpha, A.p, A.s, A.ld, B.p, B.s, B.ld, beta, C.p, C.s, C.ld);
    } else {
      hipLaunchKernelGGL((k_gemm<false, BM, BN, WM, WN>), grid, dim3(256), 0, st, tA ? 1 : 0, tB ? 1 : 0, M, N, K,
                         alpha, A.p, A.s, A.ld, B.p, B.s, B.ld, beta, C.p, C.s, C.ld);
    }
  }
  // tria in place: the blocked LQ (panels of kNB rows, trailing rows updated in one
  // pass per panel), or k_lq for matrices of at most one panel.  b1 > 0 declares a
  // banded second block: in columns b1 .. n - 1, row r is zero beyond column
  // b1 + max(r, w2).  Pivot i's reflector is then zero beyond column b1 + max(i, w2)
  // (the reflectors before it only reach that far), so a panel of first-block pivots
  // leaves the columns beyond its last pivot's band as they are and skips them: the
  // same factor, with the panel and its trailing update over fewer columns.
  void lq(V A, int m, int n, int cnt, int b1 = 0, int w2 = 0) {
    if (err || cnt <= 0) return;
    if (n > 256) {
      fail("lq: more than 256 columns", -1);
      return;
    }
    const int K = m < n ? m : n;
    const int kmax = K < n - 1 ? K : n - 1;
    if (m <= kNB || kmax <= 0) {
      hipLaunchKernelGGL(k_lq, dim3(cnt), dim3(kLqThreads), 0, st, A.p, A.s, m, n, A.ld);
      check("k_lq");
      return;
    }
    const size_t mark = arena_mark(ar);
    V Vb{arena_take(ar, (size_t)cnt * kNB * n), (long long)kNB * n, n};
    V Tb{arena_take(ar, (size_t)cnt * kNB * kNB), (long long)kNB * kNB, kNB};
    if (!Vb.p || !Tb.p) {
      fail("workspace (lq)", -2);
      return;
    }
    for (int i0 = 0; i0 < kmax && !err; i0 += kNB) {
      const int np = (kmax - i0) < kNB ? (kmax - i0) : kNB;
      int ce = n;  // end (exclusive) of the columns this panel's reflectors reach
      if (b1 > 0 && i0 + np <= b1) {
        const int last = i0 + np - 1;
        ce = std::min(n, b1 + std::max(last, w2) + 1);
      }
      const int nc = ce - i0;
      // column slots for this panel's own width: later panels are narrower
      const int nq = (nc + 63) / 64;
      if (nq <= 1) {
        hipLaunchKernelGGL(k_lq_panel<1>, dim3(cnt), dim3(64), 0, st, A.p, A.s, ce, A.ld, i0, np, Vb.p, Vb.s, Vb.ld,
                           Tb.p, Tb.s, Tb.ld);
      } else if (nq == 2) {
        hipLaunchKernelGGL(k_lq_panel<2>, dim3(cnt), dim3(64), 0, st, A.p, A.s, ce, A.ld, i0, np, Vb.p, Vb.s, Vb.ld,
                           Tb.p, Tb.s, Tb.ld);
      } else if (nq == 3) {
        hipLaunchKernelGGL(k_lq_panel<3>, dim3(cnt), dim3(64), 0, st, A.p, A.s, ce, A.ld, i0, np, Vb.p, Vb.s, Vb.ld,
                           Tb.p, Tb.s, Tb.ld);
      } else {
        hipLaunchKernelGGL(k_lq_panel<4>, dim3(cnt), dim3(64), 0, st, A.p, A.s, ce, A.ld, i0, np, Vb.p, Vb.s, Vb.ld,
                           Tb.p, Tb.s, Tb.ld);
      }
      check("k_lq_panel");
      const int mr = m - (i0 + np);
      if (mr <= 0) continue;
      // A_trail (I - V^T T V) in one pass (FP64, as the rest of the factorisation)
      lq_update(A, i0, np, mr, nc, Vb, Tb, cnt);
    }
    arena_release(ar, mark);
  }
  void lq_update(V A, int i0, int np, int mr, int nc, V Vb, V Tb, int cnt) {
    const int strips = (mr + 15) / 16;
    const int sbk = (strips + kLqUpdWaves - 1) / kLqUpdWaves;
    const dim3 grid((unsigned)(((cnt + 7) / 8) * 8 * sbk));
    const dim3 block(64 * kLqUpdWaves);
    // 16-byte loads and stores when every row of the block starts on a 16-byte boundary
    const bool vec = ((uintptr_t)A.p % 16 == 0) && (A.s % 2 == 0) && (A.ld % 2 == 0) && (i0 % 2 == 0);
    // tiles of 16 columns: the fewest that cover nc (registers and LDS scale with them; the banded
    // panels of the D = 128 combines have 144 or 160 columns)
    if (nc <= 64) {
      launch_update<4>(vec, grid, block, A, i0, np, mr, nc, Vb, Tb, cnt, sbk);
    } else if (nc <= 128) {
      launch_update<8>(vec, grid, block, A, i0, np, mr, nc, Vb, Tb, cnt, sbk);
    } else if (nc <= 144) {
      launch_update<9>(vec, grid, block, A, i0, np, mr, nc, Vb, Tb, cnt, sbk);
    } else if (nc <= 160) {
      launch_update<10>(vec, grid, block, A, i0, np, mr, nc, Vb, Tb, cnt, sbk);
    } else if (nc <= 192) {
      launch_update<12>(vec, grid, block, A, i0, np, mr, nc, Vb, Tb, cnt, sbk);
    } else {
      launch_update<16>(vec, grid, block, A, i0, np, mr, nc, Vb, Tb, cnt, sbk);
    }
    check("k_lq_update");
  }
  template <int NT>
  void launch_update(bool vec, dim3 grid, dim3 block, V A, int i0, int np, int mr, int nc, V Vb, V Tb, int cnt,
                     int sbk) {
    if (vec) {
      hipLaunchKernelGGL((k_lq_update<NT, true>), grid, block, 0, st, A.p, A.s, A.ld, i0, np, mr, nc, Vb.p, Vb.s,
                         Vb.ld, Tb.p, Tb.s, Tb.ld, cnt, sbk);
    } else {
      hipLaunchKernelGGL((k_lq_update<NT, false>), grid, block, 0, st, A.p, A.s, A.ld, i0, np, mr, nc, Vb.p, Vb.s,
                         Vb.ld, Tb.p, Tb.s, Tb.ld, cnt, sbk);
    }
  }
  // op(L) X = R (L lower n x n, X n x m): R is X itself (in place), or the separate R when given
  // (its transpose when rt)
  void trsm(bool trans, int n, int m, V L, V X, int cnt, V R = V{}, bool rt = false) {
    if (err || cnt <= 0) return;
    if (n > kTrsmMaxN || m > 16 * kTrsmWaves) {
      fail("trsm: order or right-hand sides beyond 128", -1);
      return;
    }
    if (!R.p) {
      R = X;
      rt = false;
    }
    const dim3 grid(cnt), block(64 * kTrsmWaves);
    if (trans)
      hipLaunchKernelGGL(k_trsm<true>, grid, block, 0, st, L.p, L.s, L.ld, X.p, X.s, X.ld, n, m, R.p, R.s, R.ld,
                         rt ? 1 : 0);
    else
      hipLaunchKernelGGL(k_trsm<false>, grid, block, 0, st, L.p, L.s, L.ld, X.p, X.s, X.ld, n, m, R.p, R.s, R.ld,
                         rt ? 1 : 0);
    check("k_trsm");
  }
  void copy(V dst, V src, int rows, int cols, int cnt, bool trans = false, bool rev = false, double alpha = 1.0) {
    if (err || cnt <= 0) return;
    if (trans && rows > 1 && cols > 1) {
      hipLaunchKernelGGL(k_copy_t, dim3((cols + 31) / 32, (rows + 31) / 32, cnt), dim3(256), 0, st, dst.p, dst.s,
                         dst.ld, src.p, src.s, src.ld, rows, cols, cnt, rev ? 1 : 0, alpha);
      check("k_copy_t");
      return;
    }
    const int e = rows * cols;
    hipLaunchKernelGGL(k_copy, dim3((e + 255) / 256, cnt), dim3(256), 0, st, dst.p, dst.s, dst.ld, src.p, src.s,
                       src.ld, rows, cols, cnt, trans ? 1 : 0, rev ? 1 : 0, alpha);
    check("k_copy");
  }
  void fill(V dst, int rows, int cols, int cnt, double dv) {
    if (err || cnt <= 0) return;
    const int e = rows * cols;
    hipLaunchKernelGGL(k_fill, dim3((e + 255) / 256, cnt), dim3(256), 0, st, dst.p, dst.s, dst.ld, rows, cols, cnt,
                       dv);
    check("k_fill");
  }
  // several independent copies (src.p == nullptr: dst = fill I) over cnt items in one launch
  struct CopySpec {
    V dst, src;
    int rows, cols;
    double fill;
  };
  void copies(std::initializer_list<CopySpec> ops, int cnt) {
    if (err || cnt <= 0) return;
    if (ops.size() > (size_t)kMaxCopyOps) {
      fail("copies: too many", -1);
      return;
    }
    CopyOps a{};
    int n = 0, emax = 0;
    for (const CopySpec& c : ops) {
      a.op[n++] = CopyOp{c.dst.p, c.dst.s, c.dst.ld, c.src.p, c.src.s, c.src.ld, c.rows, c.cols, c.fill};
      emax = std::max(emax, c.rows * c.cols);
    }
    if (n == 0 || emax == 0) return;
    hipLaunchKernelGGL(k_copy_multi, dim3((emax + 255) / 256, cnt, n), dim3(256), 0, st, a);
    check("k_copy_multi");
  }
  void check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(what, (int)e);
  }
};

// bump allocator over the engine's workspace
struct Arena {
  char* base;
  size_t cap, off = 0;
  double* take(size_t n) {
    const size_t bytes = ((n * sizeof(double)) + 255) & ~(size_t)255;
    if (off + bytes > cap) return nullptr;
    double* p = (double*)(base + off);
    off += bytes;
    return p;
  }
};
double* arena_take(Arena* ar, size_t n) { return ar->take(n); }
size_t arena_mark(Arena* ar) { return ar->off; }
void arena_release(Arena* ar, size_t mark) { ar->off = mark; }

// filter elements (A, b, U, eta, Z) as strided views
struct FE {
  V A, b, U, eta, Z;
  FE every(int k) const { return FE{A.every(k), b.every(k), U.every(k), eta.every(k), Z.every(k)}; }
  FE item(long long i) const { return FE{A.item(i), b.item(i), U.item(i), eta.item(i), Z.item(i)}; }
};
// smoother elements (g, E, D)
struct SE {
  V g, E, Dm;
  SE every(int k) const { return SE{g.every(k), E.every(k), Dm.every(k)}; }
  SE item(long long i) const { return SE{g.item(i), E.item(i), Dm.item(i)}; }
};

FE fe_alloc(Arena& ar, int D, long long n) {
  const long long DD = (long long)D * D;
  FE e;
  e.A = V{ar.take(n * DD), DD, D};
  e.b = V{ar.take(n * D), D, 1};
  e.U = V{ar.take(n * DD), DD, D};
  e.eta = V{ar.take(n * D), D, 1};
  e.Z = V{ar.take(n * DD), DD, D};
  return e;
}
bool fe_ok(const FE& e) { return e.A.p && e.b.p && e.U.p && e.eta.p && e.Z.p; }
SE se_alloc(Arena& ar, int D, long long n) {
  const long long DD = (long long)D * D;
  SE e;
  e.g = V{ar.take(n * D), D, 1};
  e.E = V{ar.take(n * DD), DD, D};
  e.Dm = V{ar.take(n * DD), DD, D};
  return e;
}
bool se_ok(const SE& e) { return e.g.p && e.E.p && e.Dm.p; }

// Filter combine (earlier e1, later e2) -> out, filter.py:117-142 with the
// operand order of the reference.
// With FP32 combine products the GEMMs of this function run in FP32 (the
// factorisations and triangular solves inside stay FP64).
struct CombineProducts {
  Ops& o;
  explicit CombineProducts(Ops& o_) : o(o_) { o.f32 = o.f32_combines; }
  ~CombineProducts() { o.f32 = false; }
};

void filter_fn(Ops& o, Arena& ar, int D, const FE& e1, const FE& e2, const FE& out, int cnt) {
  if (cnt <= 0 || o.err) return;
  CombineProducts prec(o);
  const size_t mark = ar.off;
  const long long DD = (long long)D * D;
  V X{ar.take((size_t)cnt * 4 * DD), 4 * DD, 2 * D};
  // [M^T | U2] for items 0 .. cnt - 1 and [A1^T Xi22 | Z1] for items cnt .. 2 cnt - 1: both are
  // factored by one LQ call over 2 cnt items
  V Y{ar.take((size_t)cnt * 4 * DD), 2 * DD, 2 * D};
  const V YZ = Y.item(cnt);
  V T2{ar.take((size_t)cnt * DD), DD, D};
  V T3{ar.take((size_t)cnt * DD), DD, D}, T4{ar.take((size_t)cnt * DD), DD, D};
  if (!X.p || !Y.p || !T2.p || !T3.p || !T4.p) {
    o.fail("workspace", -2);
    return;
  }
  // Xi = tria([[U1^T Z2, I], [Z2, 0]])
  o.gemm(true, false, D, D, D, 1.0, e1.U, e2.Z, 0.0, X.at(0, 0), cnt);
  o.copies({{X.at(0, D), V{}, D, D, 1.0}, {X.at(D, 0), e2.Z, D, D, 0.0}, {X.at(D, D), V{}, D, D, 0.0}}, cnt);
  // second block [I; 0]: row r is zero beyond column D + r
  o.lq(X, 2 * D, 2 * D, cnt, D, 0);
  const V Xi11 = X.at(0, 0), Xi21 = X.at(D, 0), Xi22 = X.at(D, D);
  // m_ = Xi11^-1 U1^T -> T3; M^T = (m_ A2^T)^T = A2 m_^T, formed in place as the left
  // block of [M^T | U2] (Y)
  o.trsm(false, D, D, Xi11, T3, cnt, e1.U, true);
  o.gemm(false, true, D, D, D, 1.0, e2.A, T3, 0.0, Y.at(0, 0), cnt);
  // A = A2 A1 - M^T (Xi21^T A1)
  o.gemm(true, false, D, D, D, 1.0, Xi21, e1.A, 0.0, T2, cnt);
  o.gemm(false, false, D, D, D, 1.0, e2.A, e1.A, 0.0, out.A, cnt);
  o.gemm(false, false, D, D, D, -1.0, Y.at(0, 0), T2, 1.0, out.A, cnt);
  // b = A2 (v2 - m_^T Xi21^T v2) + b2, v2 = b1 + U1 (U1^T eta2)
  o.vec_chain(D, e1.U, true, e1.U, false, Xi21, true, T3, true, e2.A, false, e2.eta, e1.b, 1.0, e2.b, out.b, cnt);
  // _e = Xi11^-T Xi21^T -> T4; eta = A1^T (w2 - _e^T U1^T w2) + eta1, w2 = eta2 - Z2 (Z2^T b1)
  o.trsm(true, D, D, Xi11, T4, cnt, Xi21, true);
  o.vec_chain(D, e2.Z, true, e2.Z, false, e1.U, true, T4, true, e1.A, true, e1.b, e2.eta, -1.0, e1.eta, out.eta,
              cnt);
  // A1^T Xi22 and the right blocks U2, Z1; then U = tria([M^T | U2]) and Z = tria([A1^T Xi22 | Z1]) in one call
  o.gemm(true, false, D, D, D, 1.0, e1.A, Xi22, 0.0, YZ.at(0, 0), cnt);
  o.copies({{Y.at(0, D), e2.U, D, D, 0.0}, {YZ.at(0, D), e1.Z, D, D, 0.0}}, cnt);
  // second blocks U2 (lower triangular) and Z1 (lower triangular, or [Zr | 0] at the leaves)
  o.lq(Y, D, 2 * D, 2 * cnt, D, o.zw2);
  o.copies({{out.U, Y.at(0, 0), D, D, 0.0}, {out.Z, YZ.at(0, 0), D, D, 0.0}}, cnt);
  ar.off = mark;
}

// Smoother combine (smoother.py:53-63; under reverse=True e1 is the later aggregate)
void smoother_fn(Ops& o, Arena& ar, int D, const SE& e1, const SE& e2, const SE& out, int cnt, bool with_chol) {
  if (cnt <= 0 || o.err) return;
  CombineProducts prec(o);
  const size_t mark = ar.off;
  o.copy(out.g, e2.g, D, 1, cnt);
  o.gemm(false, false, D, 1, D, 1.0, e2.E, e1.g, 1.0, out.g, cnt);
  o.gemm(false, false, D, D, D, 1.0, e2.E, e1.E, 0.0, out.E, cnt);
  if (with_chol) {
    const long long DD = (long long)D * D;
    V Y{ar.take((size_t)cnt * 2 * DD), 2 * DD, 2 * D};
    if (!Y.p) {
      o.fail("workspace", -2);
      return;
    }
    o.gemm(false, false, D, D, D, 1.0, e2.E, e1.Dm, 0.0, Y.at(0, 0), cnt);
    o.copy(Y.at(0, D), e2.Dm, D, D, cnt);
    o.lq(Y, D, 2 * D, cnt, D, 0);  // second block D2: lower triangular
    o.copy(out.Dm, Y.at(0, 0), D, D, cnt);
  }
  ar.off = mark;
}

// JAX's associative_scan (odd/even recursion) of n elements `el` into `out`
// (inclusive prefixes; out must not alias el).
template <typename E, typename Fn, typename Alloc, typename Ok>
void scan(Ops& o, Arena& ar, int D, const E& el, const E& out, long long n, Fn fn, Alloc alloc, Ok ok,
          void (*copy1)(Ops&, int, const E&, const E&)) {
  if (o.err || n <= 0) return;
  if (n == 1) {
    copy1(o, D, out, el);
    return;
  }
  const size_t mark = ar.off;
  const long long h = n / 2;
  E red = alloc(ar, D, h);
  if (!ok(red)) {
    o.fail("workspace", -2);
    return;
  }
  fn(o, ar, D, el.every(2), el.item(1).every(2), red, (int)h);
  // odd positions of out hold the scan of the pair sums
  const E odd = out.item(1).every(2);
  scan(o, ar, D, red, odd, h, fn, alloc, ok, copy1);
  // even positions: out[2j] = odd[j - 1] o el[2j], j >= 1
  const long long ne = (n % 2 == 0) ? h - 1 : h;
  fn(o, ar, D, odd, el.item(2).every(2), out.item(2).every(2), (int)ne);
  copy1(o, D, out, el);
  ar.off = mark;
}

void fe_copy1(Ops& o, int D, const FE& dst, const FE& src) {
  o.copies({{dst.A, src.A, D, D, 0.0},
            {dst.b, src.b, D, 1, 0.0},
            {dst.U, src.U, D, D, 0.0},
            {dst.eta, src.eta, D, 1, 0.0},
            {dst.Z, src.Z, D, D, 0.0}},
           1);
}
void se_copy1(Ops& o, int D, const SE& dst, const SE& src) {
  o.copies({{dst.g, src.g, D, 1, 0.0}, {dst.E, src.E, D, D, 0.0}, {dst.Dm, src.Dm, D, D, 0.0}}, 1);
}

}  // namespace

struct Engine {
  int device = 0;
  char* buf = nullptr;
  size_t cap = 0;
};

Engine* engine_create(int device) {
  Engine* e = new Engine();
  e->device = device;
  return e;
}

void engine_destroy(Engine* e) {
  if (!e) return;
  if (e->buf) (void)hipFree(e->buf);
  delete e;
}

bool supports(int kind, int D, int d) {
  // the LQ panel kernel holds up to 256 columns (2 D) in registers and the triangular solves
  // take orders up to 128: D <= 128
  return kind == kLorenz96 && d >= 4 && D % d == 0 && D / d >= 2 && D <= 128;
}

namespace {

// workspace doubles for one solve
size_t workspace_doubles(int D, int d, long long Ns, bool f32) {
  const size_t DD = (size_t)D * D;
  const size_t n1 = (size_t)Ns + 1;
  size_t w = 0;
  w += 2 * n1 * D;                          // two mean trajectories
  w += (size_t)Ns * ((size_t)d * D + d);    // H, c
  w += 2 * (size_t)Ns * (3 * DD + 2 * D);   // leaf elements + scan outputs (A, eta, Z); b, U go to fm / fL
  w += n1 * (D + DD);                       // filtered fm, fL
  w += 2 * (size_t)Ns * (3 * DD + 2 * D);   // scan recursion (pair sums over all levels)
  w += 4 * n1 * (D + 2 * DD);               // smoother elements natural / reversed / scan out / recursion
  w += n1 * (D + DD);                       // smoothed natural order
  // per-call temporaries at the widest level (the filter combine), and the post step
  const size_t h = (size_t)(Ns / 2 + 1);
  w += h * (4 * DD + 4 * DD + 5 * DD) + 4096;
  w += (size_t)(Ns + 1) * (2 * D * kNB + kNB * kNB) + 4096;  // blocked-LQ panels (V, T)
  w += (size_t)Ns * (5 * DD + (size_t)d * D + 2 * D + 3 * d + 4) + 4096;
  w += 8 * n1 + 64 * 1024;
  // FP32 staging of one GEMM (both operands and the product; at most n1 items of D x D)
  if (f32) w += n1 * 2 * DD + 3 * 4096;
  return w + w / 10;
}

}  // namespace

int solve(Engine* e, const Problem& p, const pof_opts& opts, pof_out* out, const EngineFns* small, hipStream_t st,
          std::string& err) {
  const int D = p.D, d = p.d, q1 = p.q + 1;
  const long long Ns = p.Ns;
  const long long DD = (long long)D * D;
  if (Ns > 65534) {  // batched launches put the item count in grid y / z
    err = "large engine supports at most 65534 steps";
    return POF_E_ARG;
  }
  if (Ns < 1) {
    err = "large engine needs Ns >= 1";
    return POF_E_ARG;
  }
  if (!small && !supports(p.kind, D, d)) {
    err = "large-state engine: no device linearisation for this field and (D, d)";
    return POF_E_UNSUPPORTED;
  }
  if (D > 128) {
    err = "large-state engine: D > 128";
    return POF_E_UNSUPPORTED;
  }
  // workspace
  const bool f32 = opts.products == POF_PRODUCTS_F32;
  const size_t need = workspace_doubles(D, d, Ns, f32) * sizeof(double);
  if (e->cap < need) {
    if (e->buf) (void)hipFree(e->buf);
    e->buf = nullptr;
    e->cap = 0;
    if (hipMalloc((void**)&e->buf, need) != hipSuccess) {
      err = "large-state engine: workspace allocation failed";
      return POF_E_HIP;
    }
    e->cap = need;
  }
  Arena ar{e->buf, e->cap, 0};
  Ops o{st, &ar};
  o.f32_combines = f32;
  o.zw2 = d - 1;  // leaf Z = [Zr | 0]: nonzero in its first d columns only
  // constants: F, QL, N1z, P, x0 (one block)
  double* cst = ar.take(5 * DD + 2 * D);
  double* mA = ar.take((size_t)(Ns + 1) * D);
  double* mB = ar.take((size_t)(Ns + 1) * D);
  double* H = ar.take((size_t)Ns * d * D);
  double* cc = ar.take((size_t)Ns * d);
  double* fm = ar.take((size_t)(Ns + 1) * D);
  double* fL = ar.take((size_t)(Ns + 1) * DD);
  double* sm = ar.take((size_t)(Ns + 1) * D);
  double* sL = ar.take((size_t)(Ns + 1) * DD);
  double* terms = ar.take(4 * (size_t)(Ns + 1) + 2 * (size_t)Ns * d + 8);
  if (!cst || !mA || !mB || !H || !cc || !fm || !fL || !sm || !sL || !terms) {
    err = "large-state engine: workspace too small";
    return POF_E_HIP;
  }
  double* dF = cst;
  double* dQL = cst + DD;
  double* dN1z = cst + 2 * DD;
  double* dL0 = cst + 3 * DD;
  double* dQLi = cst + 4 * DD;
  double* dP = cst + 5 * DD;
  double* dm0 = dP + D;
  double* t_ssq = terms;
  double* t_nll = t_ssq + (Ns + 1);
  double* t_obj = t_nll + (Ns + 1);
  double* t_viol = t_obj + (Ns + 1);
  double* t_tmp = t_viol + (Ns + 1);
  double* red = t_tmp + 2 * Ns * d;
  hipError_t he = hipSuccess;
  auto up = [&](double* dst, const double* src, size_t n) {
    if (he == hipSuccess) he = hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, st);
  };
  up(dF, p.F, DD);
  up(dQL, p.QL, DD);
  up(dN1z, p.N1z, DD);
  up(dL0, p.x0L, DD);
  up(dP, p.P, D);
  // QL^-1 by forward substitution (the objective's solve with the constant QL)
  std::vector<double> qli((size_t)DD, 0.0);
  for (int j = 0; j < D; ++j)
    for (int i = j; i < D; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
      for (int t = j; t < i; ++t) v -= p.QL[(size_t)i * D + t] * qli[(size_t)t * D + j];
      qli[(size_t)i * D + j] = v / p.QL[(size_t)i * D + i];
    }
  up(dQLi, qli.data(), DD);
  // a lower-triangular QL makes the second blocks [QL] and [QL; 0] banded (the LQs skip past the band)
  bool ql_lower = true;
  for (int i = 0; i < D; ++i)
    for (int j = i + 1; j < D; ++j) ql_lower = ql_lower && p.QL[(size_t)i * D + j] == 0.0;
  const int ql_band = ql_lower ? D : 0;
  up(dm0, p.x0m, D);
  up(mA, p.m0, (size_t)(Ns + 1) * D);
  if (he != hipSuccess) {
    err = std::string("large-state engine upload: ") + hipGetErrorString(he);
    return POF_E_HIP;
  }
  const V Fv{dF, 0, D}, QLv{dQL, 0, D}, N1zv{dN1z, 0, D}, QLiv{dQLi, 0, D};
  const size_t base_mark = ar.off;

  double* mcur = mA;
  double* mprev = mB;
  int k = 0;
  double nll = 0.0, obj = 0.0, ssq = 0.0;
  int status = POF_STATUS_CONVERGED;
  ModelArgs la{};
  la.Pd = dP;
  la.kind = p.kind;
  la.q = p.q;
  for (int i = 0; i < kMaxParams; ++i) la.params[i] = p.params[i];
  for (;;) {
    ar.off = base_mark;
    // ---- EK1 at the current means (step.py:12-22)
    if (small) {
      la.Ns = (int)Ns;
      he = small->linearize(la, mcur + D, (int)Ns, H, cc, st);
    } else {
      const long long nt = Ns * d;
      hipLaunchKernelGGL(k_lin_l96, dim3((unsigned)((nt + 127) / 128)), dim3(128), 0, st, mcur + D, dP,
                         p.params[0], (int)Ns, d, q1, H, cc);
      he = hipGetLastError();
    }
    if (he != hipSuccess) o.fail("linearize", (int)he);
    const V Hv{H, (long long)d * D, D}, cv{cc, d, 1};
    // ---- filter elements (filter.py:50-81): all outputs of element k in leaf[k]
    FE leaf = fe_alloc(ar, D, Ns);
    if (!fe_ok(leaf)) o.fail("workspace", -2);
    {
      const size_t mark = ar.off;
      const int n = (int)Ns;
      V Pb{ar.take((size_t)n * (d + D) * D), (long long)(d + D) * D, D};
      V HF{ar.take((size_t)n * d * D), (long long)d * D, D};
      V Zt{ar.take((size_t)n * d * D), (long long)d * D, D};
      V X2{ar.take((size_t)n * d * D), (long long)d * D, D};
      V KT{ar.take((size_t)n * d * D), (long long)d * D, D};
      V t0{ar.take((size_t)d + 8), 0, 1};
      V m1{ar.take((size_t)D + 8), 0, 1};
      V N10{ar.take((size_t)2 * DD), 2 * DD, 2 * D};
      if (!Pb.p || !HF.p || !Zt.p || !X2.p || !KT.p || !t0.p || !m1.p || !N10.p) o.fail("workspace", -2);
      // element 0: m1 = F m0, N1 = tria([F L0 | QL]); elements k >= 1: m1 = 0, N1 = tria([0 | QL])
      o.gemm(false, false, D, 1, D, 1.0, Fv, V{dm0, 0, 1}, 0.0, m1, 1);
      o.gemm(false, false, D, D, D, 1.0, Fv, V{dL0, 0, D}, 0.0, N10.at(0, 0), 1);
      o.copy(N10.at(0, D), QLv, D, D, 1);
      o.lq(N10, D, 2 * D, 1, ql_band, 0);
      for (int part = 0; part < 2; ++part) {
        const int k0 = part == 0 ? 0 : 1;
        const int cnt = part == 0 ? 1 : n - 1;
        if (cnt <= 0) continue;
        const V N1 = part == 0 ? V{N10.p, 0, 2 * D} : N1zv;
        const V Hk = Hv.item(k0), ck = cv.item(k0);
        const FE ok = leaf.item(k0);
        const V P0 = Pb.item(k0), HF0 = HF.item(k0), Z0 = Zt.item(k0), X0 = X2.item(k0);
        const V KT0 = KT.item(k0);
        // Psi = tria([[H N1], [N1]]) (the zero column block of filter.py:63-66 changes no pivot)
        o.gemm(false, false, d, D, D, 1.0, Hk, N1, 0.0, P0.at(0, 0), cnt);
        o.copy(P0.at(d, 0), N1, D, D, cnt);
        o.lq(P0, d + D, D, cnt);
        // U = Psi22 = [Psi[d:, d:D] | 0]
        o.copy(ok.U, P0.at(d, d), D, D - d, cnt);
        o.fill(ok.U.at(0, D - d), D, d, cnt, 0.0);
        // K^T = Psi11^-T Psi21^T (K = Psi21 Psi11^-1)
        o.trsm(true, d, D, P0.at(0, 0), KT0, cnt, P0.at(d, 0), true);
        // A = F - K (H F)
        o.gemm(false, false, d, D, D, 1.0, Hk, Fv, 0.0, HF0, cnt);
        o.copy(ok.A, Fv, D, D, cnt);
        o.gemm(true, false, D, D, d, -1.0, KT0, HF0, 1.0, ok.A, cnt);
        // b = m1 - K (H m1 + c)
        if (part == 0) {
          o.copy(t0, ck, d, 1, 1);
          o.gemm(false, false, d, 1, D, 1.0, Hk, m1, 1.0, t0, 1);
          o.copy(ok.b, m1, D, 1, 1);
          o.gemm(true, false, D, 1, d, -1.0, KT0, t0, 1.0, ok.b, 1);
        } else {
          o.gemm(true, false, D, 1, d, -1.0, KT0, ck, 0.0, ok.b, cnt);
        }
        // Z = [(Psi11^-1 H F)^T | 0]; eta = -(Psi11^-T Psi11^-1 H F)^T c
        o.trsm(false, d, D, P0.at(0, 0), Z0, cnt, HF0, false);
        o.copy(ok.Z, Z0, D, d, cnt, true);
        o.fill(ok.Z.at(0, d), D, D - d, cnt, 0.0);
        o.trsm(true, d, D, P0.at(0, 0), X0, cnt, Z0, false);
        o.gemm(true, false, D, 1, d, -1.0, X0, ck, 0.0, ok.eta, cnt);
      }
      ar.off = mark;
    }
    // ---- filter scan (filter.py:35): prefixes land in fm / fL (items 1..Ns) and scratch
    {
      FE outp;
      outp.b = V{fm + D, D, 1};
      outp.U = V{fL + DD, DD, D};
      outp.A = V{ar.take((size_t)Ns * DD), DD, D};
      outp.eta = V{ar.take((size_t)Ns * D), D, 1};
      outp.Z = V{ar.take((size_t)Ns * DD), DD, D};
      if (!fe_ok(outp)) o.fail("workspace", -2);
      scan<FE>(o, ar, D, leaf, outp, Ns, filter_fn, fe_alloc, fe_ok, fe_copy1);
      o.copy(V{fm, D, 1}, V{dm0, 0, 1}, D, 1, 1);
      o.copy(V{fL, DD, D}, V{dL0, 0, D}, D, D, 1);
    }
    // ---- smoother elements and observation statistics (smoother.py:25-50, filter.py:84-114)
    SE snat = se_alloc(ar, D, Ns + 1);
    if (!se_ok(snat)) o.fail("workspace", -2);
    {
      const size_t mark = ar.off;
      const int n = (int)Ns;
      V Phi{ar.take((size_t)n * 4 * DD), 4 * DD, 2 * D};
      V Fm{ar.take((size_t)n * D), D, 1};
      V Sb{ar.take((size_t)n * d * D), (long long)d * D, D};
      V yv{ar.take((size_t)n * d), d, 1};
      V Phin{ar.take((size_t)n * DD), DD, D};
      if (!Phi.p || !Fm.p || !Sb.p || !yv.p || !Phin.p) o.fail("workspace", -2);
      const V fmv{fm, D, 1}, fLv{fL, DD, D};
      o.gemm(false, false, D, D, D, 1.0, Fv, fLv, 0.0, Phi.at(0, 0), n);
      o.copies({{Phi.at(0, D), QLv, D, D, 0.0}, {Phi.at(D, 0), fLv, D, D, 0.0}, {Phi.at(D, D), V{}, D, D, 0.0}}, n);
      o.lq(Phi, 2 * D, 2 * D, n, ql_band, 0);  // second block [QL; 0]
      // E = Phi21 Phi11^-1, D = Phi22, g = m - E F m
      // E^T = Phi11^-T Phi21^T
      o.trsm(true, D, D, Phi.at(0, 0), Phin, n, Phi.at(D, 0), true);
      o.copy(snat.E, Phin, D, D, n, true);
      o.copy(snat.Dm, Phi.at(D, D), D, D, n);
      o.gemm(false, false, D, 1, D, 1.0, Fv, fmv, 0.0, Fm, n);
      o.copy(snat.g, fmv, D, 1, n);
      o.gemm(false, false, D, 1, D, -1.0, snat.E, Fm, 1.0, snat.g, n);
      // S = tria(H Phi11), yhat = H F m + c
      o.gemm(false, false, d, D, D, 1.0, Hv, Phi.at(0, 0), 0.0, Sb, n);
      o.lq(Sb, d, D, n);
      o.copy(yv, cv, d, 1, n);
      o.gemm(false, false, d, 1, D, 1.0, Hv, Fm, 1.0, yv, n);
      if (!o.err) {
        hipLaunchKernelGGL(k_obs_terms, dim3((n + 127) / 128), dim3(128), 0, st, Sb.p, Sb.s, Sb.ld, yv.p, yv.s, d, n,
                           t_ssq, t_nll, t_tmp);
        o.check("k_obs_terms");
      }
      // terminal element (m_Ns, 0, L_Ns)
      o.copy(snat.g.item(Ns), fmv.item(Ns), D, 1, 1);
      o.fill(snat.E.item(Ns), D, D, 1, 0.0);
      o.copy(snat.Dm.item(Ns), fLv.item(Ns), D, D, 1);
      ar.off = mark;
    }
    // ---- reverse scan (smoother.py:15-17): reversed inputs, scan, reversed outputs
    {
      const size_t mark = ar.off;
      const long long n = Ns + 1;
      SE srev = se_alloc(ar, D, n), sout = se_alloc(ar, D, n);
      if (!se_ok(srev) || !se_ok(sout)) o.fail("workspace", -2);
      o.copy(srev.g, snat.g, D, 1, (int)n, false, true);
      o.copy(srev.E, snat.E, D, D, (int)n, false, true);
      o.copy(srev.Dm, snat.Dm, D, D, (int)n, false, true);
      auto sfn = [](Ops& oo, Arena& aa, int DD2, const SE& a, const SE& b, const SE& c, int cnt) {
        smoother_fn(oo, aa, DD2, a, b, c, cnt, true);
      };
      scan<SE>(o, ar, D, srev, sout, n, sfn, se_alloc, se_ok, se_copy1);
      o.copy(V{sm, D, 1}, sout.g, D, 1, (int)n, false, true);
      o.copy(V{sL, DD, D}, sout.Dm, D, D, (int)n, false, true);
      ar.off = mark;
    }
    // ---- objective (utils.py:97-101) and the convergence partials
    {
      const size_t mark = ar.off;
      V r{ar.take((size_t)Ns * D), D, 1}, r2{ar.take((size_t)Ns * D), D, 1};
      if (!r.p || !r2.p) o.fail("workspace", -2);
      o.copy(r, V{sm, D, 1}, D, 1, (int)Ns);
      o.gemm(false, false, D, 1, D, -1.0, Fv, V{sm + D, D, 1}, 1.0, r, (int)Ns);
      o.gemm(false, false, D, 1, D, 1.0, QLiv, r, 0.0, r2, (int)Ns);
      if (!o.err) {
        hipLaunchKernelGGL(k_sqnorm, dim3((unsigned)((Ns + 127) / 128)), dim3(128), 0, st, r2.p, D, (int)Ns, t_obj);
        hipLaunchKernelGGL(k_viol, dim3((unsigned)((Ns + 128) / 128)), dim3(128), 0, st, sm, mcur, D, (int)(Ns + 1),
                           t_viol);
        hipLaunchKernelGGL(k_reduce4, dim3(1), dim3(256), 0, st, t_ssq, t_nll, t_obj, (int)Ns, t_viol, (int)(Ns + 1),
                           red);
        o.check("reductions");
      }
      ar.off = mark;
    }
    if (o.err) {
      err = "large-state engine: " + o.msg;
      return POF_E_HIP;
    }
    double hr[4] = {0, 0, 0, 0};
    he = hipMemcpyAsync(hr, red, 4 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) {
      err = std::string("large-state engine: ") + hipGetErrorString(he);
      return POF_E_HIP;
    }
    k += 1;
    const double obj_old = obj;
    ssq = hr[0] / (double)Ns / (double)d;
    nll = hr[1];
    obj = hr[2];
    const bool means_close = hr[3] == 0.0;
    const bool isnan_ = std::isnan(obj) || std::isnan(nll);
    // the smoothed means feed the next linearisation
    he = hipMemcpyAsync(mprev, sm, (size_t)(Ns + 1) * D * sizeof(double), hipMemcpyDeviceToDevice, st);
    std::swap(mcur, mprev);
    bool finished = false;
    if (opts.fixed_iters > 0) {
      if (k >= opts.fixed_iters) {
        status = POF_STATUS_FIXED;
        finished = true;
      }
    } else {
      const bool obj_conv = std::fabs(obj_old - obj) <= 1e-9 + 1e-6 * std::fabs(obj);
      if (!(k <= opts.maxiters)) {
        status = POF_STATUS_MAXITER;
        finished = true;
      } else if (isnan_ || obj_conv || means_close) {
        status = isnan_ ? POF_STATUS_NAN : POF_STATUS_CONVERGED;
        finished = true;
      }
    }
    if (finished) break;
  }
  // ---- outputs: factors calibrated twice (step.py:42-44 + solver.py:66-68)
  const double scale = opts.calibrate ? ssq : std::sqrt(ssq);
  const long long ne = (Ns + 1) * D;
  hipLaunchKernelGGL(k_project, dim3((unsigned)((ne + 127) / 128)), dim3(128), 0, st, sm, sL, dP, D, d,
                     (int)(Ns + 1), scale, out->means_dev, out->chols_dev, out->full_means_dev, out->full_chols_dev);
  he = hipGetLastError();
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  if (he != hipSuccess) {
    err = std::string("large-state engine output: ") + hipGetErrorString(he);
    return POF_E_HIP;
  }
  out->iterations = (double)k;
  out->nll = nll;
  out->obj = obj;
  out->sigma_squared = ssq;
  out->calibrated = opts.calibrate ? 1 : 0;
  out->status = status;
  return POF_OK;
}

}  // namespace large
}  // namespace pof
