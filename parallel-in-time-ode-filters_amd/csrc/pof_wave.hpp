// Wave-wide sums of FP64 values in registers, for 64-lane gfx950 waves: no LDS,
// no barrier.  Lanes trade values with v_permlane32_swap / v_permlane16_swap
// (gfx950) and DPP row permutations.
#pragma once

#include <hip/hip_runtime.h>

namespace pof {
namespace wave {

__device__ __forceinline__ double from_bits(unsigned lo, unsigned hi) {
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | (unsigned long long)lo);
}

// v_permlane32_swap on both halves of a double: lanes 32..63 of a trade places
// with lanes 0..31 of b, so a = [a.lo, b.lo] and b = [a.hi, b.hi] afterwards.
__device__ __forceinline__ void swap_halves(double& a, double& b) {
  const unsigned long long xa = __builtin_bit_cast(unsigned long long, a);
  const unsigned long long xb = __builtin_bit_cast(unsigned long long, b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xa, (unsigned)xb, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(xa >> 32), (unsigned)(xb >> 32), false, false);
  a = from_bits(lo[0], hi[0]);
  b = from_bits(lo[1], hi[1]);
}

// v_permlane16_swap on both halves of a double: the odd 16-lane rows of a trade
// places with the even rows of b, so a = [a.r0, b.r0, a.r2, b.r2] and
// b = [a.r1, b.r1, a.r3, b.r3] afterwards.
__device__ __forceinline__ void swap_rows(double& a, double& b) {
  const unsigned long long xa = __builtin_bit_cast(unsigned long long, a);
  const unsigned long long xb = __builtin_bit_cast(unsigned long long, b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)xa, (unsigned)xb, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(xa >> 32), (unsigned)(xb >> 32), false, false);
  a = from_bits(lo[0], hi[0]);
  b = from_bits(lo[1], hi[1]);
}

// A DPP permutation of a double (two 32-bit moves); every source lane is valid
// for the controls used here.
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, false);
  return from_bits(lo, hi);
}
constexpr int kRowMirror = 0x140;      // lane i of a 16-lane row reads lane 15 - i
constexpr int kRowHalfMirror = 0x141;  // lane i of an 8-lane half-row reads lane 7 - i
constexpr int kQuadSwap1 = 0xB1;       // quad_perm [1, 0, 3, 2]
constexpr int kQuadSwap2 = 0x4E;       // quad_perm [2, 3, 0, 1]

// One halving step over lanes paired by a DPP mirror, on the first M values:
// the first lane of a pair (low) keeps x[i], the other keeps x[i + M / 2], and
// each adds what its partner held of the value it keeps.
template <int CTRL, int M, int N>
__device__ __forceinline__ void halve_dpp(double (&x)[N], bool low) {
#pragma unroll
  for (int i = 0; i < M / 2; ++i) {
    const double keep = low ? x[i] : x[i + M / 2];
    const double send = low ? x[i + M / 2] : x[i];
    x[i] = keep + dpp<CTRL>(send);
  }
}

// Transposed wave sum of NV (1, 2, 4, 8 or 16) values: on return x[0] of
// lanes j (64 / NV) .. (j + 1) (64 / NV) - 1 holds the sum over all 64 lanes
// of x[j].  Each step halves both the lanes a partial sum covers and the
// values a lane holds (lanes 32 apart, then 16, then mirrored within a row and
// a half-row), and once a lane holds one value the lanes of its group add up
// what remains: NV + 5 - log2(NV) adds per lane.  The order of the additions
// is fixed by the lane layout, so every call with the same inputs gives the
// same bits.
template <int NV>
__device__ __forceinline__ void sum_scatter(double (&x)[NV], int lane) {
  static_assert(NV == 1 || NV == 2 || NV == 4 || NV == 8 || NV == 16, "NV: 1, 2, 4, 8 or 16");
  // lanes 32 apart
  if constexpr (NV >= 2) {
#pragma unroll
    for (int i = 0; i < NV / 2; ++i) {
      swap_halves(x[i], x[i + NV / 2]);
      x[i] += x[i + NV / 2];
    }
  } else {
    double y = x[0];
    swap_halves(x[0], y);
    x[0] += y;
  }
  // lanes 16 apart (rows 0 / 1 and 2 / 3)
  constexpr int N1 = NV >= 2 ? NV / 2 : 1;
  if constexpr (N1 >= 2) {
#pragma unroll
    for (int i = 0; i < N1 / 2; ++i) {
      swap_rows(x[i], x[i + N1 / 2]);
      x[i] += x[i + N1 / 2];
    }
  } else {
    double y = x[0];
    swap_rows(x[0], y);
    x[0] += y;
  }
  // lane i of a row with lane 15 - i
  constexpr int N2 = N1 >= 2 ? N1 / 2 : 1;
  if constexpr (N2 >= 2) {
    halve_dpp<kRowMirror, N2>(x, (lane & 15) < 8);
  } else {
    x[0] += dpp<kRowMirror>(x[0]);
  }
  // lane i of a half-row with lane 7 - i
  constexpr int N3 = N2 >= 2 ? N2 / 2 : 1;
  if constexpr (N3 >= 2) {
    halve_dpp<kRowHalfMirror, N3>(x, (lane & 7) < 4);
  } else {
    x[0] += dpp<kRowHalfMirror>(x[0]);
  }
  // the four lanes of a quad
  x[0] += dpp<kQuadSwap1>(x[0]);
  x[0] += dpp<kQuadSwap2>(x[0]);
}

// The value of v in lane L (a compile-time lane): two scalar reads.
__device__ __forceinline__ double lane_value(double v, int L) {
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, L);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), L);
  return from_bits(lo, hi);
}

}  // namespace wave
}  // namespace pof
