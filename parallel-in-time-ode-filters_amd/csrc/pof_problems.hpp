// ODE vector fields shared by device kernels (f and its Jacobian through dual
// numbers) and host set-up (Taylor-mode derivatives through truncated series).
//
// Fields follow pof/ivp.py of the reference: logistic 7-14, lotkavolterra
// 17-31, vanderpol 34-41, fitzhughnagumo 44-60, rober 63-79, rigid_body 82-90,
// seir 93-108, threebody 111-126, henonheiles 137-152.  fhn_classic, lorenz63
// and lorenz96 are textbook fields the reference lacks.  Every field is
// autonomous: the reference calls f(None, y) (pof/convenience.py:28).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace pof {

enum ProblemKind : int {
  kLogistic = 1,
  kLotkaVolterra = 2,
  kVanDerPol = 3,
  kFitzHughNagumo = 4,
  kFhnClassic = 5,
  kRober = 6,
  kRigidBody = 7,
  kSeir = 8,
  kThreeBody = 9,
  kHenonHeiles = 10,
  kLorenz63 = 11,
  kLorenz96 = 12,
};

constexpr int kMaxParams = 8;

// ODE dimension each field requires (0: any d >= 4, used by Lorenz-96).
__host__ __device__ constexpr int problem_dim(int kind) {
  switch (kind) {
    case kLogistic: return 1;
    case kLotkaVolterra: case kVanDerPol: case kFitzHughNagumo: case kFhnClassic: return 2;
    case kRober: case kRigidBody: case kLorenz63: return 3;
    case kSeir: case kThreeBody: case kHenonHeiles: return 4;
    case kLorenz96: return 0;
    default: return -1;
  }
}

// ---------------------------------------------------------------- dual numbers
template <int N>
struct Dual {
  double v;
  double g[N];
};

template <int N>
__host__ __device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = -a.g[i];
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = (a.g[i] - r.v * b.g[i]) / b.v;
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator+(const Dual<N>& a, double b) { Dual<N> r = a; r.v += b; return r; }
template <int N>
__host__ __device__ inline Dual<N> operator+(double b, const Dual<N>& a) { return a + b; }
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a, double b) { Dual<N> r = a; r.v -= b; return r; }
template <int N>
__host__ __device__ inline Dual<N> operator-(double b, const Dual<N>& a) { Dual<N> r = -a; r.v += b; return r; }
template <int N>
__host__ __device__ inline Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r; r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b;
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator*(double b, const Dual<N>& a) { return a * b; }
template <int N>
__host__ __device__ inline Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r; r.v = a.v / b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] / b;
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> dsqrt(const Dual<N>& a) {
  Dual<N> r; r.v = sqrt(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] / (2.0 * r.v);
  return r;
}
__host__ __device__ inline double dsqrt(double a) { return sqrt(a); }

// ------------------------------------------------------------- vector fields
// f[0..d) = field(y[0..d)); T is double, Dual<N> or a host power series.
template <typename T, int d>
__host__ __device__ inline void vector_field(int kind, const double* p, const T* y, T* f) {
  switch (kind) {
    case kLogistic:
      if constexpr (d == 1) f[0] = p[0] * y[0] * (1.0 - y[0] / p[1]);
      break;
    case kLotkaVolterra:
      if constexpr (d == 2) {
        f[0] = p[0] * y[0] - p[1] * y[0] * y[1];
        f[1] = -p[2] * y[1] + p[3] * y[0] * y[1];
      }
      break;
    case kVanDerPol:
      if constexpr (d == 2) {
        f[0] = y[1];
        f[1] = p[0] * ((1.0 - y[0] * y[0]) * y[1] - y[0]);
      }
      break;
    case kFitzHughNagumo:
      if constexpr (d == 2) {
        // (a, b, tinv, l) = p
        f[0] = y[0] - (y[0] * y[0] * y[0]) / 3.0 - y[1] + p[3];
        f[1] = p[2] * (y[0] + p[0] - p[1] * y[1]);
      }
      break;
    case kFhnClassic:
      if constexpr (d == 2) {
        // (a, b, c) = p
        f[0] = p[2] * (y[0] - (y[0] * y[0] * y[0]) / 3.0 + y[1]);
        f[1] = -(y[0] - p[0] + p[1] * y[1]) / p[2];
      }
      break;
    case kRober:
      if constexpr (d == 3) {
        f[0] = -p[0] * y[0] + p[2] * y[1] * y[2];
        f[1] = p[0] * y[0] - p[1] * (y[1] * y[1]) - p[2] * y[1] * y[2];
        f[2] = p[1] * (y[1] * y[1]);
      }
      break;
    case kRigidBody:
      if constexpr (d == 3) {
        f[0] = p[0] * y[1] * y[2];
        f[1] = p[1] * y[0] * y[2];
        f[2] = p[2] * y[0] * y[1];
      }
      break;
    case kSeir:
      if constexpr (d == 4) {
        f[0] = -p[1] * y[0] * y[2] / p[3];
        f[1] = p[1] * y[0] * y[2] / p[3] - p[0] * y[1];
        f[2] = p[0] * y[1] - p[2] * y[2];
        f[3] = p[2] * y[2];
      }
      break;
    case kThreeBody:
      if constexpr (d == 4) {
        const double mu = p[0];
        const double mp = 1.0 - p[0];
        T r1 = dsqrt((y[0] + mu) * (y[0] + mu) + y[1] * y[1]);
        T r2 = dsqrt((y[0] - mp) * (y[0] - mp) + y[1] * y[1]);
        T D1 = r1 * r1 * r1;
        T D2 = r2 * r2 * r2;
        f[0] = y[2];
        f[1] = y[3];
        f[2] = y[0] + 2.0 * y[3] - mp * (y[0] + mu) / D1 - mu * (y[0] - mp) / D2;
        f[3] = y[1] - 2.0 * y[2] - mp * y[1] / D1 - mu * y[1] / D2;
      }
      break;
    case kHenonHeiles:
      if constexpr (d == 4) {
        f[0] = y[2];
        f[1] = y[3];
        f[2] = -y[0] - 2.0 * p[0] * y[0] * y[1];
        f[3] = -y[1] - p[0] * (y[0] * y[0] - y[1] * y[1]);
      }
      break;
    case kLorenz63:
      if constexpr (d == 3) {
        f[0] = p[0] * (y[1] - y[0]);
        f[1] = y[0] * (p[1] - y[2]) - y[1];
        f[2] = y[0] * y[1] - p[2] * y[2];
      }
      break;
    case kLorenz96:
      if constexpr (d >= 4) {
#pragma unroll
        for (int i = 0; i < d; ++i) {
          const int ip1 = (i + 1) % d, im1 = (i + d - 1) % d, im2 = (i + d - 2) % d;
          f[i] = (y[ip1] - y[im2]) * y[im1] - y[i] + p[0];
        }
      }
      break;
    default:
      break;
  }
}

// f(y) and its Jacobian J[i][j] = df_i/dy_j by one forward-mode pass.
template <int d>
__host__ __device__ inline void field_and_jacobian(int kind, const double* p, const double* y, double* f,
                                                   double (*J)[d]) {
  Dual<d> yy[d];
  Dual<d> ff[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    yy[i].v = y[i];
#pragma unroll
    for (int j = 0; j < d; ++j) yy[i].g[j] = (i == j) ? 1.0 : 0.0;
  }
#pragma unroll
  for (int i = 0; i < d; ++i) {
    ff[i].v = 0.0;
#pragma unroll
    for (int j = 0; j < d; ++j) ff[i].g[j] = 0.0;
  }
  vector_field<Dual<d>, d>(kind, p, yy, ff);
#pragma unroll
  for (int i = 0; i < d; ++i) {
    f[i] = ff[i].v;
#pragma unroll
    for (int j = 0; j < d; ++j) J[i][j] = ff[i].g[j];
  }
}

}  // namespace pof
