// Alternative linearisations of the ODE residual r(x) = E1 P x - f(E0 P x),
// one thread per state (pof/observations.py:35-89,
// pof/linearization/unscented.py:19-70, pof/utils.py:115-185).  These are
// per-state, sequential and small (D <= 20), so each thread works through its
// state serially, with its matrices in a per-state slice of a work buffer.
#pragma once

#include "pof_core.hpp"
#include "pof_problems.hpp"
#include "pof_types.hpp"

namespace pof {

// doubles of work buffer per state
template <int D, int d>
__host__ __device__ constexpr int lin_work() {
  return lin_work_host(D, d);
}

// r(x) in preconditioned coordinates (convenience.py:28): r_i = P_{i q1 + 1} x_{i q1 + 1} - f_i(P_{l q1} x_{l q1})
template <int D, int d>
__device__ inline void lin_residual(const ModelArgs& a, const double* x, double* r) {
  constexpr int q1 = D / d;
  double y[d], f[d];
#pragma unroll
  for (int i = 0; i < d; ++i) y[i] = a.Pd[i * q1] * x[i * q1];
  vector_field<double, d>(a.kind, a.params, y, f);
#pragma unroll
  for (int i = 0; i < d; ++i) r[i] = a.Pd[i * q1 + 1] * x[i * q1 + 1] - f[i];
}

// Householder LQ of the nr x nc row-major matrix at `a` (leading dimension lda), in place:
// afterwards its first min(nr, nc) columns hold tria(A) (utils.py:33-41), LAPACK dlarfg's
// reflector convention as in the kernels' lq().
__device__ inline void lin_tria(double* a, int nr, int nc, int lda) {
  const int K = nr < nc ? nr : nc;
  for (int i = 0; i < K && i < nc - 1; ++i) {
    double* pr = a + (size_t)i * lda;
    const double alpha = pr[i];
    double sig = 0.0;
    for (int j = i + 1; j < nc; ++j) sig = fma(pr[j], pr[j], sig);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (sig != 0.0) {
      const double nrm = sqrt(fma(alpha, alpha, sig));
      beta = -copysign(nrm, alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int j = i + 1; j < nc; ++j) pr[j] *= scal;  // v (v_i = 1 implied)
    for (int r = i + 1; r < nr; ++r) {
      double* rr = a + (size_t)r * lda;
      double w = rr[i];
      for (int j = i + 1; j < nc; ++j) w = fma(rr[j], pr[j], w);
      const double tw = tau * w;
      rr[i] -= tw;
      for (int j = i + 1; j < nc; ++j) rr[j] = fma(-tw, pr[j], rr[j]);
    }
    pr[i] = beta;
    for (int j = i + 1; j < nc; ++j) pr[j] = 0.0;
  }
  for (int r = 0; r < nr; ++r)
    for (int c = r + 1; c < K; ++c) a[(size_t)r * lda + c] = 0.0;
}

// One rank-one update L L^T + mult v v^T of the n x n lower factor L (row-major, in
// place), utils.py:124-175: Algorithm 3.1 as one loop over columns, then the strict
// upper part zeroed and non-finite entries replaced by 0.  `nc` holds n * n doubles
// of scratch, `om` n doubles.
__device__ inline void lin_chol_update(double* L, int n, const double* v, double mult, double* nc, double* om) {
  for (int i = 0; i < n; ++i) om[i] = v[i];
  double b = 1.0;
  for (int i = 0; i < n; ++i) {
    const double dm = L[(size_t)i * n + i];
    const double w = om[i];
    const double nd = sqrt(dm * dm + mult / b * (w * w));
    const double gamma = dm * dm * b + mult * (w * w);
    for (int r = 0; r < n; ++r) om[r] = om[r] - (w / dm) * L[(size_t)r * n + i];
    const double k = mult * w / gamma;
    for (int r = 0; r < n; ++r) nc[(size_t)r * n + i] = nd * (L[(size_t)r * n + i] / dm + k * om[r]);
    nc[(size_t)i * n + i] = nd;
    const double t = w / dm;
    b = b + mult * (t * t);
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      const double x = (c > r) ? 0.0 : nc[(size_t)r * n + c];
      L[(size_t)r * n + c] = isfinite(x) ? x : 0.0;
    }
}

// Linearise at state k: mean m[k], factor L[k] (may be null unless uncertain /
// unscented).  Outputs H [dO, D], c [dO], R [dO, dO] with dO = d, or d + D for the
// regularised model.
template <int D, int d>
__global__ void k_linearize_alt(ModelArgs a, int method, LinParams p, const double* m, const double* Lx, int n,
                                double* Hout, double* cout, double* Rout, double* work) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  constexpr int q1 = D / d;
  const double* mk = m + (size_t)k * D;
  const double* Lk = Lx ? Lx + (size_t)k * D * D : nullptr;
  const int dO = (method == kLinRegularized) ? d + D : d;
  double* H = Hout + (size_t)k * dO * D;
  double* c = cout + (size_t)k * dO;
  double* R = Rout + (size_t)k * dO * dO;
  double* w = work ? work + (size_t)k * lin_work<D, d>() : nullptr;
  for (int i = 0; i < dO * dO; ++i) R[i] = 0.0;
  double res[d];
  lin_residual<D, d>(a, mk, res);
  if (method == kLinEK0) {
    // H = E1 of the IWP, not preconditioned (observations.py:52-61)
    for (int i = 0; i < d; ++i) {
      for (int j = 0; j < D; ++j) H[i * D + j] = (j == i * q1 + 1) ? 1.0 : 0.0;
      c[i] = res[i] - mk[i * q1 + 1];
    }
    return;
  }
  if (method == kLinUnscented) {
    // unscented.py:19-70 with n = D
    const double nn = (double)D;
    const double kap = p.kappa_set ? p.kappa : 3.0 + nn;
    const double lam = p.alpha * p.alpha * (nn + kap) - nn;
    const double wmi = 1.0 / (2.0 * (nn + lam));
    const double wm0 = lam / (nn + lam);
    const double wc0 = lam / (nn + lam) + (1.0 - p.alpha * p.alpha + p.beta);
    const double sc = sqrt(nn + lam);
    constexpr int NP = 2 * D + 1;
    double* fp = w;                  // [NP, d] residual at the sigma points
    double* sp = fp + NP * d;        // [d, NP] sqrt(wc_i) (f_i - m_f), transposed
    double* psi = sp + NP * d;       // [D, d]
    double* fx = psi + D * d;        // [d, D]
    double* fl = fx + D * d;         // [D, d] (F L)^T
    double* ch = fl + D * d;         // [d, d]
    double* nc = ch + d * d;         // [d, d]
    double* om = nc + d * d;         // [d]
    double* mf = om + d;             // [d]
    double x[D];
    for (int i = 0; i < NP; ++i) {
      // points: m, m + sc L[:, j], m - sc L[:, j]
      const int j = (i == 0) ? 0 : ((i <= D) ? i - 1 : i - 1 - D);
      const double sg = (i == 0) ? 0.0 : ((i <= D) ? 1.0 : -1.0);
#pragma unroll
      for (int r = 0; r < D; ++r) x[r] = mk[r] + sg * (sc * Lk[r * D + j]);
      double fr[d];
      lin_residual<D, d>(a, x, fr);
      for (int c2 = 0; c2 < d; ++c2) fp[i * d + c2] = fr[c2];
    }
    for (int c2 = 0; c2 < d; ++c2) {
      double acc = 0.0;
      for (int i = 0; i < NP; ++i) acc = fma((i == 0) ? wm0 : wmi, fp[i * d + c2], acc);
      mf[c2] = acc;
    }
    // Psi = ((x_i - m)^T * wc) (f_i - m_f): x_0 - m = 0
    for (int r = 0; r < D; ++r)
      for (int c2 = 0; c2 < d; ++c2) {
        double acc = 0.0;
        for (int i = 0; i < NP; ++i) {
          const int j = (i == 0) ? 0 : ((i <= D) ? i - 1 : i - 1 - D);
          const double sg = (i == 0) ? 0.0 : ((i <= D) ? 1.0 : -1.0);
          const double dx = (mk[r] + sg * (sc * Lk[r * D + j])) - mk[r];
          acc = fma(dx * ((i == 0) ? wc0 : wmi), fp[i * d + c2] - mf[c2], acc);
        }
        psi[r * d + c2] = acc;
      }
    // cho_solve((L, lower), Psi): L Y = Psi, then L^T X = Y; F = X^T
    for (int c2 = 0; c2 < d; ++c2) {
      for (int r = 0; r < D; ++r) {
        double v = psi[r * d + c2];
        for (int t = 0; t < r; ++t) v = fma(-Lk[r * D + t], psi[t * d + c2], v);
        psi[r * d + c2] = v / Lk[r * D + r];
      }
      for (int r = D - 1; r >= 0; --r) {
        double v = psi[r * d + c2];
        for (int t = r + 1; t < D; ++t) v = fma(-Lk[t * D + r], psi[t * d + c2], v);
        psi[r * d + c2] = v / Lk[r * D + r];
      }
      for (int r = 0; r < D; ++r) fx[c2 * D + r] = psi[r * d + c2];
    }
    // chol = tria(sqrt(wc)[:, None] (f - m_f))^T ... downdated by the columns of F L
    for (int c2 = 0; c2 < d; ++c2)
      for (int i = 0; i < NP; ++i) sp[c2 * NP + i] = sqrt((i == 0) ? wc0 : wmi) * (fp[i * d + c2] - mf[c2]);
    lin_tria(sp, d, NP, NP);
    for (int r = 0; r < d; ++r)
      for (int c2 = 0; c2 < d; ++c2) ch[r * d + c2] = sp[r * NP + c2];
    for (int j = 0; j < D; ++j)
      for (int c2 = 0; c2 < d; ++c2) {
        double acc = 0.0;
        for (int t = 0; t < D; ++t) acc = fma(fx[c2 * D + t], Lk[t * D + j], acc);
        fl[j * d + c2] = acc;
      }
    for (int j = 0; j < D; ++j) lin_chol_update(ch, d, fl + j * d, -1.0, nc, om);
    for (int i = 0; i < d; ++i) {
      double acc = 0.0;
      for (int j = 0; j < D; ++j) {
        H[i * D + j] = fx[i * D + j];
        acc = fma(fx[i * D + j], mk[j], acc);
      }
      c[i] = mf[i] - acc;
      for (int j = 0; j < d; ++j) R[i * d + j] = ch[i * d + j];
    }
    return;
  }
  // EK1 (observations.py:35-40): H = E1 - J E0 (preconditioned), c = r(m) - H m
  ModelArgs b = a;
  b.lin = mk - D;  // Obs::load(k = 0) reads lin[1]
  b.Hx = nullptr;
  Obs<D, d, true> ob;
  ob.load(b, 0);
  double Hk[d][D];
  ob.dense(Hk);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < D; ++j) H[i * D + j] = Hk[i][j];
  if (method == kLinRegularized) {
    // observations.py:64-89: H = [J; I], b = [r(m); -m], R = blockdiag(0, I / sqrt(lam))
    const double s = 1.0 / sqrt(p.lam);
    for (int i = 0; i < d; ++i) c[i] = res[i];
    for (int i = 0; i < D; ++i) {
      for (int j = 0; j < D; ++j) H[(d + i) * D + j] = (i == j) ? 1.0 : 0.0;
      c[d + i] = -mk[i];
      R[(d + i) * dO + d + i] = s;
    }
    return;
  }
  for (int i = 0; i < d; ++i) c[i] = ob.c[i];
  if (method == kLinUncertain) {
    // observations.py:43-49: R = tria(H L)
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < D; ++j) {
        double acc = 0.0;
        for (int t = 0; t < D; ++t) acc = fma(Hk[i][t], Lk[t * D + j], acc);
        w[i * D + j] = acc;
      }
    lin_tria(w, d, D, D);
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) R[i * d + j] = w[i * D + j];
  }
}

}  // namespace pof
