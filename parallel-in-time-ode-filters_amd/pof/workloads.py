"""Seeded synthetic IVP workloads for the batched and served benchmarks.

``mixed_ivps`` draws the mixed batch: problem uniform over
{logistic (r ~ U[0.5, 3], K = 1), FitzHugh-Nagumo classic (a, b ~ U[0.1, 0.3],
c ~ U[2, 4]), Lorenz-63 (rho ~ U[20, 30])}, y0 perturbed by N(0, 0.1^2),
Ns log-uniform in [ns_min, ns_max], order q uniform in {1, 2, 3, 4}, time
spans logistic [0, 10], FitzHugh-Nagumo [0, 100], Lorenz-63 [0, 10].
The logistic initial value is taken in absolute value after the
perturbation (a negative y0 blows up in finite time).

``poisson_arrivals`` draws the request times of the served benchmark.
"""
import numpy as np

from . import ivp as _ivp


def mixed_ivps(n=2048, seed=4, ns_min=256, ns_max=16384):
    """List of ``(f, y0, ts, order, "constant")`` tuples (the ``pof.batch.Problem`` fields)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = int(rng.integers(0, 3))
        if kind == 0:
            p = _ivp.logistic(r=float(rng.uniform(0.5, 3.0)), K=1.0)
            y0 = np.abs(np.array([0.01]) + 0.1 * rng.standard_normal(1))
            t0, t1 = 0.0, 10.0
        elif kind == 1:
            p = _ivp.fhn_classic(a=float(rng.uniform(0.1, 0.3)), b=float(rng.uniform(0.1, 0.3)),
                                 c=float(rng.uniform(2.0, 4.0)))
            y0 = np.array([1.0, 1.0]) + 0.1 * rng.standard_normal(2)
            t0, t1 = 0.0, 100.0
        else:
            p = _ivp.lorenz63(rho=float(rng.uniform(20.0, 30.0)))
            y0 = np.array([1.0, 1.0, 1.0]) + 0.1 * rng.standard_normal(3)
            t0, t1 = 0.0, 10.0
        ns = int(round(np.exp(rng.uniform(np.log(ns_min), np.log(ns_max)))))
        q = int(rng.integers(1, 5))
        out.append((p.f, y0, np.linspace(t0, t1, ns + 1), q, "constant"))
    return out


def poisson_arrivals(n, rate, seed=5):
    """Arrival times (seconds from 0) of n requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def max_rel_err(yhat, y):
    """``max_{k,i} |yhat - y| / max(|y|, 1e-8 max_k |y_{., i}|)`` and the normwise ``max|yhat - y| / max|y|``."""
    yhat = np.asarray(yhat, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    floor = 1e-8 * np.abs(y).max(axis=0, keepdims=True)
    elem = (np.abs(yhat - y) / np.maximum(np.abs(y), np.maximum(floor, 1e-300))).max()
    norm = np.abs(yhat - y).max() / max(np.abs(y).max(), 1e-300)
    return float(elem), float(norm)
