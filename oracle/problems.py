"""ODE vector fields, forward-mode derivatives and Taylor-mode initialisation (oracle).

The vector fields restate ``pof/ivp.py`` (logistic 7-14, lotkavolterra 17-31,
vanderpol 34-41, fitzhughnagumo 44-60, rober 63-79, rigid_body 82-90,
seir 93-108, threebody 111-126, henonheiles 137-152).  Three fields the
reference lacks are added with their textbook definitions: the classic
FitzHugh-Nagumo form, Lorenz-63 and Lorenz-96.

Each field is written once over the components ``y[0..d-1]`` with plain
arithmetic, so the same code evaluates on float arrays, on dual numbers
(``jax.jacfwd``, used by ``pof/observations.py:38``) and on truncated power
series (``tornadox.init.TaylorMode``, used by ``pof/initialization.py:15-22``).
"""
import math

import numpy as np


# --------------------------------------------------------------------------
# truncated power series (Taylor mode) and dual numbers (forward mode)
# --------------------------------------------------------------------------
class Jet:
    """Truncated power series ``sum_k c[k] t^k`` (coefficients may be arrays)."""

    __array_priority__ = 100

    def __init__(self, coeffs):
        self.c = list(coeffs)

    @property
    def order(self):
        return len(self.c) - 1

    @staticmethod
    def _lift(x, K):
        if isinstance(x, Jet):
            return x.c
        return [x] + [0.0 * x] * K

    def __add__(self, o):
        oc = Jet._lift(o, self.order)
        return Jet([a + b for a, b in zip(self.c, oc)])

    __radd__ = __add__

    def __neg__(self):
        return Jet([-a for a in self.c])

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if not isinstance(o, Jet):
            return Jet([a * o for a in self.c])
        K = self.order
        return Jet([sum(self.c[j] * o.c[k - j] for j in range(k + 1)) for k in range(K + 1)])

    __rmul__ = __mul__

    def __truediv__(self, o):
        if not isinstance(o, Jet):
            return Jet([a / o for a in self.c])
        K = self.order
        out = []
        for k in range(K + 1):
            acc = self.c[k] - sum(out[j] * o.c[k - j] for j in range(k))
            out.append(acc / o.c[0])
        return Jet(out)

    def __rtruediv__(self, o):
        return Jet(Jet._lift(o, self.order)) / self

    def __pow__(self, n):
        assert isinstance(n, int) and n >= 1
        out = self
        for _ in range(n - 1):
            out = out * self
        return out

    def sqrt(self):
        K = self.order
        s = [np.sqrt(self.c[0])]
        for k in range(1, K + 1):
            acc = self.c[k] - sum(s[j] * s[k - j] for j in range(1, k))
            s.append(acc / (2.0 * s[0]))
        return Jet(s)


class Dual:
    """Dual number ``v + g . eps`` with a gradient over ``n`` seeds: g has shape ``[..., n]``."""

    __array_priority__ = 100

    def __init__(self, v, g):
        self.v = v
        self.g = g

    def _lift(self, o):
        if isinstance(o, Dual):
            return o
        return Dual(o, np.zeros_like(self.g))

    def __add__(self, o):
        o = self._lift(o)
        return Dual(self.v + o.v, self.g + o.g)

    __radd__ = __add__

    def __neg__(self):
        return Dual(-self.v, -self.g)

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v * o, self.g * np.asarray(o)[..., None])
        return Dual(self.v * o.v, self.g * np.asarray(o.v)[..., None] + o.g * np.asarray(self.v)[..., None])

    __rmul__ = __mul__

    def __truediv__(self, o):
        if not isinstance(o, Dual):
            return Dual(self.v / o, self.g / np.asarray(o)[..., None])
        v = self.v / o.v
        return Dual(v, (self.g - o.g * np.asarray(v)[..., None]) / np.asarray(o.v)[..., None])

    def __rtruediv__(self, o):
        return self._lift(o) / self

    def __pow__(self, n):
        assert isinstance(n, int) and n >= 1
        out = self
        for _ in range(n - 1):
            out = out * self
        return out

    def sqrt(self):
        v = np.sqrt(self.v)
        return Dual(v, self.g / (2.0 * np.asarray(v)[..., None]))


def _sqrt(x):
    if isinstance(x, (Jet, Dual)):
        return x.sqrt()
    return np.sqrt(x)


# --------------------------------------------------------------------------
# vector fields f(y, p) over components; autonomous (called as f(None, y))
# --------------------------------------------------------------------------
def f_logistic(y, p):
    # pof/ivp.py:10-12  f = 1.0 * y * (1 - y); here r * y * (1 - y / K), (r, K) = (1, 1)
    r, K = p
    return [r * y[0] * (1 - y[0] / K)]


def f_lotkavolterra(y, p):
    # pof/ivp.py:21-29
    a, b, c, d = p
    return [a * y[0] - b * y[0] * y[1], -c * y[1] + d * y[0] * y[1]]


def f_vanderpol(y, p):
    # pof/ivp.py:37-39
    (mu,) = p
    return [y[1], mu * ((1.0 - y[0] ** 2) * y[1] - y[0])]


def f_fitzhughnagumo(y, p):
    # pof/ivp.py:48-58
    a, b, tinv, l = p
    v, w = y[0], y[1]
    return [v - (v**3) / 3 - w + l, tinv * (v + a - b * w)]


def f_fhn_classic(y, p):
    # classic form: y1' = c (y1 - y1^3/3 + y2), y2' = -(y1 - a + b y2) / c
    a, b, c = p
    return [c * (y[0] - (y[0] ** 3) / 3 + y[1]), -(y[0] - a + b * y[1]) / c]


def f_rober(y, p):
    # pof/ivp.py:67-77
    k1, k2, k3 = p
    y1, y2, y3 = y[0], y[1], y[2]
    return [-k1 * y1 + k3 * y2 * y3, k1 * y1 - k2 * y2**2 - k3 * y2 * y3, k2 * y2**2]


def f_rigid_body(y, p):
    # pof/ivp.py:86-88
    return [p[0] * y[1] * y[2], p[1] * y[0] * y[2], p[2] * y[0] * y[1]]


def f_seir(y, p):
    # pof/ivp.py:97-106
    return [
        -p[1] * y[0] * y[2] / p[3],
        p[1] * y[0] * y[2] / p[3] - p[0] * y[1],
        p[0] * y[1] - p[2] * y[2],
        p[2] * y[2],
    ]


def f_threebody(y, p):
    # pof/ivp.py:117-124 (norm written as sqrt of the sum of squares)
    mu = p[0]
    mp = 1.0 - p[0]
    D1 = _sqrt((y[0] + mu) ** 2 + y[1] ** 2) ** 3
    D2 = _sqrt((y[0] - mp) ** 2 + y[1] ** 2) ** 3
    du0p = y[0] + 2 * y[3] - mp * (y[0] + mu) / D1 - mu * (y[0] - mp) / D2
    du1p = y[1] - 2 * y[2] - mp * y[1] / D1 - mu * y[1] / D2
    return [y[2], y[3], du0p, du1p]


def f_henonheiles(y, p):
    # pof/ivp.py:141-150
    (pp,) = p
    return [y[2], y[3], -y[0] - 2 * pp * y[0] * y[1], -y[1] - pp * (y[0] ** 2 - y[1] ** 2)]


def f_lorenz63(y, p):
    sigma, rho, beta = p
    return [sigma * (y[1] - y[0]), y[0] * (rho - y[2]) - y[1], y[0] * y[1] - beta * y[2]]


def f_lorenz96(y, p):
    (F,) = p
    d = len(y)
    return [(y[(i + 1) % d] - y[(i - 2) % d]) * y[(i - 1) % d] - y[i] + F for i in range(d)]


FIELDS = {
    "logistic": f_logistic,
    "lotkavolterra": f_lotkavolterra,
    "vanderpol": f_vanderpol,
    "fitzhughnagumo": f_fitzhughnagumo,
    "fhn_classic": f_fhn_classic,
    "rober": f_rober,
    "rigid_body": f_rigid_body,
    "seir": f_seir,
    "threebody": f_threebody,
    "henonheiles": f_henonheiles,
    "lorenz63": f_lorenz63,
    "lorenz96": f_lorenz96,
}


class Problem:
    """An autonomous IVP ``y' = f(y)``: name, parameters, y0 and the time span."""

    def __init__(self, name, params, y0, t0, tmax):
        self.name = name
        self.params = tuple(float(x) for x in params)
        self.y0 = np.asarray(y0, dtype=np.float64)
        self.t0 = float(t0)
        self.tmax = float(tmax)
        self.d = self.y0.shape[0]

    @property
    def t_span(self):
        return (self.t0, self.tmax)

    def fc(self, comps):
        return FIELDS[self.name](comps, self.params)

    def f(self, t, y):
        """Evaluate on ``y[..., d]`` (the reference's ``f(None, y)`` call form)."""
        y = np.asarray(y, dtype=np.float64)
        out = self.fc([y[..., i] for i in range(self.d)])
        return np.stack([np.broadcast_to(np.asarray(o, dtype=np.float64), y.shape[:-1]) for o in out], axis=-1)

    def jac(self, y):
        """Jacobian ``[..., d, d]`` by forward-mode dual numbers (``jax.jacfwd``)."""
        y = np.asarray(y, dtype=np.float64)
        d = self.d
        comps = []
        for i in range(d):
            g = np.zeros(y.shape[:-1] + (d,))
            g[..., i] = 1.0
            comps.append(Dual(y[..., i], g))
        out = self.fc(comps)
        rows = []
        for o in out:
            if isinstance(o, Dual):
                rows.append(np.broadcast_to(o.g, y.shape[:-1] + (d,)))
            else:
                rows.append(np.zeros(y.shape[:-1] + (d,)))
        return np.stack(rows, axis=-2)


def taylor_mode(problem, q):
    """Derivatives ``y^(k)(t0)``, k = 0..q, shape ``[q+1, d]`` (``tornadox.init.TaylorMode``).

    Restates the Taylor-mode recursion ``(k+1) a_{k+1} = [f(y(t))]_k`` on
    truncated power series; returns ``a_k * k!`` (derivatives, the quantities
    the Nordsieck preconditioner of ``pof/transitions.py:53-68`` scales).
    """
    d = problem.d
    coeffs = [[float(problem.y0[i])] for i in range(d)]
    for k in range(q):
        jets = [Jet(coeffs[i]) for i in range(d)]
        fy = problem.fc(jets)
        for i in range(d):
            o = fy[i]
            ck = o.c[k] if isinstance(o, Jet) else (float(o) if k == 0 else 0.0)
            coeffs[i].append(float(ck) / (k + 1))
    out = np.zeros((q + 1, d))
    for k in range(q + 1):
        for i in range(d):
            out[k, i] = coeffs[i][k] * math.factorial(k)
    return out


# --------------------------------------------------------------------------
# problem factories (defaults as in pof/ivp.py)
# --------------------------------------------------------------------------
def logistic(t0=0.0, tmax=10.0, y0=None, r=1.0, K=1.0):
    return Problem("logistic", (r, K), y0 if y0 is not None else [0.01], t0, tmax)


def lotkavolterra(t0=0.0, tmax=7.0, y0=None, p=None):
    return Problem("lotkavolterra", p or (1.5, 1.0, 3.0, 1.0), y0 if y0 is not None else [1.0, 1.0], t0, tmax)


def vanderpol(t0=0.0, tmax=6.3, y0=None, stiffness_constant=1e1):
    return Problem("vanderpol", (stiffness_constant,), y0 if y0 is not None else [2.0, 0.0], t0, tmax)


def fitzhughnagumo(t0=0.0, tmax=100.0, y0=None, p=None):
    return Problem("fitzhughnagumo", p or (0.7, 0.8, 1 / 12.5, 0.5), y0 if y0 is not None else [1.0, 1.0], t0, tmax)


def fhn_classic(t0=0.0, tmax=20.0, y0=None, a=0.2, b=0.2, c=3.0):
    return Problem("fhn_classic", (a, b, c), y0 if y0 is not None else [-1.0, 1.0], t0, tmax)


def rober(t0=0.0, tmax=1e11, y0=None, p=None):
    return Problem("rober", p or (0.04, 3e7, 1e4), y0 if y0 is not None else [1.0, 0.0, 0.0], t0, tmax)


def rigid_body(t0=0.0, tmax=20.0, y0=None, p=None):
    return Problem("rigid_body", p or (-2.0, 1.25, -0.5), y0 if y0 is not None else [1.0, 0.0, 0.9], t0, tmax)


def seir(t0=0.0, tmax=200.0, y0=None, p=None):
    y0 = np.asarray(y0 if y0 is not None else [998.0, 1.0, 1.0, 1.0])
    return Problem("seir", p or (0.3, 0.3, 0.1, float(y0.sum())), y0, t0, tmax)


def threebody(t0=0.0, tmax=17.0652165601579625588917206249, y0=None, p=None):
    y0 = y0 if y0 is not None else [0.994, 0.0, 0.0, -2.00158510637908252240537862224]
    return Problem("threebody", p or (0.012277471,), y0, t0, tmax)


def henonheiles(t0=0.0, tmax=100.0, y0=None, p=None):
    return Problem("henonheiles", (p or 1.0,), y0 if y0 is not None else [0.5, 0.0, 0.0, 0.1], t0, tmax)


def lorenz63(t0=0.0, tmax=10.0, y0=None, sigma=10.0, rho=28.0, beta=8.0 / 3.0):
    return Problem("lorenz63", (sigma, rho, beta), y0 if y0 is not None else [1.0, 1.0, 1.0], t0, tmax)


def lorenz96(t0=0.0, tmax=5.0, y0=None, d=32, F=8.0, seed=0):
    if y0 is None:
        y0 = F + 0.01 * np.random.default_rng(seed).standard_normal(d)
    return Problem("lorenz96", (F,), y0, t0, tmax)
