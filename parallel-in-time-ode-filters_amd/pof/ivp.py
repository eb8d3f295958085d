"""Initial value problems whose vector fields run on the GPU (reference ``pof/ivp.py``).

Each factory returns an :class:`IVP` whose ``f`` is a :class:`VectorField`:
a device-registered field (kind + parameters) that the HIP kernels evaluate,
together with its Jacobian, inside the IEKS.  ``f(None, y)`` also evaluates
the field on the host through the same native code, for convenience.
Defaults follow the reference factories (``pof/ivp.py``: logistic 7-14,
lotkavolterra 17-31, vanderpol 34-41, fitzhughnagumo 44-60, rober 63-79,
rigid_body 82-90, seir 93-108, threebody 111-126, henonheiles 137-152);
``fhn_classic``, ``lorenz63`` and ``lorenz96`` are added.
"""
import ctypes
from typing import NamedTuple, Tuple

import numpy as np

from . import _lib


class VectorField:
    """A vector field ``f(t, y)`` the GPU kernels can evaluate (autonomous)."""

    def __init__(self, name, params, d):
        if name not in _lib.KINDS:
            raise ValueError(f"unknown vector field {name!r}")
        self.name = name
        self.kind = _lib.KINDS[name]
        self.params = tuple(float(p) for p in params)
        if len(self.params) > _lib.MAX_PARAMS:
            raise ValueError("too many parameters")
        self.d = int(d)

    def params_array(self):
        p = np.zeros(_lib.MAX_PARAMS)
        p[: len(self.params)] = self.params
        return p

    def __call__(self, t, y):
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float64))
        if y.shape[-1] != self.d:
            raise ValueError(f"{self.name} expects y of dimension {self.d}")
        lib = _lib.load()
        p = self.params_array()
        flat = y.reshape(-1, self.d)
        out = np.empty_like(flat)
        for i in range(flat.shape[0]):
            yi = np.ascontiguousarray(flat[i])
            oi = np.empty(self.d)
            _lib.check(lib.pof_eval_field(self.kind, self.d, _lib.hptr(p), _lib.hptr(yi), _lib.hptr(oi)),
                       what="pof_eval_field")
            out[i] = oi
        return out.reshape(y.shape)

    def __repr__(self):
        return f"VectorField({self.name}, params={self.params}, d={self.d})"


class IVP(NamedTuple):
    """Same fields as ``tornadox.ivp.InitialValueProblem`` (f, t0, tmax, y0)."""

    f: VectorField
    t0: float
    tmax: float
    y0: np.ndarray

    @property
    def t_span(self) -> Tuple[float, float]:
        return (self.t0, self.tmax)

    @property
    def dimension(self):
        return self.y0.shape[0]


def _ivp(name, params, y0, t0, tmax):
    y0 = np.asarray(y0, dtype=np.float64)
    return IVP(VectorField(name, params, y0.shape[0]), float(t0), float(tmax), y0)


def logistic(t0=0, tmax=10.0, y0=None, r=1.0, K=1.0):
    return _ivp("logistic", (r, K), [0.01] if y0 is None else y0, t0, tmax)


def lotkavolterra(t0=0.0, tmax=7, y0=None, p=None):
    return _ivp("lotkavolterra", p or (1.5, 1.0, 3.0, 1.0), [1.0, 1.0] if y0 is None else y0, t0, tmax)


def vanderpol(t0=0.0, tmax=6.3, y0=None, stiffness_constant=1e1):
    return _ivp("vanderpol", (stiffness_constant,), [2.0, 0.0] if y0 is None else y0, t0, tmax)


def fitzhughnagumo(t0=0.0, tmax=100.0, y0=None, p=None):
    return _ivp("fitzhughnagumo", p or (0.7, 0.8, 1 / 12.5, 0.5), [1.0, 1.0] if y0 is None else y0, t0, tmax)


def fhn_classic(t0=0.0, tmax=20.0, y0=None, a=0.2, b=0.2, c=3.0):
    """Classic FitzHugh-Nagumo: y1' = c (y1 - y1^3/3 + y2), y2' = -(y1 - a + b y2) / c."""
    return _ivp("fhn_classic", (a, b, c), [-1.0, 1.0] if y0 is None else y0, t0, tmax)


def rober(t0=0.0, tmax=1e11, y0=None, p=None):
    return _ivp("rober", p or (0.04, 3e7, 1e4), [1.0, 0.0, 0.0] if y0 is None else y0, t0, tmax)


def rigid_body(t0=0.0, tmax=20.0, y0=None, p=None):
    return _ivp("rigid_body", p or (-2.0, 1.25, -0.5), [1.0, 0.0, 0.9] if y0 is None else y0, t0, tmax)


def seir(t0=0.0, tmax=200.0, y0=None, p=None):
    y0 = np.asarray([998.0, 1.0, 1.0, 1.0] if y0 is None else y0, dtype=np.float64)
    return _ivp("seir", p or (0.3, 0.3, 0.1, float(y0.sum())), y0, t0, tmax)


def threebody(t0=0.0, tmax=17.0652165601579625588917206249, y0=None, p=None):
    y0 = [0.994, 0.0, 0.0, -2.00158510637908252240537862224] if y0 is None else y0
    return _ivp("threebody", p or (0.012277471,), y0, t0, tmax)


def henonheiles(t0=0.0, tmax=100.0, y0=None, p=None):
    return _ivp("henonheiles", (p or 1.0,), [0.5, 0.0, 0.0, 0.1] if y0 is None else y0, t0, tmax)


def lorenz63(t0=0.0, tmax=10.0, y0=None, sigma=10.0, rho=28.0, beta=8.0 / 3.0):
    return _ivp("lorenz63", (sigma, rho, beta), [1.0, 1.0, 1.0] if y0 is None else y0, t0, tmax)


def lorenz96(t0=0.0, tmax=5.0, y0=None, d=4, F=8.0, seed=0):
    if y0 is None:
        y0 = F + 0.01 * np.random.default_rng(seed).standard_normal(d)
    return _ivp("lorenz96", (F,), y0, t0, tmax)
