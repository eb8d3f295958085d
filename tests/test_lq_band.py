"""The banded second block of the large-state engine's LQ (CPU).

Every LQ of that engine factors ``[B1 | B2]`` whose second block (from column
``b1``) has row ``r`` zero beyond column ``b1 + max(r, w2)``: ``[I; 0]`` in the
filter combine's ``Xi`` (w2 = 0), lower-triangular ``U``, ``Z``, ``D`` and
``[L_Q; 0]`` factors (w2 = 0), and a leaf's ``Z = [Zr | 0]`` with ``d`` nonzero
columns (w2 = d - 1).  Then the Householder reflector of every pivot ``i`` of
the first block is exactly zero beyond column ``b1 + max(i, w2)``, and a panel
of 16 first-block pivots can end, with its trailing update, at column
``b1 + max(last pivot, w2) + 1`` (the column end ``Ops::lq`` uses) with the
same factor.

Checked with an explicit Householder LQ (LAPACK dlarfg's reflectors: beta =
-sign(alpha) norm, tau = (beta - alpha) / beta, v = [1, x / (alpha - beta)],
no reflector for a zero tail) in panels of 16 pivots, pinned against the
oracle's ``tria`` (numpy's dgeqrf).
"""
import numpy as np
import pytest

from oracle.linalg import tria

NB = 16  # pivots per panel, as in the engine


def householder_lq(A, b1=0, w2=0, banded=False):
    """L of A (m x n) and the reflectors (pivot, first column, v) by panels of NB pivots."""
    A = np.array(A, dtype=np.float64)
    m, n = A.shape
    kmax = min(m, n - 1)  # a length-1 tail gets no reflector
    refl = []
    for i0 in range(0, kmax, NB):
        npv = min(NB, kmax - i0)
        ce = n
        if banded and b1 > 0 and i0 + npv <= b1:
            ce = min(n, b1 + max(i0 + npv - 1, w2) + 1)
        for i in range(i0, i0 + npv):
            x = A[i, i:ce].copy()
            alpha, sig = x[0], float(np.dot(x[1:], x[1:]))
            v = np.zeros_like(x)
            v[0] = 1.0
            if sig == 0.0:
                tau, beta = 0.0, alpha
            else:
                nrm = np.sqrt(alpha * alpha + sig)
                beta = -np.copysign(nrm, alpha)
                tau = (beta - alpha) / beta
                v[1:] = x[1:] / (alpha - beta)
            refl.append((i, v))
            w = A[i + 1:, i:ce] @ v
            A[i + 1:, i:ce] -= tau * np.outer(w, v)
            A[i, i] = beta
            A[i, i + 1:ce] = 0.0
    k = min(m, n)
    return np.tril(A[:, :k]), refl


def _xi(rng, D):
    """[[R1, I], [R2, 0]]: the filter combine's Xi (b1 = D, w2 = 0)."""
    top = np.hstack([rng.standard_normal((D, D)), np.eye(D)])
    bot = np.hstack([rng.standard_normal((D, D)), np.zeros((D, D))])
    return np.vstack([top, bot]), D, 0


def _lower(rng, D):
    """[M | L] with L lower triangular: the combines' [M^T | U2] (b1 = D, w2 = 0)."""
    return np.hstack([rng.standard_normal((D, D)), np.tril(rng.standard_normal((D, D)))]), D, 0


def _leaf_z(rng, D, d):
    """[M | [Zr | 0]] with d nonzero columns: [A1^T Xi22 | Z1] at the leaves (b1 = D, w2 = d - 1)."""
    z = np.zeros((D, D))
    z[:, :d] = rng.standard_normal((D, d))
    return np.hstack([rng.standard_normal((D, D)), z]), D, d - 1


CASES = {
    "xi": _xi,
    "lower": _lower,
    "leaf_z_d5": lambda r, D: _leaf_z(r, D, 5),  # the band follows the pivots (w2 < 16)
    "leaf_z_d24": lambda r, D: _leaf_z(r, D, 24),  # the first panels end at w2 = 23
}


@pytest.mark.parametrize("case,D", [("xi", 20), ("xi", 40), ("lower", 20), ("lower", 40), ("leaf_z_d5", 20),
                                    ("leaf_z_d5", 40), ("leaf_z_d24", 40)])
def test_reflectors_end_at_the_band_and_the_panels_can_stop_there(case, D):
    rng = np.random.default_rng(D + len(case))
    A, b1, w2 = CASES[case](rng, D)
    L_full, refl = householder_lq(A)
    # the explicit factor is the oracle's tria (same reflector convention, other rounding)
    T = tria(A)
    assert np.abs(L_full - T).max() <= 1e-12 * np.abs(T).max()
    # every first-block reflector is exactly zero beyond column b1 + max(i, w2)
    for i, v in refl:
        if i < b1:
            end = b1 + max(i, w2) + 1  # exclusive, absolute column
            assert np.all(v[end - i:] == 0.0), (i, end)
    # ending each first-block panel at its band gives the same factor
    L_band, _ = householder_lq(A, b1, w2, banded=True)
    assert np.abs(L_band - L_full).max() <= 1e-13 * np.abs(L_full).max()


def test_a_band_too_narrow_is_detected():
    """Declaring [Zr | 0] with d = 24 nonzero columns as lower triangular (w2 = 0 instead of 23)
    cuts columns the first panel's reflectors reach, and the factor changes."""
    rng = np.random.default_rng(3)
    A, b1, w2 = _leaf_z(rng, 40, 24)
    L_full, _ = householder_lq(A)
    L_bad, _ = householder_lq(A, b1, 0, banded=True)
    assert np.abs(L_bad - L_full).max() > 1e-6 * np.abs(L_full).max()
