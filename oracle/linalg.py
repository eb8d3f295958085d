"""Batched small dense linear algebra in float64 (oracle).

Restates ``pof/utils.py:33-41`` (``tria``/``qr``) and the
``jax.scipy.linalg.solve_triangular`` calls used along the hot path.
"""
import numpy as np


def tria(A):
    """Lower-triangular square root: ``tria(A) = qr(A^T, mode="r")^T``.

    Follows ``pof/utils.py:33-35,38-41``.  NumPy's ``qr(mode="r")`` is LAPACK
    ``dgeqrf``, the routine jax's CPU backend calls, so the Householder sign
    convention (``R_ii = -sign(alpha) * norm``) is the reference's.
    Works on stacks ``[..., m, n]`` and returns ``[..., m, min(m, n)]``.
    """
    A = np.asarray(A, dtype=np.float64)
    R = np.linalg.qr(np.swapaxes(A, -1, -2), mode="r")
    return np.swapaxes(R, -1, -2)


def solve_triangular(L, B, lower=True, trans=False):
    """Batched triangular solve, the semantics of ``jax.scipy.linalg.solve_triangular``.

    Solves ``op(L) X = B`` with ``op(L) = L`` or ``L^T`` (``trans=True``).
    ``L``: ``[..., n, n]``; ``B``: ``[..., n]`` or ``[..., n, k]``.  Only the
    triangle named by ``lower`` is read.  Plain forward/back substitution.
    """
    L = np.asarray(L, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    vec = B.ndim == L.ndim - 1
    if vec:
        B = B[..., None]
    n = L.shape[-1]
    M = np.swapaxes(L, -1, -2) if trans else L
    # op(L) is lower iff (lower and not trans) or (not lower and trans)
    eff_lower = lower != trans
    if eff_lower:
        M = np.tril(M)
    else:
        M = np.triu(M)
    X = np.zeros(np.broadcast_shapes(M.shape[:-2], B.shape[:-2]) + B.shape[-2:])
    order = range(n) if eff_lower else range(n - 1, -1, -1)
    for i in order:
        if eff_lower:
            acc = np.einsum("...j,...jk->...k", M[..., i, :i], X[..., :i, :])
        else:
            acc = np.einsum("...j,...jk->...k", M[..., i, i + 1 :], X[..., i + 1 :, :])
        X[..., i, :] = (B[..., i, :] - acc) / M[..., i, i][..., None]
    return X[..., 0] if vec else X


def mv(A, x):
    """Batched matrix-vector product ``A @ x``."""
    return np.einsum("...ij,...j->...i", A, x)


def mT(A):
    """Swap the last two axes."""
    return np.swapaxes(A, -1, -2)
