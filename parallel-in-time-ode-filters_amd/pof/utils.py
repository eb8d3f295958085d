"""Small shared types (pof/utils.py of the reference)."""
from typing import Any, NamedTuple

import numpy as np


class MVNSqrt(NamedTuple):
    """Gaussian in square-root form ``N(mean, chol chol^T)`` (reference ``pof/utils.py:9-11``)."""

    mean: Any
    chol: Any


def as_f64_host(x):
    """Host float64 contiguous copy of a list / numpy array / torch tensor."""
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return np.ascontiguousarray(x.detach().to("cpu", torch.float64).numpy())
    except ImportError:  # pragma: no cover
        pass
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def as_f64_device(x, device=None):
    """float64 contiguous CUDA tensor view/copy of ``x``."""
    import torch

    if isinstance(x, torch.Tensor):
        t = x.to(dtype=torch.float64)
        if not t.is_cuda:
            t = t.cuda(device)
        return t.contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device if device is not None else "cuda").contiguous()
