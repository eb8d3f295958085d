"""Algorithmic cost model of one parallel IEKS iteration (per kernel class).

Flop counts are those of the reference's formulas (dense, as written in
``pof/parallel_filtsmooth/filter.py:50-142`` and ``smoother.py:37-63``),
counted once per element or combine, independent of how the kernels compute
them; bytes are the element-once traffic (inputs read once, outputs written
once).  ``D = d (q + 1)``; ``s`` = bytes per scalar.

* filter element  (``_get_element``):          2 D^3 + 4/3 (D + d)^3
* filter combine  (``sqrt_filtering_operator``): 44.3 D^3
* filter prefix   (combine restricted to (b, U)): 21 D^3
  (U1^T Z2 2D^3, tria of the 2D x 2D Xi 32/3 D^3, A2 U1 Xi11^-T 3D^3,
  W = A2 - M Xi21^T 2D^3, tria([M | U2]) 10/3 D^3)
* smoother element + observation statistics:   15.3 D^3 + 5.3 D^3
* smoother combine (``sqrt_smoothing_operator``): 7.3 D^3
* smoother prefix (E m + g, tria([E L | D])):   5.3 D^3
"""

KERNEL_CLASSES = ["filter_leaf", "filter_up", "filter_down", "filter_post", "smooth_up", "smooth_down",
                  "smooth_post", "reduce"]

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E


def flops_per_unit(cls, D, d):
    D3 = float(D) ** 3
    return {
        "filter_leaf": 2 * D3 + (4.0 / 3.0) * (D + d) ** 3,
        "filter_up": 44.3 * D3,
        "filter_down": 21.0 * D3,
        "filter_post": 15.3 * D3 + 5.3 * D3,
        "smooth_up": 7.3 * D3,
        "smooth_down": 5.3 * D3,
        "smooth_post": 3.0 * D * D,
        "reduce": 4.0,
    }[cls]


def bytes_per_unit(cls, D, d, s=8):
    elem_f = 3 * D * D + 2 * D
    elem_s = 2 * D * D + D
    state = D * D + D
    return s * {
        "filter_leaf": D + elem_f,                 # read a mean, write an element
        "filter_up": 3 * elem_f,                   # read two elements, write one
        "filter_down": state + elem_f + state,     # read prefix + element, write prefix
        "filter_post": state + elem_s,             # read a filtered state, write a smoother element
        "smooth_up": 3 * elem_s,
        "smooth_down": state + elem_s + state,
        "smooth_post": 3 * D,
        "reduce": 4,
    }[cls]


def iteration_flops(Ns, D, d):
    """Algorithmic flops of one IEKS iteration: Ns (74.3 D^3 + 4/3 (D + d)^3)."""
    return Ns * (74.3 * D**3 + (4.0 / 3.0) * (D + d) ** 3)


def iteration_bytes(Ns, D, s=8):
    """Element-once bytes per iteration: s Ns (3 D^2 + 4 D)."""
    return s * Ns * (3 * D * D + 4 * D)
