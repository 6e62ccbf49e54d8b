"""Batched PSD projection by in-register Jacobi eigen-decomposition (psd_project.py:12-47)."""

import numpy as np

from . import _lib
from ._device import dev, empty, host, ptr, stream

_METHODS = {"proj": _lib.SK_PSD_PROJ, "abs": _lib.SK_PSD_ABS}


def psd_project(H, method="proj"):
    """Floor eigenvalues at 1e-6 ('proj') or take |.| ('abs'); returns (n, d, d).

    Like the reference (numpy.linalg.eigh), only the lower triangle is read and
    a 2D input is promoted to a batch of one. Any matrix size is supported (up to 12 in registers).
    """
    H = np.asarray(H, dtype=np.float64)
    if H.ndim == 2:
        H = H[None]
    n, d, d2 = H.shape
    if d != d2:
        raise ValueError("psd_project expects square matrices")
    if n == 0:
        return H.copy()
    Hd = dev(H.reshape(-1))
    out = empty(n * d * d)
    _lib.call("sk_psd_project", ptr(Hd), n, d, _METHODS.get(method, _lib.SK_PSD_NONE), ptr(out), stream())
    return host(out).reshape(n, d, d)
