"""Per-element deformation gradient F = (D_e U[T_e])^T on the GPU (deformation_gradient.py:8-67)."""

import numpy as np

from . import _lib
from ._device import I32, dev, empty, host, ptr, stream
from ._mesh import tet_operator


def deformation_gradient(X, T, U=None):
    """F (t, 3, 3) mapping rest X to deformed U (U is required, as in the reference)."""
    if U is None:
        raise TypeError("deformation_gradient: U (deformed positions) is required")
    D, _ = tet_operator(X, T)
    T = np.asarray(T)
    nt = T.shape[0]
    Ud = dev(np.asarray(U, dtype=np.float64).reshape(-1))
    Td = dev(T.astype(np.int32).reshape(-1), I32)
    F = empty(9 * nt)
    _lib.call("sk_deformation_gradient", ptr(D), ptr(Td), ptr(Ud), nt, ptr(F), stream())
    return host(F).reshape(nt, 3, 3)
