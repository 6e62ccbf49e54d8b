"""Rotation-corrected 3x3 SVD (svd_rv.py:8-53) on the GPU."""

import numpy as np

from . import _lib
from ._device import dev, empty, host, ptr, stream


def signed_svd(F):
    """U, s (signed, (n,3)), V with U, V rotations and sign(s[:,2]) = sign(det F)."""
    F = np.asarray(F, dtype=np.float64)
    if F.shape[-2:] != (3, 3):
        raise NotImplementedError("simkit_amd SVD supports 3x3 matrices only")
    F = F.reshape(-1, 3, 3)
    n = F.shape[0]
    Fd = dev(F.reshape(-1))
    U, S, V = empty(9 * n), empty(3 * n), empty(9 * n)
    _lib.call("sk_polar_svd", ptr(Fd), n, ptr(U), ptr(S), ptr(V), None, stream())
    return host(U).reshape(n, 3, 3), host(S).reshape(n, 3), host(V).reshape(n, 3, 3)


def svd_rv(F):
    """U, S (diagonal (n,3,3)), V with F = U S V^T and det(U V^T) = +1 (the smallest
    singular value carries the sign of det F, svd_rv.py:42-51)."""
    U, s, V = signed_svd(F)
    return U, s[:, :, None] * np.eye(3)[None], V
