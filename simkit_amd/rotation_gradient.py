"""dR/dF of the polar rotation (rotation_gradient.py:12-75) on the GPU."""

import numpy as np

from . import _lib
from ._device import dev, empty, host, ptr, stream


def rotation_gradient_F(F):
    """(n, 9, 9) = sum_k 2/max(s_i + s_j, 1e-8) t_k t_k^T over the three twists."""
    F = np.asarray(F, dtype=np.float64)
    if F.shape[-1] != 3:
        raise NotImplementedError("simkit_amd rotation_gradient_F supports dim == 3 only")
    F = F.reshape(-1, 3, 3)
    n = F.shape[0]
    Fd = dev(F.reshape(-1))
    K = empty(81 * n)
    _lib.call("sk_rotation_gradient", ptr(Fd), n, ptr(K), stream())
    return host(K).reshape(n, 9, 9)
