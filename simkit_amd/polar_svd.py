"""Polar decomposition F = R S from the signed SVD (polar_svd.py:8-89)."""

import numpy as np

from .svd_rv import signed_svd


def svd_rv(F, flip=True):
    """With flip (default) as svd_rv.svd_rv; without, a plain SVD (s >= 0)."""
    U, s, V = signed_svd(F)
    if not flip:
        neg = s[:, 2] < 0
        U[neg, :, 2] *= -1.0
        s[neg, 2] *= -1.0
    return U, s[:, :, None] * np.eye(3)[None], V


def polar_svd(F, flip=True):
    """R = U V^T and the symmetric stretch V S V^T (polar_svd.py:209-221).

    The reference's flip=False branch references undefined names and raises;
    here it returns the factors of the plain SVD.
    """
    U, S, V = svd_rv(F, flip=flip)
    Vt = V.transpose(0, 2, 1)
    return U @ Vt, V @ S @ Vt
