"""Deformation Jacobian J (9t x 3n) with J @ x = stacked row-major F
(deformation_jacobian.py:9-87). The per-tet operator D is computed on the GPU
(sk_tet_operator); J is assembled from it by index bookkeeping only, and
carries the device mesh handle so the x-tier energies reuse it."""

import numpy as np
import scipy.sparse as sps

from ._device import host
from ._mesh import mesh_for_rest


def deformation_jacobian(X, T):
    """J[9e+3i+c, 3v+i] = D_e[c, v]; exact zeros of D are not stored (as in the reference)."""
    mesh, _ = mesh_for_rest(X, T)
    D = host(mesh.D)
    T = np.asarray(T)
    nt, nv = T.shape[0], np.asarray(X).shape[0]
    e, i, c, v = np.meshgrid(np.arange(nt), np.arange(3), np.arange(3), np.arange(4), indexing="ij")
    J = sps.csc_matrix((D[e, c, v].ravel(), ((9 * e + 3 * i + c).ravel(), (3 * T[e, v] + i).ravel())),
                       shape=(9 * nt, 3 * nv))
    J.eliminate_zeros()
    J._simkit_amd_mesh = mesh
    return J
