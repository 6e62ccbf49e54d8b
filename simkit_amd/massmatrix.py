"""Lumped diagonal mass matrix (massmatrix.py:16-50), summed per vertex on the GPU."""

import numpy as np
import scipy.sparse as sps

from ._device import host
from ._mesh import mesh_for_rest


def lumped_mass(X, T, rho=1):
    """(n,) vertex masses (sum_e vol_e rho_e)/4."""
    mesh, vol = mesh_for_rest(X, T)
    rho_arr = None if np.ndim(rho) == 0 and float(rho) == 1.0 else rho
    return host(mesh.lumped_mass(vol, rho_arr))


def massmatrix(X, T, rho=1):
    """scipy dia_matrix (n, n) with the lumped masses on the diagonal."""
    return sps.diags(lumped_mass(X, T, rho))
