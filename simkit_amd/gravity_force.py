"""Lumped gravity load M [0, a, 0] per vertex (gravity_force.py:8-42)."""

import numpy as np

from .massmatrix import lumped_mass


def gravity_force(X, T, a=-9.8, rho=1):
    """(n, dim) force with the lumped vertex mass times `a` in the second coordinate."""
    X = np.asarray(X)
    g = np.zeros((X.shape[0], X.shape[1]))
    g[:, 1] = lumped_mass(X, T, rho) * a
    return g
