"""Pins the oracle's linear modal analysis (shift-invert Lanczos, as the
reference's eigs.py does) against a dense generalized symmetric eigensolver
on small meshes, plus the eigenpair invariants (CPU only)."""

import numpy as np
import scipy.linalg as sl

from oracle import modal
from oracle.mesh import unit_cube


def _dense(X, T, bI, k):
    K, m = modal.stiffness_and_mass(X, T)
    free = np.ones(3 * X.shape[0], bool)
    for c in range(3):
        free[3 * np.asarray(bI) + c] = False
    Kd = K.toarray()[np.ix_(free, free)]
    Md = np.diag(m[free])
    return sl.eigh(Kd + 1e-6 * Md, Md, eigvals_only=True)[:k], K, m, free


def test_oracle_modes_match_dense_eigh():
    X, T = unit_cube(4)
    bI = np.where(X[:, 1] == 0)[0]
    E, B = modal.linear_modal_analysis(X, T, 12, bI)
    w, K, m, free = _dense(X, T, bI, 12)
    assert np.abs(E - w).max() <= 1e-10 * w.max()
    assert np.abs(B[~free]).max() == 0.0
    G = B.T @ (m[:, None] * B)
    assert np.abs(G - np.eye(12)).max() < 1e-10
    R = (K @ B + 1e-6 * m[:, None] * B)[free] - (m[:, None] * B)[free] * E[None, :]
    assert np.abs(R).max() <= 1e-8 * np.abs(K).max()


def test_linear_stiffness_has_rigid_null_space():
    """Unconstrained linear-elasticity stiffness: exactly 6 zero modes (3 translations, 3 infinitesimal
    rotations), the rest positive - a dropped or transposed term in the element Hessian breaks this."""
    X, T = unit_cube(2)
    K, m = modal.stiffness_and_mass(X, T)
    w = sl.eigh(K.toarray(), np.diag(m), eigvals_only=True)
    assert np.all(np.abs(w[:6]) < 1e-9 * w[-1]) and w[6] > 1e-3 * w[-1] / 100
