"""GPU linear modal analysis (Chebyshev-filtered subspace iteration) against
the oracle's shift-invert Lanczos: eigenvalues to 1e-8 relative, eigenpair
residuals, M-orthonormality, the spanned subspace at spectral gaps, bitwise
determinism, and the 64-mode basis of the 32^3-cell cube used by the
learned warm start."""

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import modal
from oracle.mesh import unit_cube

pytestmark = pytest.mark.gpu


def _check_pairs(E, B, K, m, free, shift, tol):
    G = B.T @ (m[:, None] * B)
    assert np.abs(G - np.eye(B.shape[1])).max() < 1e-9
    assert np.abs(B[~free]).max() == 0.0
    R = (K @ B + shift * m[:, None] * B) - (m[:, None] * B) * E[None, :]
    rel = np.linalg.norm(R[free] / np.sqrt(m[free])[:, None], axis=0) / np.maximum(E, 1e-300)
    assert rel.max() < tol, rel.max()


def _projector_gap(Eo, k):
    """Largest k' <= k with a spectral gap after it (degenerate pairs are not split)."""
    for kk in range(k, 0, -1):
        if Eo[kk] > Eo[kk - 1] * (1 + 1e-6):
            return kk
    return 0


def test_modes_match_oracle_cube5():
    from simkit_amd.linear_modal_analysis import linear_modal_analysis
    X, T = unit_cube(5)
    bI = np.where(X[:, 1] == 0)[0]
    k = 12
    E, B = linear_modal_analysis(X, T, k, bI)
    Eo, Bo = modal.linear_modal_analysis(X, T, k + 4, bI)
    assert np.abs(E - Eo[:k]).max() <= 1e-8 * Eo[k - 1]
    K, m = modal.stiffness_and_mass(X, T)
    free = np.ones(3 * X.shape[0], bool)
    for c in range(3):
        free[3 * bI + c] = False
    _check_pairs(E, B, K, m, free, 1e-6, 1e-7)
    kk = _projector_gap(Eo, k)
    P = B[:, :kk] @ (B[:, :kk].T * m[None, :])
    Po = Bo[:, :kk] @ (Bo[:, :kk].T * m[None, :])
    assert np.abs(P - Po).max() < 1e-6


def test_unconstrained_modes_include_rigid_motions():
    from simkit_amd.linear_modal_analysis import linear_modal_analysis
    X, T = unit_cube(3)
    E, B = linear_modal_analysis(X, T, 10)
    K, m = modal.stiffness_and_mass(X, T)
    w = sl.eigh(K.toarray(), np.diag(m), eigvals_only=True)[:10]
    assert np.all(np.abs(E[:6]) < 1e-8 * w[-1])
    assert np.abs(E[6:] - w[6:]).max() < 1e-8 * w[-1]


def test_modes_bitwise_deterministic_and_32cube_basis():
    import torch
    from simkit_amd import _lib
    from simkit_amd._device import dev, per_element
    from simkit_amd._mesh import mesh_for_rest
    from simkit_amd.linear_modal_analysis import modal_basis
    X, T = unit_cube(32)
    mesh, vol = mesh_for_rest(X, T)
    mass = mesh.lumped_mass(vol, 1.0).repeat_interleave(3).contiguous()
    fv = np.where((X[:, 1] == 0) | (X[:, 1] == 1))[0]
    f = np.zeros(X.size, np.uint8)
    for c in range(3):
        f[3 * fv + c] = 1
    fixed = dev(f, torch.uint8)
    ones = per_element(1.0, mesh.nt)
    x = dev(X.reshape(-1))
    B1, E1, info = modal_basis(mesh, x, ones, ones, vol, mass, fixed, 64, shift=1e-6, tol=1e-6)
    assert info["max_rel_residual"] <= 1e-6 and info["block"] == 128
    B2, E2, _ = modal_basis(mesh, x, ones, ones, vol, mass, fixed, 64, shift=1e-6, tol=1e-6)
    assert torch.equal(B1, B2) and np.array_equal(E1, E2)
    Bh = B1.cpu().numpy()
    mh = mass.cpu().numpy()
    G = Bh.T @ (mh[:, None] * Bh)
    assert np.abs(G - np.eye(64)).max() < 1e-9
    assert np.all(np.diff(E1) >= -1e-9 * E1[-1]) and E1[0] > 0
    print(f"\n32^3 modes: {info}")
