"""Linear modal analysis on the GPU (linear_modal_analysis.py:18-73).

Same name, signature and return convention as the reference: the lowest k
generalized eigenpairs of the linear-elasticity stiffness (mu = lam = 1)
under the lumped mass (rho = 1), with the DOFs of the vertices in bI
eliminated (and K + 1e-6 M solved, as the reference does) and re-inserted as
zero rows. The eigen-solve is the library's Chebyshev-filtered subspace
iteration (sk_modal_basis) instead of a UMFPACK shift-invert Lanczos.

``modal_basis`` is the device-level form used by the learned warm start: any
density and positions, per-DOF mass and fixed mask, device outputs.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._device import dev, empty, per_element, ptr, stream
from ._mesh import mesh_for_rest


def modal_basis(mesh, x, mu, lam, vol, mass, fixed, k, shift=0.0, material=_lib.SK_LINEAR, tol=1e-8, degree=20,
                max_sweeps=400, seed=0, check=True):
    """Device (3nv, k) M-orthonormal eigenvectors and host (k,) eigenvalues of (K + shift M, M).

    mesh: DeviceMesh; x, mass (3nv) and fixed (3nv uint8 or None) device tensors; mu, lam, vol
    per-element device tensors. Returns (basis, eigenvalues, info dict). check=False returns the
    best pairs found when max_sweeps is reached (info["max_rel_residual"] says how good)."""
    basis = empty(3 * mesh.nv * k)
    w = np.zeros(k)
    info = _lib.SkModalInfo()
    st = _lib.load().sk_modal_basis(mesh.handle, material, ptr(x), ptr(mu), ptr(lam), ptr(vol), ptr(mass),
                                    ptr(fixed), int(k), float(shift), float(tol), int(degree), int(max_sweeps),
                                    ctypes.c_uint64(seed), ptr(basis),
                                    w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(info), stream())
    if st != _lib.SK_OK and (check or st != _lib.SK_ERR_EIG_MAXIT):
        _lib.check(st)
    return basis.view(3 * mesh.nv, k), w, info.as_dict()


def linear_modal_analysis(X, T, k, bI=None):
    """E (k,) generalized eigenvalues and B (n*3, k) M-orthonormal modes, as the reference returns."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T)
    n, dim = X.shape
    if dim != 3:
        raise NotImplementedError("simkit_amd linear_modal_analysis: 3D tetrahedral meshes only")
    if k == 0:
        return np.zeros(0), np.zeros((n * dim, 0))
    mesh, vol = mesh_for_rest(X, T)
    ones = per_element(1.0, mesh.nt)
    mass = mesh.lumped_mass(vol, 1.0).repeat_interleave(3).contiguous()
    fixed = None
    shift = 0.0
    if bI is not None:
        bI = np.asarray(bI)
        assert bI.ndim == 1
        f = np.zeros(n * dim, np.uint8)
        for c in range(dim):
            f[dim * bI + c] = 1
        fixed = dev(f, torch.uint8)
        shift = 1e-6
    B, E, _ = modal_basis(mesh, dev(X.reshape(-1)), ones, ones, vol, mass, fixed, k, shift=shift)
    return E, B.cpu().numpy()
