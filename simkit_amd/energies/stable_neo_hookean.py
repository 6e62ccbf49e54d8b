"""Stable neo-Hookean energy (Smith, de Goes, Kim 2018) on the GPU.

Same tiers, names and signatures as simkit/energies/stable_neo_hookean.py:
element tier ``*_element_F`` (:65-443), global x tier ``*_x`` (:449-538),
displacement tier ``*_u`` (:544-648) and self-contained tier (:654-737).
The x/u-tier Hessian is the vol-weighted, 9x9-PSD-projected element Hessian
assembled into a CSR matrix with sorted indices over the full vertex-pair
block pattern (entries that cancel to exactly zero are kept).
"""

from .. import _lib
from ..deformation_jacobian import deformation_jacobian
from ..volume import volume
from . import _tiers

_M = _lib.SK_SNH


def stable_neo_hookean_energy_element_F(F, mu, lam):
    """psi (t, 1) = mu/2 (I_C-3) - mu/2 log(I_C+1) + lam/2 (J-alpha)^2."""
    return _tiers.element(_M, F, mu, lam, "psi", "Stable Neo-Hookean")


def stable_neo_hookean_gradient_element_F(F, mu, lam):
    """P (t, 3, 3) = mu I_C/(I_C+1) F + lam (J-alpha) cof F."""
    return _tiers.element(_M, F, mu, lam, "P", "Stable Neo-Hookean")


def stable_neo_hookean_hessian_element_F(F, mu, lam):
    """H (t, 9, 9) over row-major vec(F); not projected."""
    return _tiers.element(_M, F, mu, lam, "H", "Stable Neo-Hookean")


def stable_neo_hookean_energy_x(X, J, mu, lam, vol):
    return _tiers.energy(_M, X, J, mu, lam, vol, name="Stable Neo-Hookean")


def stable_neo_hookean_gradient_x(X, J, mu, lam, vol):
    return _tiers.gradient(_M, X, J, mu, lam, vol, name="Stable Neo-Hookean")


def stable_neo_hookean_hessian_x(X, J, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, X, J, mu, lam, vol, psd, name="Stable Neo-Hookean")


def stable_neo_hookean_energy_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.energy(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name="Stable Neo-Hookean")


def stable_neo_hookean_gradient_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.gradient(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name="Stable Neo-Hookean")


def stable_neo_hookean_hessian_u(u, J, Jx_bar, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, u, J, mu, lam, vol, psd, Jx_bar=Jx_bar, name="Stable Neo-Hookean")


def stable_neo_hookean_energy(X, T, mu, lam, U=None):
    U = X if U is None else U
    return stable_neo_hookean_energy_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def stable_neo_hookean_gradient(X, T, mu, lam, U=None):
    U = X if U is None else U
    return stable_neo_hookean_gradient_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def stable_neo_hookean_hessian(X, T, mu, lam, U=None, psd=True):
    U = X if U is None else U
    return stable_neo_hookean_hessian_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T), psd=psd)
