"""Fixed-corotational energy (twice ARAP plus lam/2 (J - 1)^2) on the GPU.

Same F-representation tiers, names and signatures as simkit/energies/fcr.py:
element tier ``*_element_F``, global x tier ``*_x``, displacement tier ``*_u``
and self-contained tier.
"""

from .. import _lib
from ..deformation_jacobian import deformation_jacobian
from ..volume import volume
from . import _tiers

_M = _lib.SK_FCR
_NAME = "FCR"


def fcr_energy_element_F(F, mu, lam):
    """psi (t, 1) = mu ||F - R||^2 + lam/2 (J - 1)^2."""
    return _tiers.element(_M, F, mu, lam, "psi", _NAME)


def fcr_gradient_element_F(F, mu, lam):
    """P (t, 3, 3) = 2 mu (F - R) + lam (J - 1) cof F."""
    return _tiers.element(_M, F, mu, lam, "P", _NAME)


def fcr_hessian_element_F(F, mu, lam):
    """H (t, 9, 9) over row-major vec(F); not projected."""
    return _tiers.element(_M, F, mu, lam, "H", _NAME)


def fcr_energy_x(X, J, mu, lam, vol):
    return _tiers.energy(_M, X, J, mu, lam, vol, name=_NAME)


def fcr_gradient_x(X, J, mu, lam, vol):
    return _tiers.gradient(_M, X, J, mu, lam, vol, name=_NAME)


def fcr_hessian_x(X, J, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, X, J, mu, lam, vol, psd, name=_NAME)


def fcr_energy_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.energy(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name=_NAME)


def fcr_gradient_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.gradient(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name=_NAME)


def fcr_hessian_u(u, J, Jx_bar, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, u, J, mu, lam, vol, psd, Jx_bar=Jx_bar, name=_NAME)


def fcr_energy(X, T, mu, lam, U=None):
    U = X if U is None else U
    return fcr_energy_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def fcr_gradient(X, T, mu, lam, U=None):
    U = X if U is None else U
    return fcr_gradient_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def fcr_hessian(X, T, mu, lam, U=None, psd=True):
    U = X if U is None else U
    return fcr_hessian_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T), psd=psd)
