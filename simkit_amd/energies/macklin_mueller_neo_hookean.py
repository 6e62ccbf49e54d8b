"""Macklin-Mueller neo-Hookean energy on the GPU.

Same F-representation tiers, names and signatures as simkit/energies/macklin_mueller_neo_hookean.py:
element tier ``*_element_F``, global x tier ``*_x``, displacement tier ``*_u``
and self-contained tier. The stretch (S) tiers are not on the position path and are not provided.
"""

from .. import _lib
from ..deformation_jacobian import deformation_jacobian
from ..volume import volume
from . import _tiers

_M = _lib.SK_MMNH
_NAME = "Macklin-Mueller Neo-Hookean"


def macklin_mueller_neo_hookean_energy_element_F(F, mu, lam):
    """psi (t, 1) = mu (1 - J) + lam/2 (1 - J)^2 + mu/2 (I_C - 3)."""
    return _tiers.element(_M, F, mu, lam, "psi", _NAME)


def macklin_mueller_neo_hookean_gradient_element_F(F, mu, lam):
    """P (t, 3, 3) = mu F - (mu + lam (1 - J)) cof F."""
    return _tiers.element(_M, F, mu, lam, "P", _NAME)


def macklin_mueller_neo_hookean_hessian_element_F(F, mu, lam):
    """H (t, 9, 9) over row-major vec(F); not projected."""
    return _tiers.element(_M, F, mu, lam, "H", _NAME)


def macklin_mueller_neo_hookean_energy_x(X, J, mu, lam, vol):
    return _tiers.energy(_M, X, J, mu, lam, vol, name=_NAME)


def macklin_mueller_neo_hookean_gradient_x(X, J, mu, lam, vol):
    return _tiers.gradient(_M, X, J, mu, lam, vol, name=_NAME)


def macklin_mueller_neo_hookean_hessian_x(X, J, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, X, J, mu, lam, vol, psd, name=_NAME)


def macklin_mueller_neo_hookean_energy_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.energy(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name=_NAME)


def macklin_mueller_neo_hookean_gradient_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.gradient(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name=_NAME)


def macklin_mueller_neo_hookean_hessian_u(u, J, Jx_bar, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, u, J, mu, lam, vol, psd, Jx_bar=Jx_bar, name=_NAME)


def macklin_mueller_neo_hookean_energy(X, T, mu, lam, U=None):
    U = X if U is None else U
    return macklin_mueller_neo_hookean_energy_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def macklin_mueller_neo_hookean_gradient(X, T, mu, lam, U=None):
    U = X if U is None else U
    return macklin_mueller_neo_hookean_gradient_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def macklin_mueller_neo_hookean_hessian(X, T, mu, lam, U=None, psd=True):
    U = X if U is None else U
    return macklin_mueller_neo_hookean_hessian_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T), psd=psd)
