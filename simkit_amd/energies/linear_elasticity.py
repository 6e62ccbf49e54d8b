"""Linear (small-strain) elasticity on the GPU.

Same F-representation tiers, names and signatures as simkit/energies/linear_elasticity.py:
element tier ``*_element_F``, global x tier ``*_x``, displacement tier ``*_u``
and self-contained tier. The Hessian is constant and PSD; as in the reference, the ``psd``
flag is accepted and has no effect (no projection is applied).
"""

from .. import _lib
from ..deformation_jacobian import deformation_jacobian
from ..volume import volume
from . import _tiers

_M = _lib.SK_LINEAR
_NAME = "Linear elasticity"


def linear_elasticity_energy_element_F(F, mu, lam):
    """psi (t, 1) = mu ||eps||^2 + lam/2 tr(eps)^2, eps = (F + F^T)/2 - I."""
    return _tiers.element(_M, F, mu, lam, "psi", _NAME)


def linear_elasticity_gradient_element_F(F, mu, lam):
    """P (t, 3, 3) = mu (F + F^T - 2 I) + lam tr(F - I) I."""
    return _tiers.element(_M, F, mu, lam, "P", _NAME)


def linear_elasticity_hessian_element_F(F, mu, lam):
    """H (t, 9, 9) = mu (I + T) + lam vec(I) vec(I)^T over row-major vec(F)."""
    return _tiers.element(_M, F, mu, lam, "H", _NAME)


def linear_elasticity_energy_x(X, J, mu, lam, vol):
    return _tiers.energy(_M, X, J, mu, lam, vol, name=_NAME)


def linear_elasticity_gradient_x(X, J, mu, lam, vol):
    return _tiers.gradient(_M, X, J, mu, lam, vol, name=_NAME)


def linear_elasticity_hessian_x(X, J, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, X, J, mu, lam, vol, False, name=_NAME)


def linear_elasticity_energy_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.energy(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name=_NAME)


def linear_elasticity_gradient_u(u, J, Jx_bar, mu, lam, vol):
    return _tiers.gradient(_M, u, J, mu, lam, vol, Jx_bar=Jx_bar, name=_NAME)


def linear_elasticity_hessian_u(u, J, Jx_bar, mu, lam, vol, psd=True):
    return _tiers.hessian(_M, u, J, mu, lam, vol, False, Jx_bar=Jx_bar, name=_NAME)


def linear_elasticity_energy(X, T, mu, lam, U=None):
    U = X if U is None else U
    return linear_elasticity_energy_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def linear_elasticity_gradient(X, T, mu, lam, U=None):
    U = X if U is None else U
    return linear_elasticity_gradient_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T))


def linear_elasticity_hessian(X, T, mu, lam, U=None, psd=True):
    U = X if U is None else U
    return linear_elasticity_hessian_x(U, deformation_jacobian(X, T), mu, lam, volume(X, T), psd=psd)
