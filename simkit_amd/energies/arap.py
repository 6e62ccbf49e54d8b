"""As-rigid-as-possible energy psi = mu/2 ||F - R||^2 on the GPU.

Same F-representation tiers, names and signatures as simkit/energies/arap.py:
element tier (:71-145), x tier (:242-325), u tier (:331-424) and
self-contained tier (:505-582). The stretch (S) tiers are not on the
position path and are not provided.
"""

from .. import _lib
from ..deformation_jacobian import deformation_jacobian
from ..volume import volume
from . import _tiers

_M = _lib.SK_ARAP


def arap_energy_element_F(F, mu):
    return _tiers.element(_M, F, mu, None, "psi", "ARAP")


def arap_gradient_element_F(F, mu):
    return _tiers.element(_M, F, mu, None, "P", "ARAP")


def arap_hessian_element_F(F, mu):
    return _tiers.element(_M, F, mu, None, "H", "ARAP")


def arap_energy_x(X, J, mu, vol):
    return _tiers.energy(_M, X, J, mu, None, vol, name="ARAP")


def arap_gradient_x(X, J, mu, vol):
    return _tiers.gradient(_M, X, J, mu, None, vol, name="ARAP")


def arap_hessian_x(X, J, mu, vol, psd=True):
    return _tiers.hessian(_M, X, J, mu, None, vol, psd, name="ARAP")


def arap_energy_u(u, J, Jx_bar, mu, vol):
    return _tiers.energy(_M, u, J, mu, None, vol, Jx_bar=Jx_bar, name="ARAP")


def arap_gradient_u(u, J, Jx_bar, mu, vol):
    return _tiers.gradient(_M, u, J, mu, None, vol, Jx_bar=Jx_bar, name="ARAP")


def arap_hessian_u(u, J, Jx_bar, mu, vol, psd=True):
    return _tiers.hessian(_M, u, J, mu, None, vol, psd, Jx_bar=Jx_bar, name="ARAP")


def arap_energy(X, T, mu, U=None):
    U = X if U is None else U
    return arap_energy_x(U, deformation_jacobian(X, T), mu, volume(X, T))


def arap_gradient(X, T, mu, U=None):
    U = X if U is None else U
    return arap_gradient_x(U, deformation_jacobian(X, T), mu, volume(X, T))


def arap_hessian(X, T, mu, U=None, psd=True):
    U = X if U is None else U
    return arap_hessian_x(U, deformation_jacobian(X, T), mu, volume(X, T), psd=psd)
