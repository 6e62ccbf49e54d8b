"""Shared plumbing for the element, x, u and self-contained energy tiers."""

import numpy as np

from .. import _lib
from .._device import dev, empty, host, per_element, ptr, stream
from .._mesh import mesh_for_jacobian

PSD = {True: _lib.SK_PSD_PROJ, False: _lib.SK_PSD_NONE}


def check_dim(dim, name):
    if dim == 2:
        raise NotImplementedError(f"simkit_amd {name}: the 2D tier is outside this package's scope")
    if dim != 3:
        raise ValueError(f"{name} supports dim=2 or dim=3")


def element(material, F, mu, lam, want, name):
    """Element tier on (t, 3, 3) F: returns psi (t,1), P (t,3,3) or H (t,9,9)."""
    F = np.asarray(F, dtype=np.float64)
    check_dim(F.shape[-1], name)
    F = F.reshape(-1, 3, 3)
    nt = F.shape[0]
    if nt == 0:
        return {"psi": np.zeros((0, 1)), "P": np.zeros((0, 3, 3)), "H": np.zeros((0, 9, 9))}[want]
    Fd = dev(F.reshape(-1))
    mu_d = per_element(mu, nt)
    lam_d = per_element(lam, nt) if lam is not None else None
    out = empty(nt * {"psi": 1, "P": 9, "H": 81}[want])
    args = {"psi": (ptr(out), None, None), "P": (None, ptr(out), None), "H": (None, None, ptr(out))}[want]
    _lib.call("sk_element_F", material, ptr(Fd), ptr(mu_d), ptr(lam_d), nt, *args, stream())
    shape = {"psi": (nt, 1), "P": (nt, 3, 3), "H": (nt, 9, 9)}[want]
    return host(out).reshape(shape)


def _setup(X_or_u, J, mu, lam, vol, Jx_bar, name):
    X_or_u = np.asarray(X_or_u, dtype=np.float64)
    check_dim(X_or_u.shape[-1] if X_or_u.ndim == 2 else 3, name)
    mesh = mesh_for_jacobian(J)
    x = dev(X_or_u.reshape(-1))
    mu_d = per_element(mu, mesh.nt)
    lam_d = per_element(lam, mesh.nt) if lam is not None else None
    vol_d = per_element(vol, mesh.nt)
    F0 = None if Jx_bar is None else dev(np.asarray(Jx_bar, dtype=np.float64).reshape(-1))
    return mesh, x, mu_d, lam_d, vol_d, F0


def energy(material, X, J, mu, lam, vol, Jx_bar=None, name="energy"):
    mesh, x, mu_d, lam_d, vol_d, F0 = _setup(X, J, mu, lam, vol, Jx_bar, name)
    return float(mesh.energy(material, x, mu_d, lam_d, vol_d, F0))


def gradient(material, X, J, mu, lam, vol, Jx_bar=None, name="gradient"):
    mesh, x, mu_d, lam_d, vol_d, F0 = _setup(X, J, mu, lam, vol, Jx_bar, name)
    return host(mesh.gradient(material, x, mu_d, lam_d, vol_d, F0)).reshape(-1, 1)


def hessian(material, X, J, mu, lam, vol, psd, Jx_bar=None, name="hessian"):
    mesh, x, mu_d, lam_d, vol_d, F0 = _setup(X, J, mu, lam, vol, Jx_bar, name)
    return mesh.to_csr(mesh.hessian(material, x, mu_d, lam_d, vol_d, PSD[bool(psd)], F0))
