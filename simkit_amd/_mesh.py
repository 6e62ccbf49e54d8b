"""DeviceMesh: the library's mesh handle (CSR pattern, element -> nnz map,
staging buffers) plus helpers to obtain one for a rest mesh or for a
deformation Jacobian built by either this package or the reference."""

import ctypes
import hashlib
import weakref

import numpy as np
import scipy.sparse as sps

from . import _lib
from ._device import F64, I32, dev, empty, host, per_element, ptr, require_gpu, stream

MATERIAL_SNH = _lib.SK_SNH
MATERIAL_ARAP = _lib.SK_ARAP


def tet_operator(X, T):
    """Device D (nt, 3, 4) and signed volumes (nt,) of a tet mesh (sk_tet_operator)."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T)
    if X.ndim != 2 or X.shape[1] != 3 or T.ndim != 2 or T.shape[1] != 4:
        raise NotImplementedError("simkit_amd supports 3D tetrahedral meshes (X (n,3), T (t,4)) only")
    nt = T.shape[0]
    Xd = dev(X.reshape(-1))
    Td = dev(T.astype(np.int32).reshape(-1), I32)
    D = empty(12 * nt)
    vol = empty(nt)
    _lib.call("sk_tet_operator", ptr(Xd), ptr(Td), X.shape[0], nt, ptr(D), ptr(vol), stream())
    return D.view(nt, 3, 4), vol


class DeviceMesh:
    """Owns an sk_mesh (sorted block-CSR pattern and assembly maps) for one tet mesh."""

    def __init__(self, T, D_dev, nv):
        require_gpu()
        T = np.ascontiguousarray(np.asarray(T), dtype=np.int32)
        self.T = T
        self.nt = T.shape[0]
        self.nv = int(nv)
        self.D = D_dev.contiguous()
        handle = ctypes.c_void_p()
        _lib.call("sk_mesh_create", T.ctypes.data, ptr(self.D), self.nv, self.nt, ctypes.byref(handle))
        self._h = handle
        self._finalizer = weakref.finalize(self, _destroy, handle.value)
        nnz = ctypes.c_int64()
        nblk = ctypes.c_int64()
        _lib.call("sk_mesh_info", self._h, None, None, ctypes.byref(nnz), ctypes.byref(nblk))
        self.nnz = nnz.value
        self.nblocks = nblk.value
        self._pattern = None

    @property
    def handle(self):
        return self._h

    # -- pattern -------------------------------------------------------------
    def csr_pattern(self):
        """(indptr, indices) of the sorted 3nv x 3nv CSR pattern, as host int32 arrays."""
        if self._pattern is None:
            ip = empty(3 * self.nv + 1, I32)
            ix = empty(self.nnz, I32)
            _lib.call("sk_mesh_csr_pattern", self._h, ptr(ip), ptr(ix), stream())
            self._pattern = (host(ip), host(ix))
        return self._pattern

    def element_nnz(self):
        """(nt, 144) CSR positions of every K_e entry (row-major 12x12 per element)."""
        out = empty(144 * self.nt, I32)
        _lib.call("sk_mesh_element_nnz", self._h, ptr(out), stream())
        return host(out).reshape(self.nt, 144)

    def to_csr(self, values_dev):
        indptr, indices = self.csr_pattern()
        n = 3 * self.nv
        return sps.csr_matrix((host(values_dev), indices.copy(), indptr.copy()), shape=(n, n))

    # -- rest measures -----------------------------------------------------
    def lumped_mass(self, vol_dev, rho=None):
        m = empty(self.nv)
        rho_d = None if rho is None else per_element(rho, self.nt)
        _lib.call("sk_mesh_lumped_mass", self._h, ptr(vol_dev), ptr(rho_d), ptr(m), stream())
        return m

    # -- x tier -------------------------------------------------------------
    def energy(self, material, x_dev, mu_dev, lam_dev, vol_dev, F0_dev=None, per_element_out=False):
        e = ctypes.c_double()
        elem = empty(self.nt) if per_element_out else None
        _lib.call("sk_mesh_energy", self._h, material, ptr(x_dev), ptr(F0_dev), ptr(mu_dev), ptr(lam_dev),
                  ptr(vol_dev), ctypes.byref(e), ptr(elem), stream())
        return (e.value, elem) if per_element_out else e.value

    def gradient(self, material, x_dev, mu_dev, lam_dev, vol_dev, F0_dev=None):
        g = empty(3 * self.nv)
        _lib.call("sk_mesh_gradient", self._h, material, ptr(x_dev), ptr(F0_dev), ptr(mu_dev), ptr(lam_dev),
                  ptr(vol_dev), ptr(g), stream())
        return g

    def hessian(self, material, x_dev, mu_dev, lam_dev, vol_dev, psd, F0_dev=None):
        vals = empty(self.nnz)
        _lib.call("sk_mesh_hessian", self._h, material, ptr(x_dev), ptr(F0_dev), ptr(mu_dev), ptr(lam_dev),
                  ptr(vol_dev), int(psd), ptr(vals), stream())
        return vals


def _destroy(h):
    try:
        _lib.load().sk_mesh_destroy(h)
    except Exception:  # interpreter shutdown
        pass


# ---------------------------------------------------------------------------
# Mesh lookup
# ---------------------------------------------------------------------------
_REST_CACHE = {}


def _digest(*arrays):
    h = hashlib.blake2b(digest_size=16)
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str((a.dtype, a.shape)).encode())
        h.update(a.view(np.uint8).reshape(-1) if a.size else b"")
    return h.hexdigest()


def mesh_for_rest(X, T):
    """DeviceMesh (and device volumes) of a rest mesh, cached by content."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T)
    key = _digest(X, T.astype(np.int64))
    hit = _REST_CACHE.get(key)
    if hit is None:
        D, vol = tet_operator(X, T)
        hit = (DeviceMesh(T, D, X.shape[0]), vol)
        if len(_REST_CACHE) > 8:
            _REST_CACHE.clear()
        _REST_CACHE[key] = hit
    return hit


def mesh_for_jacobian(J):
    """DeviceMesh for a deformation Jacobian (9t x 3n).

    Jacobians built by ``simkit_amd.deformation_jacobian`` carry their mesh.
    Any other J with the reference's layout (J[9e+3i+c, 3v+i] = D_e[c, v],
    deformation_jacobian.py:65-86) is decoded once: the vertex set of tet e is
    read from the rows of its first spatial component, D from their values.
    The decoded handle is cached on the matrix object.
    """
    m = getattr(J, "_simkit_amd_mesh", None)
    if m is not None:
        return m
    if not sps.issparse(J):
        raise TypeError("J must be a scipy sparse deformation Jacobian")
    nrow, ncol = J.shape
    if nrow % 9 or ncol % 3:
        raise NotImplementedError("simkit_amd supports 3D tetrahedral deformation Jacobians only")
    nt, nv = nrow // 9, ncol // 3
    C = sps.coo_matrix(J)
    r, c, v = C.row.astype(np.int64), C.col.astype(np.int64), C.data
    first = (r % 9) < 3                      # rows of spatial component i = 0
    e, comp, vert, val = r[first] // 9, r[first] % 9, c[first] // 3, v[first]
    key = e * nv + vert
    ukeys = np.unique(key)
    if ukeys.size != 4 * nt or np.any(np.bincount(ukeys // nv, minlength=nt) != 4):
        raise ValueError("J is not a tetrahedral deformation Jacobian (expected 4 vertices per element)")
    T = (ukeys % nv).reshape(nt, 4)
    slot = np.searchsorted(ukeys, key) - 4 * e
    D = np.zeros((nt, 3, 4))
    np.add.at(D, (e, comp, slot), val)
    m = DeviceMesh(T, dev(D.reshape(-1)).view(nt, 3, 4), nv)
    try:
        J._simkit_amd_mesh = m
    except AttributeError:
        pass
    return m
