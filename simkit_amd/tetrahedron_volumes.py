"""Signed tet volumes det([x1-x0, x2-x0, x3-x0])/6 on the GPU (tetrahedron_volumes.py:10-27)."""

from ._device import host
from ._mesh import tet_operator


def tetrahedron_volumes(X, T):
    """(m,) signed volumes."""
    _, vol = tet_operator(X, T)
    return host(vol)
