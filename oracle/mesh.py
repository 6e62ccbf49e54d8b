"""Oracle restatement of the reference's tet-grid generator (test infrastructure).

Follows ``examples/interactive_demos/utils.py:52-90`` (``tetrahedralized_grid``;
duplicate at ``tests/test_p2_beam_3d.py:47-80``): a brick of ``nx*ny*nz``
vertices on ``linspace(-w/2, w/2)`` per axis, vertex id ``(i*ny+j)*nz+k``,
each hex cell cut into 5 tets with a parity-flipped pattern so that faces
match between neighbouring cells.

Pinned by: vertex/tet counts, every tet positive with total volume equal to
the box volume, exact corner coordinates, and bit-equality with the product
generator (``tests/test_oracle_mesh.py``).
"""

import numpy as np

# Corner numbering of one hex cell: 0=(i,j,k) 1=(i+1,j,k) 2=(i+1,j+1,k)
# 3=(i,j+1,k) 4..7 = same with k+1 (utils.py:80-89).
_EVEN = np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]])
_ODD = np.array([[0, 1, 2, 5], [0, 2, 3, 7], [0, 5, 7, 4], [2, 5, 6, 7], [0, 2, 7, 5]])


def tetrahedralized_grid(nx, ny, nz, width=1.0, height=1.0, depth=1.0):
    """utils.py:52-90. Returns X (nx*ny*nz, 3) f64 and T (5*(nx-1)(ny-1)(nz-1), 4) int64."""
    axes = [np.linspace(-s / 2.0, s / 2.0, n) for s, n in ((width, nx), (height, ny), (depth, nz))]
    grid = np.meshgrid(*axes, indexing="ij")
    X = np.column_stack([g.reshape(-1) for g in grid])

    ci, cj, ck = (a.reshape(-1) for a in np.meshgrid(
        np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij"))

    def vid(i, j, k):
        return (i * ny + j) * nz + k

    corners = np.column_stack([
        vid(ci, cj, ck), vid(ci + 1, cj, ck), vid(ci + 1, cj + 1, ck), vid(ci, cj + 1, ck),
        vid(ci, cj, ck + 1), vid(ci + 1, cj, ck + 1), vid(ci + 1, cj + 1, ck + 1),
        vid(ci, cj + 1, ck + 1),
    ])
    even = (ci + cj + ck) % 2 == 0
    T = np.where(even[:, None, None], corners[:, _EVEN], corners[:, _ODD])
    return X, T.reshape(-1, 4)


def unit_cube(n):
    """n^3-cell cube on [0,1]^3: the reference generator shifted by +0.5 on each axis."""
    X, T = tetrahedralized_grid(n + 1, n + 1, n + 1)
    return X + 0.5, T
