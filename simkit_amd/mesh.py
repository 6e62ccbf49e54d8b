"""Workload meshes: the reference's 5-tet-per-cell brick, unit cubes on [0,1]^3,
seeded rest-vertex jitter and cantilever-beam batches.

``tetrahedralized_grid`` has the vertex numbering and cell split of the
reference's generator (examples/interactive_demos/utils.py:52-90, duplicated
at tests/test_p2_beam_3d.py:47-80): vertex id (i*ny + j)*nz + k on
linspace(-w/2, w/2) axes, each hex cut into 5 tets with a parity-flipped
pattern so neighbouring cells share faces.
"""

import numpy as np

# Hex corners: 0=(i,j,k) 1=(i+1,j,k) 2=(i+1,j+1,k) 3=(i,j+1,k), 4..7 = same at k+1.
_SPLIT = {
    0: np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]]),
    1: np.array([[0, 1, 2, 5], [0, 2, 3, 7], [0, 5, 7, 4], [2, 5, 6, 7], [0, 2, 7, 5]]),
}
_CORNER = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def tetrahedralized_grid(nx, ny, nz, width=1.0, height=1.0, depth=1.0):
    """Brick of nx*ny*nz vertices centred on the origin; returns X (n,3) f64, T (t,4) int64."""
    xs = np.linspace(-width / 2.0, width / 2.0, nx)
    ys = np.linspace(-height / 2.0, height / 2.0, ny)
    zs = np.linspace(-depth / 2.0, depth / 2.0, nz)
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), axis=-1).reshape(-1, 3)
    cells = np.stack(np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij"),
                     axis=-1).reshape(-1, 3)
    c = cells[:, None, :] + _CORNER[None, :, :]                  # (ncell, 8, 3)
    corner_id = (c[..., 0] * ny + c[..., 1]) * nz + c[..., 2]    # (ncell, 8)
    parity = cells.sum(axis=1) % 2
    T = np.empty((cells.shape[0], 5, 4), dtype=np.int64)
    for p, split in _SPLIT.items():
        sel = parity == p
        T[sel] = corner_id[sel][:, split]
    return X, T.reshape(-1, 4)


def unit_cube(n):
    """n^3-cell cube on [0,1]^3 (the centred generator shifted by +0.5); y=0 is the bottom face."""
    X, T = tetrahedralized_grid(n + 1, n + 1, n + 1)
    return X + 0.5, T


def jitter(X, cell, amount=0.2, seed=0, keep=None):
    """Rest vertices displaced by U(-amount*cell, +amount*cell)^3 (seeded); rows in `keep` untouched."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amount * cell, amount * cell, size=X.shape)
    if keep is not None:
        d[keep] = 0.0
    return X + d


def beam(length_cells, section_cells=4, cell=0.025):
    """Cantilever beam along +x: length_cells x section x section cells of size `cell`, x=0 face at origin."""
    nx, ny, nz = length_cells + 1, section_cells + 1, section_cells + 1
    X, T = tetrahedralized_grid(nx, ny, nz, width=length_cells * cell, height=section_cells * cell,
                                depth=section_cells * cell)
    X = X - X.min(axis=0)
    return X, T
