"""Pins the oracle's mesh generator and checks the product generator against it."""

import numpy as np

from oracle import fem
from oracle.mesh import tetrahedralized_grid as oracle_grid, unit_cube as oracle_cube
from simkit_amd import mesh as product_mesh


def test_counts_and_volume():
    for n in (1, 2, 3, 10):
        X, T = oracle_cube(n)
        assert X.shape == ((n + 1) ** 3, 3)
        assert T.shape == (5 * n ** 3, 4)
        vol = fem.tet_volumes(X, T)
        assert np.all(vol > 0)
        assert abs(vol.sum() - 1.0) < 1e-12
        assert X.min() == 0.0 and X.max() == 1.0


def test_vertex_ordering_and_corners():
    X, T = oracle_grid(3, 4, 5, width=2.0, height=3.0, depth=4.0)
    i, j, k = 2, 1, 3
    vid = (i * 4 + j) * 5 + k
    assert np.allclose(X[vid], [-1.0 + i * 1.0, -1.5 + j * 1.0, -2.0 + k * 1.0])
    # every cell's 5 tets cover exactly the 8 corners of the cell
    for c in range(T.shape[0] // 5):
        assert len(np.unique(T[5 * c:5 * c + 5])) == 8


def test_shared_faces_are_conforming():
    # Each interior face is shared by exactly two tets, boundary faces by one.
    X, T = oracle_cube(3)
    faces = np.sort(np.concatenate([T[:, [0, 1, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 3]], T[:, [1, 2, 3]]]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert set(np.unique(counts)) <= {1, 2}
    assert (counts == 1).sum() == 6 * 3 * 3 * 2


def test_product_generator_bit_equal():
    for n in (1, 4, 7):
        Xo, To = oracle_cube(n)
        Xp, Tp = product_mesh.unit_cube(n)
        assert np.array_equal(Xo, Xp) and np.array_equal(To, Tp)
    Xo, To = oracle_grid(9, 5, 5, width=0.2, height=0.1, depth=0.1)
    Xp, Tp = product_mesh.tetrahedralized_grid(9, 5, 5, width=0.2, height=0.1, depth=0.1)
    assert np.array_equal(Xo, Xp) and np.array_equal(To, Tp)


def test_jitter_and_beam():
    X, T = product_mesh.unit_cube(6)
    Xj = product_mesh.jitter(X, 1.0 / 6, 0.2, seed=0)
    assert np.all(np.abs(Xj - X) <= 0.2 / 6 + 1e-15)
    assert np.all(fem.tet_volumes(Xj, T) > 0)
    Xb, Tb = product_mesh.beam(8, 4, 0.025)
    assert Xb.shape == (9 * 5 * 5, 3) and Tb.shape == (5 * 8 * 16, 4)
    assert np.isclose(fem.tet_volumes(Xb, Tb).sum(), 8 * 4 * 4 * 0.025 ** 3)
