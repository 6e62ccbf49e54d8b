"""GPU parity: CSR pattern and element -> nnz map (bit-exact against brute
force), and the x / u tier energy, gradient and PSD-projected Hessian against
the oracle (1e-10 relative, FP64) and the reference's golden CSR."""

import numpy as np
import pytest
import scipy.sparse as sps

from conftest import rel_err
from oracle import fem
from oracle.mesh import unit_cube

pytestmark = pytest.mark.gpu
TOL = 1e-10


def brute_force_pattern(T, nv):
    """Sorted structural CSR pattern: every vertex pair sharing a tet, as full 3x3 blocks."""
    a = np.repeat(T, 4, axis=1).reshape(-1)
    b = np.tile(T, (1, 4)).reshape(-1)
    keys = np.unique(np.concatenate([a * nv + b, np.arange(nv) * nv + np.arange(nv)]))
    rows, cols = keys // nv, keys % nv
    R = (3 * rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, axis=2)
    C = (3 * cols[:, None, None] + np.arange(3)[None, None, :]).repeat(3, axis=1)
    M = sps.csr_matrix((np.ones(R.size), (R.reshape(-1), C.reshape(-1))), shape=(3 * nv, 3 * nv))
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32)


def problem(n, seed=0, jitter=0.0, deform=0.05):
    X, T = unit_cube(n)
    rng = np.random.default_rng(seed)
    if jitter:
        X = X + rng.uniform(-jitter / n, jitter / n, X.shape)
    x = X + deform * rng.standard_normal(X.shape) / n
    t = T.shape[0]
    mu = rng.uniform(2e5, 4e5, t)
    lam = mu * rng.uniform(3, 10, t)
    return X, T, x, mu, lam


def test_pattern_and_element_map_bit_exact():
    from simkit_amd._mesh import mesh_for_rest
    for n in (1, 3, 7):
        X, T = unit_cube(n)
        mesh, _ = mesh_for_rest(X, T)
        ip, ix = mesh.csr_pattern()
        bip, bix = brute_force_pattern(T, X.shape[0])
        assert np.array_equal(ip, bip) and np.array_equal(ix, bix)
        enz = mesh.element_nnz()
        rows = (3 * T[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 12)
        for e in range(0, T.shape[0], max(1, T.shape[0] // 50)):
            for a in range(12):
                r = rows[e, a]
                for b in range(12):
                    c = rows[e, b]
                    pos = ip[r] + np.searchsorted(ix[ip[r]:ip[r + 1]], c)
                    assert ix[pos] == c and enz[e, 12 * a + b] == pos


@pytest.mark.parametrize("mat", ["snh", "arap"])
def test_x_tier_matches_reference_golden(ref_assembly, mat):
    import simkit_amd as sk
    from simkit_amd.energies import arap, stable_neo_hookean as snh
    d = ref_assembly
    J = sk.deformation_jacobian(d["X"], d["T"])
    vol = sk.volume(d["X"], d["T"])
    for name, x in (("rest", d["X"]), ("def", d["U"])):
        if mat == "snh":
            Q = snh.stable_neo_hookean_hessian_x(x, J, d["mu"], d["lam"], vol)
            g = snh.stable_neo_hookean_gradient_x(x, J, d["mu"], d["lam"], vol)
            E = snh.stable_neo_hookean_energy_x(x, J, d["mu"], d["lam"], vol)
        else:
            Q = arap.arap_hessian_x(x, J, d["mu"], vol)
            g = arap.arap_gradient_x(x, J, d["mu"], vol)
            E = arap.arap_energy_x(x, J, d["mu"], vol)
        ref = sps.csr_matrix((d[f"{mat}_{name}_data"], d[f"{mat}_{name}_indices"], d[f"{mat}_{name}_indptr"]),
                             shape=Q.shape)
        assert rel_err(Q.toarray(), ref.toarray()) < TOL
        # At rest the gradient is zero up to roundoff on both sides; measure it on the
        # force scale of the same mesh under the deformed state.
        gscale = max(np.abs(d[f"{mat}_{name}_g"]).max(), np.abs(d[f"{mat}_def_g"]).max())
        assert np.abs(g - d[f"{mat}_{name}_g"]).max() <= TOL * gscale
        Eref = float(d[f"{mat}_{name}_E"])
        assert abs(E - Eref) <= TOL * max(abs(Eref), abs(float(d[f"{mat}_def_E"])))
        # reference pattern is the structural pattern minus exact cancellations
        Qs = Q.tocoo()
        present = set(zip(ref.tocoo().row.tolist(), ref.tocoo().col.tolist()))
        extra = np.array([v for r, c, v in zip(Qs.row, Qs.col, Qs.data) if (r, c) not in present])
        assert len(present) <= Q.nnz
        assert extra.size == 0 or np.abs(extra).max() <= 1e-12 * np.abs(ref.data).max()
        if name == "def":
            assert np.array_equal(Q.indptr, ref.indptr) and np.array_equal(Q.indices, ref.indices)
        else:
            # At rest the reference's pattern is set by LAPACK roundoff: it keeps ~1,800 mathematically
            # zero entries as values between 1e-30 and 1e-8 of the scale and drops a few dozen that
            # happen to round to exactly 0.0. The entries above roundoff (> 1e-9 of the scale, the
            # smallest real entry is ~7e-3 of it, the largest roundoff one ~1e-11) form the same pattern,
            # index for index.
            def significant(A):
                A = A.tocsr().copy()
                A.data[np.abs(A.data) <= 1e-9 * np.abs(ref.data).max()] = 0.0
                A.eliminate_zeros()
                A.sort_indices()
                return A
            Qs_, Rs_ = significant(Q), significant(ref)
            assert np.array_equal(Qs_.indptr, Rs_.indptr) and np.array_equal(Qs_.indices, Rs_.indices)


_DISPATCH = {"arap": "arap", "linear": "linear-elasticity", "mmnh": "macklin-mueller-neo-hookean", "fcr": "fcr"}


def x_tier(mat):
    """(energy_x, gradient_x, hessian_x) of the drop-in API for an oracle material key."""
    if mat == "snh":
        from simkit_amd.energies import stable_neo_hookean as snh
        return (snh.stable_neo_hookean_energy_x, snh.stable_neo_hookean_gradient_x,
                snh.stable_neo_hookean_hessian_x)
    from simkit_amd.energies import elastic
    name = _DISPATCH[mat]
    return (lambda x, J, mu, lam, vol: elastic.elastic_energy_x(x, J, mu, lam, vol, name),
            lambda x, J, mu, lam, vol: elastic.elastic_gradient_x(x, J, mu, lam, vol, name),
            lambda x, J, mu, lam, vol, psd=True: elastic.elastic_hessian_x(x, J, mu, lam, vol, name, psd=psd))


@pytest.mark.parametrize("mat", ["snh", "arap", "linear", "mmnh", "fcr"])
@pytest.mark.parametrize("jit", [0.0, 0.2])
def test_x_tier_vs_oracle_cube10(mat, jit):
    import simkit_amd as sk
    X, T, x, mu, lam = problem(10, seed=3, jitter=jit, deform=0.3)
    J = sk.deformation_jacobian(X, T)
    vol = sk.volume(X, T)
    D = fem.tet_operator(X, T)
    Jo = fem.deformation_jacobian(D, T, X.shape[0])
    volo = fem.tet_volumes(X, T)
    xf = x.reshape(-1)
    energy_x, gradient_x, hessian_x = x_tier(mat)
    for psd in (True, False):
        Q = hessian_x(x, J, mu, lam, vol, psd=psd)
        # linear elasticity never projects (its Hessian is constant and PSD)
        Qo = fem.hessian_x(mat, xf, Jo, mu, lam, volo, psd=psd and mat != "linear")
        diff = (Q - Qo).tocsr()
        # row-wise relative error against each row's largest entry
        rowmax = np.maximum(abs(Qo).max(axis=1).toarray().ravel(), 1e-300)
        assert (abs(diff).max(axis=1).toarray().ravel() / rowmax).max() < TOL
    E = energy_x(x, J, mu, lam, vol)
    g = gradient_x(x, J, mu, lam, vol)
    Eo = fem.energy_x(mat, xf, Jo, mu, lam, volo)
    assert abs(E - Eo) <= TOL * abs(Eo)
    assert rel_err(g, fem.gradient_x(mat, xf, Jo, mu, lam, volo)) < TOL


def test_per_element_energies_and_u_tier():
    import simkit_amd as sk
    from simkit_amd._device import dev, host, per_element
    from simkit_amd._mesh import mesh_for_jacobian
    from simkit_amd.energies import stable_neo_hookean as snh
    X, T, x, mu, lam = problem(8, seed=4, deform=0.2)
    J = sk.deformation_jacobian(X, T)
    vol = sk.volume(X, T)
    mesh = mesh_for_jacobian(J)
    E, elem = mesh.energy(0, dev(x.reshape(-1)), per_element(mu, mesh.nt), per_element(lam, mesh.nt),
                          per_element(vol, mesh.nt), per_element_out=True)
    D = fem.tet_operator(X, T)
    Jo = fem.deformation_jacobian(D, T, X.shape[0])
    eo = fem.element_energies_x("snh", x.reshape(-1), Jo, mu, lam, fem.tet_volumes(X, T))
    assert np.all(np.abs(host(elem) - eo) <= TOL * np.abs(eo))
    # u tier: F = J u + J x_bar equals the x tier at x_bar + u
    rng = np.random.default_rng(5)
    x_bar = X + 0.01 * rng.standard_normal(X.shape)
    u = x - x_bar
    Jx_bar = J @ x_bar.reshape(-1, 1)
    assert abs(snh.stable_neo_hookean_energy_u(u, J, Jx_bar, mu, lam, vol) - E) <= 1e-12 * abs(E)
    assert rel_err(snh.stable_neo_hookean_gradient_u(u, J, Jx_bar, mu, lam, vol),
                   snh.stable_neo_hookean_gradient_x(x, J, mu, lam, vol)) < 1e-11
    Qu = snh.stable_neo_hookean_hessian_u(u, J, Jx_bar, mu, lam, vol)
    Qx = snh.stable_neo_hookean_hessian_x(x, J, mu, lam, vol)
    assert rel_err(Qu.data, Qx.data) < 1e-11


def test_foreign_jacobian_and_self_contained_tier():
    """A J built outside this package (reference layout) is decoded and gives the same results."""
    import simkit_amd as sk
    from simkit_amd.energies import stable_neo_hookean as snh
    X, T, x, mu, lam = problem(5, seed=6, deform=0.2)
    D = fem.tet_operator(X, T)
    Jo = fem.deformation_jacobian(D, T, X.shape[0])   # plain scipy matrix, no device mesh attached
    vol = fem.tet_volumes(X, T).reshape(-1, 1)
    Q1 = snh.stable_neo_hookean_hessian_x(x, Jo, mu, lam, vol)
    Q2 = snh.stable_neo_hookean_hessian(X, T, mu, lam, U=x)
    assert rel_err(Q1.toarray(), Q2.toarray()) < 1e-12
    E1 = snh.stable_neo_hookean_energy_x(x, Jo, mu, lam, vol)
    assert abs(E1 - snh.stable_neo_hookean_energy(X, T, mu, lam, U=x)) <= 1e-12 * abs(E1)


def test_assembly_is_bitwise_deterministic():
    import simkit_amd as sk
    from simkit_amd.energies import stable_neo_hookean as snh
    X, T, x, mu, lam = problem(12, seed=7, deform=0.3)
    J = sk.deformation_jacobian(X, T)
    vol = sk.volume(X, T)
    Q1 = snh.stable_neo_hookean_hessian_x(x, J, mu, lam, vol)
    Q2 = snh.stable_neo_hookean_hessian_x(x, J, mu, lam, vol)
    g1 = snh.stable_neo_hookean_gradient_x(x, J, mu, lam, vol)
    g2 = snh.stable_neo_hookean_gradient_x(x, J, mu, lam, vol)
    assert np.array_equal(Q1.data, Q2.data) and np.array_equal(g1, g2)
