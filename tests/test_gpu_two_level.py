"""Two-level PCG preconditioner (block-Jacobi + piecewise-rigid coarse space):
the same trajectories as the block-Jacobi path and the oracle, in FP64 and in
mixed precision, with fewer PCG iterations; the coarse space's shape on a cube
and on a slender beam (one aggregate across the section); determinism; the
coarse-size limit.

Position tolerance: 1e-6 relative per kept step (max |x - x_ref| / max |x_ref|),
as for the block-Jacobi path: PCG stops at the same relative residual either
way, so only the Newton iterates' rounding-level differences remain."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from oracle import fem, solver
from oracle.mesh import unit_cube

pytestmark = pytest.mark.gpu


def _cube_sim(n, precision="fp64", **kw):
    from simkit_amd.sim import BackwardEulerSim
    X, T = unit_cube(n)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, fixed_vertices=np.concatenate([bottom, top]),
                           pcg_rtol=1e-8, precision=precision, **kw)
    return X, top, sim


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_level_cube20_matches_oracle_with_fewer_iterations(precision):
    path = os.path.join(GOLDEN, "oracle_c2_cube20.npz")
    if not os.path.exists(path):
        pytest.skip("oracle_c2_cube20.npz not generated")
    d = golden("oracle_c2_cube20.npz")
    from simkit_amd.sim import top_face_sinusoid
    X, top, sim = _cube_sim(20, precision, preconditioner="two_level", coarse_vertices=125)
    info0 = sim.coarse_info()
    assert info0["aggregates"] > 1 and not info0["built"]
    assert 0 < info0["coarse_unknowns"] <= 6 * info0["aggregates"]
    kept, infos = sim.rollout(20, bc=top_face_sinusoid(X, top, 0.1, 1.0), keep={1, 10, 20})
    assert all(i["status"] == 0 for i in infos), infos
    assert sim.coarse_info()["built"]
    for s in (1, 10, 20):
        ref = d[f"x_{s}"]
        assert np.abs(kept[s] - ref).max() / np.abs(ref).max() < 1e-6, s
    _, _, bj = _cube_sim(20, precision)
    _, binfos = bj.rollout(20, bc=top_face_sinusoid(X, top, 0.1, 1.0), keep=set())
    it2, it1 = sum(i["pcg_iters"] for i in infos), sum(i["pcg_iters"] for i in binfos)
    assert it2 < 0.6 * it1, (it2, it1)


def test_two_level_jittered_arap_mixed_matches_oracle():
    """ARAP on a jittered 12^3 cube (irregular aggregates' rigid motions) in mixed precision."""
    from simkit_amd.mesh import jitter
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    n = 12
    X, T = unit_cube(n)
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    X = jitter(X, 1.0 / n, 0.2, seed=0, keep=np.concatenate([bottom, top]))
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    P = lambda t: X + np.outer(np.isin(np.arange(X.shape[0]), top), [0.1 * np.sin(2 * np.pi * t), 0, 0])
    steps = 6
    ref = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, np.concatenate([bottom, top]), material="arap",
                           prescribe=P, linsolve="spsolve").rollout(steps)
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, material="arap",
                           fixed_vertices=np.concatenate([bottom, top]), precision="mixed", pcg_rtol=1e-8,
                           preconditioner="two_level", coarse_vertices=64)
    kept, infos = sim.rollout(steps, bc=top_face_sinusoid(X, top, 0.1, 1.0))
    assert all(i["status"] == 0 for i in infos), infos
    for s in range(1, steps + 1):
        err = np.abs(kept[s] - ref[s - 1]).max() / np.abs(ref[s - 1]).max()
        assert err < 1e-6, (s, err)


def test_two_level_beam_gravity_matches_block_jacobi():
    """A slender cantilever under gravity: the bins span the whole cross-section."""
    from simkit_amd.mesh import beam
    from simkit_amd.sim import BackwardEulerSim
    X, T = beam(40, 4, 0.025)
    fixed = np.where(X[:, 0] == 0.0)[0]
    mu, lam = fem.ympr_to_lame(1e6, 0.4)
    runs = {}
    for pc in ("bjacobi", "two_level"):
        sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, fixed_vertices=fixed, gravity=-9.8,
                               preconditioner=pc, coarse_vertices=100)
        runs[pc] = sim.rollout(10, keep={5, 10})
        if pc == "two_level":
            ci = sim.coarse_info()
    (kb, ib), (kt, it) = runs["bjacobi"], runs["two_level"]
    assert all(i["status"] == 0 for i in it), it
    for s in (5, 10):
        assert np.abs(kt[s] - kb[s]).max() / np.abs(kb[s]).max() < 1e-6, s
    assert ci["aggregates"] >= 8 and ci["coarse_unknowns"] <= 6 * ci["aggregates"]
    assert sum(i["pcg_iters"] for i in it) < sum(i["pcg_iters"] for i in ib)


def test_two_level_deterministic_and_size_limit():
    """Bitwise determinism; the C entry point rejects a coarse space above its cap, and the
    simulator raises its aggregate size so that it never asks for one."""
    import ctypes
    from simkit_amd import _lib
    _, _, sim = _cube_sim(10, preconditioner="two_level", coarse_vertices=27)
    a, ia = sim.rollout(3)
    b, ib = sim.rollout(3)
    assert all(np.array_equal(a[k], b[k]) for k in a) and ia == ib
    _, _, big = _cube_sim(20, preconditioner="bjacobi")
    X = np.ascontiguousarray(unit_cube(20)[0])
    with pytest.raises(_lib.SimkitAmdError, match="coarse unknowns"):
        _lib.call("sk_sim_set_coarse", big._h, X.ctypes.data_as(ctypes.c_void_p), 1)
    _, _, auto = _cube_sim(20, preconditioner="two_level", coarse_vertices=1)
    info = auto.coarse_info()
    assert 1 < info["coarse_unknowns"] <= 4096 and info["vertices_per_aggregate"] >= 6 * 9261 / 4096
