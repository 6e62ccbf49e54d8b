"""Batched steps (sk_batch_*): every problem of a batch of uneven cantilever
beams (and of a batch of differently shaped meshes) against the same mesh
advanced alone by the single-mesh simulator
(mixed precision and FP64) and by the oracle's FP64 harness, per kept step
at 1e-6 relative (max |x - x_ref| / max |x_ref| of that beam); every
per-problem kernel shape (vectors in global memory or in LDS), with and
without the hipGraph replay; bitwise determinism; the autotuner and its
cache."""

import json

import numpy as np
import pytest

from oracle import fem, solver

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _beams(seed, lengths):
    from simkit_amd.mesh import beam
    rng = np.random.default_rng(seed)
    meshes, fixed = [], []
    for L in lengths:
        X, T = beam(int(L), 4, 0.025)
        meshes.append((X, T))
        fixed.append(np.where(X[:, 0] == 0.0)[0])
    E = np.exp(rng.uniform(np.log(1e5), np.log(1e7), len(lengths)))
    nu = rng.uniform(0.30, 0.48, len(lengths))
    mu, lam = fem.ympr_to_lame(E, nu)
    return meshes, fixed, mu, lam


def _batch(meshes, fixed, mu, lam, **kw):
    from simkit_amd.batch import BatchSim
    return BatchSim(meshes, mu, lam, rho=1000.0, fixed_vertices=fixed, gravity=-9.8, **kw)


@pytest.mark.parametrize("block,lanes,lds,graph", [(256, 1, 0, True), (128, 4, 0, True), (1024, 4, 0, False),
                                                    (512, 1, 0, False), (256, 1, 1, True), (128, 4, 1, False),
                                                    (1024, 1, 1, True), (512, 4, 1, True)])
def test_batch_matches_single_mesh_sims(block, lanes, lds, graph):
    from simkit_amd.sim import BackwardEulerSim
    lengths = [8, 13, 9, 16, 11, 10]
    meshes, fixed, mu, lam = _beams(1, lengths)
    steps = 8
    b = _batch(meshes, fixed, mu, lam, block=block, lanes=lanes, lds=lds, use_graph=graph)
    kept, infos, rounds = b.rollout(steps)
    assert all(i["status"] == 0 for step in infos for i in step), infos
    assert max(rounds) == max(i["newton_iters"] for i in infos[-1])
    for p, (X, T) in enumerate(meshes):
        s = BackwardEulerSim(X, T, mu[p], lam[p], rho=1000.0, fixed_vertices=fixed[p], gravity=-9.8,
                             precision="mixed")
        ref, sinfos = s.rollout(steps)
        sl = b.problem_slice(p)
        for n in range(1, steps + 1):
            err = np.abs(kept[n][sl] - ref[n]).max() / np.abs(ref[n]).max()
            assert err < TOL, (p, n, err)


def test_batch_one_beam_matches_oracle():
    lengths = [8, 12]
    meshes, fixed, mu, lam = _beams(2, lengths)
    steps = 5
    b = _batch(meshes, fixed, mu, lam)
    kept, infos, _ = b.rollout(steps)
    for p in range(2):
        X, T = meshes[p]
        ref = solver.BEProblem(X, T, mu[p], lam[p], 1000.0, 1 / 60, fixed[p], gravity=-9.8).rollout(steps)
        sl = b.problem_slice(p)
        for n in range(steps):
            err = np.abs(kept[n + 1][sl] - ref[n]).max() / np.abs(ref[n]).max()
            assert err < TOL, (p, n, err)


def test_batch_deterministic_and_autotune(tmp_path):
    meshes, fixed, mu, lam = _beams(3, [8, 20, 9, 14])
    b = _batch(meshes, fixed, mu, lam)
    cache = tmp_path / "tune.json"
    tuned = b.autotune(reps=1, cache=str(cache))
    assert tuned["block"] in (128, 256, 512, 1024) and tuned["lanes"] in (1, 4) and tuned["lds"] in (0, 1)
    assert len(tuned["ms"]) == 16 and all(v > 0 for v in tuned["ms"].values())
    table = json.loads(cache.read_text())
    assert len(table) == 1 and list(table.values())[0]["block"] == tuned["block"]
    k1, i1, r1 = b.rollout(4)
    k2, i2, r2 = b.rollout(4)
    assert all(np.array_equal(k1[n], k2[n]) for n in k1) and i1 == i2 and r1 == r2
    b2 = _batch(meshes, fixed, mu, lam)
    assert b2.autotune(cache=str(cache)) == tuned      # served from the cache
    k3, _, _ = b2.rollout(4)
    assert all(np.array_equal(k1[n], k3[n]) for n in k1)


def test_damped_stiff_beam_stops_at_position_resolution():
    """Short stiff beams come to rest; the single-mesh simulator, the batch and the oracle all stop
    on the step-length test there (status OK, stalled) and agree per step."""
    from simkit_amd.sim import BackwardEulerSim
    from simkit_amd.mesh import beam
    X, T = beam(8, 4, 0.025)
    mu, lam = fem.ympr_to_lame(6.22e6, 0.479)
    fixed = np.where(X[:, 0] == 0.0)[0]
    steps = 12
    ref, rinfo = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, fixed, gravity=-9.8).rollout(steps, return_info=True)
    assert any(i["stalled"] for i in rinfo)
    s = BackwardEulerSim(X, T, mu, lam, rho=1000.0, fixed_vertices=fixed, gravity=-9.8, precision="mixed")
    kept, infos = s.rollout(steps)
    assert all(i["status"] == 0 for i in infos) and any(i["stalled"] for i in infos)
    b = _batch([(X, T), (X, T)], [fixed, fixed], np.array([mu, mu]), np.array([lam, lam]))
    bk, binfo, _ = b.rollout(steps)
    assert all(i["status"] == 0 for st in binfo for i in st) and any(st[0]["stalled"] for st in binfo)
    for n in range(1, steps + 1):
        for x in (kept[n], bk[n][b.problem_slice(0)], bk[n][b.problem_slice(1)]):
            assert np.abs(x - ref[n - 1]).max() <= TOL * np.abs(ref[n - 1]).max()


@pytest.mark.parametrize("block,lanes,lds,graph,rows", [(256, 1, 0, True, 64), (1024, 4, 1, True, 64),
                                                         (128, 4, 0, False, 128), (512, 1, 1, False, 64),
                                                         (1024, 1, 0, True, 192)])
def test_batch_two_level_matches_single_mesh_sims(block, lanes, lds, graph, rows):
    """Per-problem two-level PCG (rigid-motion aggregates of consecutive vertices): every beam still
    matches the same beam advanced alone by the single-mesh simulator, per step at 1e-6."""
    from simkit_amd.sim import BackwardEulerSim
    lengths = [8, 13, 9, 16, 11, 10]
    meshes, fixed, mu, lam = _beams(1, lengths)
    steps = 6
    b = _batch(meshes, fixed, mu, lam, block=block, lanes=lanes, lds=lds, use_graph=graph, coarse_rows=rows)
    kept, infos, rounds = b.rollout(steps)
    assert all(i["status"] == 0 for step in infos for i in step), infos
    for p, (X, T) in enumerate(meshes):
        s = BackwardEulerSim(X, T, mu[p], lam[p], rho=1000.0, fixed_vertices=fixed[p], gravity=-9.8,
                             precision="mixed")
        ref, _ = s.rollout(steps)
        sl = b.problem_slice(p)
        for n in range(1, steps + 1):
            err = np.abs(kept[n][sl] - ref[n]).max() / np.abs(ref[n]).max()
            assert err < TOL, (p, n, err)


def test_batch_two_level_fewer_iterations_deterministic():
    meshes, fixed, mu, lam = _beams(4, [40, 64, 25])
    plain = _batch(meshes, fixed, mu, lam)
    _, pinfo, _ = plain.rollout(4)
    k1, i1, r1 = _batch(meshes, fixed, mu, lam, coarse_rows=64).rollout(4)
    k2, i2, r2 = _batch(meshes, fixed, mu, lam, coarse_rows=64).rollout(4)
    assert all(np.array_equal(k1[n], k2[n]) for n in k1) and i1 == i2 and r1 == r2
    it_plain = sum(i["pcg_iters"] for st in pinfo for i in st)
    it_two = sum(i["pcg_iters"] for st in i1 for i in st)
    assert it_two < 0.5 * it_plain, (it_two, it_plain)


def test_batch_two_level_coarse_period():
    """Coarse operators re-formed once, every step or every third step: the same positions to the
    PCG tolerance; a second rollout of a period-3 batch starts from the operators left by the first
    and still agrees; a negative period is rejected."""
    import ctypes
    from simkit_amd import _lib
    meshes, fixed, mu, lam = _beams(6, [30, 12, 21])
    steps = 7
    runs = {k: _batch(meshes, fixed, mu, lam, coarse_rows=64, coarse_period=k) for k in (0, 1, 3)}
    out = {k: b.rollout(steps) for k, b in runs.items()}
    again = runs[3].rollout(steps)
    for k, (kept, infos, _) in list(out.items()) + [("3 again", again)]:
        assert all(i["status"] == 0 for st in infos for i in st), (k, infos)
        for n in range(1, steps + 1):
            ref = out[1][0][n]
            assert np.abs(kept[n] - ref).max() <= TOL * np.abs(ref).max(), (k, n)
    b = runs[3]
    X = np.ascontiguousarray(b.X, dtype=np.float64)
    with pytest.raises(_lib.SimkitAmdError, match="period"):
        _lib.call("sk_batch_set_coarse", b._h, X.ctypes.data_as(ctypes.c_void_p), 64, -1)


@pytest.mark.parametrize("rows,lds", [(0, 0), (0, 1), (64, 1)])
def test_batch_mixed_shapes_match_single_mesh_sims(rows, lds):
    """Problems of different shapes in one batch (a cube, a beam, a jittered cube, a second beam):
    rows of 5 to 19 blocks share the sliced matrix's 64-row chunks, pinned faces differ per
    problem, and each problem still matches the same mesh advanced alone, per step at 1e-6."""
    from simkit_amd.mesh import beam, jitter, unit_cube
    from simkit_amd.sim import BackwardEulerSim
    meshes, fixed = [], []
    for kind in ("cube", "beam", "jcube", "beam2"):
        if kind == "cube":
            X, T = unit_cube(5)
            X = 0.2 * X
        elif kind == "jcube":
            X, T = unit_cube(4)
            X = 0.15 * jitter(X, 1.0 / 4, 0.2, seed=2, keep=np.where(X[:, 0] == 0.0)[0])
        else:
            X, T = beam(9 if kind == "beam" else 14, 4 if kind == "beam" else 3, 0.025)
        meshes.append((X, T))
        fixed.append(np.where(X[:, 0] == 0.0)[0])
    mu, lam = fem.ympr_to_lame(np.array([2e5, 1e6, 5e5, 3e6]), np.array([0.35, 0.45, 0.4, 0.3]))
    steps = 5
    b = _batch(meshes, fixed, mu, lam, block=256, lanes=1, lds=lds, coarse_rows=rows)
    kept, infos, _ = b.rollout(steps)
    assert all(i["status"] == 0 for st in infos for i in st), infos
    for p, (X, T) in enumerate(meshes):
        s = BackwardEulerSim(X, T, mu[p], lam[p], rho=1000.0, fixed_vertices=fixed[p], gravity=-9.8,
                             precision="mixed")
        ref, _ = s.rollout(steps)
        sl = b.problem_slice(p)
        for n in range(1, steps + 1):
            err = np.abs(kept[n][sl] - ref[n]).max() / np.abs(ref[n]).max()
            assert err < TOL, (p, n, err)


def test_batch_two_level_bad_rows():
    from simkit_amd import _lib
    meshes, fixed, mu, lam = _beams(5, [8, 9])
    with pytest.raises(_lib.SimkitAmdError, match="agg_rows"):
        _batch(meshes, fixed, mu, lam, coarse_rows=100)
