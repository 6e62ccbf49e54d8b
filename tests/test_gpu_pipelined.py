"""Pipelined block-Jacobi PCG (one launch per iteration, Ghysels-Vanroose
recurrences on A B^-1, FP64 path): the same trajectories as the classic
two-launch PCG and the oracle; iteration counts within a few percent of the
classic form (equal in exact arithmetic); determinism; the breakdown status on
an indefinite system. The mixed path keeps the classic form when the pipelined
one is asked for (its FP64 residual replacement breaks the pipelined
recurrences in FP32), bit for bit.

Position tolerance: 1e-6 relative per kept step, as for the classic path."""

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


@pytest.mark.parametrize("precision", ["fp64"])
def test_pipelined_cube20_matches_oracle_and_classic(precision):
    path = os.path.join(GOLDEN, "oracle_c2_cube20.npz")
    if not os.path.exists(path):
        pytest.skip("oracle_c2_cube20.npz not generated")
    d = golden("oracle_c2_cube20.npz")
    from simkit_amd.sim import top_face_sinusoid
    X, top, sim = _cube_sim(20, precision, pcg="pipelined")
    kept, infos = sim.rollout(20, bc=top_face_sinusoid(X, top, 0.1, 1.0), keep={1, 10, 20})
    assert all(i["status"] == 0 for i in infos), [(i["status"], i["newton_iters"], i["pcg_iters"],
                                                   i["refine_steps"]) for i in infos]
    for s in (1, 10, 20):
        ref = d[f"x_{s}"]
        assert np.abs(kept[s] - ref).max() / np.abs(ref).max() < 1e-6, s
    _, _, cl = _cube_sim(20, precision)
    ckept, cinfos = cl.rollout(20, bc=top_face_sinusoid(X, top, 0.1, 1.0), keep={20})
    assert np.abs(kept[20] - ckept[20]).max() / np.abs(ckept[20]).max() < 1e-6
    itp, itc = sum(i["pcg_iters"] for i in infos), sum(i["pcg_iters"] for i in cinfos)
    print(f"pipelined {precision}: pcg {itp} vs classic {itc}", flush=True)
    assert abs(itp - itc) <= 0.05 * itc, (itp, itc)
    assert [i["newton_iters"] for i in infos] == [i["newton_iters"] for i in cinfos]


@pytest.mark.parametrize("mat", ["snh", "arap"])
def test_pipelined_fp64_jittered_matches_oracle(mat):
    from simkit_amd.mesh import jitter
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    n = 12
    X, T = unit_cube(n)
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    X = jitter(X, 1.0 / n, 0.2, seed=0, keep=np.concatenate([bottom, top]))
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    P = lambda t: X + np.outer(np.isin(np.arange(X.shape[0]), top), [0.1 * np.sin(2 * np.pi * t), 0, 0])
    steps = 6
    ref = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, np.concatenate([bottom, top]), material=mat,
                           prescribe=P, linsolve="spsolve").rollout(steps)
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, material=mat,
                           fixed_vertices=np.concatenate([bottom, top]), pcg_rtol=1e-8, pcg="pipelined")
    kept, infos = sim.rollout(steps, bc=top_face_sinusoid(X, top, 0.1, 1.0))
    assert all(i["status"] == 0 for i in infos), infos
    for s in range(1, steps + 1):
        err = np.abs(kept[s] - ref[s - 1]).max() / np.abs(ref[s - 1]).max()
        assert err < 1e-6, (s, err)


def test_mixed_path_keeps_the_classic_form():
    from simkit_amd.sim import top_face_sinusoid
    X, top, a = _cube_sim(8, "mixed", pcg="pipelined")
    _, _, b = _cube_sim(8, "mixed")
    ka, ia = a.rollout(4, bc=top_face_sinusoid(X, top, 0.1, 1.0))
    kb, ib = b.rollout(4, bc=top_face_sinusoid(X, top, 0.1, 1.0))
    assert all(np.array_equal(ka[k], kb[k]) for k in ka)
    assert [i["pcg_iters"] for i in ia] == [i["pcg_iters"] for i in ib]


@pytest.mark.parametrize("precision", ["fp64"])
def test_pipelined_gravity_cube_deterministic_and_breakdown(precision):
    import torch
    from simkit_amd import _lib
    from simkit_amd._device import dev
    from simkit_amd.sim import BackwardEulerSim
    X, T = unit_cube(10)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    mk = lambda **kw: BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60,
                                       fixed_vertices=np.where(X[:, 1] == 0)[0], gravity=-9.8, precision=precision,
                                       pcg="pipelined", **kw)
    sim = mk()
    a, ia = sim.rollout(3)
    b, ib = sim.rollout(3)
    assert all(np.array_equal(a[k], b[k]) for k in a) and ia == ib
    bad = mk(psd=False)
    half = dev(0.5 * X.reshape(-1))
    _, info = bad.step(half, half)
    assert info["status"] == _lib.SK_ERR_PCG_BREAKDOWN, info
    x = bad.X_dev.clone()
    y1, i1 = bad.step(x, x)
    y2, i2 = mk(psd=False).step(x, x)
    assert i1["status"] == 0 and torch.equal(y1, y2)
