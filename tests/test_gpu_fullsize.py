"""Full-size workloads on the GPU, checked through size-independent properties.

* The 48^3-cell prescribed-motion cube (117,649 vertices, 552,960 tets; plain
  FP64 path, block-Jacobi PCG at rtol 1e-8): two backward-Euler steps. The
  Newton stopping test ||g_free(x_next)|| <= 1e-6 ||g_free(x0)|| is recomputed
  independently (the x-tier gradient on a separately built mesh in the
  caller's vertex numbering, plus the kinetic term in numpy), and the
  incremental potential must have decreased (backward_euler.py:76-91).
* The 512-beam batch (lengths U{8..64} cells, 1.5 M tets, FP32 + FP64
  refinement, two-level per-problem PCG, graph-captured Newton iteration):
  one step from rest; every problem's status is 0 and its stopping test holds
  when recomputed the same way per beam.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _potential_parts(mesh, material, x, mu, lam, vol, mass, xt, h, fext=None):
    """(free-DOF gradient pieces, energy) of V(x) + 1/(2h^2) |x - xt|_M^2 - f.x, V from the x tier."""
    from simkit_amd._device import dev, host
    xd = dev(x)
    g = host(mesh.gradient(material, xd, mu, lam, vol)) + mass * (x - xt) / h ** 2
    E = mesh.energy(material, xd, mu, lam, vol) + 0.5 / h ** 2 * float(((x - xt) ** 2 * mass).sum())
    if fext is not None:
        g = g - fext
        E -= float(fext @ x)
    return g, E


def test_cube48_two_fp64_steps_meet_the_newton_test():
    import torch
    from simkit_amd import _lib
    from simkit_amd._device import host, per_element
    from simkit_amd._mesh import mesh_for_rest
    from simkit_amd.mesh import unit_cube
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    from simkit_amd.ympr_to_lame import ympr_to_lame
    n, h = 48, 1.0 / 60.0
    X, T = unit_cube(n)
    assert X.shape[0] == 117649 and T.shape[0] == 552960
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    fixed_v = np.concatenate([bottom, top])
    mu, lam = ympr_to_lame(1e6, 0.45)
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=h, fixed_vertices=fixed_v, pcg_rtol=1e-8)
    bc = top_face_sinusoid(X, top, 0.1, 1.0)
    mesh, vol = mesh_for_rest(X, T)                      # the caller's numbering, built separately
    mass = np.repeat(host(mesh.lumped_mass(vol, 1000.0)), 3)
    mu_d, lam_d = per_element(mu, T.shape[0]), per_element(lam, T.shape[0])
    free = np.ones(X.size, bool)
    for c in range(3):
        free[3 * fixed_v + c] = False
    prev = sim.X_dev.clone()
    curr = prev.clone()
    for step in range(2):
        xb = bc(step, (step + 1) * h)
        x_next, info = sim.step(curr, prev, xb)
        print(f"cube48 step {step}: newton {info['newton_iters']}, pcg {info['pcg_iters']}", flush=True)
        assert info["status"] == _lib.SK_OK, info
        c_, p_, b_ = host(curr), host(prev), host(xb)
        xt = c_ + h * ((c_ - p_) / h)
        x0 = np.where(free, xt, b_)
        x1 = host(x_next)
        assert np.array_equal(x1[~free], b_[~free])           # fixed DOFs at their prescribed values
        g0, E0 = _potential_parts(mesh, _lib.SK_SNH, x0, mu_d, lam_d, vol, mass, xt, h)
        g1, E1 = _potential_parts(mesh, _lib.SK_SNH, x1, mu_d, lam_d, vol, mass, xt, h)
        n0, n1 = np.linalg.norm(g0[free]), np.linalg.norm(g1[free])
        assert abs(n0 - info["g0_norm"]) <= 1e-9 * n0
        assert n1 <= 1e-6 * n0 * (1 + 1e-6), (n1 / n0)
        assert E1 < E0
        prev, curr = curr, x_next
    torch.cuda.synchronize()


def test_batch512_one_step_meets_every_beams_newton_test():
    from simkit_amd import _lib
    from simkit_amd._device import host
    from simkit_amd.batch import cantilever_batch
    sim, desc = cantilever_batch(512, coarse_rows=64)
    assert desc["problems"] == 512 and max(desc["lengths"]) <= 64 and min(desc["lengths"]) >= 8
    print(f"batch: {desc['vertices']} vertices, {desc['tets']} tets", flush=True)
    x = sim.X_dev.clone()
    x_next, infos, rounds = sim.step(x, x)
    assert rounds > 0 and all(i["status"] == _lib.SK_OK for i in infos), \
        sorted(set(i["status"] for i in infos))
    h = sim.h
    X = sim.X.reshape(-1)
    x1 = host(x_next)
    mass = host(sim.mass)
    fext = host(sim.f_ext)
    fixed = host(sim.fixed).astype(bool)
    g0, _ = _potential_parts(sim.mesh, sim.material, X, sim.mu, sim.lam, sim.vol, mass, X, h, fext)
    g1, _ = _potential_parts(sim.mesh, sim.material, x1, sim.mu, sim.lam, sim.vol, mass, X, h, fext)
    g0[fixed] = 0.0
    g1[fixed] = 0.0
    assert np.array_equal(x1[fixed], X[fixed])
    for p in range(sim.P):
        s = sim.problem_slice(p)
        n0, n1 = np.linalg.norm(g0[s]), np.linalg.norm(g1[s])
        assert abs(n0 - infos[p]["g0_norm"]) <= 1e-9 * n0, p
        assert infos[p]["stalled"] or n1 <= 1e-6 * n0 * (1 + 1e-6), (p, n1 / n0, infos[p])


@pytest.mark.parametrize("mat", ["snh", "arap"])
def test_jittered_cube48_mixed_and_fp64_match_the_oracle(mat):
    """The mixed-precision workload at full size: the jittered 48^3 prescribed-motion cube (every rest
    vertex off the held faces moved by U(-0.2, 0.2) cells, seed 0), 20 steps in FP32 + FP64 refinement
    and in plain FP64, against the FP64 oracle's trajectory (tests/golden/make_oracle_golden.py
    jittered <material>: block-Jacobi PCG at rtol 1e-10) at a seeded sample of 1,500 vertices after steps
    1, 5, 10, 15 and 20, within 1e-6 of the sampled positions' scale per step."""
    import os
    from conftest import GOLDEN, golden
    from simkit_amd.mesh import jitter, unit_cube
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    from simkit_amd.ympr_to_lame import ympr_to_lame
    name = f"oracle_c3_jittered48_{mat}.npz"
    if not os.path.exists(os.path.join(GOLDEN, name)):
        pytest.skip(f"{name} not generated")
    d = golden(name)
    X, T = unit_cube(48)
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    X = jitter(X, 1.0 / 48, 0.2, seed=0, keep=np.concatenate([bottom, top]))
    mu, lam = ympr_to_lame(1e6, 0.45)
    steps = [int(s) for s in d["steps"]]
    idx = d["idx"]
    for prec in ("mixed", "fp64"):
        sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1.0 / 60.0, material=mat,
                               fixed_vertices=np.concatenate([bottom, top]), pcg_rtol=1e-8, precision=prec)
        kept, infos = sim.rollout(max(steps), bc=top_face_sinusoid(X, top, 0.1, 1.0), keep=set(steps))
        assert all(i["status"] == 0 for i in infos), [i["status"] for i in infos]
        for s in steps:
            ref = d[f"x_{s}"]
            got = kept[s].reshape(-1, 3)[idx]
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"jittered48 {mat} {prec} step {s}: rel err {err:.2e}", flush=True)
            assert err < 1e-6, (prec, s, err)
        del sim
