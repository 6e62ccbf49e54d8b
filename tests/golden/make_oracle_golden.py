"""Oracle trajectories for the prescribed-motion cube (run in the build container).

Also, with `jittered <material> [cells]`, the jittered 48^3 cube of the
mixed-precision workload (sampled vertices, see jittered_cube).

Workload: 20^3-cell cube on [0,1]^3 (9261 vertices, 40000 tets), stable
neo-Hookean E=1e6, nu=0.45, rho=1000, dt=1/60, no gravity; vertices with y=0
held at rest, vertices with y=1 held at X + (0.1 sin(2 pi t), 0, 0). Newton
to ||g_free|| <= 1e-6 ||g_free(x0)||; the linear solves use the oracle's
block-Jacobi PCG at rtol 1e-12 (the same Newton iterates as a direct solve
to ~1e-11, and minutes instead of an hour of SuperLU at this size).
Writes tests/golden/oracle_c2_cube20.npz with the states after steps 1, 10, 20.
"""

import os
import sys
import time

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(OUT, "..", "..")))

from oracle import fem, solver  # noqa: E402
from oracle.mesh import unit_cube  # noqa: E402


def prescribed_cube(n, steps=20, keep=(1, 10, 20), material="snh"):
    X, T = unit_cube(n)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    bottom = np.where(X[:, 1] == 0.0)[0]
    top = np.where(X[:, 1] == 1.0)[0]

    def prescribe(t):
        P = X.copy()
        P[top, 0] += 0.1 * np.sin(2.0 * np.pi * t)
        return P

    p = solver.BEProblem(X, T, mu, lam, 1000.0, 1.0 / 60.0, np.concatenate([bottom, top]), material=material,
                         prescribe=prescribe, linsolve="pcg", pcg_rtol=1e-12)
    t0 = time.time()
    states, infos = p.rollout(steps, return_info=True)
    print(f"{n}^3 {material}: {time.time() - t0:.1f} s, newton/step {[i['newton'] for i in infos]}", flush=True)
    return {f"x_{s}": states[s - 1] for s in keep}, infos


JIT_STEPS = (1, 5, 10, 15, 20)
JIT_SAMPLE = 1500


def jittered_cube(n, material, steps=20, amount=0.2, seed=0):
    """The mixed-precision workload: every rest vertex off the two held faces displaced by
    U(-0.2, 0.2) cells per axis (numpy default_rng(seed)), then the prescribed-motion cube above.
    PCG at rtol 1e-10. Returns the states after JIT_STEPS at a seeded sample of JIT_SAMPLE
    vertices."""
    X, T = unit_cube(n)
    bottom = np.where(X[:, 1] == 0.0)[0]
    top = np.where(X[:, 1] == 1.0)[0]
    d = np.random.default_rng(seed).uniform(-amount / n, amount / n, size=X.shape)
    d[np.concatenate([bottom, top])] = 0.0
    X = X + d
    mu, lam = fem.ympr_to_lame(1e6, 0.45)

    def prescribe(t):
        P = X.copy()
        P[top, 0] += 0.1 * np.sin(2.0 * np.pi * t)
        return P

    p = solver.BEProblem(X, T, mu, lam, 1000.0, 1.0 / 60.0, np.concatenate([bottom, top]), material=material,
                         prescribe=prescribe, linsolve="pcg", pcg_rtol=1e-10)
    t0 = time.time()
    states, infos = p.rollout(steps, return_info=True)
    print(f"jittered {n}^3 {material}: {time.time() - t0:.1f} s, newton/step {[i['newton'] for i in infos]}",
          flush=True)
    idx = np.sort(np.random.default_rng(1).choice(X.shape[0], JIT_SAMPLE, replace=False))
    kept = {f"x_{s}": states[s - 1].reshape(-1, 3)[idx] for s in JIT_STEPS}
    return kept, idx, infos


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "jittered":
        # python make_oracle_golden.py jittered snh|arap [cells]: oracle_c3_jittered48_<material>.npz
        mat = sys.argv[2]
        cells = int(sys.argv[3]) if len(sys.argv) > 3 else 48
        kept, idx, infos = jittered_cube(cells, mat)
        np.savez_compressed(os.path.join(OUT, f"oracle_c3_jittered{cells}_{mat}.npz"), steps=np.array(JIT_STEPS),
                            idx=idx, newton=np.array([i["newton"] for i in infos]), **kept)
    else:
        kept, infos = prescribed_cube(20)
        np.savez_compressed(os.path.join(OUT, "oracle_c2_cube20.npz"), steps=np.array([1, 10, 20]),
                            newton=np.array([i["newton"] for i in infos]), **kept)
