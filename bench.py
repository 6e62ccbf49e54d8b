#!/usr/bin/env python
"""Backward-Euler steps/s on the 48^3-cell stable neo-Hookean cube, one MI355X.

Workload (prescribed-motion cube): 48^3 cells on [0,1]^3 (117,649 vertices,
552,960 tets, 352,947 DOF), stable neo-Hookean E=1e6, nu=0.45, rho=1000,
dt=1/60, no gravity; vertices with y=0 held at rest, vertices with y=1 held at
X + (0.1 sin(2 pi t), 0, 0). Each step: Newton with the 9x9-PSD-projected
Hessian to ||g_free|| <= 1e-6 ||g_free(x0)||, block-Jacobi PCG at rtol 1e-8,
Armijo backtracking. A "step" is one full backward-Euler step; the timed
region runs the first --steps steps from rest with the mesh resident in HBM.

Prints one JSON line (rank 0). Also measures, on the same stream with HIP
events, the average duration of the dominant kernel (the block-CSR PCG SpMV)
for the roofline object, and times the CPU oracle on a bounded sample of the
same workload for cpu_baseline.
"""

import argparse
import json
import os
import sys
import time

# The CPU baseline times single-threaded numpy/scipy (the reference path's sparse products, SuperLU and
# batched 9x9 eigh are serial anyway): pin the BLAS/OpenMP pools before numpy loads so the reported
# core count is the one used. The GPU legs do no host-side math that these pools would serve.
for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# element pass of the stable neo-Hookean density: the closed-form eigensystem launch (the Jacobi
# fallback launch that follows it runs only on the elements the first declines)
HESS_KERNEL = {"fp64": "elem_hess_kernel<SNH, 1, double, 1>", "mixed": "elem_hess_kernel<SNH, 1, float, 1>"}
# dominant kernel of each path (the FP32 SpMV walks chunk ranges with 1024 workgroups)
SPMV_KERNEL = {"fp64": "bpcg_spmv_kernel<double, false>", "mixed": "bpcg_spmv_kernel<float, false>"}
METRIC = "BE steps/s, 48^3-cell neo-Hookean (353k DOF), learned+FP32 vs plain FP64"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(cells):
    from simkit_amd.mesh import unit_cube
    X, T = unit_cube(cells)
    bottom = np.where(X[:, 1] == 0.0)[0]
    top = np.where(X[:, 1] == 1.0)[0]
    return X, T, bottom, top


def spmv_bytes(nv, nblk, value_bytes):
    """Algorithmic HBM bytes of one SpMV launch of the PCG: the matrix values
    (9 per 3x3 block), one int32 column index per block, the z and previous-p
    vectors read, the new p and A p written (vector entries of the same width as
    the values), and the 1-byte fixed-DOF mask."""
    n3 = 3 * nv
    return 9 * nblk * value_bytes + 4 * nblk + 4 * n3 * value_bytes + n3


def elem_bytes(nv, nt, value_bytes):
    """Algorithmic HBM bytes of one element pass: per tet its 4 vertex ids, 12 shape-function
    derivatives, mu, lambda and volume (FP64), a flag byte, and the 10 staged 3x3 blocks (90 values)
    written; the positions read once per vertex."""
    return nt * (16 + 96 + 24 + 1 + 90 * value_bytes) + 24 * nv


def pmc_traffic(path_key):
    """HBM bytes per launch of a kernel ("spmv_fp64", "spmv_fp32", "elem_hess_fp64", ...) from the
    committed hardware-counter summary (profiles/pmc_spmv_cube48.md: FETCH_SIZE and WRITE_SIZE passes scaled
    by a copy of known size), or None when absent."""
    path = os.path.join(ROOT, "profiles", "pmc_spmv_cube48.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return int(json.loads(block)[path_key]["traffic_bytes"])
    except (IndexError, KeyError, ValueError):
        return None


def pmc_flops(key):
    """FP64 FLOPs per element-pass launch ("elem_hess_fp64") from the committed hardware-counter
    summary (profiles/pmc_elem_hess_cube48.md: the FP64 VALU instruction counters of the kernel), or
    None when absent."""
    path = os.path.join(ROOT, "profiles", "pmc_elem_hess_cube48.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return float(json.loads(block)[key]["flops_per_launch"])
    except (IndexError, KeyError, ValueError):
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def reference_method_cube10(steps=20):
    """The reference's own method, run in full on the host: the 10^3-cell gravity cube (1,331 vertices,
    5,000 tets; y = 0 pinned; E = 1e6, nu = 0.45, rho = 1000; dt = 1/60; 20 steps from rest), each step
    Newton with the PSD-projected Hessian, scipy spsolve (SuperLU, newton.py:52) and Armijo
    backtracking, to ||g|| <= 1e-6 ||g0|| (the oracle's restatement, single thread)."""
    from oracle import fem, solver
    from oracle.mesh import unit_cube
    Xc, Tc = unit_cube(10)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    prob = solver.BEProblem(Xc, Tc, mu, lam, 1000.0, 1.0 / 60.0, np.where(Xc[:, 1] == 0.0)[0], gravity=-9.8,
                            linsolve="spsolve")
    prob.timers = {}
    t0 = time.perf_counter()
    _, infos = prob.rollout(steps, return_info=True)
    dt = time.perf_counter() - t0
    return {"steps_per_s": steps / dt, "seconds": round(dt, 2), "steps": steps,
            "newton_iters": sum(i["newton"] for i in infos), "ls_trials": sum(i["ls"] for i in infos),
            "phase_ms_total": {k: round(1e3 * v, 1) for k, v in prob.timers.items()}}


def gpu_cube10(steps=20, warmup=2):
    """The plain FP64 device path on the same 10^3 gravity cube, for the same-config ratio."""
    import torch
    from simkit_amd.mesh import unit_cube
    from simkit_amd.sim import BackwardEulerSim
    from simkit_amd.ympr_to_lame import ympr_to_lame
    Xc, Tc = unit_cube(10)
    mu, lam = ympr_to_lame(1e6, 0.45)
    sim = BackwardEulerSim(Xc, Tc, mu, lam, rho=1000.0, h=1.0 / 60.0, fixed_vertices=np.where(Xc[:, 1] == 0.0)[0],
                           gravity=-9.8, pcg_rtol=1e-8)
    sim.rollout(warmup, keep=set())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, infos = sim.rollout(steps, keep=set())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if any(i["status"] != 0 for i in infos):
        raise RuntimeError(f"cube10: status {[i['status'] for i in infos]}")
    return {"steps_per_s": steps / dt, "newton_iters": sum(i["newton_iters"] for i in infos),
            "pcg_iters": sum(i["pcg_iters"] for i in infos)}


def cpu_baseline(X, T, bottom, top, mu, lam, gpu_infos, steps, pcg_sample, gpu10=None):
    """Time the CPU oracle (numpy/scipy FP64, one thread) on a bounded sample of the 48^3 workload
    and compose the per-step cost with the GPU run's own Newton / PCG / line-search counts (value);
    also run the reference's own method (spsolve) in full on the 10^3 cube (reference_method_cube10)."""
    from oracle import fem, solver
    threads = os.environ.get("OMP_NUM_THREADS") or os.environ.get("OPENBLAS_NUM_THREADS") or "unset"
    P = X.copy()
    P[top, 0] += 0.1 * np.sin(2.0 * np.pi / 60.0)
    prob = solver.BEProblem(X, T, mu, lam, 1000.0, 1.0 / 60.0, np.concatenate([bottom, top]),
                            prescribe=lambda t: P, linsolve="pcg")
    x = X.reshape(-1).copy()
    x[prob.fixed] = P.reshape(-1)[prob.fixed]
    xt = X.reshape(-1).copy()
    t0 = time.perf_counter()
    g = prob.potential_gradient(x, xt)
    t1 = time.perf_counter()
    H = prob.potential_hessian(x)
    t2 = time.perf_counter()
    fem.element_energies_x("snh", x, prob.J, prob.mu, prob.lam, prob.vol)
    t3 = time.perf_counter()
    solver.block_jacobi_pcg(H, -g, prob.free, rtol=0.0, maxit=pcg_sample)
    t4 = time.perf_counter()
    t_grad, t_hess, t_energy, t_pcg = t1 - t0, t2 - t1, t3 - t2, (t4 - t3) / pcg_sample
    n_newton = sum(i["newton_iters"] for i in gpu_infos)
    n_pcg = sum(i["pcg_iters"] for i in gpu_infos)
    n_ls = sum(i["ls_trials"] for i in gpu_infos)
    total = (n_newton + steps) * t_grad + n_newton * (t_hess + t_energy) + n_ls * t_energy + n_pcg * t_pcg
    ref10 = reference_method_cube10()
    if gpu10 is not None:
        ref10["gpu_fp64_steps_per_s"] = round(gpu10["steps_per_s"], 3)
        ref10["gpu_newton_iters"] = gpu10["newton_iters"]
        ref10["gpu_over_cpu"] = round(gpu10["steps_per_s"] / ref10["steps_per_s"], 1)
    cells = round((len(T) / 5) ** (1 / 3))
    return {
        "value": steps / total, "unit": "steps/s", "cores": 1,
        "kind": "port",
        "sample": (f"composed, not a full run: oracle (numpy/scipy FP64, one thread) at {cells}^3 with the CPU "
                   f"block-Jacobi PCG in place of the reference's spsolve (SuperLU at this size is ~1.5 h per "
                   f"Newton system, so this is generous to the CPU): 1 gradient {t_grad:.2f}s + 1 PSD Hessian "
                   f"assembly {t_hess:.2f}s + 1 energy {t_energy:.2f}s + {pcg_sample} PCG iterations "
                   f"{t_pcg * 1e3:.1f} ms/it, composed with the GPU run's counts ({n_newton} Newton, {n_pcg} PCG, "
                   f"{n_ls} line-search trials over {steps} steps); the reference method itself is run in full "
                   f"on the 10^3 cube (reference_method_cube10)"),
        "host": {"cpu": cpu_model(), "nproc": os.cpu_count(), "blas_threads": threads},
        "reference_method_cube10": ref10,
    }


# solver paths: name -> (precision, PCG preconditioner, PCG form, PSD eigen-solver)
PATHS = {"fp64": ("fp64", "bjacobi", "classic", "closed_form"), "mixed": ("mixed", "bjacobi", "classic", "closed_form"),
         "fp64_pipelined": ("fp64", "bjacobi", "pipelined", "closed_form"),
         "fp64_jacobi": ("fp64", "bjacobi", "classic", "jacobi"), "mixed_jacobi": ("mixed", "bjacobi", "classic", "jacobi"),
         "fp64_two_level": ("fp64", "two_level", "classic", "closed_form"),
         "mixed_two_level": ("mixed", "two_level", "classic", "closed_form")}


def run_path(args, X, T, bottom, top, mu, lam, name):
    """Warm up, then time the first --steps steps of one solver path; returns a result dict."""
    import torch
    from simkit_amd import _lib
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    precision, pc, form, psd_solver = PATHS[name]
    t_setup = time.perf_counter()
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1.0 / 60.0, material="snh",
                           fixed_vertices=np.concatenate([bottom, top]), pcg_rtol=1e-8, newton_rtol=1e-6,
                           timing=True, precision=precision, inner_rtol=args.inner_rtol, preconditioner=pc,
                           pcg=form, psd_solver=psd_solver)
    bc = top_face_sinusoid(X, top, 0.1, 1.0)
    torch.cuda.synchronize()
    log(f"bench[{name}]: setup {time.perf_counter() - t_setup:.1f}s; warmup {args.warmup} steps")

    def run(nsteps):
        x_prev = sim.X_dev.clone()
        x_curr = x_prev.clone()
        infos = []
        for n in range(nsteps):
            x_next, info = sim.step(x_curr, x_prev, bc(n, (n + 1) * sim.h))
            infos.append(info)
            if info["status"] != 0:
                raise RuntimeError(f"{name} step {n}: status {info['status']} ({info})")
            x_prev, x_curr = x_curr, x_next
        return x_curr, infos

    run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x_final, infos = run(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # dominant kernel: the PCG SpMV, average over back-to-back launches on this stream
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    _lib.call("sk_sim_spmv_reps", sim._h, args.spmv_reps, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    spmv_ms = e0.elapsed_time(e1) / args.spmv_reps
    # the element pass (F, 9x9 Hessian, PSD projection, staged blocks), same stream
    e0.record(stream)
    _lib.call("sk_sim_hess_reps", sim._h, args.hess_reps, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    hess_ms = e0.elapsed_time(e1) / args.hess_reps
    traffic = spmv_bytes(sim.nv, sim.mesh.nblocks, 8 if precision == "fp64" else 4)
    res = {
        "steps_per_s": args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
        "newton_iters": sum(i["newton_iters"] for i in infos), "pcg_iters": sum(i["pcg_iters"] for i in infos),
        "ls_trials": sum(i["ls_trials"] for i in infos), "refine_steps": sum(i["refine_steps"] for i in infos),
        "phase_ms_total": {k: round(sum(i[k] for i in infos), 2)
                           for k in ("ms_gradient", "ms_hessian", "ms_pcg", "ms_line_search")},
        "spmv_ms": spmv_ms, "spmv_bytes": traffic, "hess_ms": hess_ms, "infos": infos,
        "x_final": x_final.cpu().numpy(),
        "coarse": sim.coarse_info() if pc == "two_level" else None,
    }
    log(f"bench[{name}]: {elapsed:.2f}s, {res['steps_per_s']:.3f} steps/s, newton {res['newton_iters']}, "
        f"pcg {res['pcg_iters']}, spmv {spmv_ms * 1e3:.1f} us, element pass {hess_ms * 1e3:.1f} us, "
        f"phases {res['phase_ms_total']}")
    del sim
    return res


def learned_leg(cells, steps, warmup, preconditioner="bjacobi"):
    """Learned warm start (boundary lift + extrapolation of its error + the network's modal correction) on
    the FP32 + FP64-refinement path, on the same trajectory as the headline, with the committed model for
    this mesh (trained on random top-face sinusoids, so this trajectory is held out); also the same start
    without the network and the lift alone, to isolate each part's share. None when no model exists for
    `cells`."""
    import torch
    from simkit_amd import learned
    model = learned.load_model(cells)
    if model is None:
        return None
    t0 = time.perf_counter()
    sim, coords = learned.cube_problem(cells, precision="mixed", k=model.k, preconditioner=preconditioner)
    setup = time.perf_counter() - t0
    A, f = np.array([0.1, 0.0, 0.0]), np.array([1.0, 1.0, 1.0])   # d(t) = (0.1 sin 2 pi t, 0, 0)
    res = {}
    for start in ("learned", "extrapolated", "lift"):
        ro = learned.Rollout(sim, coords, model if start == "learned" else None, start=start)
        ro.run(A, f, warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = ro.run(A, f, steps, keep_last=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        infos = out["infos"]
        used = [i for i in infos if i["guess_used"]]
        res[start] = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
                      "newton_iters": sum(i["newton_iters"] for i in infos),
                      "pcg_iters": sum(i["pcg_iters"] for i in infos),
                      "guess_accepted": len(used) / len(infos),
                      "prediction_taken": sum(1 for i in infos if i["guess_used"] == 1) / len(infos),
                      "start_used": (np.bincount([int(i["guess_used"]) for i in infos]) / len(infos)).round(4).tolist(),
                      "guess_g_ratio": float(np.exp(np.mean([np.log(i["guess_g_norm"] / i["g0_norm"])
                                                             for i in used]))) if used else None,
                      "x_final": out["x"].cpu().numpy()}
        log(f"bench[{start}/{preconditioner}]: {dt:.2f}s, {res[start]['steps_per_s']:.3f} steps/s, newton "
            f"{res[start]['newton_iters']}, pcg {res[start]['pcg_iters']}, guess taken {res[start]['guess_accepted']:.2f}")
    out = res["learned"]
    out["setup_s_modes_and_lift"] = round(setup, 2)
    out["modes"] = coords.modal_info
    out["extrapolated_only"] = {k: v for k, v in res["extrapolated"].items() if k != "x_final"}
    out["lift_only"] = {k: v for k, v in res["lift"].items() if k != "x_final"}
    out["network_speedup_over_extrapolated"] = round(out["steps_per_s"] / res["extrapolated"]["steps_per_s"], 3)
    out["extrapolation_speedup_over_lift"] = round(res["extrapolated"]["steps_per_s"] / res["lift"]["steps_per_s"], 3)
    return out


def warmstart_eval_leg(cells, episodes, steps):
    """Learned warm start on its own workload: held-out random sinusoids of the driven cube from the plain
    start x~, the boundary lift alone, the lift plus the extrapolation of its error, and that plus the
    network, with the committed model."""
    from simkit_amd import learned
    model = learned.load_model(cells)
    if model is None:
        return None
    sim, coords = learned.cube_problem(cells, precision="mixed", k=model.k)
    res = learned.evaluate(sim, coords, model, episodes, steps, seed=2)
    for v in res.values():
        if isinstance(v, dict):
            v.pop("newton_per_step_list", None)
    log(f"bench[warmstart{cells}]: speedup {res['speedup']:.3f}, newton/step {res['plain']['newton_per_step']:.2f}"
        f" -> lift {res['lift']['newton_per_step']:.2f} -> extrapolated {res['extrapolated']['newton_per_step']:.2f}"
        f" -> learned {res['learned']['newton_per_step']:.2f}")
    return {"workload": f"driven cube {cells}^3 cells, random top-face sinusoids (held out), stable neo-Hookean, "
                        f"FP32 + FP64 refinement, block-Jacobi PCG", **res}


def mixed_jittered_leg(cells, steps):
    """Mixed-precision workload: the prescribed-motion cube with every interior rest vertex jittered
    by U(-0.2, 0.2) cells per axis (seed 0; the pinned bottom and driven top faces stay planar), for
    stable neo-Hookean and ARAP. Each energy runs `steps` steps in plain FP64 and in FP32 with FP64
    refinement; reports the per-step relative position error max|x32 - x64| / max|x64| and the rate
    of each path (with one state read back per step)."""
    import torch
    from simkit_amd.mesh import jitter
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    from simkit_amd.ympr_to_lame import ympr_to_lame
    X, T, bottom, top = build_workload(cells)
    X = jitter(X, 1.0 / cells, 0.2, seed=0, keep=np.concatenate([bottom, top]))
    e = X[T[:, 1:]] - X[T[:, :1]]
    assert np.all(np.einsum("ij,ij->i", e[:, 0], np.cross(e[:, 1], e[:, 2])) > 0), "jitter inverted a tet"
    mu, lam = ympr_to_lame(1e6, 0.45)
    bc = top_face_sinusoid(X, top, 0.1, 1.0)
    out = {"workload": f"jittered prescribed-motion cube {cells}^3 cells (+-0.2 cell, seed 0), {steps} steps",
           "steps": steps}
    for mat in ("snh", "arap"):
        runs = {}
        for prec in ("fp64", "mixed"):
            sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1.0 / 60.0, material=mat,
                                   fixed_vertices=np.concatenate([bottom, top]), pcg_rtol=1e-8, newton_rtol=1e-6,
                                   precision=prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            kept, infos = sim.rollout(steps, bc=bc)
            dt = time.perf_counter() - t0
            if any(i["status"] != 0 for i in infos):
                raise RuntimeError(f"{mat}/{prec}: status {[i['status'] for i in infos]}")
            runs[prec] = (kept, infos, dt)
            del sim
        (k64, i64, t64), (k32, i32, t32) = runs["fp64"], runs["mixed"]
        err = [float(np.abs(k32[s] - k64[s]).max() / np.abs(k64[s]).max()) for s in range(1, steps + 1)]
        # against the FP64 oracle's trajectory of the same workload (sampled vertices; the fixture
        # tests/golden/make_oracle_golden.py writes), where it exists for this size
        vs_oracle = None
        fx = os.path.join(ROOT, "tests", "golden", f"oracle_c3_jittered{cells}_{mat}.npz")
        if os.path.exists(fx):
            with np.load(fx) as d:
                idx = d["idx"]
                vs_oracle = {"steps": [int(s) for s in d["steps"] if s <= steps], "sampled_vertices": int(idx.size)}
                for prec, kk in (("mixed", k32), ("fp64", k64)):
                    vs_oracle[prec] = [float(np.abs(kk[s].reshape(-1, 3)[idx] - d[f"x_{s}"]).max() /
                                             np.abs(d[f"x_{s}"]).max()) for s in vs_oracle["steps"]]
        out[mat] = {"per_step_rel_err_vs_fp64": err, "max_rel_err": max(err), "rel_err_vs_oracle": vs_oracle,
                    "fp64_steps_per_s": round(steps / t64, 4), "mixed_steps_per_s": round(steps / t32, 4),
                    "newton_fp64": sum(i["newton_iters"] for i in i64),
                    "newton_mixed": sum(i["newton_iters"] for i in i32),
                    "pcg_fp64": sum(i["pcg_iters"] for i in i64), "pcg_mixed": sum(i["pcg_iters"] for i in i32)}
        log(f"bench[jittered {mat}]: max rel err {max(err):.2e}, fp64 {steps / t64:.3f} / mixed {steps / t32:.3f} "
            f"steps/s")
    return out


def batch_leg(n, steps, coarse_rows=64):
    """Batched solve of n cantilever beams, one launch per Newton phase; coarse_rows > 0: two-level
    per-problem PCG (rigid-motion aggregates of that many consecutive vertices, re-formed every step)."""
    import torch
    from simkit_amd.batch import cantilever_batch
    sim, desc = cantilever_batch(n, coarse_rows=coarse_rows)
    tuned = sim.autotune()
    x = sim.X_dev.clone()
    sim.step(x, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, infos, rounds = sim.rollout(steps, keep=set())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    flat = [i for st in infos for i in st]
    res = {"workload": f"{n} cantilever beams (4x4 cells x U{{8..64}}, cell {desc['cell']} m), stable neo-Hookean,"
                       f" FP32 + FP64 refinement, gravity", "beams": n, "vertices": desc["vertices"],
           "preconditioner": f"two-level, aggregates of {coarse_rows} consecutive vertices" if coarse_rows
           else "block-Jacobi",
           "tets": desc["tets"], "steps": steps, "batch_steps_per_s": round(steps / dt, 4),
           "beam_steps_per_s": round(steps * n / dt, 1), "newton_rounds_per_step": float(np.mean(rounds)),
           "newton_per_beam_step": float(np.mean([i["newton_iters"] for i in flat])),
           "pcg_per_beam_step": float(np.mean([i["pcg_iters"] for i in flat])),
           "statuses": sorted(set(int(i["status"]) for i in flat)), "autotune": tuned, "graph": True}
    log(f"bench[batch{coarse_rows}]: {dt:.2f}s, {res['beam_steps_per_s']} beam-steps/s")
    return res


def pmc_probe(args):
    """Short run for hardware-counter passes: a 256 MiB device-to-device copy (known traffic, the
    calibration kernel) and a few SpMV launches of each path on the assembled 48^3 system. Prints
    the algorithmic bytes the counters are compared with."""
    import torch
    from simkit_amd import _lib
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    from simkit_amd.ympr_to_lame import ympr_to_lame
    X, T, bottom, top = build_workload(args.cells)
    mu, lam = ympr_to_lame(1e6, 0.45)
    nbytes = 256 << 20
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    info = {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes}
    for prec in [p for p in args.paths.split(",") if p in ("fp64", "mixed")]:
        sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, fixed_vertices=np.concatenate([bottom, top]),
                               precision=prec)
        x = sim.X_dev.clone()
        sim.step(x, x, top_face_sinusoid(X, top, 0.1, 1.0)(0, sim.h))
        _lib.call("sk_sim_spmv_reps", sim._h, 5, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        info[f"spmv_{prec}_algorithmic_bytes"] = spmv_bytes(sim.nv, sim.mesh.nblocks, 8 if prec == "fp64" else 4)
        del sim
    print(json.dumps(info), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=48)
    ap.add_argument("--spmv-reps", type=int, default=200)
    ap.add_argument("--hess-reps", type=int, default=20)
    ap.add_argument("--cpu-pcg-sample", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--inner-rtol", type=float, default=1e-2)
    ap.add_argument("--paths", default="fp64,mixed,fp64_two_level,mixed_two_level",
                    help="comma list of " + ", ".join(PATHS))
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start legs")
    ap.add_argument("--warmstart-episodes", type=int, default=2)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-beams leg")
    ap.add_argument("--no-jittered", action="store_true",
                    help="skip the jittered-cube mixed-precision leg (SNH and ARAP vs FP64)")
    ap.add_argument("--jittered-steps", type=int, default=20)
    ap.add_argument("--batch-beams", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only run the short hardware-counter probe (copy calibration + SpMV launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implementation", file=sys.stderr)
        sys.exit(2)

    import torch
    from simkit_amd.ympr_to_lame import ympr_to_lame

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.pmc_probe:
        pmc_probe(args)
        return
    X, T, bottom, top = build_workload(args.cells)
    mu, lam = ympr_to_lame(1e6, 0.45)
    log(f"bench: {args.cells}^3 cube, {X.shape[0]} vertices, {T.shape[0]} tets")
    paths = [p for p in args.paths.split(",") if p]
    results = {p: run_path(args, X, T, bottom, top, mu, lam, p) for p in paths}
    head = results.get("fp64") or results[paths[0]]
    head_name = "fp64" if "fp64" in results else paths[0]
    achieved = head["spmv_bytes"] / (head["spmv_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(head["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if PATHS[head_name][0] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"prescribed-motion cube {args.cells}^3 cells, stable neo-Hookean, "
                        f"{'plain FP64' if head_name == 'fp64' else head_name} path",
            "vertices": int(X.shape[0]), "tets": int(T.shape[0]), "dof": int(3 * X.shape[0]),
            "newton_iters": head["newton_iters"], "pcg_iters": head["pcg_iters"], "ls_trials": head["ls_trials"],
            "phase_ms_total": head["phase_ms_total"],
        },
        "roofline": {"bound": "hbm", "kernel": SPMV_KERNEL[PATHS[head_name][0]], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic("spmv_fp64" if PATHS[head_name][0] == "fp64" else "spmv_fp32")
                     if args.cells == 48 else None,
                     "bytes_per_launch": int(head["spmv_bytes"]), "avg_launch_ms": round(head["spmv_ms"], 5)},
    }
    if PATHS[head_name][0] == "fp64":
        # the element pass against HBM on its algorithmic bytes (it is neither stream- nor FP64-bound:
        # the FP64 rate from the committed instruction counters is reported beside it), launch time
        # measured live above (both launches of the pass)
        eb = elem_bytes(int(X.shape[0]), int(T.shape[0]), 8)
        ach = eb / (head["hess_ms"] * 1e-3) / 1e9
        fl = pmc_flops("elem_hess_fp64") if args.cells == 48 else None
        tf = fl / (head["hess_ms"] * 1e-3) / 1e12 if fl else None
        out["element_roofline"] = {
            "bound": "hbm", "kernel": HESS_KERNEL["fp64"], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic("elem_hess_fp64") if args.cells == 48 else None,
            "bytes_per_launch": int(eb), "elements": int(T.shape[0]), "avg_launch_ms": round(head["hess_ms"], 4),
            "fp64": {"flops_per_launch": fl, "achieved_tflops": round(tf, 2) if tf else None,
                     "peak_tflops": FP64_PEAK_TFLOPS, "frac": round(tf / FP64_PEAK_TFLOPS, 4) if tf else None}}
    if "mixed" in results and head_name == "fp64":
        mx = results["mixed"]
        dev = float(np.abs(mx["x_final"] - head["x_final"]).max() / np.abs(head["x_final"]).max())
        ach = mx["spmv_bytes"] / (mx["spmv_ms"] * 1e-3) / 1e9
        out["fp32_refined"] = {
            "value": round(mx["steps_per_s"], 4), "ms_per_step": round(mx["ms_per_step"], 3),
            "speedup_vs_fp64": round(mx["steps_per_s"] / head["steps_per_s"], 3),
            "final_position_rel_dev_vs_fp64": dev, "inner_rtol": args.inner_rtol,
            "newton_iters": mx["newton_iters"], "pcg_iters": mx["pcg_iters"], "refine_steps": mx["refine_steps"],
            "phase_ms_total": mx["phase_ms_total"],
            "spmv": {"kernel": SPMV_KERNEL["mixed"], "avg_launch_ms": round(mx["spmv_ms"], 5),
                     "bytes_per_launch": int(mx["spmv_bytes"]), "achieved_GBs": round(ach, 1),
                     "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic("spmv_fp32") if args.cells == 48 else None},
        }
    for name in ("fp64_two_level", "mixed_two_level"):
        if name in results and head_name == "fp64":
            r = results[name]
            dev = float(np.abs(r["x_final"] - head["x_final"]).max() / np.abs(head["x_final"]).max())
            out.setdefault("two_level", {})[PATHS[name][0]] = {
                "value": round(r["steps_per_s"], 4), "ms_per_step": round(r["ms_per_step"], 3),
                "speedup_vs_fp64": round(r["steps_per_s"] / head["steps_per_s"], 3),
                "final_position_rel_dev_vs_fp64": dev, "newton_iters": r["newton_iters"],
                "pcg_iters": r["pcg_iters"], "refine_steps": r["refine_steps"], "phase_ms_total": r["phase_ms_total"],
                "coarse": r["coarse"]}
    if "two_level" in out:
        out["two_level"]["preconditioner"] = ("block-Jacobi + additive coarse correction over piecewise-rigid "
                                              "aggregates (6 rigid motions each), coarse operator from the first "
                                              "Newton system")
    if not args.no_learned and head_name == "fp64":
        for pc, key, ref in (("bjacobi", "learned_fp32", "mixed"), ("two_level", "learned_fp32_two_level",
                                                                     "mixed_two_level")):
            lr = learned_leg(args.cells, args.steps, args.warmup, pc)
            if lr is None:
                continue
            dev = float(np.abs(lr["x_final"] - head["x_final"]).max() / np.abs(head["x_final"]).max())
            lr.pop("x_final")
            out[key] = {"value": round(lr["steps_per_s"], 4), "preconditioner": pc,
                        "speedup_vs_fp64": round(lr["steps_per_s"] / head["steps_per_s"], 3),
                        f"speedup_vs_{ref}": round(lr["steps_per_s"] / results[ref]["steps_per_s"], 3)
                        if ref in results else None,
                        "final_position_rel_dev_vs_fp64": dev, **lr}
        ws = warmstart_eval_leg(32, args.warmstart_episodes, 100)
        if ws is not None:
            out["learned_warm_start_32"] = ws
    if not args.no_jittered and head_name == "fp64":
        out["mixed_jittered"] = mixed_jittered_leg(args.cells, args.jittered_steps)
    if not args.no_batch:
        out["batched_beams"] = batch_leg(args.batch_beams, args.batch_steps, coarse_rows=64)
        out["batched_beams"]["block_jacobi"] = {
            k: v for k, v in batch_leg(args.batch_beams, args.batch_steps, coarse_rows=0).items()
            if k in ("beam_steps_per_s", "batch_steps_per_s", "pcg_per_beam_step", "newton_rounds_per_step")}
    if not args.no_cpu_baseline:
        g10 = gpu_cube10()
        log(f"bench: 10^3 gravity cube on the GPU {g10['steps_per_s']:.1f} steps/s; CPU baseline sample")
        out["cpu_baseline"] = cpu_baseline(X, T, bottom, top, mu, lam, head["infos"], args.steps,
                                           args.cpu_pcg_sample, g10)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
