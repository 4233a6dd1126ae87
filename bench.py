"""Benchmark: explicit second-order RK5 steps of the explicit path.

    python bench.py [--gpus 1] [--steps 400] [--warmup 20] [--precision f64]
                    [--workload step|airfoil] [--converge]

One step = one RK5 step (5 stages: gradient+limiter, flux+update+boundary)
plus the time-step pass, inputs resident in HBM.  The default workload is
the 975,744-cell Mach-3 forward step in FP64; ``--workload airfoil`` is the
2.0M-cell transonic NACA 0012 O-mesh with local time steps, and with
``--converge`` it runs from the freestream until the density residual has
dropped 6 orders and reports the time to get there.  Prints one JSON line
with the throughput in Mcell-stage/s, the HBM roofline of the dominant
kernel (measured with HIP events on the solver's stream over the timed
region) and a CPU baseline (the reference library on a bounded sample of
the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "finite-volume-2d-cpp_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(C, E, B, s, g, V=4):
    """Per-stage algorithmic bytes split between the two stage kernels.

    Whole stage (second order): B2 = B1 + C(7Vs + 2g) + E(2g), with
    B1 = C(3Vs + s + g) + E(3g + 8) + B(2Vs + 2g + 4).  The flux kernel owns
    B1 plus the reconstruction inputs C(3Vs + 2g) + E(2g); the gradient/
    limiter kernel owns the remaining C(4Vs)."""
    b1 = C * (3 * V * s + s + g) + E * (3 * g + 8) + B * (2 * V * s + 2 * g + 4)
    flux = b1 + C * (3 * V * s + 2 * g) + E * (2 * g)
    gradlim = C * (4 * V * s)
    return {"stage": flux + gradlim, "flux": flux, "gradlim": gradlim}


def cpu_baseline(mesh_arrays, cfg, steps: int, workload: str = "step"):
    """Time the reference library (oracle/_ref/fvh_ref, kind 'reference') on
    `steps` RK5 steps of the same mesh; falls back to the C restatement.
    For the airfoil the reference's farfield has its own hard-coded
    freestream (Mach 0.8 at 1.2532 degrees), which costs the same."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fvhyper_amd import meshgen

    C = mesh_arrays.n_cells
    if O.ref_available():
        try:
            work = tempfile.mkdtemp(prefix="fvh_cpu_")
            meshgen.write_msh(os.path.join(work, "case.msh"), mesh_arrays)
            cfg2 = dict(cfg)
            cfg2["K"] = 7.5   # the reference library fixes K; the cost is the same
            kv = O.ref_args(cfg2)
            if workload == "step":
                kv.update(ic="fs", mach="3")
            else:
                kv.update(ic="nc")
            kv.update(max_step=str(steps), print=str(steps))
            r = O.run_ref("run", work, "case", flags=O.ref_flags(cfg2), timeout=900, **kv)
            last = r["stdout"].strip().splitlines()[-1].split(",")
            secs = float(last[2])
            one = {"value": C * 5 * steps / secs / 1e6, "unit": "Mcell-stage/s", "cores": 1,
                   "kind": "reference",
                   "sample": "%d RK5 steps (%d stages) of the %d-cell %s workload, fvhyper library "
                             "built from its sources (-O3), 1 rank, RealTime loop clock"
                             % (steps, 5 * steps, C, workload)}
            ranks = min(16, os.cpu_count() or 1)
            if ranks < 2 or not O.mpi_available():
                return one
            # the same library on every core of the host: the mesh partitioned by
            # recursive bisection, one ghost layer per rank, as the reference reads
            # it (mesh.cpp:803-855), run under mpirun
            try:
                part = meshgen.partition_rcb(mesh_arrays, ranks)
                meshgen.write_partitioned_msh(work, "case", mesh_arrays, part)
                steps_p = min(steps * ranks, 100)
                kv.update(max_step=str(steps_p), print=str(steps_p))
                rp = O.run_ref_mpi("run", work, "case", ranks, flags=O.ref_flags(cfg2), timeout=900, **kv)
                secs_p = float(rp["stdout"].strip().splitlines()[-1].split(",")[2])
                return {"value": C * 5 * steps_p / secs_p / 1e6, "unit": "Mcell-stage/s", "cores": ranks,
                        "kind": "reference",
                        "sample": "%d RK5 steps (%d stages) of the %d-cell %s workload, fvhyper library built "
                                  "from its sources (-O3), %d MPI ranks (mpirun) on a recursive-bisection "
                                  "partition with one ghost layer, RealTime loop clock of rank 0"
                                  % (steps_p, 5 * steps_p, C, workload, ranks),
                        "single_rank": {"value": one["value"], "sample": one["sample"]}}
            except Exception as e:  # pragma: no cover - reported, the single rank stands
                print("multi-rank reference baseline unavailable: %s" % e, file=sys.stderr)
                return one
        except Exception as e:  # pragma: no cover - reported, then the port is timed
            print("reference baseline unavailable: %s" % e, file=sys.stderr)
    om = O.Mesh(mesh_arrays)
    from fvhyper_amd import workloads
    _, rho, u, v, p = cfg["bcs"]["inlet" if workload == "step" else "farfield"]
    q0 = workloads.initial_state(om, workloads.uniform_ic(rho, u, v, p))
    t0 = time.perf_counter()
    om.run(cfg, q0, steps)
    secs = time.perf_counter() - t0
    return {"value": C * 5 * steps / secs / 1e6, "unit": "Mcell-stage/s", "cores": 1, "kind": "port",
            "sample": "%d RK5 steps of the %d-cell %s workload, C restatement, 1 thread" % (steps, C, workload)}


# rocprofv3 PMC summaries committed under profiles/: (workload, precision,
# kernel) -> (file, row prefix).  Each file's traffic table gives HBM bytes per
# dispatch as 2.0 x FETCH_SIZE + WRITE_SIZE (factors from a plain device copy).
PMC_PROFILES = {
    ("step", "f64", "stage"): ("k_stage__mach3_step_975k_fp64.md", "k_stage<double"),
    ("step", "f32", "stage"): ("k_stage__mach3_step_975k_fp32.md", "k_stage<float"),
    ("step", "f64", "flux"): ("k_flux_tile__mach3_step_975k_fp64.md", "k_flux_tile<double"),
    ("airfoil", "f64", "stage"): ("k_stage__naca0012_2m_fp64.md", "k_stage<double"),
    ("sweep", "f32", "flux"): ("k_flux_tile__naca0012_sweep256_fp32.md", "k_flux_tile<float"),
    ("sweep", "f32", "stage"): ("k_stage__naca0012_sweep256_fp32.md", "k_stage<float"),
}


def achievable(roof):
    """Beside the published peak, the bandwidth an in-repo device copy
    kernel reaches on this GPU (fvh_stream_copy: 1 GB, read + write)."""
    from fvhyper_amd import stream_copy_gbs
    g = stream_copy_gbs(1024.0, 20)
    roof["peak_achievable"] = round(g, 1)
    roof["frac_achievable"] = round(roof["achieved"] / g, 4)


def calibrate_copy(a):
    """A plain device copy of --calibrate-copy-mb MiB (three times), whose
    FETCH_SIZE / WRITE_SIZE counters scale the PMC traffic of a profiled run."""
    if not a.calibrate_copy_mb:
        return
    import torch
    n = a.calibrate_copy_mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def pmc_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary of this
    workload, or None when there is none."""
    key = PMC_PROFILES.get((workload, precision, kernel))
    if key is None:
        return None
    path = os.path.join(ROOT, "profiles", key[0])
    if not os.path.exists(path):
        return None
    section = False
    tot = n = 0.0
    for line in open(path):
        if line.startswith("## HBM traffic"):
            section = True
        elif line.startswith("## "):
            section = False
        elif section and line.startswith("| `" + key[1]):
            # dispatch-weighted over the kernel's instantiations
            f = [v.strip() for v in line.rstrip().rstrip("|").split("|")]
            tot += float(f[-1]) * float(f[2])
            n += float(f[2])
    if n == 0:
        return None
    return {"bytes": round(tot / n * 1e6), "source": "profiles/" + key[0]}


def converge(s, C, stages, tol, stop_var, chunk=2000, limit=400000):
    """Run from the freestream until R[stop_var] < tol (the reference's
    max-over-variables rule for stop_var=-1); returns a dict of the run."""
    import numpy as np
    s.run_begin()
    s.synchronize()
    t0 = time.perf_counter()
    while True:
        st = s.run_continue(chunk, max_step=limit, tol=tol, stop_var=stop_var, print_interval=100)
        print("converge: step %d R %s" % (st["step"], np.array2string(np.asarray(st["R"]), precision=3)),
              file=sys.stderr, flush=True)
        if st["stopped"] or not np.all(np.isfinite(st["R"])) or st["step"] >= limit:
            break
    secs = time.perf_counter() - t0
    return {"steps": st["step"], "seconds": round(secs, 3), "converged": bool(st["stopped"]) and st["step"] < limit,
            "tolerance": tol, "rule": "density" if stop_var == 0 else "max_over_variables",
            "R": [float(x) for x in st["R"]], "Mcell_stage_per_s": round(C * stages * st["step"] / secs / 1e6, 2)}


def sweep(a):
    """Parameter sweep: a batch of coarsened airfoil instances (sizes
    log-uniform in 5k-200k cells, Mach 0.5-0.85, angle -2..4 degrees), FP32
    state with FP64 residual sums, every step replayed from a HIP graph;
    against the same instances run one after another through the plain
    single-mesh path (own solver each, eager launches, synchronised per
    instance)."""
    import numpy as np
    import torch

    from fvhyper_amd import Batch, Mesh, Solver, workloads

    t0 = time.perf_counter()
    items = workloads.naca_sweep(a.instances)
    built = {}
    meshes = []
    for it in items:
        key = id(it[0])
        if key not in built:
            built[key] = Mesh.from_arrays(it[0])
        meshes.append(built[key])
    cfgs = [it[1] for it in items]
    q0 = [workloads.initial_state(m, it[2]) for m, it in zip(meshes, items)]
    # stage form and tile size from the autotuner, timed on the whole batch: a
    # sample of the instances fits the 256 MB Infinity Cache better than the
    # batch does and can rank the forms the other way round
    from fvhyper_amd import autotune
    sub = list(range(0, len(meshes), a.tune_every))
    tuned = autotune([meshes[k] for k in sub], [cfgs[k] for k in sub], np.concatenate([q0[k] for k in sub]),
                     precision=a.precision, steps=3)
    b = Batch(meshes, cfgs, precision=a.precision, flux_mode=tuned["flux_mode"], tile_cells=tuned["tile_cells"])
    b.set_state(b.join(q0))
    b.apply_bounds()
    b.use_graphs(True)
    setup = time.perf_counter() - t0
    C = sum(m.n_real for m in meshes)
    E = sum(m.n_edges for m in meshes)
    B = sum(m.n_bnd for m in meshes)
    stages = len(cfgs[0]["alpha"])
    b.advance(a.warmup)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance(a.steps)
    b.synchronize()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    if not np.all(np.isfinite(b.get_state())):
        print("non-finite batch state after the timed steps", file=sys.stderr)
        sys.exit(3)
    # kernel timing (events, eager launches) over a few more steps
    b.use_graphs(False)
    b.profile(True)
    b.advance(max(2, a.steps // 10))
    b.synchronize()
    kt = b.kernel_times()
    b.profile(False)
    # the same instances one by one
    k_pm = a.per_mesh_steps or a.steps
    solvers = []
    for m, c, q in zip(meshes, cfgs, q0):
        x = Solver(m, c, precision=a.precision, flux_mode="tile")
        x.set_state(q)
        x.apply_bounds()
        x.advance(2)
        x.synchronize()
        solvers.append(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in solvers:
        x.advance(k_pm)
        x.synchronize()
    secs_pm = time.perf_counter() - t0
    calibrate_copy(a)
    value = C * stages * a.steps / secs / 1e6
    value_pm = C * stages * k_pm / secs_pm / 1e6
    sb = 4 if a.precision == "f32" else 8
    ab = algorithmic_bytes(C, E, B, sb, sb)
    roof = None
    key = "stage" if kt.get("stage", {}).get("count", 0) else "flux"
    if kt.get(key, {}).get("count", 0):
        avg_ms = kt[key]["total_ms"] / kt[key]["count"]
        nbytes = ab["stage"] if key == "stage" else ab["flux"]
        ach = nbytes / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": ("k_stage (fused stage pass)" if key == "stage" else
                           "k_flux (edge-flux gather + RK update + boundary)") + ", all instances",
                "bytes_per_launch": nbytes, "avg_launch_us": round(avg_ms * 1e3, 2),
                "kernels_us": {k: round(v["total_ms"] / v["count"] * 1e3, 2) for k, v in kt.items() if v["count"]}}
        achievable(roof)
        tr = pmc_traffic("sweep", a.precision, key) if a.instances == 256 else None
        if tr is not None:
            roof["traffic"] = tr["bytes"]
            roof["traffic_source"] = tr["source"]
    out = {
        "metric": "Mcell-stage/s", "value": round(value, 2), "unit": "Mcell-stage/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(secs / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "f64" else "f32", "data": "synthetic",
        "config": {"workload": "naca0012_sweep", "instances": len(meshes), "cells": C, "edges": E,
                   "boundary_edges": B, "cells_range": [5000, 200000], "mach_range": [0.5, 0.85],
                   "aoa_deg_range": [-2.0, 4.0], "order": 2, "limiter": "michalak", "dt": "local",
                   "rk_stages": stages, "precision": a.precision, "residual_sums": "f64",
                   "launch": "batched, hipGraph per step", "parallelism": "single_gpu",
                   "autotuned": {"flux_mode": tuned["flux_mode"], "tile_cells": tuned["tile_cells"],
                                 "candidates_ms_per_step": [list(c) for c in tuned["candidates"]]}},
        "per_mesh": {"value": round(value_pm, 2), "steps": k_pm, "seconds": round(secs_pm, 3),
                     "launch": "one solver per instance (two-pass tiled stage kernels), eager, "
                               "synchronised per instance"},
        "speedup_vs_per_mesh": round(value / value_pm, 3), "setup_s": round(setup, 1),
        "roofline": roof, "cpu_baseline": None,
    }
    print(json.dumps(out))


def adjoint(a):
    """Discrete adjoint: gradient of the step-face pressure force after
    --adj-steps RK5 steps of the second-order forward-step solver (h = 1/220,
    243,936 cells) from a seeded smooth perturbation of the inflow, with
    respect to every initial cell state and, in a second timed run, also to
    every node coordinate.  Reports the adjoint/forward cost ratios, the
    checkpoint memory, and the agreement with central differences (eps 1e-6
    and 1e-8) on --adj-dirs random directions and with the complex-step
    derivative of the CPU restatement on a smaller mesh (states and interior
    node coordinates)."""
    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from fvhyper_amd import Mesh, Solver, workloads

    m, cfg, ic = workloads.forward_step(n_per_unit=220)
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg)
    s.use_graphs(True)           # forward steps replayed from HIP graphs, in both runs
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    N = a.adj_steps
    s.objective(q0, 2, wx=w)                                 # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = s.objective(q0, N, wx=w)
    t_fwd = time.perf_counter() - t0
    s.adjoint(q0, 2, wx=w)                                   # warm-up
    t0 = time.perf_counter()
    Ja, grad, st = s.adjoint(q0, N, wx=w, mem_cap_bytes=8e9)
    t_adj = time.perf_counter() - t0
    s.adjoint(q0, 2, wx=w, nodes=True)                       # warm-up
    t0 = time.perf_counter()
    _, _, gxy, _ = s.adjoint(q0, N, wx=w, mem_cap_bytes=8e9, nodes=True)
    t_adj_nodes = time.perf_counter() - t0
    rng = np.random.default_rng(20240606)
    rel6, rel8 = [], []
    for _ in range(a.adj_dirs):
        v = rng.standard_normal((pm.n_cells, 4)) * np.array([1.0, 3.0, 3.0, 10.0])
        v[pm.n_real:] = 0.0
        v = v.ravel()
        ad = float(grad @ v)
        for eps, out in ((1e-6, rel6), (1e-8, rel8)):
            fd = (s.objective(q0 + eps * v, N, wx=w) - s.objective(q0 - eps * v, N, wx=w)) / (2 * eps)
            out.append(abs(ad - fd) / max(abs(fd), 1e-300))
    # complex step of the CPU restatement, on a smaller mesh (CPU time)
    cs = None
    try:
        import oracle as O
        mc, cfgc, _ = workloads.forward_step(n_per_unit=40)
        pmc = Mesh.from_arrays(mc)
        omc = O.Mesh(mc)
        wc = workloads.step_face_weights(pmc)
        qc = workloads.bump_state(pmc)
        Nc = min(N, 40)
        _, gc, _ = Solver(pmc, cfgc).adjoint(qc, Nc, wx=wc)
        errs = []
        for _ in range(4):
            v = rng.standard_normal((pmc.n_cells, 4)) * np.array([1.0, 3.0, 3.0, 10.0])
            v[pmc.n_real:] = 0.0
            v = v.ravel()
            _, dJ = O.cstep_derivative(omc, cfgc, qc, v, Nc, wx=wc)
            errs.append(abs(float(gc @ v) - dJ) / abs(dJ))
        # node coordinates: interior displacements of a tenth of the spacing
        _, _, gxyc, _ = Solver(pmc, cfgc).adjoint(qc, Nc, wx=wc, nodes=True)
        inner = np.ones(mc.x.size, dtype=bool)
        inner[np.asarray(mc.bnd).ravel()] = False
        nerrs = []
        for _ in range(2):
            dx = np.where(inner, rng.standard_normal(mc.x.size), 0.0) * 0.1 / 40
            dy = np.where(inner, rng.standard_normal(mc.x.size), 0.0) * 0.1 / 40
            _, dJ = O.cstep_derivative_nodes(omc, cfgc, qc, Nc, dx=dx, dy=dy, wx=wc)
            nerrs.append(abs(float(gxyc[:, 0] @ dx + gxyc[:, 1] @ dy) - dJ) / abs(dJ))
        cs = {"cells": pmc.n_real, "steps": Nc, "directions": 4, "max_rel_err": max(errs),
              "node_directions": 2, "node_max_rel_err": max(nerrs)}
    except Exception as e:  # pragma: no cover - reported
        print("complex-step check skipped: %s" % e, file=sys.stderr)
    C = pm.n_real
    stages = len(cfg["alpha"])
    out = {
        "metric": "Mcell-stage/s", "value": round(C * stages * N / t_fwd / 1e6, 2), "unit": "Mcell-stage/s",
        "n_gpus": 1, "steps": N, "warmup": 2, "ms_per_step": round(t_fwd / N * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "forward_step_adjoint", "cells": C, "steps": N, "stages": N * stages,
                   "objective": "pressure force on the step face", "parallelism": "single_gpu"},
        "adjoint": {"J": J, "forward_s": round(t_fwd, 4), "adjoint_s": round(t_adj, 4),
                    "adjoint_over_forward": round(t_adj / t_fwd, 3),
                    "with_node_gradient_s": round(t_adj_nodes, 4),
                    "with_node_gradient_over_forward": round(t_adj_nodes / t_fwd, 3),
                    "node_gradient_norm": float(np.linalg.norm(gxy)), "snaps": st["snaps"],
                    "forward_steps_executed": st["forward_steps"], "recorded_steps": st["recorded_steps"],
                    "library_ms": {"forward_sweep": round(st["forward_ms"], 2), "total": round(st["total_ms"], 2)},
                    "checkpoint_GB": round(st["checkpoint_bytes"] / 1e9, 3),
                    "work_GB": round(st["work_bytes"] / 1e9, 3),
                    "fd_eps_1e-6": None if not rel6 else {
                        "directions": len(rel6), "median_rel_err": float(np.median(rel6)),
                        "max_rel_err": float(np.max(rel6)), "frac_below_1e-4": float(np.mean(np.array(rel6) < 1e-4))},
                    "fd_eps_1e-8": None if not rel8 else {
                        "directions": len(rel8), "median_rel_err": float(np.median(rel8)),
                        "max_rel_err": float(np.max(rel8))},
                    "complex_step": cs},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default=None, choices=["f64", "f32", "f32_f64"],
                    help="default: f64 (f32 for the sweep, whose state is FP32)")
    ap.add_argument("--workload", default="step", choices=["step", "airfoil", "sweep", "adjoint", "plate"])
    ap.add_argument("--smooth", action="store_true", help="plate, airfoil: with residual smoothing")
    ap.add_argument("--cfl", type=float, default=0.0, help="airfoil: CFL number instead of the workload's 2")
    ap.add_argument("--adj-steps", type=int, default=80, help="adjoint: RK5 steps of the objective's run")
    ap.add_argument("--adj-dirs", type=int, default=64, help="adjoint: random directions checked")
    ap.add_argument("--instances", type=int, default=256, help="sweep: number of airfoil instances")
    ap.add_argument("--per-mesh-steps", type=int, default=0,
                    help="sweep: steps of the one-by-one baseline (default: --steps)")
    ap.add_argument("--tune-every", type=int, default=1,
                    help="sweep: the autotuner times every n-th instance (1: the whole batch)")
    ap.add_argument("--n-per-unit", type=int, default=440, help="step mesh resolution")
    ap.add_argument("--flux-mode", default="auto", choices=["auto", "fused", "tile", "cell"],
                    help="step/airfoil: stage form (auto: fused where it fits)")
    ap.add_argument("--cells", type=int, default=2_000_000, help="airfoil mesh size")
    ap.add_argument("--stage-threads", type=int, default=0, choices=[0, 256, 512],
                    help="fused stage pass workgroup size (0: the library's choice)")
    ap.add_argument("--tile-cells", type=int, default=0,
                    help="step/airfoil/plate: cells per tile (32..256; 0: the library's automatic choice)")
    ap.add_argument("--converge", action="store_true",
                    help="airfoil: also run from the freestream until the density residual drops 6 orders")
    ap.add_argument("--stop-var", type=int, default=0, help="with --converge: 0 density, -1 max over variables")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-launch events")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run a plain torch copy of this size (for rocprof counter scaling)")
    a = ap.parse_args()
    if a.precision is None:
        a.precision = "f32" if a.workload == "sweep" else "f64"
    if a.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from fvhyper_amd import Mesh, Solver, workloads

    torch.cuda.set_device(0)
    if a.workload == "sweep":
        return sweep(a)
    if a.workload == "adjoint":
        return adjoint(a)
    if a.workload == "step":
        m, cfg, ic = workloads.forward_step(n_per_unit=a.n_per_unit)
    elif a.workload == "plate":
        m, cfg, ic = workloads.flat_plate(smooth=a.smooth)
    else:
        m, cfg, ic = workloads.naca_airfoil(a.cells)
        if a.smooth:
            cfg["smooth"] = True
        if a.cfl > 0:
            cfg["cfl"] = a.cfl
    if a.stage_threads:
        cfg["stage_threads"] = a.stage_threads
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg, precision=a.precision, flux_mode=a.flux_mode, tile_cells=a.tile_cells)
    q0 = workloads.initial_state(pm, ic)
    s.set_state(q0)
    s.apply_bounds()
    C, E, B = pm.n_real, pm.n_edges, pm.n_bnd
    stages = len(cfg["alpha"])

    # timed region: the steps replayed from HIP graphs, no per-launch events
    s.use_graphs(True)
    s.advance(a.warmup)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.advance(a.steps)
    s.synchronize()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    # kernel durations: HIP events around every launch on the solver's
    # stream, over a further window of the same run (eager launches)
    kt = {}
    if not a.no_events:
        s.use_graphs(False)
        s.profile(True)
        s.advance(max(20, a.steps // 10))
        s.synchronize()
        kt = s.kernel_times()
        s.profile(False)
    q = s.get_state()
    if not np.all(np.isfinite(q)):
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)

    calibrate_copy(a)

    value = C * stages * a.steps / secs / 1e6
    sb = 4 if a.precision == "f32" else 8
    ab = algorithmic_bytes(C, E, B, sb, sb)
    roof = None
    if kt and kt.get("stage", {}).get("count", 0):
        # fused form: one launch is one whole stage, priced at the stage's bytes
        avg_ms = kt["stage"]["total_ms"] / kt["stage"]["count"]
        ach = ab["stage"] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_stage (fused gradient + limiter + flux + RK update + boundary, one launch per stage)",
                "bytes_per_launch": ab["stage"], "avg_launch_us": round(avg_ms * 1e3, 2)}
    elif kt and a.workload == "plate":
        # viscous two-pass stage: the gradient pass also writes gx, gy; the
        # flux pass reads them and the cell centres (face gradients)
        V = 4
        H = 2 * E - B
        per = {"gradlim": ab["gradlim"] + C * 2 * V * sb,
               "flux": ab["flux"] + C * (2 * V * sb + 2 * sb)}
        if a.smooth:
            # two sweeps: qt in and out, the half-edge lists; the second also
            # finishes the stage (q0 read, q written, dt, area)
            per["smooth"] = (C * (2 * V * sb + 4) + H * 8 + C * (4 * V * sb + 4 + 2 * sb) + H * 8) // 2
        tot = {k: kt[k]["total_ms"] for k in per if kt.get(k, {}).get("count", 0)}
        key = max(tot, key=tot.get)
        avg_ms = kt[key]["total_ms"] / kt[key]["count"]
        ach = per[key] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_" + key + " (viscous two-pass stage" + (", residual smoothing" if a.smooth else "") + ")",
                "bytes_per_launch": per[key], "avg_launch_us": round(avg_ms * 1e3, 2),
                "others": {k: {"bytes_per_launch": per[k],
                               "avg_launch_us": round(kt[k]["total_ms"] / kt[k]["count"] * 1e3, 2),
                               "frac": round(per[k] / (kt[k]["total_ms"] / kt[k]["count"] * 1e-3) / 1e9
                                             / HBM_PEAK_GBS, 4)} for k in tot if k != key}}
    elif kt and kt.get("flux", {}).get("count", 0):
        avg_ms = kt["flux"]["total_ms"] / kt["flux"]["count"]
        ach = ab["flux"] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_flux (edge-flux gather + RK update + boundary)",
                "bytes_per_launch": ab["flux"], "avg_launch_us": round(avg_ms * 1e3, 2)}
    if roof is not None:
        achievable(roof)
        roof["events_window_steps"] = max(20, a.steps // 10)
        tr = pmc_traffic(a.workload, a.precision, "stage" if "stage" in roof["kernel"] else "flux") \
            if a.workload != "plate" else None
        if tr is not None and (a.workload != "step" or a.n_per_unit == 440) and (
                a.workload != "airfoil" or a.cells == 2_000_000):
            roof["traffic"] = tr["bytes"]
            roof["traffic_source"] = tr["source"]
        stage_ms = sum(v["total_ms"] for v in kt.values()) / (roof["events_window_steps"] * stages)
        roof["stage_achieved"] = round(ab["stage"] / (stage_ms * 1e-3) / 1e9, 1)
        roof["kernels_us"] = {k: round(v["total_ms"] / v["count"] * 1e3, 2) for k, v in kt.items() if v["count"]}
    conv = None
    if a.converge:
        s.set_state(q0)
        conv = converge(s, C, stages, 1e-6, a.stop_var)
    cpu = None if (a.no_cpu or a.workload == "plate") else cpu_baseline(m, cfg, a.cpu_steps, a.workload)
    if a.workload == "plate":
        wl = {"workload": "laminar_flat_plate", "cells": C, "edges": E, "boundary_edges": B, "order": 2,
              "limiter": "michalak", "flux": "roe + laminar viscous (diffusive face gradients)",
              "residual_smoothing": a.smooth, "rk_stages": stages, "cfl": cfg["cfl"], "dt": "local viscous",
              "precision": a.precision, "stage_form": "tile", "parallelism": "single_gpu"}
    elif a.workload == "step":
        wl = {"workload": "mach3_forward_step", "cells": C, "edges": E, "boundary_edges": B,
              "order": 2, "limiter": "venkatakrishnan", "K": cfg["K"], "rk_stages": stages,
              "cfl": cfg["cfl"], "dt": "global", "precision": a.precision, "stage_form": a.flux_mode,
              "parallelism": "single_gpu"}
    else:
        wl = {"workload": "naca0012_transonic", "cells": C, "edges": E, "boundary_edges": B,
              "order": 2, "limiter": "michalak", "mach": 0.8, "aoa_deg": 1.25, "rk_stages": stages,
              "cfl": cfg["cfl"], "dt": "local", "precision": a.precision, "stage_form": a.flux_mode,
              "residual_smoothing": bool(cfg.get("smooth")), "parallelism": "single_gpu"}
    out = {
        "metric": "Mcell-stage/s", "value": round(value, 2), "unit": "Mcell-stage/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(secs / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "f64" else "f32", "data": "synthetic",
        "config": wl, "roofline": roof, "cpu_baseline": cpu,
    }
    wl["tile_cells"] = s.tile_cells
    if conv is not None:
        out["converge"] = conv
    print(json.dumps(out))


if __name__ == "__main__":
    main()
