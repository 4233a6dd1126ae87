"""Benchmark of the explicit path on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
    python bench.py --workload step|airfoil|sweep|adjoint|plate [...]

With no --workload, one JSON line carries the metric's three parts:
  * top level: the Mach-3 forward step (975,744 cells, second order,
    Venkatakrishnan, RK5, global dt) in FP64, bit-identical to the reference
    -- Mcell-stage/s over --steps RK5 steps replayed from HIP graphs, the HBM
    roofline of the fused stage kernel, and the reference library itself on
    the host's CPU share as the CPU baseline;
  * "sweep": 256 airfoil instances (5k-200k cells, Mach 0.5-0.85, angle
    -2..4 degrees) with FP32 states (FP64 arithmetic, residual and sums) in
    one launch per stage, against the same instances one by one through the
    per-mesh path, the error against the FP64 oracle on a sample of the
    instances, and 256 single-rank reference processes as the CPU baseline;
  * "adjoint": dJ/dq0 of the step-face pressure force after 80 RK5 steps
    (400 stages) at 243,936 cells, its cost over the forward run, central
    differences on 64 directions and the complex-step derivative of the CPU
    restatement on 16 directions at that size (computed by CPU worker
    processes while the GPU runs the rest).
One step = one RK5 step (5 stages) plus the time-step pass; inputs are
resident in HBM when the timed region starts.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import re
import subprocess
import sys
import tempfile
import time

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "finite-volume-2d-cpp_amd"))

HBM_PEAK_GBS = 8000.0
# a GPU box lends one GPU 16 of its host's cores, whatever the host has
CPU_SHARE = 16


def cpu_info():
    """The host CPU and the cores this run may use (its affinity, at most the
    16-core share of one GPU)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:   # pragma: no cover - not Linux
        aff = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff, "ranks": max(1, min(CPU_SHARE, aff))}


def algorithmic_bytes(C, E, B, s, g, V=4):
    """Per-stage algorithmic bytes split between the two stage kernels
    (s: state bytes, g: geometry bytes).

    Whole stage (second order): B2 = B1 + C(7Vs + 2g) + E(2g), with
    B1 = C(3Vs + s + g) + E(3g + 8) + B(2Vs + 2g + 4).  The flux kernel owns
    B1 plus the reconstruction inputs C(3Vs + 2g) + E(2g); the gradient/
    limiter kernel owns the remaining C(4Vs)."""
    b1 = C * (3 * V * s + s + g) + E * (3 * g + 8) + B * (2 * V * s + 2 * g + 4)
    flux = b1 + C * (3 * V * s + 2 * g) + E * (2 * g)
    gradlim = C * (4 * V * s)
    return {"stage": flux + gradlim, "flux": flux, "gradlim": gradlim}


def state_bytes(precision):
    """(state bytes, geometry bytes) of a precision mode."""
    return {"f64": (8, 8), "f32": (4, 4), "f32_f64": (4, 8)}[precision]


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def cpu_baseline(mesh_arrays, cfg, steps: int, workload: str = "step", ranks: int = 0):
    """Time the reference library (oracle/_ref/fvh_ref, kind 'reference') on
    `steps` RK5 steps of the same mesh, on one rank and on `ranks` MPI ranks
    (the mesh partitioned by recursive bisection with one ghost layer, as the
    reference reads partitioned meshes); falls back to the C restatement on
    one thread.  The timing is the reference's own RealTime column (steady
    clock around the loop).  For the airfoil the reference's farfield has its
    own hard-coded freestream (Mach 0.8 at 1.2532 degrees), which costs the
    same."""
    O = _oracle()
    from fvhyper_amd import meshgen
    info = cpu_info()
    ranks = ranks or info["ranks"]
    C = mesh_arrays.n_cells
    if O.ref_available():
        try:
            work = tempfile.mkdtemp(prefix="fvh_cpu_")
            meshgen.write_msh(os.path.join(work, "case.msh"), mesh_arrays)
            cfg2 = dict(cfg)
            cfg2["K"] = 7.5   # the reference library fixes K; the cost is the same
            cfg2.pop("positivity", None)
            kv = O.ref_args(cfg2)
            kv.update(ic="fs", mach="3") if workload == "step" else kv.update(ic="nc")
            kv.update(max_step=str(steps), print=str(steps))
            r = O.run_ref("run", work, "case", flags=O.ref_flags(cfg2), timeout=900, **kv)
            secs = float(r["stdout"].strip().splitlines()[-1].split(",")[2])
            one = {"value": C * 5 * steps / secs / 1e6, "unit": "Mcell-stage/s", "cores": 1, "kind": "reference",
                   "cpu": info["model"],
                   "sample": "%d RK5 steps (%d stages) of the %d-cell %s workload, fvhyper library built from its "
                             "sources (-O3), 1 rank, RealTime loop clock" % (steps, 5 * steps, C, workload)}
            if ranks < 2 or not O.mpi_available():
                return one
            try:
                part = meshgen.partition_rcb(mesh_arrays, ranks)
                meshgen.write_partitioned_msh(work, "case", mesh_arrays, part)
                steps_p = max(steps, min(steps * ranks, 96))
                kv.update(max_step=str(steps_p), print=str(steps_p))
                rp = O.run_ref_mpi("run", work, "case", ranks, flags=O.ref_flags(cfg2), timeout=900, **kv)
                secs_p = float(rp["stdout"].strip().splitlines()[-1].split(",")[2])
                return {"value": C * 5 * steps_p / secs_p / 1e6, "unit": "Mcell-stage/s", "cores": ranks,
                        "kind": "reference", "cpu": info["model"], "host_cores": info["nproc"],
                        "sample": "%d RK5 steps (%d stages) of the %d-cell %s workload, fvhyper library built "
                                  "from its sources (-O3), %d MPI ranks (mpirun, the CPU share one GPU has on "
                                  "the host) on a recursive-bisection partition with one ghost layer, RealTime "
                                  "loop clock of rank 0" % (steps_p, 5 * steps_p, C, workload, ranks),
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
            "cpu": info["model"],
            "sample": "%d RK5 steps of the %d-cell %s workload, C restatement, 1 thread" % (steps, C, workload)}


def sweep_mesh_names(items):
    """File name of each instance's mesh (one file per distinct mesh)."""
    return ["m%06d" % it[0].cells.shape[0] for it in items]


def sweep_write_meshes(items, work, part=0, parts=1):
    """The distinct instance meshes of a sweep as MSH 4.1 in `work` (the
    physical names the reference's naca plugin binds: airfoil, farfield);
    this call writes every parts-th of them, from the part-th."""
    from fvhyper_amd import meshgen
    seen = []
    for it, nm in zip(items, sweep_mesh_names(items)):
        if nm not in seen:
            seen.append(nm)
            if (len(seen) - 1) % parts == part:
                meshgen.write_msh(os.path.join(work, nm + ".msh"), it[0])
    return sweep_mesh_names(items)


def wait_for_release(a):
    """A gated worker blocks until its parent closes its stdin (release);
    it exits if the parent ended without releasing it."""
    if a.wait_stdin:
        parent = os.getppid()
        sys.stdin.read()
        if os.getppid() != parent:
            sys.exit(1)


def release(procs):
    """Let gated workers start: close the stdin each one waits on."""
    for p in procs:
        if p.stdin is not None:
            p.stdin.close()
            p.stdin = None


def mesh_writer_worker(a):
    """CPU only: write this worker's share of the sweep meshes."""
    wait_for_release(a)
    from fvhyper_amd import workloads
    part, parts = (int(x) for x in a.mesh_writer.split("/"))
    sweep_write_meshes(workloads.naca_sweep(a.instances), a.mesh_dir, part, parts)


def start_mesh_writers(a, work, procs):
    """Spawn `procs` mesh writers (lowest priority), gated: each waits for
    release()."""
    out = []
    for k in range(procs):
        cmd = ["nice", "-n", "19", sys.executable, os.path.abspath(__file__), "--mesh-writer", "%d/%d" % (k, procs),
               "--mesh-dir", work, "--instances", str(a.instances), "--wait-stdin"]
        out.append(subprocess.Popen(cmd, stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                    text=True))
    return out


def sweep_cpu_baseline(items, files, work, steps: int, ranks: int):
    """The sweep on the CPU as the reference runs a sweep: one single-rank
    reference process per instance (the naca plugin's farfield has a
    hard-coded freestream, which costs the same for every Mach and angle),
    `ranks` processes at a time.  Each process reports its RealTime loop
    time; the aggregate rate is ranks x (total cell-stages / total loop
    time)."""
    O = _oracle()
    info = cpu_info()
    if not O.ref_available():
        return None
    cfg2 = dict(items[0][1])
    cfg2["K"] = 7.5
    kv = O.ref_args(cfg2)
    kv.update(ic="nc", max_step=str(steps), print=str(steps))
    flags = O.ref_flags(cfg2)

    def one(k):
        # instances share meshes: a dump file of its own per instance
        r = O.run_ref("run", work, files[k], flags=flags, timeout=900, tag="i%d" % k, **kv)
        return float(r["stdout"].strip().splitlines()[-1].split(",")[2])

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=ranks) as ex:
        secs = list(ex.map(one, range(len(items))))
    wall = time.perf_counter() - t0
    cells = [it[0].cells.shape[0] for it in items]
    return {"value": ranks * sum(cells) * 5 * steps / sum(secs) / 1e6, "unit": "Mcell-stage/s", "cores": ranks,
            "kind": "reference", "cpu": info["model"], "host_cores": info["nproc"],
            "sample": "the %d instances, %d RK5 steps each, one single-rank fvhyper reference process per instance "
                      "(-O3), %d processes at a time; rate = processes x cell-stages / summed RealTime loop time"
                      % (len(items), steps, ranks),
            "wall_s": round(wall, 2)}


# rocprofv3 PMC summaries committed under profiles/: (workload, precision,
# kernel) -> (file, row prefix).  Each file's traffic table gives HBM bytes per
# dispatch as 2.0 x FETCH_SIZE + WRITE_SIZE (factors from a plain device copy).
PMC_PROFILES = {
    ("step", "f64", "stage"): ("k_stage__mach3_step_975k_fp64.md", "k_stage<double"),
    ("airfoil", "f64", "stage"): ("k_stage__naca0012_2m_fp64.md", "k_stage<double"),
    ("sweep", "f32_f64", "stage"): ("k_stage__naca0012_sweep256_f32_f64.md", "k_stage<double"),
}


def note(msg):
    """A progress line on stderr (the JSON line stays alone on stdout)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


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


def rel_err(a, b) -> float:
    """max_{i,k} |a-b| / max(|a_ik|, max_i |a_.k|): the error measure of the
    FP32 paths against FP64 (a per-variable infinity-norm floor)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64).reshape(-1, 4)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 4)
    scale = np.maximum(np.abs(a), np.max(np.abs(a), axis=0, keepdims=True))
    scale = np.where(scale == 0, 1.0, scale)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def stage_roofline(kt, nbytes, key="stage", kernel="k_stage (fused gradient + limiter + flux + RK update + "
                   "boundary, one launch per stage)"):
    """roofline object of the kernel `key` from event timings kt."""
    if not kt.get(key, {}).get("count", 0):
        return None
    avg_ms = kt[key]["total_ms"] / kt[key]["count"]
    ach = nbytes / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel, "bytes_per_launch": nbytes,
            "avg_launch_us": round(avg_ms * 1e3, 2),
            "kernels_us": {k: round(v["total_ms"] / v["count"] * 1e3, 2) for k, v in kt.items() if v["count"]}}


def event_times(s, steps):
    """Kernel durations: HIP events around every launch on the solver's
    stream over `steps` more steps (eager launches)."""
    s.use_graphs(False)
    s.profile(True)
    s.advance(steps)
    s.synchronize()
    kt = s.kernel_times()
    s.profile(False)
    return kt


def sweep(a, with_cpu=True):
    """Parameter sweep: 256 coarsened airfoil instances (sizes log-uniform in
    5k-200k cells, Mach 0.5-0.85, angle -2..4 degrees), FP32 states with FP64
    arithmetic and residual sums by default, every step of the batch
    replayed from a HIP graph, stage form and tile size from the autotuner
    timed on the whole batch; against the same instances one by one through
    the per-mesh path (the library's default form, each step replayed from
    its graph, one instance after the other).  Returns the result dict."""
    import numpy as np
    import torch

    O = _oracle()
    from fvhyper_amd import Batch, Mesh, Solver, autotune, workloads

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
    t_tune = time.perf_counter()
    sub = list(range(0, len(meshes), a.tune_every))
    tuned = autotune([meshes[k] for k in sub], [cfgs[k] for k in sub], np.concatenate([q0[k] for k in sub]),
                     precision=a.precision, steps=2)
    t_tune = time.perf_counter() - t_tune
    note("sweep: autotuned %s in %.1f s" % ({k: tuned[k] for k in ("flux_mode", "tile_cells")}, t_tune))
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
    note("sweep: %d batched steps in %.3f s" % (a.steps, secs))
    ev_steps = max(2, a.steps // 10)
    kt = event_times(b, ev_steps)
    qb = b.split(b.get_state())
    # the persistent form of the fused stage pass on the same batch: resident
    # workgroups pulling the tiles from per-XCD work queues
    persist = None
    if tuned["flux_mode"] == "fused":
        bp = Batch(meshes, [dict(c, stage_launch=1) for c in cfgs], precision=a.precision, flux_mode="fused",
                   tile_cells=tuned["tile_cells"])
        bp.set_state(bp.join(q0))
        bp.apply_bounds()
        bp.use_graphs(True)
        bp.advance(a.warmup)
        bp.synchronize()
        torch.cuda.synchronize()
        t0p = time.perf_counter()
        bp.advance(a.steps)
        bp.synchronize()
        torch.cuda.synchronize()
        secs_p = time.perf_counter() - t0p
        persist = {"value": round(C * stages * a.steps / secs_p / 1e6, 2),
                   "ms_per_step": round(secs_p / a.steps * 1e3, 4),
                   "launch": "persistent: resident workgroups (as many as the LDS lets a CU hold, times the CUs) pull the tiles "
                             "from per-XCD work queues with stealing, one launch per stage, hipGraph per step"}
        del bp
        note("sweep: persistent batched leg done")
    if not all(np.all(np.isfinite(x)) for x in qb):
        print("non-finite batch state after the timed steps", file=sys.stderr)
        sys.exit(3)
    # error against the FP64 oracle after every step the batch took, on the
    # 8 smallest instances and the median one
    total = a.warmup + a.steps + ev_steps
    order = sorted(range(len(items)), key=lambda k: meshes[k].n_real)
    sample = order[:8] + [order[len(order) // 2]]
    errs = []
    for k in sample:
        qo, *_ = O.Mesh(items[k][0]).run(cfgs[k], q0[k], total)
        errs.append(rel_err(qo, qb[k]))
    note("sweep: oracle check of %d instances done" % len(sample))
    # the same instances one by one: the library's default per-mesh path
    k_pm = a.per_mesh_steps or a.steps

    def per_mesh(flux_mode, graphs, sync_each):
        solvers = []
        for m, c, q in zip(meshes, cfgs, q0):
            x = Solver(m, c, precision=a.precision, flux_mode=flux_mode)
            x.set_state(q)
            x.apply_bounds()
            x.use_graphs(graphs)
            x.advance(2)
            x.synchronize()
            solvers.append(x)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for x in solvers:
            x.advance(k_pm)
            if sync_each:
                x.synchronize()
        for x in solvers:
            x.synchronize()
        dt = time.perf_counter() - t
        del solvers
        return C * stages * k_pm / dt / 1e6, dt

    v_pm, s_pm = per_mesh("auto", True, True)
    note("sweep: per-mesh leg done")
    v_old, s_old = per_mesh("tile", False, True)
    note("sweep: per-mesh two-pass leg done")
    value = C * stages * a.steps / secs / 1e6
    sb, gb = state_bytes(a.precision)
    ab = algorithmic_bytes(C, E, B, sb, gb)
    key = "stage" if kt.get("stage", {}).get("count", 0) else "flux"
    roof = stage_roofline(kt, ab["stage"] if key == "stage" else ab["flux"], key,
                          ("k_stage (fused stage pass)" if key == "stage" else
                           "k_flux (edge-flux gather + RK update + boundary)") + ", all instances")
    if roof is not None:
        achievable(roof)
        tr = pmc_traffic("sweep", a.precision, key) if a.instances == 256 else None
        if tr is not None:
            roof["traffic"] = tr["bytes"]
            roof["traffic_source"] = tr["source"]
    out = {
        "value": round(value, 2), "unit": "Mcell-stage/s", "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(secs / a.steps * 1e3, 4),
        "dtype": "f32" if a.precision == "f32" else "f64",
        "config": {"workload": "naca0012_sweep", "instances": len(meshes), "cells": C, "edges": E,
                   "boundary_edges": B, "cells_range": [5000, 200000], "mach_range": [0.5, 0.85],
                   "aoa_deg_range": [-2.0, 4.0], "order": 2, "limiter": "michalak", "dt": "local",
                   "rk_stages": stages, "precision": a.precision,
                   "state": "f32" if a.precision != "f64" else "f64",
                   "arithmetic": "f32" if a.precision == "f32" else "f64", "residual_sums": "f64",
                   "launch": "batched (one launch per stage for all instances), hipGraph per step",
                   "autotuned": {"flux_mode": tuned["flux_mode"], "tile_cells": tuned["tile_cells"],
                                 "seconds": round(t_tune, 1),
                                 "candidates_ms_per_step": [list(c) for c in tuned["candidates"]]}},
        "per_mesh": {"value": round(v_pm, 2), "steps": k_pm, "seconds": round(s_pm, 3),
                     "launch": "one solver per instance, the library's default form (fused stage pass, "
                               "automatic tiles), each step replayed from its HIP graph, instances one after "
                               "the other (synchronised per instance)"},
        "per_mesh_two_pass_eager": {"value": round(v_old, 2), "seconds": round(s_old, 3),
                                    "launch": "one solver per instance, two-pass tiled stage kernels, eager "
                                              "launches, synchronised per instance"},
        "persistent": persist,
        "speedup_vs_per_mesh": round(value / v_pm, 3),
        "speedup_vs_per_mesh_two_pass_eager": round(value / v_old, 3),
        "rel_err_vs_fp64": {"max": float(max(errs)), "instances": len(sample), "steps": total,
                            "cells": [int(meshes[k].n_real) for k in sample],
                            "measure": "max_ik |a-b| / max(|a_ik|, max_i |a_.k|) against the FP64 oracle"},
        "setup_s": round(setup, 1), "roofline": roof, "cpu_baseline": None,
    }
    if with_cpu and not a.no_cpu:
        info = cpu_info()
        work = tempfile.mkdtemp(prefix="fvh_sweep_")
        files = sweep_write_meshes(items, work)
        out["cpu_baseline"] = sweep_cpu_baseline(items, files, work, a.sweep_cpu_steps, info["ranks"])
    return out


# ---------------------------------------------------------------------------
# complex step at the adjoint's size, in CPU worker processes
# ---------------------------------------------------------------------------
ADJ_N_PER_UNIT = 220          # 243,936 cells
ADJ_DIR_SEED = 20240606
ADJ_SCALE = (1.0, 3.0, 3.0, 10.0)


def adj_direction(n_cells, n_real, k):
    """Direction k of the adjoint checks: seeded normal values scaled per
    variable, zero on the ghost cells (run() overwrites them)."""
    import numpy as np
    v = np.random.default_rng(ADJ_DIR_SEED + k).standard_normal((n_cells, 4)) * np.array(ADJ_SCALE)
    v[n_real:] = 0.0
    return v.ravel()


def adj_smooth_direction(cx, cy, n_real, k):
    """Smooth direction k of the adjoint checks: per variable, eight Gaussian
    bumps of seeded centres (in the cells' bounding box), widths (0.1-0.3)
    and normal amplitudes at the cell centres, scaled to unit RMS over the
    real cells and then per variable as the white-noise directions; zero on
    the ghost cells."""
    import numpy as np
    rng = np.random.default_rng(ADJ_DIR_SEED + 5000 + k)
    x0, x1, y0, y1 = cx[:n_real].min(), cx[:n_real].max(), cy[:n_real].min(), cy[:n_real].max()
    v = np.zeros((cx.size, 4))
    for var in range(4):
        f = np.zeros(cx.size)
        for _ in range(8):
            bx, by, wd = rng.uniform(x0, x1), rng.uniform(y0, y1), rng.uniform(0.1, 0.3)
            f += rng.standard_normal() * np.exp(-((cx - bx) ** 2 + (cy - by) ** 2) / (wd * wd))
        v[:, var] = f / np.sqrt(np.mean(f[:n_real] ** 2)) * ADJ_SCALE[var]
    v[n_real:] = 0.0
    return v.ravel()


def cstep_worker(a):
    """CPU only (no HIP call): the complex-step derivative of J along the
    directions in --cstep-worker on the adjoint workload (an entry "s<k>":
    direction k of the scheme with smoothed kinks), from the state and
    weights the parent wrote to --cstep-dir; prints one JSON line per
    direction."""
    wait_for_release(a)
    import numpy as np
    O = _oracle()
    from fvhyper_amd import workloads
    m, cfg, _ = workloads.forward_step(n_per_unit=ADJ_N_PER_UNIT)
    om = O.Mesh(m)
    q0 = np.load(os.path.join(a.cstep_dir, "q0.npy"))
    w = np.load(os.path.join(a.cstep_dir, "w.npy"))
    for item in a.cstep_worker.split(","):
        smooth = item.startswith("s")
        k = int(item.lstrip("s"))
        t0 = time.perf_counter()
        J, dJ = O.cstep_derivative(om, dict(cfg, smooth_kinks=smooth), q0, adj_direction(om.n_cells, om.n_real, k),
                                   a.adj_steps, wx=w)
        print(json.dumps({"dir": k, "smooth": smooth, "J": J, "dJ": dJ, "seconds": time.perf_counter() - t0}),
              flush=True)


def start_cstep_workers(a, q0, w, n_dirs, procs):
    """Spawn `procs` CPU worker processes (lowest priority) for n_dirs
    directions of the scheme and n_dirs of its smoothed-kink form, gated:
    each waits for release(), so that no GPU leg is timed beside them;
    returns the handles."""
    import numpy as np
    d = tempfile.mkdtemp(prefix="fvh_cstep_")
    np.save(os.path.join(d, "q0.npy"), q0)
    np.save(os.path.join(d, "w.npy"), w)
    items = [str(k) for k in range(n_dirs)] + ["s%d" % k for k in range(n_dirs)]
    groups = [items[i::procs] for i in range(min(procs, len(items)))]
    out = []
    for g in groups:
        cmd = ["nice", "-n", "19", sys.executable, os.path.abspath(__file__), "--cstep-worker",
               ",".join(str(k) for k in g), "--cstep-dir", d, "--adj-steps", str(a.adj_steps), "--wait-stdin"]
        out.append(subprocess.Popen(cmd, stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                    text=True))
    return out


def collect_cstep_workers(procs, timeout=900):
    res = []
    for p in procs:
        o, e = p.communicate(timeout=timeout)
        if p.returncode != 0:
            print("complex-step worker failed: %s" % e[-1000:], file=sys.stderr)
            continue
        res += [json.loads(l) for l in o.strip().splitlines() if l.startswith("{")]
    return sorted(res, key=lambda r: (r.get("smooth", False), r["dir"]))


def adjoint(a):
    """Discrete adjoint: gradient of the step-face pressure force after
    --adj-steps RK5 steps of the second-order forward-step solver (h = 1/220,
    243,936 cells) from a seeded smooth perturbation of the inflow, with
    respect to every initial cell state and, in a second timed run, also to
    every node coordinate.  Reports the adjoint/forward cost ratios, the
    checkpoint memory and central differences (eps 1e-6 and 1e-8) on
    --adj-dirs directions.  Returns the result dict (with the gradient, for
    the complex-step check of the CPU workers)."""
    import numpy as np
    import torch

    from fvhyper_amd import Mesh, Solver, workloads

    m, cfg, ic = workloads.forward_step(n_per_unit=ADJ_N_PER_UNIT)
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
    rel6, rel8 = [], []
    for k in range(a.adj_dirs):
        v = adj_direction(pm.n_cells, pm.n_real, 1000 + k)
        ad = float(grad @ v)
        for eps, out in ((1e-6, rel6), (1e-8, rel8)):
            fd = (s.objective(q0 + eps * v, N, wx=w) - s.objective(q0 - eps * v, N, wx=w)) / (2 * eps)
            out.append(abs(ad - fd) / max(abs(fd), 1e-300))
    # the same run with smoothed kinks (FVH_FLAG_SMOOTH_KINKS): its own
    # adjoint/forward ratio, and fourth- and second-order central differences
    # at eps 1e-6 on the same directions
    cfg_s = dict(cfg, smooth_kinks=True)
    ss = Solver(pm, cfg_s)
    ss.use_graphs(True)
    ss.objective(q0, 2, wx=w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Js = ss.objective(q0, N, wx=w)
    t_fwd_s = time.perf_counter() - t0
    ss.adjoint(q0, 2, wx=w)
    t0 = time.perf_counter()
    _, grad_s, st_s = ss.adjoint(q0, N, wx=w, mem_cap_bytes=8e9)
    t_adj_s = time.perf_counter() - t0
    eps_list = (1e-6, 5e-7, 2.5e-7, 1e-7)

    def fd4_err(v, eps):
        ad = float(grad_s @ v)
        Jv = {j: ss.objective(q0 + j * eps * v, N, wx=w) for j in (-2, -1, 1, 2)}
        d4 = (8 * (Jv[1] - Jv[-1]) - (Jv[2] - Jv[-2])) / (12 * eps)
        d2 = (Jv[1] - Jv[-1]) / (2 * eps)
        return abs(ad - d4) / max(abs(d4), 1e-300), abs(ad - d2) / max(abs(d2), 1e-300)
    fd4 = {e: [] for e in eps_list}
    fd2 = []
    for k in range(a.adj_dirs):
        v = adj_direction(pm.n_cells, pm.n_real, 1000 + k)
        for eps in eps_list:
            e4, e2 = fd4_err(v, eps)
            fd4[eps].append(e4)
            if eps == 1e-6:
                fd2.append(e2)
    # smooth random directions (Gaussian bumps): no grid-scale content
    cx, cy = pm.export("cellsCentersX"), pm.export("cellsCentersY")
    fd4_sm = {e: [] for e in (1e-6, 1e-7)}
    for k in range(min(a.adj_dirs, 16)):
        v = adj_smooth_direction(cx, cy, pm.n_real, k)
        for eps in fd4_sm:
            fd4_sm[eps].append(fd4_err(v, eps)[0])

    def fd_stats(errs):
        return None if not errs else {
            "directions": len(errs), "median_rel_err": float(np.median(errs)), "max_rel_err": float(np.max(errs)),
            "frac_below_1e-6": float(np.mean(np.array(errs) <= 1e-6))}
    kw, kwv, kwl = _oracle().kink_widths(cfg_s)
    smoothed = {
        "what": "the scheme with smoothed kinks (|.|_w, min_w, max_w in the entropy cure, wave speeds and "
                "limiter; sign blend and Venkatakrishnan's regularised form in the limiter; p-norm global step)",
        "widths": {"rho_rhou_rhov_rhoE": kw, "wave_speed": kwv, "limiter_min": kwl},
        "J": Js, "J_minus_reference_scheme": Js - J,
        "forward_s": round(t_fwd_s, 4), "adjoint_s": round(t_adj_s, 4),
        "adjoint_over_forward": round(t_adj_s / t_fwd_s, 3), "recorded_steps": st_s["recorded_steps"],
        "fd4_stencil": "(8 (J(+e) - J(-e)) - (J(+2e) - J(-2e))) / (12 e)",
        "fd4_eps_1e-6": fd_stats(fd4[1e-6]),
        "fd4_eps_5e-7": fd_stats(fd4[5e-7]),
        "fd4_eps_2.5e-7": fd_stats(fd4[2.5e-7]),
        "fd4_eps_1e-7": fd_stats(fd4[1e-7]),
        "fd2_eps_1e-6": fd_stats(fd2),
        "smooth_directions": {
            "what": "Gaussian-bump fields (8 per variable, widths 0.1-0.3, unit RMS, scaled as the white-noise "
                    "directions)",
            "fd4_eps_1e-6": fd_stats(fd4_sm[1e-6]), "fd4_eps_1e-7": fd_stats(fd4_sm[1e-7])},
        "complex_step": None,
    }
    cs = None
    C = pm.n_real
    stages = len(cfg["alpha"])
    return {
        "value": round(C * stages * N / t_fwd / 1e6, 2), "unit": "Mcell-stage/s", "steps": N,
        "ms_per_step": round(t_fwd / N * 1e3, 4), "dtype": "f64",
        "config": {"workload": "forward_step_adjoint", "cells": C, "steps": N, "stages": N * stages,
                   "objective": "pressure force on the step face", "checkpoint_cap_GB": 8.0},
        "J": J, "forward_s": round(t_fwd, 4), "adjoint_s": round(t_adj, 4),
        "adjoint_over_forward": round(t_adj / t_fwd, 3),
        "with_node_gradient_s": round(t_adj_nodes, 4),
        "with_node_gradient_over_forward": round(t_adj_nodes / t_fwd, 3),
        "node_gradient_norm": float(np.linalg.norm(gxy)), "snaps": st["snaps"],
        "forward_steps_executed": st["forward_steps"], "recorded_steps": st["recorded_steps"],
        "library_ms": {"forward_sweep": round(st["forward_ms"], 2), "total": round(st["total_ms"], 2)},
        "checkpoint_GB": round(st["checkpoint_bytes"] / 1e9, 3), "work_GB": round(st["work_bytes"] / 1e9, 3),
        "fd_eps_1e-6": None if not rel6 else {
            "directions": len(rel6), "median_rel_err": float(np.median(rel6)),
            "max_rel_err": float(np.max(rel6)), "frac_below_1e-4": float(np.mean(np.array(rel6) < 1e-4))},
        "fd_eps_1e-8": None if not rel8 else {
            "directions": len(rel8), "median_rel_err": float(np.median(rel8)), "max_rel_err": float(np.max(rel8))},
        "complex_step": cs, "smoothed_kinks": smoothed, "cpu_baseline": None, "_mesh": m, "_cfg": cfg,
        "_grad": grad, "_grad_s": grad_s, "_J_s": Js,
        "_n_cells": pm.n_cells, "_n_real": pm.n_real,
    }


def step_workload(a, precision):
    """The Mach-3 forward step (or another single-mesh workload): the timed
    region, kernel durations and roofline.  Returns (result dict, mesh
    arrays, cfg)."""
    import numpy as np
    import torch

    from fvhyper_amd import Mesh, Solver, workloads

    if a.workload == "plate":
        m, cfg, ic = workloads.flat_plate(smooth=a.smooth)
    elif a.workload == "airfoil":
        m, cfg, ic = workloads.naca_airfoil(a.cells)
        if a.smooth:
            cfg["smooth"] = True
        if a.cfl > 0:
            cfg["cfl"] = a.cfl
    else:
        m, cfg, ic = workloads.forward_step(n_per_unit=a.n_per_unit)
    if a.stage_threads:
        cfg["stage_threads"] = a.stage_threads
    cfg["stage_launch"] = a.stage_launch
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg, precision=precision, flux_mode=a.flux_mode, tile_cells=a.tile_cells)
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
    ev_steps = max(20, a.steps // 10)
    kt = {} if a.no_events else event_times(s, ev_steps)
    q = s.get_state()
    if not np.all(np.isfinite(q)):
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)
    calibrate_copy(a)
    value = C * stages * a.steps / secs / 1e6
    sb, gb = state_bytes(precision)
    ab = algorithmic_bytes(C, E, B, sb, gb)
    roof = None
    if kt and kt.get("stage", {}).get("count", 0):
        roof = stage_roofline(kt, ab["stage"])
    elif kt and a.workload == "plate":
        # viscous two-pass stage: the gradient pass also writes gx, gy; the
        # flux pass reads them and the cell centres (face gradients)
        V = 4
        H = 2 * E - B
        per = {"gradlim": ab["gradlim"] + C * 2 * V * sb, "flux": ab["flux"] + C * (2 * V * sb + 2 * gb)}
        if a.smooth:
            per["smooth"] = (C * (2 * V * sb + 4) + H * 8 + C * (4 * V * sb + 4 + 2 * sb) + H * 8) // 2
        tot = {k: kt[k]["total_ms"] for k in per if kt.get(k, {}).get("count", 0)}
        key = max(tot, key=tot.get)
        roof = stage_roofline(kt, per[key], key, "k_" + key + " (viscous two-pass stage" +
                              (", residual smoothing" if a.smooth else "") + ")")
    elif kt and kt.get("flux", {}).get("count", 0):
        roof = stage_roofline(kt, ab["flux"], "flux", "k_flux (edge-flux gather + RK update + boundary)")
    if roof is not None:
        achievable(roof)
        roof["events_window_steps"] = ev_steps
        # the default line's step leg runs the step workload
        wname = "step" if a.workload == "headline" else a.workload
        tr = pmc_traffic(wname, precision, "stage" if "stage" in roof["kernel"] else "flux") \
            if a.workload != "plate" else None
        if tr is not None and (wname != "step" or a.n_per_unit == 440) and (
                a.workload != "airfoil" or a.cells == 2_000_000):
            roof["traffic"] = tr["bytes"]
            roof["traffic_source"] = tr["source"]
            roof["frac_traffic"] = round(tr["bytes"] / (roof["avg_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
        stage_ms = sum(v["total_ms"] for v in kt.values()) / (ev_steps * stages)
        roof["stage_achieved"] = round(ab["stage"] / (stage_ms * 1e-3) / 1e9, 1)
    conv = None
    if a.converge:
        s.set_state(q0)
        conv = converge(s, C, stages, 1e-6, a.stop_var)
    if a.workload == "plate":
        wl = {"workload": "laminar_flat_plate", "cells": C, "edges": E, "boundary_edges": B, "order": 2,
              "limiter": "michalak", "flux": "roe + laminar viscous (diffusive face gradients)",
              "residual_smoothing": a.smooth, "rk_stages": stages, "cfl": cfg["cfl"], "dt": "local viscous",
              "precision": precision, "stage_form": "tile"}
    elif a.workload == "airfoil":
        wl = {"workload": "naca0012_transonic", "cells": C, "edges": E, "boundary_edges": B,
              "order": 2, "limiter": "michalak", "mach": 0.8, "aoa_deg": 1.25, "rk_stages": stages,
              "cfl": cfg["cfl"], "dt": "local", "precision": precision, "stage_form": a.flux_mode,
              "residual_smoothing": bool(cfg.get("smooth"))}
    else:
        wl = {"workload": "mach3_forward_step", "cells": C, "edges": E, "boundary_edges": B,
              "order": 2, "limiter": "venkatakrishnan", "K": cfg["K"], "rk_stages": stages,
              "cfl": cfg["cfl"], "dt": "global", "precision": precision, "stage_form": a.flux_mode}
    wl["parallelism"] = "single_gpu"
    wl["stage_launch"] = "persistent" if a.stage_launch else "grid"
    wl["tile_cells"] = s.tile_cells
    wl["stage_layout"] = s.stage_info
    out = {"value": round(value, 2), "ms_per_step": round(secs / a.steps * 1e3, 4),
           "dtype": "f32" if precision == "f32" else "f64", "config": wl, "roofline": roof}
    if precision == "f64":
        out["parity"] = "FP64: bit-identical to the oracle, itself bit-identical to the reference (GPU tests)"
    if conv is not None:
        out["converge"] = conv
    return out, m, cfg


def line(result, a):
    """The top-level JSON fields around a workload result."""
    out = {"metric": "Mcell-stage/s", "value": result["value"], "unit": "Mcell-stage/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": result["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": result["dtype"], "data": "synthetic"}
    for k, v in result.items():
        if k not in out and not k.startswith("_"):
            out[k] = v
    return out


def headline(a):
    """The default line: the step workload at top level, the sweep and the
    adjoint as sub-objects; the CPU work (complex-step workers, sweep mesh
    writers, CPU baselines) after the GPU legs, so that none of them is
    timed beside it."""
    import numpy as np
    import torch

    from fvhyper_amd import Mesh, workloads
    t_start = time.perf_counter()
    info = cpu_info()
    # the CPU workers (complex step, sweep meshes for the reference runs)
    # are spawned before any HIP call of this process and wait until the
    # GPU legs are done
    sweep_dir = tempfile.mkdtemp(prefix="fvh_sweep_")
    writers = [] if a.no_cpu else start_mesh_writers(a, sweep_dir, 4)
    workers = []
    if a.cstep_dirs > 0:
        ma, _, _ = workloads.forward_step(n_per_unit=ADJ_N_PER_UNIT)
        pma = Mesh.from_arrays(ma)
        workers = start_cstep_workers(a, workloads.bump_state(pma), workloads.step_face_weights(pma), a.cstep_dirs,
                                      min(a.cstep_dirs, info["ranks"]))
    torch.cuda.set_device(0)
    step_res, m, cfg = step_workload(a, "f64")
    sw = sweep(argparse.Namespace(**dict(vars(a), precision=a.sweep_precision)), with_cpu=False)
    adj = adjoint(a)
    t_gpu = time.perf_counter() - t_start
    release(writers + workers)
    # the workers' results, then the CPU baselines on the freed cores
    cs = collect_cstep_workers(workers) if workers else []
    if cs:
        adj["complex_step"] = adjoint_complex_step_check(adj, cs)
        adj["smoothed_kinks"]["complex_step"] = adjoint_complex_step_check(adj, cs, smooth=True)
    if not a.no_cpu:
        step_res["cpu_baseline"] = cpu_baseline(m, cfg, a.cpu_steps, "step", info["ranks"])
        items = workloads.naca_sweep(a.instances)
        for p in writers:
            _, e = p.communicate(timeout=900)
            if p.returncode != 0:
                raise RuntimeError("sweep mesh writer failed: %s" % e[-1000:])
        sw["cpu_baseline"] = sweep_cpu_baseline(items, sweep_mesh_names(items), sweep_dir, a.sweep_cpu_steps,
                                                info["ranks"])
        adj["cpu_baseline"] = cpu_baseline(adj["_mesh"], adj["_cfg"], a.cpu_steps, "step", info["ranks"])
        if adj["cpu_baseline"] is not None:
            adj["cpu_baseline"]["note"] = ("the reference's forward solve of the adjoint's mesh (the reference "
                                           "has no adjoint); one complex-step direction costs cpu_s_per_direction")
    out = line(step_res, a)
    out["sweep"] = sw
    out["adjoint"] = {k: v for k, v in adj.items() if not k.startswith("_")}
    out["headline"] = {"batched_mcell_stage_s": sw["value"], "per_mesh_mcell_stage_s": sw["per_mesh"]["value"],
                       "batched_vs_per_mesh": sw["speedup_vs_per_mesh"],
                       "sweep_rel_err_vs_fp64": sw["rel_err_vs_fp64"]["max"],
                       "adjoint_over_forward": adj["adjoint_over_forward"],
                       "adjoint_complex_step_max_rel_err": (adj["complex_step"] or {}).get("max_rel_err"),
                       "adjoint_smoothed_fd4_eps_1e-6_max_rel_err":
                           (adj["smoothed_kinks"]["fd4_eps_1e-6"] or {}).get("max_rel_err"),
                       "adjoint_smoothed_fd4_eps_1e-7_max_rel_err":
                           (adj["smoothed_kinks"]["fd4_eps_1e-7"] or {}).get("max_rel_err"),
                       "adjoint_smoothed_complex_step_max_rel_err":
                           (adj["smoothed_kinks"]["complex_step"] or {}).get("max_rel_err"),
                       "cpu": info["model"], "cpu_cores_used": info["ranks"]}
    out["wall_s"] = {"gpu_part": round(t_gpu, 1), "total": round(time.perf_counter() - t_start, 1)}
    print(json.dumps(out))


def adjoint_complex_step_check(adj, cs, smooth=False):
    """Compare the complex-step derivatives of the CPU workers with the
    adjoint gradient along the same directions (the scheme, or its
    smoothed-kink form)."""
    import numpy as np
    cs = [r for r in cs if r.get("smooth", False) == smooth]
    if not cs:
        return None
    grad = adj["_grad_s"] if smooth else adj["_grad"]
    J = adj["_J_s"] if smooth else adj["J"]
    n_cells, n_real = adj["_n_cells"], adj["_n_real"]
    errs = [abs(float(grad @ adj_direction(n_cells, n_real, r["dir"])) - r["dJ"]) / abs(r["dJ"]) for r in cs]
    return {"cells": n_real, "steps": adj["steps"], "directions": len(errs), "max_rel_err": float(max(errs)),
            "median_rel_err": float(np.median(errs)),
            "J_cpu_minus_gpu": float(max(abs(r["J"] - J) for r in cs)),
            "cpu_s_per_direction": round(float(np.median([r["seconds"] for r in cs])), 1),
            "where": "CPU worker processes, the templated restatement on complex numbers (h = 1e-30), run after "
                     "the GPU legs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default=None, choices=["f64", "f32", "f32_f64"],
                    help="single workloads: default f64 (f32_f64 for the sweep: FP32 states, FP64 arithmetic)")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "step", "airfoil", "sweep", "adjoint", "plate"])
    ap.add_argument("--sweep-precision", default="f32_f64", choices=["f64", "f32", "f32_f64"],
                    help="headline: precision of the sweep")
    ap.add_argument("--smooth", action="store_true", help="plate, airfoil: with residual smoothing")
    ap.add_argument("--cfl", type=float, default=0.0, help="airfoil: CFL number instead of the workload's 2")
    ap.add_argument("--adj-steps", type=int, default=80, help="adjoint: RK5 steps of the objective's run")
    ap.add_argument("--adj-dirs", type=int, default=64, help="adjoint: random directions of the central differences")
    ap.add_argument("--cstep-dirs", type=int, default=16,
                    help="headline: directions of the complex-step check at the adjoint's size (CPU workers)")
    ap.add_argument("--cstep-worker", default="", help=argparse.SUPPRESS)
    ap.add_argument("--cstep-dir", default="", help=argparse.SUPPRESS)
    ap.add_argument("--mesh-writer", default="", help=argparse.SUPPRESS)
    ap.add_argument("--mesh-dir", default="", help=argparse.SUPPRESS)
    ap.add_argument("--wait-stdin", action="store_true", help=argparse.SUPPRESS)
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
    ap.add_argument("--stage-launch", type=int, default=0, choices=[0, 1],
                    help="single workloads: fused stage pass as one workgroup per tile (0) or persistent (1)")
    ap.add_argument("--tile-cells", type=int, default=0,
                    help="step/airfoil/plate: cells per tile (32..256; 0: the library's automatic choice)")
    ap.add_argument("--converge", action="store_true",
                    help="airfoil: also run from the freestream until the density residual drops 6 orders")
    ap.add_argument("--stop-var", type=int, default=0, help="with --converge: 0 density, -1 max over variables")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--sweep-cpu-steps", type=int, default=2, help="steps of each reference process of the sweep")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-launch events")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run a plain torch copy of this size (for rocprof counter scaling)")
    a = ap.parse_args()
    if a.cstep_worker:
        return cstep_worker(a)
    if a.mesh_writer:
        return mesh_writer_worker(a)
    if a.precision is None:
        a.precision = "f32_f64" if a.workload == "sweep" else "f64"
    if a.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.workload == "headline":
        return headline(a)
    if a.workload == "adjoint":
        workers = []
        if a.cstep_dirs > 0:
            from fvhyper_amd import Mesh, workloads
            ma, _, _ = workloads.forward_step(n_per_unit=ADJ_N_PER_UNIT)
            pma = Mesh.from_arrays(ma)
            workers = start_cstep_workers(a, workloads.bump_state(pma), workloads.step_face_weights(pma),
                                          a.cstep_dirs, min(a.cstep_dirs, cpu_info()["ranks"]))
        import torch
        torch.cuda.set_device(0)
        r = adjoint(a)
        release(workers)
        cs = collect_cstep_workers(workers) if workers else []
        if cs:
            r["complex_step"] = adjoint_complex_step_check(r, cs)
            r["smoothed_kinks"]["complex_step"] = adjoint_complex_step_check(r, cs, smooth=True)
        print(json.dumps(line({k: v for k, v in r.items() if not k.startswith("_")}, a)))
        return

    import torch
    torch.cuda.set_device(0)
    if a.workload == "sweep":
        r = sweep(a)
        print(json.dumps(line(r, a)))
        return
    r, m, cfg = step_workload(a, a.precision)
    r["cpu_baseline"] = None if (a.no_cpu or a.workload == "plate") else \
        cpu_baseline(m, cfg, a.cpu_steps, a.workload, cpu_info()["ranks"])
    print(json.dumps(line(r, a)))


if __name__ == "__main__":
    main()
