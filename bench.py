"""Benchmark: explicit second-order RK5 on the Mach-3 forward-step workload.

    python bench.py [--gpus 1] [--steps 400] [--warmup 20] [--precision f64]

One step = one RK5 step (5 stages: gradient+limiter, flux+update+boundary)
plus the time-step pass, on the 975,744-cell forward-step mesh, FP64,
inputs resident in HBM.  Prints one JSON line with the throughput in
Mcell-stage/s, the HBM roofline of the dominant kernel (measured with HIP
events on the solver's stream over the timed region) and a CPU baseline
(the reference library on a bounded sample of the same workload).
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


def cpu_baseline(mesh_arrays, cfg, steps: int):
    """Time the reference library (oracle/_ref/fvh_ref, kind 'reference') on
    `steps` RK5 steps of the same mesh; falls back to the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fvhyper_amd import meshgen

    C = mesh_arrays.n_cells
    if O.ref_available():
        try:
            work = tempfile.mkdtemp(prefix="fvh_cpu_")
            meshgen.write_msh(os.path.join(work, "step.msh"), mesh_arrays)
            cfg2 = dict(cfg)
            cfg2["K"] = 7.5   # the reference library fixes K; the cost is the same
            kv = O.ref_args(cfg2)
            kv.update(ic="fs", mach="3", max_step=str(steps), print=str(steps))
            r = O.run_ref("run", work, "step", flags=O.ref_flags(cfg2), timeout=900, **kv)
            last = r["stdout"].strip().splitlines()[-1].split(",")
            secs = float(last[2])
            return {"value": C * 5 * steps / secs / 1e6, "unit": "Mcell-stage/s", "cores": 1,
                    "kind": "reference",
                    "sample": "%d RK5 steps (%d stages) of the %d-cell step workload, fvhyper library "
                              "built from its sources (-O3), 1 rank, RealTime loop clock" % (steps, 5 * steps, C)}
        except Exception as e:  # pragma: no cover - reported, then the port is timed
            print("reference baseline unavailable: %s" % e, file=sys.stderr)
    om = O.Mesh(mesh_arrays)
    from fvhyper_amd import workloads
    q0 = workloads.initial_state(om, workloads.uniform_ic(1.4, 3.0, 0.0, 1.0))
    t0 = time.perf_counter()
    om.run(cfg, q0, steps)
    secs = time.perf_counter() - t0
    return {"value": C * 5 * steps / secs / 1e6, "unit": "Mcell-stage/s", "cores": 1, "kind": "port",
            "sample": "%d RK5 steps of the %d-cell step workload, C restatement, 1 thread" % (steps, C)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32"])
    ap.add_argument("--n-per-unit", type=int, default=440)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-launch events")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run a plain torch copy of this size (for rocprof counter scaling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from fvhyper_amd import Mesh, Solver, workloads

    torch.cuda.set_device(0)
    m, cfg, ic = workloads.forward_step(n_per_unit=a.n_per_unit)
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg, precision=a.precision)
    q0 = workloads.initial_state(pm, ic)
    s.set_state(q0)
    s.apply_bounds()
    C, E, B = pm.n_real, pm.n_edges, pm.n_bnd
    stages = len(cfg["alpha"])

    s.advance(a.warmup)
    s.synchronize()
    if not a.no_events:
        s.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.advance(a.steps)
    s.synchronize()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    kt = s.kernel_times() if not a.no_events else {}
    s.profile(False)
    q = s.get_state()
    if not np.all(np.isfinite(q)):
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)

    if a.calibrate_copy_mb:
        n = a.calibrate_copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    value = C * stages * a.steps / secs / 1e6
    sb = 4 if a.precision == "f32" else 8
    ab = algorithmic_bytes(C, E, B, sb, sb)
    roof = None
    if kt and kt.get("flux", {}).get("count", 0):
        avg_ms = kt["flux"]["total_ms"] / kt["flux"]["count"]
        ach = ab["flux"] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_flux (edge-flux gather + RK update + boundary)",
                "bytes_per_launch": ab["flux"], "avg_launch_us": round(avg_ms * 1e3, 2)}
        stage_ms = sum(v["total_ms"] for v in kt.values()) / (a.steps * stages)
        roof["stage_achieved"] = round(ab["stage"] / (stage_ms * 1e-3) / 1e9, 1)
        roof["kernels_us"] = {k: round(v["total_ms"] / v["count"] * 1e3, 2) for k, v in kt.items() if v["count"]}
    cpu = None if a.no_cpu else cpu_baseline(m, cfg, a.cpu_steps)
    out = {
        "metric": "Mcell-stage/s", "value": round(value, 2), "unit": "Mcell-stage/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(secs / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "f64" else "f32", "data": "synthetic",
        "config": {"workload": "mach3_forward_step", "cells": C, "edges": E, "boundary_edges": B,
                   "order": 2, "limiter": "venkatakrishnan", "K": cfg["K"], "rk_stages": stages,
                   "cfl": cfg["cfl"], "precision": a.precision, "parallelism": "single_gpu"},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
