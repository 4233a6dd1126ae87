"""The persistent form of the fused stage pass (stage_launch = 1): resident
workgroups pulling tiles from per-XCD work queues must give the bits of the
one-workgroup-per-tile grid -- single meshes and batches, eager launches and
HIP-graph replays (the queues reset themselves between launches), FP64 and
the FP32-state mode, the adjoint's replays -- and the oracle's."""
import numpy as np
import pytest

import oracle as O
from fvhyper_amd import Batch, Mesh, Solver, workloads
from helpers import bits_equal

pytestmark = pytest.mark.gpu


def _pair(mesh, cfg, q0, steps, graphs, precision="f64"):
    out = []
    for launch in (0, 1):
        s = Solver(mesh, dict(cfg, stage_launch=launch), precision=precision, flux_mode="fused")
        s.use_graphs(graphs)
        s.set_state(q0)
        st = s.run(max_step=steps)
        out.append((s.get_state(), st))
    return out


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("precision", ["f64", "f32_f64"])
def test_persistent_stage_equals_one_shot_grid(graphs, precision):
    print("case persistent vs grid", graphs, precision, flush=True)
    m, cfg, ic = workloads.forward_step(n_per_unit=60)
    pm = Mesh.from_arrays(m)
    q0 = workloads.initial_state(pm, ic)
    (qa, sa), (qb, sb) = _pair(pm, cfg, q0, 12, graphs, precision)
    assert bits_equal(qa, qb)
    assert sa["step"] == sb["step"] == 12 and sa["time"] == sb["time"]
    np.testing.assert_array_equal(sa["R"], sb["R"])
    if precision == "f64":
        om = O.Mesh(m)
        assert bits_equal(qb, om.run(cfg, q0, 12)[0])


def test_persistent_batch_equals_one_shot_grid():
    """A 12-instance airfoil sweep of uneven sizes, 25 steps replayed from
    the step graph (the queues must reset themselves after every launch):
    every instance equals its one-shot twin bit for bit."""
    print("case persistent batch", flush=True)
    items = workloads.naca_sweep(12, seed=5, min_cells=2000, max_cells=30000)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    cfgs = [it[1] for it in items]
    q0 = [workloads.initial_state(m, it[2]) for m, it in zip(meshes, items)]
    res = []
    for launch in (0, 1):
        b = Batch(meshes, [dict(c, stage_launch=launch) for c in cfgs], flux_mode="fused")
        b.use_graphs(True)
        b.set_state(b.join(q0))
        b.apply_bounds()
        b.advance(25)
        res.append(b.split(b.get_state()))
    for k in range(len(items)):
        assert bits_equal(res[0][k], res[1][k]), k


def test_persistent_smoothed_kinks_and_adjoint_replays():
    """The adjoint's replays and the smoothed-kink kernels through the
    persistent form: the same gradient bit for bit."""
    m, cfg, ic = workloads.forward_step(n_per_unit=30)
    pm = Mesh.from_arrays(m)
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    for c in (cfg, dict(cfg, smooth_kinks=True)):
        a = Solver(pm, dict(c, stage_launch=0), flux_mode="fused").adjoint(q0, 8, wx=w, snaps=3)
        b = Solver(pm, dict(c, stage_launch=1), flux_mode="fused").adjoint(q0, 8, wx=w, snaps=3)
        assert a[0] == b[0] and np.array_equal(a[1], b[1])
