"""Batches on the GPU: many meshes advanced in the same launches must give,
instance by instance, the bit-identical results of the single-mesh solver
(and so of the oracle), with per-instance boundary tables, time steps,
residuals and stop tests; HIP-graph replay must not change a bit."""
import numpy as np
import pytest

import oracle as O
from fvhyper_amd import Batch, Mesh, Solver, meshgen, workloads
from helpers import bits_equal, rel_err

pytestmark = pytest.mark.gpu


def _sweep(n, lo=1500, hi=12000, seed=7):
    return workloads.naca_sweep(n, seed=seed, min_cells=lo, max_cells=hi)


def _states(meshes, items):
    return [workloads.initial_state(m, it[2]) for m, it in zip(meshes, items)]


@pytest.mark.parametrize("precision", ["f64", "f32", "f32_f64"])
def test_batch_equals_single_solvers(precision):
    """6 airfoil instances (different sizes, Mach, angle): 15 steps in one
    batch == 15 steps of each mesh alone, bit for bit."""
    print("case batch vs single", precision, flush=True)
    items = _sweep(6)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    cfgs = [it[1] for it in items]
    q0 = _states(meshes, items)
    b = Batch(meshes, cfgs, precision=precision)
    b.set_state(b.join(q0))
    st = b.run(max_step=15, print_interval=5)
    assert st["step"] == 15 and st["stopped"]
    qb = b.split(b.get_state())
    for k, (m, c) in enumerate(zip(meshes, cfgs)):
        s = Solver(m, c, precision=precision)
        s.set_state(q0[k])
        ss = s.run(max_step=15, print_interval=5)
        assert bits_equal(qb[k], s.get_state()), k
        bs = b.instance_status(k)
        assert bs["step"] == ss["step"] == 15 and bs["time"] == ss["time"]
        np.testing.assert_allclose(bs["R"], ss["R"], rtol=1e-12)


def test_batch_instance_matches_oracle():
    items = _sweep(3, seed=11)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    b = Batch(meshes, [it[1] for it in items])
    q0 = _states(meshes, items)
    b.set_state(b.join(q0))
    b.run(max_step=8)
    qb = b.split(b.get_state())
    for k, it in enumerate(items):
        om = O.Mesh(it[0])
        qo, *_ = om.run(it[1], q0[k], 8)
        assert bits_equal(qb[k], qo), k


@pytest.mark.parametrize("precision", ["f32_f64", "f32"])
def test_fp32_batch_instances_against_fp64_oracle(precision):
    """The sweep's precisions, instance by instance against the FP64
    oracle after 25 steps (the default bench's step count): FP32 states
    with FP64 arithmetic within 1e-6 (and bit for bit the oracle with
    FP32-rounded stored states), FP32 arithmetic within 4e-6."""
    print("case fp32 batch", precision, flush=True)
    items = _sweep(4, lo=3000, hi=9000, seed=13)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    b = Batch(meshes, [it[1] for it in items], precision=precision)
    q0 = _states(meshes, items)
    b.set_state(b.join(q0))
    b.run(max_step=25)
    qb = b.split(b.get_state())
    for k, it in enumerate(items):
        om = O.Mesh(it[0])
        qo, *_ = om.run(it[1], q0[k], 25)
        err = rel_err(qo, qb[k])
        print("instance", k, precision, "rel err", err, flush=True)
        if precision == "f32_f64":
            assert bits_equal(qb[k], om.run_f32_state(it[1], q0[k], 25)), k
            assert err < 1e-6, (k, err)
        else:
            assert err < 4e-6, (k, err)


def test_batch_instances_stop_on_their_own():
    """Density-residual stop per instance: each instance's step count and
    final state equal its own single-mesh run with the same options."""
    items = _sweep(4, lo=1500, hi=6000, seed=5)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    cfgs = [it[1] for it in items]
    q0 = _states(meshes, items)
    b = Batch(meshes, cfgs)
    b.set_state(b.join(q0))
    st = b.run(max_step=3000, tol=0.05, stop_var=0)
    assert st["stopped"]
    qb = b.split(b.get_state())
    steps = []
    for k, (m, c) in enumerate(zip(meshes, cfgs)):
        s = Solver(m, c)
        s.set_state(q0[k])
        ss = s.run(max_step=3000, tol=0.05, stop_var=0)
        bs = b.instance_status(k)
        assert bs["stopped"] and bs["step"] == ss["step"] < 3000
        assert bits_equal(qb[k], s.get_state()), k
        steps.append(ss["step"])
    assert len(set(steps)) > 1, "the instances should stop at different steps"


def test_batch_global_dt_is_per_instance():
    """Global time step: the minimum is taken per instance (segmented)."""
    ms = [meshgen.forward_step(n_per_unit=10), meshgen.forward_step(n_per_unit=20),
          meshgen.forward_step(n_per_unit=15, seed=3)]
    m, cfg, ic = workloads.forward_step(n_per_unit=10)
    meshes = [Mesh.from_arrays(a) for a in ms]
    q0 = [workloads.initial_state(x, ic) for x in meshes]
    b = Batch(meshes, [cfg] * 3)
    b.set_state(b.join(q0))
    b.run(max_step=6)
    qb = b.split(b.get_state())
    for k, x in enumerate(meshes):
        s = Solver(x, cfg)
        s.set_state(q0[k])
        ss = s.run(max_step=6)
        assert bits_equal(qb[k], s.get_state()), k
        assert b.instance_status(k)["time"] == ss["time"]


def test_batch_sums_fp64():
    items = _sweep(3, seed=13)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    b = Batch(meshes, [it[1] for it in items], precision="f32")
    q0 = _states(meshes, items)
    b.set_state(b.join(q0))
    b.run(max_step=5)
    sums = b.sums()
    for k, (m, q) in enumerate(zip(meshes, b.split(b.get_state()))):
            # an FP32 solver holds FP32 areas; the sums themselves are FP64
        A = m.export("cellsAreas")[:m.n_real].astype(np.float32).astype(np.float64)
        ref = (A[:, None] * q.reshape(-1, 4)[:m.n_real]).sum(0)
        np.testing.assert_allclose(sums[k], ref, rtol=1e-12)


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replay_is_bit_exact(batched):
    """Steps replayed from captured HIP graphs, checked after every chunk
    (eager synchronisations between replays) against eager launches."""
    print("case graphs", batched, flush=True)
    items = _sweep(3 if batched else 1, seed=17)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    cfgs = [it[1] for it in items]
    q0 = _states(meshes, items)

    def make():
        if batched:
            x = Batch(meshes, cfgs)
            x.set_state(x.join(q0))
        else:
            x = Solver(meshes[0], cfgs[0])
            x.set_state(q0[0])
        return x

    g, e = make(), make()
    g.use_graphs(True)
    g.run_begin()
    e.run_begin()
    for chunk in range(4):
        sg = g.run_continue(7, max_step=100, print_interval=3)
        se = e.run_continue(7, max_step=100, print_interval=3)
        assert sg["step"] == se["step"] == 7 * (chunk + 1)
        assert bits_equal(g.get_state(), e.get_state()), chunk
        np.testing.assert_array_equal(sg["R"], se["R"])
    g.advance(5)
    e.advance(5)
    g.synchronize()
    assert bits_equal(g.get_state(), e.get_state())


@pytest.mark.parametrize("extra", ["smooth", "viscous"])
def test_batch_with_smoothing_or_viscous_terms_matches_oracle(extra):
    """Residual smoothing and the viscous flux in a batch: each instance's
    sweeps and face gradients stay inside it (bit-identical to the oracle)."""
    print("case batch", extra, flush=True)
    items = _sweep(3, seed=13)
    cfgs = []
    for it in items:
        c = dict(it[1])
        if extra == "smooth":
            c["smooth"] = True
        else:
            c.update(flux="roe_viscous", diffusive=True, mu=0.01, pr=0.72, cp=1.0, r_gas=2.0)
        cfgs.append(c)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    b = Batch(meshes, cfgs)
    q0 = _states(meshes, items)
    b.set_state(b.join(q0))
    b.run(max_step=6)
    qb = b.split(b.get_state())
    for k, it in enumerate(items):
        qo, *_ = O.Mesh(it[0]).run(cfgs[k], q0[k], 6)
        assert bits_equal(qb[k], qo), k
