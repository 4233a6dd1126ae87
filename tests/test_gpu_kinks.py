"""The smoothed-kink extension (FVH_FLAG_SMOOTH_KINKS) on the GPU: forward
runs bit for bit against the oracle's smoothed mode in every stage form and
in a batch, and the discrete adjoint of the smoothed scheme against the
complex-step derivative of the restatement (states and node coordinates) and
against fourth-order central differences: at eps = 1e-6 on small runs, at
eps = 1e-7 on the full-size adjoint workload."""
import numpy as np
import pytest

import oracle as O
from fvhyper_amd import Batch, Mesh, Solver, meshgen, workloads
from helpers import CONFIGS, MESHES, bits_equal, random_state

pytestmark = pytest.mark.gpu

SMOOTH = dict(CONFIGS["second_venkat_global"], smooth_kinks=True)


def _step(n_per_unit=20):
    m, cfg, _ = workloads.forward_step(n_per_unit=n_per_unit)
    pm = Mesh.from_arrays(m)
    return m, pm, dict(cfg, smooth_kinks=True), workloads.step_face_weights(pm), workloads.bump_state(pm)


def _directions(pm, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, pm.n_cells, 4))
    v[:, pm.n_real:, :] = 0.0
    v *= np.array([1.0, 3.0, 3.0, 10.0])
    return v.reshape(n, -1)


@pytest.mark.parametrize("mesh_name", ["mixed", "step", "naca"])
@pytest.mark.parametrize("form", ["fused", "tile", "cell"])
def test_smoothed_run_bit_exact(mesh_name, form):
    print("case smoothed run", mesh_name, form, flush=True)
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    q0 = random_state(om.n_cells, seed=606, mach=0.6)
    cfg = dict(SMOOTH, positivity=True)
    qo, steps, t, *_ = om.run(cfg, q0, 6)
    s = Solver(Mesh.from_arrays(m), cfg, flux_mode=form)
    s.set_state(q0)
    st = s.run(max_step=6)
    assert st["step"] == steps == 6 and st["time"] == t
    assert bits_equal(s.get_state(), qo)


def test_smoothed_batch_equals_single_runs():
    """Two instances of different freestream in one batch: each equals its
    own oracle run bit for bit (the p-norm step is per instance)."""
    m = MESHES["step"]()
    om = O.Mesh(m)
    pm = Mesh.from_arrays(m)
    cfgs = [dict(SMOOTH, positivity=True),
            dict(SMOOTH, positivity=True, bcs=dict(SMOOTH["bcs"], inlet=("inlet_outlet", 1.4, 2.5, 0.1, 1.0)))]
    # the widths come from the first instance's freestream; both must share them
    cfgs[1]["kink_scale"] = cfgs[0]["kink_scale"] = [1.4, 5.6, 5.6, 8.8, 4.0]
    q0 = [random_state(om.n_cells, seed=s, mach=0.4) for s in (1, 2)]
    b = Batch([pm, pm], cfgs)
    b.set_state(b.join(q0))
    b.apply_bounds()
    b.advance(5)
    out = b.split(b.get_state())
    for k in range(2):
        assert bits_equal(out[k], om.run(cfgs[k], q0[k], 5)[0])


def test_smoothed_adjoint_matches_complex_step():
    """The full forward-step workload with smoothed kinks, 6 steps: the
    adjoint gradient against the complex-step derivative of the
    restatement, states and interior node coordinates, to 1e-8."""
    print("case smoothed adjoint complex step", flush=True)
    m, pm, cfg, w, q0 = _step()
    om = O.Mesh(m)
    s = Solver(pm, cfg)
    J, grad, gxy, st = s.adjoint(q0, 6, wx=w, nodes=True)
    assert J == s.objective(q0, 6, wx=w)
    for v in _directions(pm, 4, 71):
        Jc, dJ = O.cstep_derivative(om, cfg, q0, v, 6, wx=w)
        assert abs(Jc - J) <= 1e-12 * abs(J)
        print("  states: complex step %.15e adjoint %.15e" % (dJ, grad @ v), flush=True)
        assert abs(grad @ v - dJ) <= 1e-8 * abs(dJ)
    rng = np.random.default_rng(72)
    keep = np.ones(m.x.size, dtype=bool)
    keep[np.asarray(m.bnd).ravel()] = False
    for _ in range(2):
        dx = np.where(keep, rng.standard_normal(m.x.size), 0.0) * 0.005
        dy = np.where(keep, rng.standard_normal(m.x.size), 0.0) * 0.005
        Jc, dJ = O.cstep_derivative_nodes(om, cfg, q0, 6, dx=dx, dy=dy, wx=w)
        ad = float(gxy[:, 0] @ dx + gxy[:, 1] @ dy)
        print("  nodes: complex step %.15e adjoint %.15e" % (dJ, ad), flush=True)
        assert abs(ad - dJ) <= 1e-8 * abs(dJ)


def test_smoothed_adjoint_matches_fourth_order_differences():
    """An 8,064-cell forward step, 20 steps (100 stages): the adjoint along
    random directions against fourth-order central differences of the GPU
    forward solver at eps = 1e-6, to 1e-6; records, checkpoints and replays
    give the same gradient bit for bit."""
    print("case smoothed adjoint fd", flush=True)
    m, pm, cfg, w, q0 = _step(40)
    s = Solver(pm, cfg)
    J, grad, st = s.adjoint(q0, 20, wx=w)
    _, grad3, st3 = s.adjoint(q0, 20, wx=w, snaps=3)
    assert np.array_equal(grad, grad3) and st3["forward_steps"] > 20
    eps = 1e-6
    for v in _directions(pm, 6, 73):
        Jv = {k: s.objective(q0 + k * eps * v, 20, wx=w) for k in (-2, -1, 1, 2)}
        fd4 = (8 * (Jv[1] - Jv[-1]) - (Jv[2] - Jv[-2])) / (12 * eps)
        print("  fd4 %.12e adjoint %.12e rel %.2e" % (fd4, grad @ v, abs(grad @ v - fd4) / abs(fd4)), flush=True)
        assert abs(grad @ v - fd4) <= 1e-6 * abs(fd4)


def _smooth_directions(pm, n, seed):
    """Per variable, eight Gaussian bumps (seeded centres in the cells'
    bounding box, widths 0.1-0.3, normal amplitudes) at the cell centres,
    unit RMS over the real cells, scaled per variable as _directions."""
    cx, cy = pm.export("cellsCentersX"), pm.export("cellsCentersY")
    nr = pm.n_real
    x0, x1, y0, y1 = cx[:nr].min(), cx[:nr].max(), cy[:nr].min(), cy[:nr].max()
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v = np.zeros((pm.n_cells, 4))
        for var in range(4):
            f = np.zeros(pm.n_cells)
            for _ in range(8):
                bx, by, wd = rng.uniform(x0, x1), rng.uniform(y0, y1), rng.uniform(0.1, 0.3)
                f += rng.standard_normal() * np.exp(-((cx - bx) ** 2 + (cy - by) ** 2) / (wd * wd))
            v[:, var] = f / np.sqrt(np.mean(f[:nr] ** 2)) * [1.0, 3.0, 3.0, 10.0][var]
        v[nr:] = 0.0
        out.append(v.ravel())
    return out


def test_smoothed_adjoint_full_size_fourth_order_differences():
    """The adjoint workload at full size with smoothed kinks: the
    243,936-cell forward step, 80 steps (400 stages).  Fourth-order central
    differences of the GPU forward solver at eps = 1e-7 agree with the
    adjoint to 1e-6 along white-noise directions (every cell perturbed
    independently) and along smooth ones (Gaussian bumps).  At eps = 1e-6
    the Taylor remainder along white-noise directions exceeds 1e-6 at this
    step count; the error falls as eps shrinks, which a kink would not
    allow."""
    print("case smoothed adjoint fd full size", flush=True)
    m, pm, cfg, w, q0 = _step(220)
    assert pm.n_real == 243936
    s = Solver(pm, cfg)
    s.use_graphs(True)
    J, grad, st = s.adjoint(q0, 80, wx=w, mem_cap_bytes=8e9)
    eps = 1e-7
    for v in list(_directions(pm, 4, 91)) + _smooth_directions(pm, 4, 92):
        Jv = {k: s.objective(q0 + k * eps * v, 80, wx=w) for k in (-2, -1, 1, 2)}
        fd4 = (8 * (Jv[1] - Jv[-1]) - (Jv[2] - Jv[-2])) / (12 * eps)
        print("  fd4 %.12e adjoint %.12e rel %.2e" % (fd4, grad @ v, abs(grad @ v - fd4) / abs(fd4)), flush=True)
        assert abs(grad @ v - fd4) <= 1e-6 * abs(fd4)
