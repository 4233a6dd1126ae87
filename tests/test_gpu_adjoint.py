"""The discrete adjoint on the GPU: the gradient of the step-face pressure
force after N RK5 steps of the second-order forward-step solver with respect
to the initial state, against central differences of the forward solver, and
independent of the checkpoint schedule; and its torch.autograd face."""
import numpy as np
import pytest

from fvhyper_amd import Mesh, Solver, meshgen, workloads

pytestmark = pytest.mark.gpu


def _case(n_per_unit=20, reference_compatible=False):
    m, cfg, ic = workloads.forward_step(n_per_unit=n_per_unit, reference_compatible=reference_compatible)
    pm = Mesh.from_arrays(m)
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    return pm, cfg, w, q0


def _directions(pm, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, pm.n_cells, 4))
    v[:, pm.n_real:, :] = 0.0            # ghost entries are overwritten by run()
    v *= np.array([1.0, 3.0, 3.0, 10.0])  # about the size of each conserved variable
    return v.reshape(n, -1)


@pytest.mark.parametrize("steps", [1, 6])
def test_adjoint_matches_central_differences(steps):
    """Step size 1e-8: with 1e-6 the perturbation is large against the
    limiter's structure in this nearly uniform flow (edge limiter values a
    fraction of a percent apart) and central differences straddle the
    piecewise limiter's kinks; the schemes without a limiter agree at 1e-6
    as well."""
    print("case adjoint fd", steps, flush=True)
    pm, cfg, w, q0 = _case()
    s = Solver(pm, cfg)
    J, grad, st = s.adjoint(q0, steps, wx=w)
    assert J == s.objective(q0, steps, wx=w)
    eps = 1e-8
    worst = 0.0
    for v in _directions(pm, 8, 11 + steps):
        fd = (s.objective(q0 + eps * v, steps, wx=w) - s.objective(q0 - eps * v, steps, wx=w)) / (2 * eps)
        ad = grad @ v
        err = abs(ad - fd) / max(abs(fd), 1e-12 * abs(J))
        worst = max(worst, err)
        print("  fd %.12e adjoint %.12e rel %.2e" % (fd, ad, err), flush=True)
    assert worst < 1e-5, worst


@pytest.mark.parametrize("variant", ["first_order_fixed_dt", "reconstruction_global_dt"])
def test_adjoint_without_limiter_matches_differences_at_1e_6(variant):
    m, cfg, ic = workloads.forward_step(n_per_unit=20)
    pm = Mesh.from_arrays(m)
    free = ("inlet_outlet", 1.4, 3.0, 0.0, 1.0)
    base = dict(flux="roe_joined", gamma=1.4, bcs={"wall": ("mirror",), "inlet": free, "outlet": free})
    cfg = dict(base, dt="fixed", dt_fixed=1e-4, gdt=True) if variant == "first_order_fixed_dt" else \
        dict(base, dt="sum_cfl", cfl=0.8, gdt=True, grad=True, lin=True)
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    s = Solver(pm, cfg)
    J, grad, st = s.adjoint(q0, 6, wx=w)
    eps = 1e-6
    for v in _directions(pm, 6, 5):
        fd = (s.objective(q0 + eps * v, 6, wx=w) - s.objective(q0 - eps * v, 6, wx=w)) / (2 * eps)
        assert abs(grad @ v - fd) <= 1e-5 * abs(fd), (grad @ v, fd)


def test_checkpoint_schedule_does_not_change_the_gradient():
    pm, cfg, w, q0 = _case()
    s = Solver(pm, cfg)
    res = [s.adjoint(q0, 12, wx=w, snaps=k) for k in (1, 2, 3, 12)]
    for J, g, st in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g, res[0][1])
    steps = [r[2]["forward_steps"] for r in res]
    print("forward steps per schedule", steps)
    assert steps[3] < steps[2] <= steps[1] <= steps[0]
    assert res[0][2]["snaps"] == 1 and res[3][2]["snaps"] == 12


def test_recorded_steps_do_not_change_the_gradient():
    """With the memory budget left after one checkpoint per step, the
    forward sweep keeps the last steps' stage records (stage states, qt and
    time steps; the reverse recomputes gradients and limiters from them) and
    the reverse reads them instead of replaying: same J and gradients, bit
    for bit, with fewer forward steps; and with a cap that fits every
    record, none is replayed."""
    print("case recorded steps", flush=True)
    pm, cfg, w, q0 = _case()
    s = Solver(pm, cfg)
    n = 12
    ref = s.adjoint(q0, n, wx=w, snaps=n)
    J0, g0, x0, st0 = s.adjoint(q0, n, wx=w, snaps=n, nodes=True)
    assert ref[2]["recorded_steps"] == 0 and st0["recorded_steps"] == 0
    rec = ref[2]["work_bytes"]
    per_step = 11 * pm.n_cells * 32   # about one record: 11 state-sized arrays (6 states, 5 qt)
    for cap, want in ((ref[2]["work_bytes"] + n * pm.n_cells * 32 + 4.5 * per_step, "some"), (8e9, "all")):
        J, g, x, st = s.adjoint(q0, n, wx=w, mem_cap_bytes=cap, nodes=True)
        print("  cap %.3g: recorded %d of %d, forward steps %d" % (cap, st["recorded_steps"], n, st["forward_steps"]),
              flush=True)
        assert J == J0 and np.array_equal(g, g0) and np.array_equal(g, ref[1])
        assert np.allclose(x, x0, rtol=0, atol=1e-13 * np.abs(x0).max())
        if want == "all":
            assert st["recorded_steps"] == n and st["forward_steps"] == n
        else:
            assert 0 < st["recorded_steps"] < n
            assert st["forward_steps"] == 2 * n - st["recorded_steps"]


def test_adjoint_with_step_graphs_is_unchanged():
    pm, cfg, w, q0 = _case()
    a = Solver(pm, cfg).adjoint(q0, 5, wx=w)
    s = Solver(pm, cfg)
    s.use_graphs(True)
    b = s.adjoint(q0, 5, wx=w)
    c = s.adjoint(q0, 5, wx=w, snaps=2)
    assert a[0] == b[0] == c[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], c[1])


@pytest.mark.parametrize("cfg_over", [{}, {"gdt": False, "dt": "fixed", "dt_fixed": 2e-4}])
def test_fused_and_two_pass_replays_agree(cfg_over):
    """The reverse sweep replays steps with the fused stage pass (storing
    gradients, limiters and qt) when the solver runs fused, with the
    two-pass kernels otherwise: the same records, so the same gradients
    bit for bit, with global and with fixed time steps, for every
    checkpoint schedule.  (Same tiles in both: the global time step's
    adjoint sums per-tile partials.)"""
    print("case fused vs two-pass replays", cfg_over, flush=True)
    pm, cfg, w, q0 = _case()
    cfg = dict(cfg, **cfg_over)
    fused = Solver(pm, cfg, flux_mode="fused", tile_cells=256)
    tile = Solver(pm, cfg, flux_mode="tile", tile_cells=256)
    J = fused.objective(q0, 7, wx=w)
    assert J == tile.objective(q0, 7, wx=w)
    for kw in ({}, {"snaps": 3}, {"snaps": 7}):
        a = fused.adjoint(q0, 7, wx=w, nodes=True, **kw)
        b = tile.adjoint(q0, 7, wx=w, nodes=True, **kw)
        assert a[0] == b[0] == J
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_adjoint_refuses_what_it_does_not_cover():
    from fvhyper_amd import FvhError
    m, cfg, ic = workloads.naca_airfoil(2000)          # local time step
    pm = Mesh.from_arrays(m)
    with pytest.raises(FvhError, match="global or fixed"):
        Solver(pm, cfg).adjoint(workloads.initial_state(pm, ic), 2)
    pm2, cfg2, w, q0 = _case()
    with pytest.raises(FvhError, match="FP64"):
        Solver(pm2, cfg2, precision="f32").adjoint(q0, 2, wx=w)


def test_adjoint_matches_complex_step_of_the_restatement():
    """Against the complex-step derivative of the templated CPU restatement
    (exact on the branch taken, like the adjoint): agreement to 1e-8, the
    full workload (limiter, positivity fallback, global time step), 6 steps."""
    import oracle as O
    m, cfg, ic = workloads.forward_step(n_per_unit=20)
    pm = Mesh.from_arrays(m)
    om = O.Mesh(m)
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    s = Solver(pm, cfg)
    J, grad, st = s.adjoint(q0, 6, wx=w)
    for v in _directions(pm, 4, 21):
        Jc, dJ = O.cstep_derivative(om, cfg, q0, v, 6, wx=w)
        assert abs(Jc - J) <= 1e-12 * abs(J)
        print("  complex step %.15e adjoint %.15e" % (dJ, grad @ v), flush=True)
        assert abs(grad @ v - dJ) <= 1e-8 * abs(dJ)


def _node_directions(m, n, seed, interior=True):
    """Random node displacements of about a tenth of the spacing, interior
    nodes only by default (moving a boundary node moves the boundary)."""
    rng = np.random.default_rng(seed)
    keep = np.ones(m.x.size, dtype=bool)
    if interior:
        keep[np.asarray(m.bnd).ravel()] = False
    h = 0.1 / 20
    return [(np.where(keep, rng.standard_normal(m.x.size), 0.0) * h,
             np.where(keep, rng.standard_normal(m.x.size), 0.0) * h) for _ in range(n)]


@pytest.mark.parametrize("variant", ["workload", "first_order_fixed_dt", "edge_min_dt"])
def test_node_gradient_matches_complex_step_of_the_restatement(variant):
    """dJ/d(node coordinates) against the complex-step derivative of the CPU
    restatement, whose geometry is itself computed over complex node
    coordinates: agreement to 1e-8 along random interior displacements.  In
    the first-order variant also along a displacement of every node,
    boundary included; with the limiter that direction is not
    differentiable here: the boundary nodes are not jittered, so congruent
    boundary cells in the uniform inflow tie exactly in the limiter's
    minimum and one-sided differences disagree (a kink, not a rounding
    effect)."""
    import oracle as O
    print("case node gradient", variant, flush=True)
    m, cfg, ic = workloads.forward_step(n_per_unit=20)
    free = ("inlet_outlet", 1.4, 3.0, 0.0, 1.0)
    if variant == "first_order_fixed_dt":
        cfg = dict(flux="roe_joined", gamma=1.4, dt="fixed", dt_fixed=1e-4, gdt=True,
                   bcs={"wall": ("mirror",), "inlet": free, "outlet": free})
    elif variant == "edge_min_dt":
        cfg = dict(cfg, dt="edge_min", cfl=0.5)
    pm = Mesh.from_arrays(m)
    om = O.Mesh(m)
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    s = Solver(pm, cfg)
    J, grad, gxy, st = s.adjoint(q0, 6, wx=w, nodes=True)
    J2, grad2, _ = s.adjoint(q0, 6, wx=w)
    assert J == J2 and np.array_equal(grad, grad2)   # the state gradient does not change
    dirs = _node_directions(m, 3, 31)
    if variant == "first_order_fixed_dt":
        dirs += _node_directions(m, 1, 32, interior=False)
    for dx, dy in dirs:
        Jc, dJ = O.cstep_derivative_nodes(om, cfg, q0, 6, dx=dx, dy=dy, wx=w)
        ad = float(gxy[:, 0] @ dx + gxy[:, 1] @ dy)
        print("  complex step %.15e adjoint %.15e" % (dJ, ad), flush=True)
        assert abs(Jc - J) <= 1e-12 * abs(J)
        assert abs(ad - dJ) <= 1e-8 * abs(dJ)


def test_node_gradient_is_translation_invariant():
    """The scheme sees only differences of coordinates: moving every node
    by the same vector changes nothing, so the node gradient sums to zero
    in x and in y."""
    print("case node gradient translation", flush=True)
    pm, cfg, w, q0 = _case()
    J, grad, gxy, st = Solver(pm, cfg).adjoint(q0, 4, wx=w, nodes=True)
    scale = np.abs(gxy).sum(axis=0)
    assert np.all(np.abs(gxy.sum(axis=0)) <= 1e-10 * scale), (gxy.sum(axis=0), scale)


def test_node_gradient_matches_central_differences_of_moved_meshes():
    """Central differences of J over meshes rebuilt with displaced interior
    nodes (first-order scheme with a fixed step, where J is smooth in the
    coordinates)."""
    import dataclasses
    print("case node gradient fd", flush=True)
    m, cfg, ic = workloads.forward_step(n_per_unit=20)
    free = ("inlet_outlet", 1.4, 3.0, 0.0, 1.0)
    cfg = dict(flux="roe_joined", gamma=1.4, dt="fixed", dt_fixed=1e-4, gdt=True,
               bcs={"wall": ("mirror",), "inlet": free, "outlet": free})
    pm = Mesh.from_arrays(m)
    w = workloads.step_face_weights(pm)
    q0 = workloads.bump_state(pm)
    J, grad, gxy, st = Solver(pm, cfg).adjoint(q0, 6, wx=w, nodes=True)
    eps = 1e-6
    for dx, dy in _node_directions(m, 2, 41):
        def J_at(t):
            mm = Mesh.from_arrays(dataclasses.replace(m, x=m.x + t * dx, y=m.y + t * dy))
            return Solver(mm, cfg).objective(q0, 6, wx=w)
        fd = (J_at(eps) - J_at(-eps)) / (2 * eps)
        ad = float(gxy[:, 0] @ dx + gxy[:, 1] @ dy)
        print("  fd %.12e adjoint %.12e" % (fd, ad), flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_autograd_function_carries_the_adjoint(device):
    """flow_objective inside a torch graph: J and the chain-ruled gradients
    (through a state parameterised by a scalar, and to the node
    coordinates) equal what Solver.adjoint returns; without grad it runs the
    plain objective."""
    import torch
    from fvhyper_amd.autograd import flow_objective
    print("case autograd function", device, flush=True)
    pm, cfg, w, q0 = _case(n_per_unit=20)
    s = Solver(pm, cfg)
    J0, g0, gxy0, _ = s.adjoint(q0, 6, wx=w, nodes=True)
    v = _directions(pm, 1, seed=77)[0]
    t = torch.zeros((), dtype=torch.float64, device=device, requires_grad=True)
    q = torch.from_numpy(q0).to(device) + t * torch.from_numpy(v).to(device)
    xy = torch.from_numpy(np.stack([pm.export("nodesX"), pm.export("nodesY")], axis=1)).to(device)
    xy.requires_grad_(True)
    J = flow_objective(s, q, 6, wx=w, xy=xy)
    (2.0 * J).backward()
    assert J.item() == J0
    assert t.grad.device.type == device and xy.grad.device.type == device
    assert abs(t.grad.item() - 2.0 * float(g0 @ v)) <= 1e-12 * abs(2.0 * float(g0 @ v))
    assert np.array_equal(xy.grad.cpu().numpy(), 2.0 * gxy0)
    with torch.no_grad():
        J2 = flow_objective(s, torch.from_numpy(q0).to(device), 6, wx=w)
    assert J2.item() == s.objective(q0, 6, wx=w)
    with pytest.raises(ValueError):
        flow_objective(s, q.detach(), 6, wx=w, xy=xy.detach() + 1.0)
