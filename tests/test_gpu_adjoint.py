"""The discrete adjoint on the GPU: the gradient of the step-face pressure
force after N RK5 steps of the second-order forward-step solver with respect
to the initial state, against central differences of the forward solver, and
independent of the checkpoint schedule."""
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


def test_adjoint_with_step_graphs_is_unchanged():
    pm, cfg, w, q0 = _case()
    a = Solver(pm, cfg).adjoint(q0, 5, wx=w)
    s = Solver(pm, cfg)
    s.use_graphs(True)
    b = s.adjoint(q0, 5, wx=w)
    c = s.adjoint(q0, 5, wx=w, snaps=2)
    assert a[0] == b[0] == c[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], c[1])


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
