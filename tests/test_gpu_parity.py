"""GPU parity: libfvh.so (HIP on gfx950, through the C ABI) against the CPU
oracle on the same meshes and states.  The FP64 path must be bit-exact."""
import numpy as np
import pytest

import oracle as O
from fvhyper_amd import Mesh, Solver, meshgen, workloads
from helpers import AIRFOIL_MESHES, CONFIGS, MESHES, airfoil_bump_state, bits_equal, random_state, rel_err, same_values

pytestmark = pytest.mark.gpu


def _pair(mesh_name, cfg, precision="f64", renumber="tiles", flux_mode="auto"):
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg, precision=precision, renumber=renumber, flux_mode=flux_mode)
    return m, om, pm, s


@pytest.mark.parametrize("flux_mode", ["tile", "cell"])
@pytest.mark.parametrize("mesh_name", sorted(MESHES))
@pytest.mark.parametrize("cfg_name", sorted(CONFIGS))
def test_phases_bit_exact(mesh_name, cfg_name, flux_mode):
    print("case", mesh_name, cfg_name, flux_mode, flush=True)
    cfg = CONFIGS[cfg_name]
    m, om, pm, s = _pair(mesh_name, cfg, flux_mode=flux_mode)
    q = random_state(om.n_cells, seed=101)
    s.set_state(q)
    g = s.phases(alpha=0.2898)
    o = om.phases(cfg, q, alpha=0.2898)
    for key in ("dt", "gx", "gy", "lim", "qt", "qk"):
        assert bits_equal(g[key], o[key]), (key, np.max(np.abs(g[key] - o[key])))


@pytest.mark.parametrize("order,flux_mode", [("tiles", "tile"), ("rcm", "tile"), ("file", "cell"), ("tiles", "fused"),
                                             ("rcm", "fused")])
@pytest.mark.parametrize("mesh_name", sorted(MESHES))
@pytest.mark.parametrize("cfg_name", sorted(CONFIGS))
def test_run_bit_exact(mesh_name, cfg_name, order, flux_mode):
    print("case", mesh_name, cfg_name, order, flux_mode, flush=True)
    cfg = CONFIGS[cfg_name]
    if flux_mode == "fused" and not cfg.get("lin"):
        pytest.skip("the fused stage pass exists for linear reconstruction only")
    if flux_mode == "fused" and cfg.get("smooth"):
        pytest.skip("residual smoothing needs every qt before any update: no fused stage pass")
    if flux_mode == "fused" and cfg.get("flux") == "roe_viscous":
        pytest.skip("the viscous flux runs in the two-pass forms")
    m, om, pm, s = _pair(mesh_name, cfg, renumber=order, flux_mode=flux_mode)
    if mesh_name in AIRFOIL_MESHES:
        q0 = airfoil_bump_state(om.cell_cx, om.cell_cy)
    else:
        q0 = random_state(om.n_cells, seed=202, mach=0.3)
    s.set_state(q0)
    from fvhyper_amd import FvhError
    try:
        st = s.run(max_step=12, print_interval=5)
        stopped = False
    except FvhError as e:
        # some stress configs go non-physical: the library stops after the
        # group of steps under way, and the oracle must have gone
        # non-physical by then too
        assert "non-physical state: instance 0" in str(e)
        st = s.status()
        stopped = True
        bad = om.first_unphysical_step(cfg, q0, 12)
        assert bad is not None and bad < st["step"], (bad, st["step"])
    qg = s.get_state()
    n = st["step"] if stopped else 12
    qo, steps, t, R, R0, _ = om.run(cfg, q0, n, print_interval=5)
    assert st["step"] == steps == n
    assert same_values([st["time"]], [t])
    # a run that blows up must blow up identically
    assert same_values(qg, qo), rel_err(qo, qg)
    if not stopped:
        assert om.first_unphysical_step(cfg, q0, 12) is None
        # residual sums run in another order: equal to rounding
        np.testing.assert_allclose(st["R0"], R0, rtol=1e-12)
        if np.all(np.isfinite(R)):
            np.testing.assert_allclose(st["R"], R, rtol=1e-10)


def test_run_tolerance_stop_matches():
    """Stop on the residual test (explicit.cpp:718-726): same step count."""
    cfg = CONFIGS["first_split_sum"]
    m, om, pm, s = _pair("sod", cfg)
    q0 = workloads.initial_state(om, workloads.sod_ic())
    s.set_state(q0)
    st = s.run(max_step=400, tol=0.5)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 400, tol=0.5)
    assert st["step"] == steps < 400
    assert bits_equal(s.get_state(), qo)


def test_run_max_time_stop_matches():
    cfg = CONFIGS["second_venkat_global"]
    m, om, pm, s = _pair("step", cfg)
    q0 = workloads.initial_state(om, workloads.uniform_ic(1.4, 3.0, 0.0, 1.0))
    s.set_state(q0)
    st = s.run(max_time=0.01)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 100000, max_time=0.01)
    assert st["step"] == steps and st["time"] == t
    assert bits_equal(s.get_state(), qo)


def test_sod_full_size_explicit_euler():
    """The 32,000-cell Sod workload, explicit Euler, CFL 0.5, to t = 0.2."""
    m, cfg, ic = workloads.sod()
    om = O.Mesh(m)
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg)
    q0 = workloads.initial_state(om, ic)
    s.set_state(q0)
    st = s.run(max_time=0.2, print_interval=100)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 100000, max_time=0.2, print_interval=100)
    print("sod steps", steps, "time", t, flush=True)
    assert st["step"] == steps and st["time"] == t
    assert bits_equal(s.get_state(), qo), rel_err(qo, s.get_state())


@pytest.mark.timeout(900)
def test_forward_step_full_size_ninety_steps():
    """The 975,744-cell Mach-3 step workload, 90 RK5 steps (450 stages),
    bit-exact against the oracle.  The positivity fallback takes over faces
    at the step corner between steps 70 and 80 (the same solver without it
    differs by then), so the run covers the fallback at full size."""
    m, cfg, ic = workloads.forward_step()
    assert cfg["positivity"]
    om = O.Mesh(m)
    pm = Mesh.from_arrays(m)
    q0 = workloads.initial_state(om, ic)
    s = Solver(pm, cfg)
    bare = Solver(pm, dict(cfg, positivity=False))
    for x in (s, bare):
        x.set_state(q0)
        x.run_begin()
    a70, b70 = s.run_continue(70), bare.run_continue(70)
    assert a70["step"] == b70["step"] == 70
    assert bits_equal(s.get_state(), bare.get_state())
    s.run_continue(20)
    bare.run_continue(20)
    qg = s.get_state()
    assert not same_values(qg, bare.get_state()), "the fallback never triggered"
    print("gpu done", flush=True)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 90)
    print("oracle done", flush=True)
    assert steps == 90 and s.status()["time"] == t
    assert np.all(np.isfinite(qo))
    assert bits_equal(qg, qo), rel_err(qo, qg)


def _face_states(om, q, gx, gy, lim):
    """Reconstructed face state of each edge side (side 0, side 1)."""
    q4, gx4, gy4, l4 = (a.reshape(-1, 4) for a in (q, gx, gy, lim))
    out = []
    for side in (0, 1):
        c = om.edge_cells[:, side]
        dx = (om.edge_cx - om.cell_cx[c])[:, None]
        dy = (om.edge_cy - om.cell_cy[c])[:, None]
        out.append(q4[c] + (gx4[c] * dx + gy4[c] * dy) * l4[c])
    return out


@pytest.mark.parametrize("flux_mode", ["tile", "cell", "fused"])
def test_positivity_fallback_phases_bit_exact(flux_mode):
    """A state with strong density and pressure contrasts and a limiter that
    never acts (K = 1e4, so sigma = 1 and the reconstruction is the plain
    linear one): a third of the faces reconstruct to rho <= 0 or p <= 0 and
    take the cell value.  GPU phases (two-pass forms) and a 3-step run (every
    form) equal the oracle bit for bit."""
    print("case positivity phases", flux_mode, flush=True)
    cfg = dict(CONFIGS["second_venkat_global"], positivity=True, K=1e4)
    m = MESHES["mixed"]()
    om = O.Mesh(m)
    u = meshgen.uniform01(4242, 2 * om.n_cells).reshape(-1, 2)
    rho = 0.02 + 2.0 * u[:, 0] ** 3
    p = 0.02 + 2.0 * u[:, 1] ** 3
    q0 = np.stack([rho, 0.1 * rho, -0.05 * rho, p / 0.4 + 0.5 * rho * 0.0125], 1).ravel()
    o = om.phases(cfg, q0, alpha=0.2898)
    faces = _face_states(om, q0, o["gx"], o["gy"], o["lim"])
    bad = 0
    for f in faces:
        pf = 0.4 * (f[:, 3] - 0.5 * (f[:, 1] ** 2 + f[:, 2] ** 2) / f[:, 0])
        bad += int(np.sum(~((f[:, 0] > 0) & (pf > 0))))
    print("non-physical faces", bad, flush=True)
    assert bad > 0
    s = Solver(Mesh.from_arrays(m), cfg, flux_mode=flux_mode)
    s.set_state(q0)
    if flux_mode != "fused":
        g = s.phases(alpha=0.2898)
        for key in ("dt", "gx", "gy", "lim", "qt", "qk"):
            assert bits_equal(g[key], o[key]), key
    s.set_state(q0)
    s.run(max_step=3)
    qo, *_ = om.run(cfg, q0, 3)
    assert same_values(s.get_state(), qo)


def test_positivity_fallback_bit_exact():
    """The forward-step workload at 40 cells per unit for 150 RK5 steps: the
    positivity fallback triggers at the step corner from step ~56 on; GPU and
    oracle must take it identically."""
    m, cfg, ic = workloads.forward_step(n_per_unit=40)
    assert cfg["positivity"]
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg)
    q0 = workloads.initial_state(om, ic)
    s.set_state(q0)
    st = s.run(max_step=150)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 150)
    assert st["step"] == steps == 150 and st["time"] == t
    assert np.all(np.isfinite(qo))
    assert bits_equal(s.get_state(), qo), rel_err(qo, s.get_state())


def test_reference_scheme_blows_up_identically():
    """Without the fallback the reference scheme goes non-finite at the step
    corner; the GPU reproduces the same trajectory up to and through it, and
    the run stops with the non-physical-state error once a stage leaves a
    cell with rho <= 0, p <= 0 or NaN."""
    from fvhyper_amd import FvhError
    m, cfg, ic = workloads.forward_step(n_per_unit=40, reference_compatible=True)
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg)
    q0 = workloads.initial_state(om, ic)
    s.set_state(q0)
    with pytest.raises(FvhError, match="non-physical state: instance 0 by step"):
        s.run(max_step=130)
    st = s.status()
    assert 0 < st["step"] <= 130
    qo, *_ = om.run(cfg, q0, st["step"])
    q = s.get_state()
    assert not np.all(np.isfinite(qo)) or np.any(qo.reshape(-1, 4)[:om.n_real, 0] <= 0) or \
        np.any(0.4 * (qo[3::4] - 0.5 * (qo[1::4] ** 2 + qo[2::4] ** 2) / qo[0::4])[:om.n_real] <= 0)
    assert same_values(q, qo)


def test_negative_density_stops_the_run():
    """A negative density injected through set_state: the first stage leaves
    it non-physical, the run returns the non-physical-state error after its
    first group of steps, and the GPU keeps working (a new run on a physical
    state is bit-exact)."""
    from fvhyper_amd import FvhError
    cfg = CONFIGS["second_venkat_global"]
    m = MESHES["step"]()
    om = O.Mesh(m)
    q0 = random_state(om.n_cells, seed=91, mach=0.3)
    bad = q0.copy()
    bad[4 * 7] = -0.5
    s = Solver(Mesh.from_arrays(m), cfg)
    s.set_state(bad)
    with pytest.raises(FvhError, match="non-physical state: instance 0 by step"):
        s.run(max_step=50)
    assert s.status()["step"] <= 8
    s.set_state(q0)
    st = s.run(max_step=6)
    qo, *_ = om.run(cfg, q0, 6)
    assert st["step"] == 6 and bits_equal(s.get_state(), qo)


def test_batch_names_the_failing_instance():
    from fvhyper_amd import Batch, FvhError
    items = workloads.naca_sweep(3, seed=11, min_cells=1500, max_cells=4000)
    meshes = [Mesh.from_arrays(it[0]) for it in items]
    b = Batch(meshes, [it[1] for it in items])
    q0 = [workloads.initial_state(mm, it[2]) for mm, it in zip(meshes, items)]
    q0[1] = q0[1].copy()
    q0[1][4 * 3 + 3] = -1.0   # negative energy: p < 0
    b.set_state(b.join(q0))
    with pytest.raises(FvhError, match="instance 1 by step"):
        b.run(max_step=20)


# FP32 arithmetic (states, geometry and every operation in FP32): error
# against the FP64 oracle after 10 steps, max_ik |a-b| / max(|a_ik|,
# max_i |a_.k|).  Measured 1.7e-6 to 3.4e-6 (the most with residual
# smoothing); the bound sits just above.
FP32_ARITH_BOUND = 4e-6
# FP32 states with FP64 arithmetic: the only FP32 rounding is that of the
# stored states.  1e-6 against the FP64 oracle, except where the edge-minimum
# time step of the naca example is ill-conditioned: its eigenvalue form
# (c+|u|)n_x + (c+|v|)n_y can nearly vanish, and dt ~ 1/|lambda| magnifies
# the state's rounding (1.02e-6 after 10 steps on the 2,866-cell O-mesh, the
# same in the CPU restatement with rounded states).
F32_STATE_BOUND = 1e-6
F32_STATE_BOUND_EDGE_MIN = 1.2e-6

FP32_CASES = [("sod", "first_split_sum"), ("step", "second_venkat_global"), ("naca", "second_michalak_local"),
              ("mixed", "second_minmod_farfield"), ("su2m", "second_venkat_global"),
              ("step", "smooth_second_michalak"), ("naca", "visc_strong")]


def _fp32_case(mesh_name, cfg_name):
    cfg = CONFIGS[cfg_name]
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    if mesh_name == "sod":
        q0 = workloads.initial_state(om, workloads.sod_ic())
    elif mesh_name in AIRFOIL_MESHES:
        q0 = airfoil_bump_state(om.cell_cx, om.cell_cy)
    else:
        q0 = random_state(om.n_cells, seed=303, mach=0.3)
    return cfg, m, om, q0


@pytest.mark.parametrize("mesh_name,cfg_name", FP32_CASES)
def test_fp32_arithmetic_against_fp64_oracle(mesh_name, cfg_name):
    """FP32 states, geometry and arithmetic, same kernels: 10 steps."""
    print("case fp32", mesh_name, cfg_name, flush=True)
    cfg, m, om, q0 = _fp32_case(mesh_name, cfg_name)
    s = Solver(Mesh.from_arrays(m), cfg, precision="f32")
    s.set_state(q0)
    st = s.run(max_step=10)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 10)
    assert st["step"] == steps == 10
    assert np.all(np.isfinite(qo))
    err = rel_err(qo, s.get_state())
    print("fp32 rel err", mesh_name, cfg_name, err)
    assert err < FP32_ARITH_BOUND


@pytest.mark.parametrize("flux_mode", ["tile", "fused", "cell"])
@pytest.mark.parametrize("mesh_name,cfg_name", FP32_CASES)
def test_fp32_states_fp64_arithmetic(mesh_name, cfg_name, flux_mode):
    """FP32 states with FP64 arithmetic and geometry, 10 steps: bit for bit
    the oracle run whose stored states are rounded to FP32 after every
    stage (so the kernels round nothing else), in every stage form; and
    within 1e-6 of the FP64 oracle."""
    print("case f32_f64", mesh_name, cfg_name, flux_mode, flush=True)
    cfg, m, om, q0 = _fp32_case(mesh_name, cfg_name)
    if flux_mode == "fused" and (cfg.get("smooth") or cfg.get("flux") == "roe_viscous" or not cfg.get("lin")):
        pytest.skip("no fused stage pass for this configuration")
    s = Solver(Mesh.from_arrays(m), cfg, precision="f32_f64", flux_mode=flux_mode)
    s.set_state(q0)
    st = s.run(max_step=10)
    assert st["step"] == 10
    qg = s.get_state()
    assert bits_equal(qg, qg.astype(np.float32).astype(np.float64))
    qr = om.run_f32_state(cfg, q0, 10)
    assert bits_equal(qg, qr), rel_err(qr, qg)
    qo, *_ = om.run(cfg, q0, 10)
    err = rel_err(qo, qg)
    print("f32_f64 rel err", mesh_name, cfg_name, err)
    bound = F32_STATE_BOUND_EDGE_MIN if cfg.get("dt") == "edge_min" else F32_STATE_BOUND
    assert err < bound


def test_airfoil_workload_bit_exact():
    """The transonic NACA 0012 workload (Mach 0.8, 1.25 degrees) on a 20k-cell
    O-mesh from the freestream, 30 steps, tile order: bit-exact."""
    m, cfg, ic = workloads.naca_airfoil(20000)
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg)
    q0 = workloads.initial_state(om, ic)
    s.set_state(q0)
    st = s.run(max_step=30, print_interval=10)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 30, print_interval=10)
    assert st["step"] == steps == 30 and np.all(np.isfinite(qo))
    assert bits_equal(s.get_state(), qo)


def test_airfoil_density_stop_matches():
    """Stop when the density residual alone has dropped below the tolerance
    (stop_var=0): same step count and state as the oracle."""
    m, cfg, ic = workloads.naca_airfoil(3000)
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg)
    q0 = workloads.initial_state(om, ic)
    s.set_state(q0)
    st = s.run(max_step=5000, tol=0.05, stop_var=0)
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 5000, tol=0.05, stop_var=0)
    assert st["stopped"] and st["step"] == steps < 5000
    assert bits_equal(s.get_state(), qo)
    np.testing.assert_allclose(st["R"], R, rtol=1e-10)


def test_fused_mode_is_refused_without_reconstruction():
    from fvhyper_amd import FvhError
    m = MESHES["sod"]()
    with pytest.raises(FvhError, match="fused"):
        Solver(Mesh.from_arrays(m), CONFIGS["first_split_sum"], flux_mode="fused")


def test_fused_mode_is_refused_with_residual_smoothing():
    from fvhyper_amd import FvhError
    m = MESHES["step"]()
    with pytest.raises(FvhError, match="smoothing"):
        Solver(Mesh.from_arrays(m), CONFIGS["smooth_second_michalak"], flux_mode="fused")


def test_viscous_terms_change_the_update():
    """The strong-viscosity config really adds the laminar terms: the same
    run with the split Roe flux, or without face gradients, differs."""
    cfg = CONFIGS["visc_strong"]
    m = MESHES["step"]()
    om = O.Mesh(m)
    q0 = random_state(om.n_cells, seed=707, mach=0.3)
    out = []
    for c in (cfg, dict(cfg, diffusive=False), dict(cfg, flux="roe_split")):
        s = Solver(Mesh.from_arrays(m), c)
        s.set_state(q0)
        s.run(max_step=3)
        out.append(s.get_state())
    qo, *_ = om.run(cfg, q0, 3)
    assert bits_equal(out[0], qo)
    assert not bits_equal(out[0], out[1]) and not bits_equal(out[0], out[2])
    # without face gradients the viscous terms vanish: the split flux
    assert bits_equal(out[1], out[2])


@pytest.mark.parametrize("cfg_name", ["visc_strong", "visc_transient", "smooth_second_michalak"])
@pytest.mark.parametrize("precision", ["f32", "f32_f64"])
def test_viscous_fp32_against_fp64_oracle(precision, cfg_name):
    print("case viscous", precision, cfg_name, flush=True)
    cfg = CONFIGS[cfg_name]
    m = MESHES["naca"]()
    om = O.Mesh(m)
    q0 = airfoil_bump_state(om.cell_cx, om.cell_cy)
    s = Solver(Mesh.from_arrays(m), cfg, precision=precision)
    s.set_state(q0)
    s.run(max_step=5)
    qo, *_ = om.run(cfg, q0, 5)
    # FP32 state against the FP64 oracle (per-variable infinity-norm floor)
    err = rel_err(qo, s.get_state())
    print(precision, "rel err", cfg_name, err)
    if precision == "f32":
        bound = FP32_ARITH_BOUND
    else:
        bound = F32_STATE_BOUND_EDGE_MIN if cfg.get("dt") == "edge_min" else F32_STATE_BOUND
        assert bits_equal(s.get_state(), om.run_f32_state(cfg, q0, 5))
    assert np.all(np.isfinite(qo)) and err < bound


def test_run_log_rows_follow_the_print_interval():
    """run_log: a row after step 0 (R = R0) and after every 4th step, and
    the same end state as run()."""
    cfg = CONFIGS["second_venkat_global"]
    m = MESHES["step"]()
    om = O.Mesh(m)
    pm = Mesh.from_arrays(m)
    q0 = random_state(om.n_cells, seed=808, mach=0.3)
    a = Solver(pm, cfg)
    a.set_state(q0)
    st, rows = a.run_log(max_step=13, print_interval=4)
    assert [r[0] for r in rows] == [0, 4, 8, 12]
    assert bits_equal(rows[0][2], st["R0"])
    b = Solver(pm, cfg)
    b.set_state(q0)
    st2 = b.run(max_step=13, print_interval=4)
    assert bits_equal(a.get_state(), b.get_state()) and st["step"] == st2["step"] == 13
    qo, steps, t, R, R0, _ = om.run(cfg, q0, 13, print_interval=4)
    assert bits_equal(a.get_state(), qo)


def test_stream_copy_reaches_hbm_rates():
    from fvhyper_amd import stream_copy_gbs
    g = stream_copy_gbs(1024.0, 3)
    print("stream copy GB/s", g)
    assert 1000.0 < g < 12000.0


def test_residual_smoothing_changes_the_update():
    """The smoothing configs really smooth: the same run without the flag
    differs (guards against a flag that never reaches the kernels)."""
    cfg = CONFIGS["smooth_second_michalak"]
    plain = dict(cfg, smooth=False)
    m = MESHES["step"]()
    om = O.Mesh(m)
    q0 = random_state(om.n_cells, seed=606, mach=0.3)
    out = []
    for c in (cfg, plain):
        s = Solver(Mesh.from_arrays(m), c)
        s.set_state(q0)
        s.run(max_step=3)
        out.append(s.get_state())
    qo, *_ = om.run(cfg, q0, 3)
    assert bits_equal(out[0], qo)
    assert not bits_equal(out[0], out[1])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32_f64"])
@pytest.mark.parametrize("cfg_name", ["second_venkat_global", "smooth_second_michalak"])
def test_restart_continues_the_run_bit_for_bit(tmp_path, precision, cfg_name):
    """save() after 5 steps, load() into a fresh solver, run on: the same
    bits, step, time and residuals as the uninterrupted 12-step run."""
    print("case restart", precision, cfg_name, flush=True)
    cfg = CONFIGS[cfg_name]
    m = MESHES["step"]()
    om = O.Mesh(m)
    pm = Mesh.from_arrays(m)
    q0 = random_state(om.n_cells, seed=303, mach=0.3)
    s = Solver(pm, cfg, precision=precision)
    s.set_state(q0)
    full = s.run(max_step=12, print_interval=5)
    q_full = s.get_state()
    s2 = Solver(pm, cfg, precision=precision)
    s2.set_state(q0)
    s2.run_begin()
    mid = s2.run_continue(5, max_step=12, print_interval=5)
    assert mid["step"] == 5
    path = str(tmp_path / "state.fvhs")
    s2.save(path)
    s3 = Solver(pm, cfg, precision=precision)
    s3.load(path)
    assert s3.status()["step"] == 5
    end = s3.run_continue(100, max_step=12, print_interval=5)
    assert bits_equal(s3.get_state(), q_full)
    for k in ("step", "stopped", "time", "time_prev", "R", "R0"):
        assert bits_equal(np.asarray(end[k], dtype=np.float64), np.asarray(full[k], dtype=np.float64)), k


def test_restart_rejects_another_mesh(tmp_path):
    from fvhyper_amd import FvhError
    cfg = CONFIGS["first_split_sum"]
    a = Solver(Mesh.from_arrays(MESHES["sod"]()), cfg)
    a.set_state(random_state(a.mesh.n_cells, seed=1))
    path = str(tmp_path / "a.fvhs")
    a.save(path)
    b = Solver(Mesh.from_arrays(MESHES["mixed"]()), cfg)
    with pytest.raises(FvhError, match="restart file holds"):
        b.load(path)
    with open(path, "r+b") as f:
        f.write(b"XXXX")
    with pytest.raises(FvhError, match="not a restart file"):
        a.load(path)


@pytest.mark.parametrize("flux_mode", ["tile", "fused"])
@pytest.mark.parametrize("mesh_name,cfg_name", [("step", "second_venkat_global"), ("naca", "second_michalak_local"),
                                               ("mixed", "second_minmod_farfield")])
def test_tile_size_128_bit_exact(mesh_name, cfg_name, flux_mode):
    """128-cell tiles (the autotuner's other size): same bits."""
    print("case tile128", mesh_name, cfg_name, flux_mode, flush=True)
    cfg = CONFIGS[cfg_name]
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg, flux_mode=flux_mode, tile_cells=128)
    q0 = airfoil_bump_state(om.cell_cx, om.cell_cy) if mesh_name in AIRFOIL_MESHES else \
        random_state(om.n_cells, seed=404, mach=0.3)
    s.set_state(q0)
    s.run(max_step=8)
    qo, *_ = om.run(cfg, q0, 8)
    assert same_values(s.get_state(), qo)


@pytest.mark.parametrize("threads", [256, 512])
@pytest.mark.parametrize("mesh_name,cfg_name", [("step", "second_venkat_global"), ("su2m", "second_michalak_local"),
                                               ("mixed", "grad_only_lin")])
def test_fused_workgroup_sizes_bit_exact(mesh_name, cfg_name, threads):
    """The fused stage pass with 256 and with 512 lanes per tile: same bits."""
    print("case stage threads", mesh_name, cfg_name, threads, flush=True)
    cfg = dict(CONFIGS[cfg_name], stage_threads=threads)
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg, flux_mode="fused")
    q0 = airfoil_bump_state(om.cell_cx, om.cell_cy) if mesh_name in AIRFOIL_MESHES else \
        random_state(om.n_cells, seed=505, mach=0.3)
    s.set_state(q0)
    s.run(max_step=8)
    qo, *_ = om.run(CONFIGS[cfg_name], q0, 8)
    assert same_values(s.get_state(), qo)


@pytest.mark.parametrize("gdt", [True, False])
@pytest.mark.parametrize("flux_mode", ["fused", "tile"])
def test_second_order_fixed_step(flux_mode, gdt):
    """A fixed time step in a second-order run, with and without the global
    minimum: the time-step kernel fills every cell's step (the gradient pass
    computes only the CFL kinds); the oracle's bits after 8 steps."""
    print("case second order fixed step", flux_mode, gdt, flush=True)
    cfg = dict(CONFIGS["second_venkat_global"], dt="fixed", dt_fixed=1e-4, gdt=gdt)
    m = MESHES["step"]()
    om = O.Mesh(m)
    q0 = random_state(om.n_cells, seed=909, mach=0.3)
    s = Solver(Mesh.from_arrays(m), cfg, flux_mode=flux_mode)
    s.set_state(q0)
    s.run(max_step=8)
    qo, *_ = om.run(cfg, q0, 8)
    assert not bits_equal(qo, q0)
    assert bits_equal(s.get_state(), qo)


@pytest.mark.parametrize("precision", ["f64", "f32", "f32_f64"])
def test_automatic_tile_size(precision):
    """tile_cells 0 (the default): on the forward-step mesh a 256-cell tile
    with its ring-1 cells outnumbers the fused pass's 256 lanes, so the
    library shrinks the tiles; the state keeps its bits (FP64: also against
    the oracle), and an explicit size is kept as given."""
    print("case automatic tile size", precision, flush=True)
    m, cfg, ic = workloads.forward_step(n_per_unit=100)
    pm = Mesh.from_arrays(m)
    auto = Solver(pm, cfg, precision=precision)
    fixed = Solver(pm, cfg, precision=precision, tile_cells=256)
    assert 128 <= auto.tile_cells < 256 and auto.tile_cells % 16 == 0
    assert fixed.tile_cells == 256
    q0 = workloads.bump_state(pm)
    for s in (auto, fixed):
        s.set_state(q0)
        s.run(max_step=6)
    assert bits_equal(auto.get_state(), fixed.get_state())
    if precision == "f64":
        qo, *_ = O.Mesh(m).run(cfg, q0, 6)
        assert bits_equal(auto.get_state(), qo)


def test_oversized_tiles_are_bisected():
    """200,000-cell airfoil with the default layout: 208-cell tiles, of which
    a few in the quad band have halos needing more than 40 KB of LDS (three
    256-lane workgroups per CU); the library bisects those tiles, so every
    tile fits four workgroups per CU, and the state keeps the oracle's bits."""
    print("case oversized tiles", flush=True)
    m, cfg, ic = workloads.naca_airfoil(200_000)
    pm = Mesh.from_arrays(m)
    s = Solver(pm, cfg)
    info = s.stage_info
    print("layout", info, flush=True)
    assert info["fused"] == 1 and info["lanes"] == 256 and info["tile_cells"] == 208
    assert info["lds_bytes"] + 1024 <= 160 * 1024 // 4
    full = -(-pm.n_real // info["tile_cells"])
    assert full < info["tiles"] <= full + full * 3 // 100
    q0 = workloads.bump_state(pm, u=0.8, v=0.0175)
    s.set_state(q0)
    s.run(max_step=3)
    qo, *_ = O.Mesh(m).run(cfg, q0, 3)
    assert bits_equal(s.get_state(), qo)


def test_autotune_returns_a_working_choice():
    from fvhyper_amd import autotune
    m, cfg, ic = workloads.naca_airfoil(20000)
    pm = Mesh.from_arrays(m)
    q0 = workloads.initial_state(pm, ic)
    r = autotune([pm], [cfg], q0, steps=3)
    print("autotune", r)
    assert r["flux_mode"] in ("tile", "fused", "cell") and r["tile_cells"] in (128, 192, 256)
    assert len(r["candidates"]) == 9 and all(c[2] > 0 for c in r["candidates"])
    s = Solver(pm, cfg, flux_mode=r["flux_mode"], tile_cells=r["tile_cells"])
    s.set_state(q0)
    s.run(max_step=5)
    qo, *_ = O.Mesh(m).run(cfg, q0, 5)
    assert bits_equal(s.get_state(), qo)
