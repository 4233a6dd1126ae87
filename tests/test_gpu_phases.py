"""The per-phase entry points (explicit.h:88-149 one at a time, on caller
arrays in the reference layout) against the oracle's phases, which the
golden fixtures pin to the compiled reference: bit for bit, except the
residual norms, which are summed in another order."""
import numpy as np
import pytest

import oracle as O
from fvhyper_amd import Mesh, Solver
from helpers import AIRFOIL_MESHES, CONFIGS, MESHES, airfoil_bump_state, bits_equal, random_state

pytestmark = pytest.mark.gpu

PHASE_CONFIGS = ["first_split_sum", "second_michalak_local", "second_venkat_global", "second_minmod_farfield",
                 "grad_only_lin", "smooth_second_michalak", "visc_strong", "visc_transient"]


@pytest.mark.parametrize("flux_mode", ["auto", "cell"])
@pytest.mark.parametrize("mesh_name", sorted(MESHES))
@pytest.mark.parametrize("cfg_name", PHASE_CONFIGS)
def test_phase_functions_bit_exact(mesh_name, cfg_name, flux_mode):
    print("case phases", mesh_name, cfg_name, flux_mode, flush=True)
    cfg = CONFIGS[cfg_name]
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    s = Solver(Mesh.from_arrays(m), cfg, flux_mode=flux_mode)
    q = airfoil_bump_state(om.cell_cx, om.cell_cy) if mesh_name in AIRFOIL_MESHES else \
        random_state(om.n_cells, seed=111, mach=0.3)
    o = om.phases(cfg, q, alpha=0.506)
    # calc_gradients
    gx, gy = s.phase_gradients(q)
    if cfg.get("grad"):
        assert bits_equal(gx, o["gx"]) and bits_equal(gy, o["gy"])
    # calc_limiters from the oracle's gradients
    if cfg.get("lim"):
        lim, qmin, qmax = s.phase_limiters(q, o["gx"], o["gy"])
        assert bits_equal(lim, o["lim"]) and bits_equal(qmin, o["qmin"]) and bits_equal(qmax, o["qmax"])
    # calc_time_derivatives from the oracle's gradients and limiters (before
    # any residual smoothing)
    qt = s.phase_time_derivatives(q, o["gx"], o["gy"], o["lim"])
    assert bits_equal(qt, o.get("qt_raw", o["qt"]))
    # update_cells with the reference's dt layout (one value per variable)
    dt4 = np.repeat(o["dt"], 4)
    qk = s.phase_update_cells(q, o["qt"], dt4, 0.506)
    assert bits_equal(qk, o["qk_cells"])
    # update_bounds on the updated cells with this stage's gradients
    assert bits_equal(s.phase_update_bounds(o["qk_cells"], o["gx"], o["gy"], o["lim"]), o["qk"])
    # min_dt over the local steps
    dmin = s.phase_min_dt(np.repeat(o["dt_local"], 4))
    assert bits_equal(dmin, np.full(dmin.size, np.min(o["dt_local"])))
    # calc_residuals: another summation order
    np.testing.assert_allclose(s.phase_residuals(o["qt"]), o["R"], rtol=1e-12, atol=0)


def test_min_dt_keeps_the_first_value_rule():
    """min_dt starts from dt[0] with std::min: a NaN in dt[0] wins, any other
    NaN is ignored (explicit.cpp:555-564)."""
    m = MESHES["sod"]()
    s = Solver(Mesh.from_arrays(m), CONFIGS["first_split_sum"])
    d = np.linspace(1.0, 2.0, 4 * m.cells.shape[0])
    d[7] = np.nan
    assert np.all(s.phase_min_dt(d) == 1.0)
    d[0] = np.nan
    assert np.all(np.isnan(s.phase_min_dt(d)))
    big = np.linspace(3.0, 5.0, 4 * (m.cells.shape[0] + 40))
    from fvhyper_amd import FvhError
    with pytest.raises(FvhError, match="more values"):
        s.phase_min_dt(np.concatenate([big, big]))


def test_phase_functions_leave_the_solver_state():
    m = MESHES["step"]()
    om = O.Mesh(m)
    cfg = CONFIGS["second_venkat_global"]
    s = Solver(Mesh.from_arrays(m), cfg)
    q0 = random_state(om.n_cells, seed=3, mach=0.3)
    s.set_state(q0)
    q1 = random_state(om.n_cells, seed=4, mach=0.3)
    gx, gy = s.phase_gradients(q1)
    s.phase_time_derivatives(q1, gx, gy, np.ones_like(gx))
    st = s.run(max_step=3)
    qo, *_ = om.run(cfg, q0, 3)
    assert st["step"] == 3 and bits_equal(s.get_state(), qo)
