"""The transonic NACA 0012 workload on the CPU: the O-mesh generator, the
density-only stop rule of the run options, and the naca example's own plugin
(examples/naca/main.cpp, compiled from the reference tree) against the oracle
from the plugin's initial solution."""
import numpy as np
import pytest

import oracle as O
from fvhyper_amd import Mesh, meshgen, workloads
from helpers import bits_equal


def _signed_areas(m):
    x, y, c = m.x, m.y, m.cells

    def tri(i, j, k):
        return 0.5 * ((x[j] - x[i]) * (y[k] - y[i]) - (x[k] - x[i]) * (y[j] - y[i]))

    q = c[:, 3] >= 0
    a = tri(c[:, 0], c[:, 1], c[:, 2])
    a[q] += tri(c[q, 0], c[q, 2], c[q, 3])
    return a, q


def test_naca_section_matches_the_4digit_formula():
    """NACA 0012 half-thickness: the published stations forward of 30% chord
    (where closing the trailing edge changes y by < 1e-5), a 12% maximum near
    30% chord, and a trailing edge closed to rounding."""
    x = np.array([0.0, 0.0125, 0.05, 0.1, 0.3])
    np.testing.assert_allclose(meshgen.naca4_thickness(x), [0.0, 0.01894, 0.03555, 0.04683, 0.06002], atol=2e-5)
    xs = np.linspace(0.0, 1.0, 100001)
    y = meshgen.naca4_thickness(xs)
    assert 0.29 < xs[np.argmax(y)] < 0.31 and abs(2 * y.max() - 0.12) < 2e-4
    assert abs(meshgen.naca4_thickness(np.array([1.0]))[0]) < 1e-15
    assert np.all(y[1:-1] > 0)


@pytest.mark.parametrize("cells", [1500, 20000, 200000])
def test_coarsened_omesh_is_valid(cells):
    m = meshgen.naca0012_sized(cells)
    assert abs(m.n_cells - cells) < max(0.02 * cells, 100)    # one ring is ~100 cells at 1.5k
    a, q = _signed_areas(m)
    assert np.all(a > 0), "every cell counter-clockwise and non-degenerate"
    assert m.names == ["airfoil", "farfield"]
    far = m.bnd[m.bnd_tag == 1]
    r = np.hypot(m.x[far.ravel()] - 0.5, m.y[far.ravel()])
    np.testing.assert_allclose(r, 50.0, rtol=1e-12)
    # every boundary element is an edge of exactly one cell: the host mesh sees
    # one ghost per element and no unlisted boundary edge
    pm = Mesh.from_arrays(m)
    assert pm.n_cells - pm.n_real == m.n_bnd
    ec = pm.export("edgesCells").reshape(-1, 2)
    assert np.sum(ec[:, 1] >= pm.n_real) == m.n_bnd


def test_full_size_omesh_layout():
    """The 2M-cell mesh: 512 surface points, 64 quad layers with first
    spacing 1e-3 and growth 1.1, triangles to R = 50."""
    m = meshgen.naca0012_omesh()
    assert 1.95e6 < m.n_cells < 2.05e6
    quads = m.cells[:, 3] >= 0
    assert quads.sum() == 512 * 64 and np.all(quads[:512 * 64])
    a, _ = _signed_areas(m)
    assert np.all(a > 0)
    surf = m.bnd[m.bnd_tag == 0]
    assert surf.shape[0] == 512
    # wall distance of layer 1 and layer 64 at mid-chord on the upper surface
    k = 128                                   # beta = pi/2: x = 0.5
    d1 = np.hypot(m.x[512 + k] - m.x[k], m.y[512 + k] - m.y[k])
    np.testing.assert_allclose(d1, 1e-3, rtol=1e-3)
    d64 = 1e-3 * (1.1 ** 64 - 1) / 0.1
    rb = np.hypot(m.x[64 * 512 + k] - 0.5, m.y[64 * 512 + k])
    np.testing.assert_allclose(rb, d64 + 0.5, rtol=1e-12)


def test_density_stop_rule():
    """stop_var = 0 stops at the first step whose normalised density residual
    is below the tolerance: one step earlier it was not, and the state then is
    the state of a plain max_step run of that length."""
    m, cfg, ic = workloads.naca_airfoil(1500)
    om = O.Mesh(m)
    q0 = workloads.initial_state(om, ic)
    tol = 0.2
    q, n, _, R, _, _ = om.run(cfg, q0, 10000, tol=tol, stop_var=0)
    assert 1 < n < 10000 and R[0] < tol
    q1, n1, _, R1, _, _ = om.run(cfg, q0, n - 1, tol=0.0, stop_var=0)
    assert n1 == n - 1 and R1[0] >= tol
    q2, n2, _, R2, _, _ = om.run(cfg, q0, n, tol=0.0)
    assert n2 == n and bits_equal(q2, q)
    # the reference's rule (max over variables) waits for every residual
    _, nmax, _, Rm, _, _ = om.run(cfg, q0, 10000, tol=tol)
    assert nmax >= n and np.max(Rm) < tol


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (needs the reference tree)")
def test_naca_plugin_run_matches_oracle(tmp_path):
    """examples/naca/main.cpp's plugin (initial solution, flux, Michalak,
    edge-min dt, wall + farfield) on a 5k-cell O-mesh, 40 steps, against the
    oracle with workloads.naca_config(reference_compatible=True)."""
    m, cfg, ic = workloads.naca_airfoil(5000, reference_compatible=True)
    meshgen.write_msh(str(tmp_path / "naca.msh"), m)
    kv = O.ref_args(cfg)
    kv.update(ic="nc", max_step="40", print="1000")
    r = O.run_ref("run", str(tmp_path), "naca", flags=O.ref_flags(cfg), **kv)
    om = O.Mesh(m)
    q0 = workloads.initial_state(om, ic)
    q, steps, t, R, R0, _ = om.run(cfg, q0, 40, print_interval=1000)
    assert steps == 40 and np.all(np.isfinite(q))
    assert bits_equal(q, r["q"])


def test_example_time_step_is_unstable_on_the_boundary_band():
    """Why the workload does not use the naca example's time step: on a 50k
    O-mesh the example's edge-minimum dt at CFL 1.0 drives the run
    non-finite within 60 steps, while the summed form the workload uses
    (CFL 2.0, per cell) stays finite."""
    m, cfg, ic = workloads.naca_airfoil(50000)
    assert cfg["dt"] == "sum_cfl" and not cfg["gdt"]
    om = O.Mesh(m)
    q0 = workloads.initial_state(om, ic)
    q_ex, *_ = om.run(dict(cfg, dt="edge_min", cfl=1.0), q0, 60)
    q_wl, *_ = om.run(cfg, q0, 60)
    assert not np.all(np.isfinite(q_ex))
    assert np.all(np.isfinite(q_wl))
