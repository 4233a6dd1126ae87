"""Drop-in parity: the reference's own example programs
(examples/{shockTube,forwardStep,naca,flatPlate,nacaTransient}/main.cpp), compiled unchanged once
against the reference library and once against this repository's fvhyper
headers + libfvhyper.so (oracle/Makefile target `dropin`), run on the same
meshes.  The printed step/time columns and every VTU file they write must be
identical (the VTU files print 6 decimals, the residuals 6 digits)."""
import os
import subprocess

import numpy as np
import pytest

from fvhyper_amd import meshgen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def _renamed(m, names):
    tag = np.array([names.index(n) if n in names else 0 for n in m.names], dtype=np.int32)
    return meshgen.MeshArrays(m.x, m.y, m.cells, m.bnd, tag[m.bnd_tag], list(names))


CASES = {
    # example: (mesh file name, mesh)
    "shockTube": ("square", lambda: _renamed(meshgen.sod_tube(nx=20, ny=2), ["wall"])),
    "forwardStep": ("step", lambda: meshgen.forward_step(n_per_unit=10)),
    "naca": ("naca", lambda: meshgen.MeshArrays(*(lambda m: (m.x, m.y, m.cells, m.bnd, m.bnd_tag,
                                                             ["airfoil", "farfield"]))(meshgen.sod_tube(nx=24, ny=4)))),
    # the viscous examples: laminar flux on diffusive face gradients, viscous
    # time steps, no-slip wall and their farfield (30,000 and 15,000 steps)
    "flatPlate": ("square", lambda: meshgen.flat_plate(nx=20, ny=6)),
    "nacaTransient": ("naca0012", lambda: (lambda m: meshgen.MeshArrays(
        m.x, m.y, m.cells, m.bnd, m.bnd_tag, ["wall" if n == "airfoil" else n for n in m.names]))(
        meshgen.naca0012_sized(800))),
}


def _run(binary, workdir, mesh_name, m, timeout=600):
    os.makedirs(os.path.join(workdir, "times"), exist_ok=True)
    meshgen.write_msh(os.path.join(workdir, mesh_name + ".msh"), m)
    r = subprocess.run([binary], cwd=workdir, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (binary, r.stdout[-2000:], r.stderr[-2000:])
    rows = [l.split(",") for l in r.stdout.strip().splitlines()[1:]]
    files = {}
    for dirpath, _, names in os.walk(workdir):
        for n in names:
            if n.endswith(".vtu"):
                p = os.path.join(dirpath, n)
                files[os.path.relpath(p, workdir)] = open(p, "rb").read()
    return rows, files


@pytest.mark.parametrize("example", sorted(CASES))
def test_reference_example_runs_unchanged_on_the_gpu(example, tmp_path):
    ref_bin = os.path.join(REF, "ref_" + example)
    dropin_bin = os.path.join(REF, "dropin_" + example)
    if not (os.path.exists(ref_bin) and os.path.exists(dropin_bin)):
        pytest.skip("example binaries not built (make -C oracle dropin needs the reference tree)")
    print("example", example, flush=True)
    mesh_name, fac = CASES[example]
    m = fac()
    rows_ref, files_ref = _run(ref_bin, str(tmp_path / "ref"), mesh_name, m)
    rows_gpu, files_gpu = _run(dropin_bin, str(tmp_path / "gpu"), mesh_name, m)
    assert len(rows_ref) == len(rows_gpu) > 1
    for a, b in zip(rows_ref, rows_gpu):
        assert a[0] == b[0] and a[1] == b[1]          # step, time
        np.testing.assert_allclose([float(x) for x in a[3:]], [float(x) for x in b[3:]], rtol=2e-5)
    assert sorted(files_ref) == sorted(files_gpu) and files_ref
    for k in files_ref:
        assert files_ref[k] == files_gpu[k], k


def _vtu_arrays(text):
    """name -> list of number strings of every DataArray of a VTU file."""
    import re
    out = {}
    for m in re.finditer(r'<DataArray type="(\w+)"(?: Name="(\w+)")?[^>]*>(.*?)</DataArray>', text, re.S):
        out[m.group(2) or "points"] = (m.group(1), m.group(3).split())
    return out


def test_full_precision_vtu_rounds_to_the_reference_text(tmp_path):
    """FVH_VTK_FULL_PRECISION=1: the same files with Float64 arrays whose
    %.17g numbers print, through std::to_string's %f, as the default text."""
    dropin_bin = os.path.join(REF, "dropin_shockTube")
    if not os.path.exists(dropin_bin):
        pytest.skip("example binaries not built (make -C oracle dropin needs the reference tree)")
    mesh_name, fac = CASES["shockTube"]
    m = fac()
    _, plain = _run(dropin_bin, str(tmp_path / "plain"), mesh_name, m)
    os.environ["FVH_VTK_FULL_PRECISION"] = "1"
    try:
        _, full = _run(dropin_bin, str(tmp_path / "full"), mesh_name, m)
    finally:
        del os.environ["FVH_VTK_FULL_PRECISION"]
    assert sorted(plain) == sorted(full) and plain
    for k in plain:
        a, b = _vtu_arrays(plain[k].decode()), _vtu_arrays(full[k].decode())
        assert sorted(a) == sorted(b)
        for name in a:
            (ta, va), (tb, vb) = a[name], b[name]
            if ta == "Int32" or name == "rank":
                assert va == vb, name
                continue
            assert ta == "Float32" and tb == "Float64", name
            assert len(va) == len(vb)
            for x, y in zip(va, vb):
                assert x == y or x == "%f" % float(y), (name, x, y)


PHASES_BIN = os.path.join(ROOT, "tests", "dropin_phases", "dropin_phases")


@pytest.mark.parametrize("mesh_name", ["step", "mixed"])
def test_user_program_calling_the_phase_functions(mesh_name, tmp_path):
    """A user program against the drop-in headers that calls the phases of
    one stage itself (tests/dropin_phases/main.cpp, built against
    libfvhyper.so): calc_gradients and calc_limiters in both the header's and
    the source's forms, calc_time_derivatives, min_dt, update_cells,
    update_bounds, calc_residuals and complete_calc_qt, each against the
    oracle's phase on the same state."""
    import ctypes as C

    import oracle as O
    from helpers import CONFIGS, MESHES, bits_equal, random_state
    if not os.path.exists(PHASES_BIN):
        pytest.skip("tests/dropin_phases is not built (make -C tests/dropin_phases)")
    print("case dropin phases", mesh_name, flush=True)
    m = MESHES[mesh_name]()
    om = O.Mesh(m)
    cfg = CONFIGS["second_venkat_global"]
    q = random_state(om.n_cells, seed=515, mach=0.3)
    meshgen.write_msh(str(tmp_path / "case.msh"), m)
    q.tofile(str(tmp_path / "q.f64"))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([PHASES_BIN, "case", str(tmp_path / "q.f64"), str(out)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    got = {p[:-4]: np.fromfile(str(out / p)) for p in os.listdir(str(out))}
    o = om.phases(cfg, q, alpha=0.506)
    for a, b in (("gx", "gx"), ("gy", "gy"), ("gx_const", "gx"), ("gy_const", "gy"), ("lim", "lim"),
                 ("lim_header", "lim"), ("qmin", "qmin"), ("qmax", "qmax"), ("qt", "qt"), ("qt_complete", "qt"),
                 ("lim_complete", "lim")):
        assert bits_equal(got[a], o[b]), a
    # min_dt over the program's own time step
    assert np.all(got["dt"] == np.min(got["dt_local"]))
    # update_cells with that step, then update_bounds with this stage's
    # gradients and limiters, as the oracle's phases do them
    L = O.lib()
    c = O.make_config(cfg, om.names)
    qk = np.zeros_like(q)
    L.orc_update_cells(om.n_cells, qk, q, o["qt"], np.ascontiguousarray(got["dt"][::4]), 0.506)
    assert bits_equal(got["qk_cells"], qk)
    L.orc_update_bounds(om._p, C.byref(c), qk, o["gx"], o["gy"], o["lim"])
    assert bits_equal(got["qk"], qk)
    R = np.zeros(4)
    L.orc_calc_residuals(om._p, o["qt"], R)
    np.testing.assert_allclose(got["R"], R, rtol=1e-12, atol=0)
