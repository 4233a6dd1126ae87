"""The oracle against the golden vectors produced by the compiled reference
(tests/golden/gen_fixtures.py): mesh numbering and geometry, one stage of
every phase, and 10-step run() trajectories must be bit-identical."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from fvhyper_amd import meshgen
from helpers import bits_equal

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mesh_arrays(d):
    return meshgen.MeshArrays(d["x"], d["y"], d["cells"], d["bnd"], d["bnd_tag"], [str(s) for s in d["names"]])


def _case_mesh(mesh_name, cfg_name):
    import gen_fixtures as G
    d = np.load(os.path.join(GOLDEN, "mesh__%s.npz" % mesh_name))
    m = _mesh_arrays(d)
    return G.renamed(m, G.ALL_CONFIGS[cfg_name])


@pytest.fixture(scope="module", autouse=True)
def _golden_path():
    import sys
    sys.path.insert(0, GOLDEN)
    yield


MESH_FIELDS = [("edgesCells", "edge_cells"), ("edgesNodes", "edge_nodes"), ("edgesLengths", "edge_len"),
               ("edgesNormalsX", "edge_nx"), ("edgesNormalsY", "edge_ny"), ("edgesCentersX", "edge_cx"),
               ("edgesCentersY", "edge_cy"), ("cellsAreas", "cell_area"), ("cellsCentersX", "cell_cx"),
               ("cellsCentersY", "cell_cy"), ("boundaryEdges", "bnd_edge"), ("cellsNodes", "cell_nodes")]


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "mesh__*.npz"))))
def test_oracle_mesh_matches_reference(path):
    d = np.load(path)
    om = O.Mesh(_mesh_arrays(d))
    assert int(d["ref_nReal"][0]) == om.n_real
    for ref_name, attr in MESH_FIELDS:
        ref = d["ref_" + ref_name]
        mine = getattr(om, attr).ravel()
        if ref.dtype == np.float64:
            assert bits_equal(ref, mine), ref_name
        else:
            assert np.array_equal(ref.astype(np.int64), mine.astype(np.int64)), ref_name


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "phases__*.npz"))))
def test_oracle_phases_match_reference(path):
    import gen_fixtures as G
    _, mesh_name, cfg_name = os.path.basename(path)[:-4].split("__")
    m2, cfg2 = _case_mesh(mesh_name, cfg_name)
    d = np.load(path)
    om = O.Mesh(m2)
    out = om.phases(cfg2, d["q"], alpha=float(d["alpha"]))
    assert bits_equal(out["dt"], d["dt"][::4])
    for key in ("gx", "gy", "lim", "qmin", "qmax", "qt", "qt_raw", "qk_cells", "qk", "R"):
        if key in ("qmin", "qmax") and not cfg2.get("lim"):
            continue
        if key == "qt_raw" and not cfg2.get("smooth"):
            continue
        assert bits_equal(out[key], d[key]), key


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "run__*.npz"))))
def test_oracle_run_matches_reference(path):
    _, mesh_name, cfg_name = os.path.basename(path)[:-4].split("__")
    m2, cfg2 = _case_mesh(mesh_name, cfg_name)
    d = np.load(path)
    om = O.Mesh(m2)
    q, steps, t, R, R0, _ = om.run(cfg2, d["q0"], 10, print_interval=1)
    assert steps == int(d["steps"]) == 10
    assert bits_equal(q, d["q"])
    # the reference prints R with 6 significant digits
    np.testing.assert_allclose(R, d["R"], rtol=1e-5)


def test_fixture_set_is_complete():
    import gen_fixtures as G
    for mn in G.FIX_MESHES:
        assert os.path.exists(os.path.join(GOLDEN, "mesh__%s.npz" % mn))
        for cn in G.configs_of(mn):
            assert os.path.exists(os.path.join(GOLDEN, "phases__%s__%s.npz" % (mn, cn)))
            assert os.path.exists(os.path.join(GOLDEN, "run__%s__%s.npz" % (mn, cn)))


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (needs the reference tree)")
def test_oracle_matches_live_reference_on_larger_mesh(tmp_path):
    """Beyond the fixtures: a 2,016-cell step mesh, 20 second-order steps."""
    import gen_fixtures as G
    m = meshgen.forward_step(n_per_unit=20)
    cfg = G.FIX_CONFIGS["second_venkat_global"]
    m2, cfg2 = G.renamed(m, cfg)
    meshgen.write_msh(str(tmp_path / "case.msh"), m2)
    om = O.Mesh(m2)
    from fvhyper_amd import workloads
    q0 = workloads.initial_state(om, workloads.uniform_ic(1.4, 3.0, 0.0, 1.0))
    q0.tofile(str(tmp_path / "q0"))
    kv = O.ref_args(cfg2)
    kv.update(mach="3", ic="file:" + str(tmp_path / "q0"), max_step="20", print="1000")
    r = O.run_ref("run", str(tmp_path), "case", flags=O.ref_flags(cfg2), **kv)
    q, steps, t, R, R0, _ = om.run(cfg2, q0, 20, print_interval=1000)
    assert bits_equal(q, r["q"])


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (needs the reference tree)")
def test_concurrent_reference_runs_on_one_mesh_keep_their_dumps(tmp_path):
    """Runs of the reference harness that overlap on one mesh (as the
    sweep's CPU baseline starts them, several instances sharing a mesh)
    write dump files of their own when tagged, and every one reads back
    whole: the same final state from each of eight concurrent runs."""
    import concurrent.futures as cf
    import gen_fixtures as G
    m = meshgen.forward_step(n_per_unit=20)
    cfg = G.FIX_CONFIGS["second_venkat_global"]
    m2, cfg2 = G.renamed(m, cfg)
    meshgen.write_msh(str(tmp_path / "case.msh"), m2)
    kv = O.ref_args(cfg2)
    kv.update(mach="3", ic="fs", max_step="3", print="1000")

    def one(k):
        return O.run_ref("run", str(tmp_path), "case", flags=O.ref_flags(cfg2), tag="i%d" % k, **kv)["q"]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        qs = list(ex.map(one, range(8)))
    assert len(set(os.listdir(tmp_path)) & {"case.run.i%d.dump" % k for k in range(8)}) == 8
    for q in qs[1:]:
        assert bits_equal(q, qs[0])
