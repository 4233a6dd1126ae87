"""Partitioned MSH 4.1 files for several ranks (meshgen.write_partitioned_msh),
in the layout the reference reads (mesh.cpp:316-340, 574-617, 803-855), and
the reference library run on them under mpirun: the CPU baseline on every
core of a host.  With a fixed time step the partitioned reference run equals
the single-rank run to rounding (local edge orders differ), which pins the
writer's ghost layer and tags.  The reference's global time step over ranks
(validate_dt, explicit.cpp:567-621) keeps only the last rank's minimum and
leaves rank 0's cells but the first at rank 0's own minimum, so summed-CFL
runs differ by more than rounding; the cost per step is the same."""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from fvhyper_amd import meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_partition_is_balanced_and_complete():
    m = meshgen.forward_step(n_per_unit=20)
    for P in (2, 3, 5, 16):
        part = meshgen.partition_rcb(m, P)
        counts = np.bincount(part, minlength=P)
        assert counts.sum() == m.n_cells and counts.max() - counts.min() <= 1


def test_partition_files_hold_own_cells_ghost_layer_and_global_tags(tmp_path):
    m = meshgen.mixed_channel(nx=12, ny=6)
    P = 3
    part = meshgen.partition_rcb(m, P)
    meshgen.write_partitioned_msh(str(tmp_path), "case", m, part)
    src, dst = meshgen.cell_neighbours(m)
    seen = np.zeros(m.n_cells, dtype=int)
    for r in range(P):
        text = open(str(tmp_path / ("case_%d.msh" % (r + 1)))).read()
        els = text.split("$Elements\n")[1].split("$EndElements")[0].splitlines()
        tags = []
        i = 1
        while i < len(els):
            dim, _, _, n = (int(v) for v in els[i].split())
            if dim == 2:
                tags += [int(l.split()[0]) for l in els[i + 1:i + 1 + n]]
            i += 1 + n
        cells = np.array(tags) - 1 - m.n_bnd
        gl = text.split("$GhostElements\n")[1].split("$EndGhostElements")[0].splitlines()[1:]
        ghosts = np.array([int(l.split()[0]) for l in gl], dtype=int) - 1 - m.n_bnd
        owners = np.array([int(l.split()[1]) for l in gl], dtype=int) - 1
        own = np.setdiff1d(cells, ghosts)
        assert np.array_equal(np.sort(own), np.nonzero(part == r)[0])
        want = np.unique(dst[(part[src] == r) & (part[dst] != r)])
        assert np.array_equal(np.sort(ghosts), want)
        assert np.array_equal(owners, part[ghosts])
        seen[own] += 1
    assert np.all(seen == 1)


@pytest.mark.skipif(not O.mpi_available(), reason="reference harness or mpirun absent")
@pytest.mark.parametrize("order", ["first", "second"])
def test_partitioned_reference_matches_one_rank_with_a_fixed_step(order, tmp_path):
    import sys
    sys.path.insert(0, GOLDEN)
    import gen_fixtures as G
    FS = ("inlet_outlet", 1.4, 3.0, 0.0, 1.0)
    bcs = {"wall": ("mirror",), "transmissive": ("copy",), "inlet": FS, "outlet": FS}
    cfg = dict(flux="roe_joined", dt="fixed", dt_fixed=2e-5, gdt=True, bcs=bcs)
    if order == "second":
        cfg.update(phi="venkat", grad=True, lim=True, lin=True)
    m2, cfg2 = G.renamed(meshgen.forward_step(n_per_unit=20), cfg)
    d = str(tmp_path)
    meshgen.write_msh(os.path.join(d, "case.msh"), m2)
    kv = O.ref_args(cfg2)
    kv.update(mach="3", ic="fs", max_step="12", print="12")
    q1 = O.run_ref("run", d, "case", flags=O.ref_flags(cfg2), **kv)["q"].reshape(-1, 4)[:m2.n_cells]
    for P in (2, 3):
        meshgen.write_partitioned_msh(d, "case", m2, meshgen.partition_rcb(m2, P))
        rp = O.run_ref_mpi("run", d, "case", P, flags=O.ref_flags(cfg2), **kv)
        q = O.gather_ranks(rp, m2.n_bnd, m2.n_cells)
        assert not np.any(np.isnan(q))
        assert np.abs(q - q1).max() <= 1e-12 * np.abs(q1).max(), (order, P)
