"""The C ABI: every entry point include/fvh.h declares is exported by
libfvh.so, the Python binding covers them, and the host-only entry points
behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np

from fvhyper_amd import _lib, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "fvh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvh_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_abi_version_and_defaults():
    L = _lib.load()
    assert L.fvh_abi_version() == 3
    c = _lib.Config()
    L.fvh_config_default(C.byref(c))
    assert c.n_alpha == 5 and list(c.alpha)[:5] == [0.0695, 0.1602, 0.2898, 0.506, 1.0]
    assert c.K == 7.5 and c.gamma == 1.4 and c.phi_yt == 2.0 and c.dt_fixed == 2e-5
    assert c.renumber == 2 and c.flux_mode == 0
    # the laminar constants of the flat-plate example
    assert (c.mu, c.pr, c.cp, c.r_gas) == (2e-5, 0.72, 1.0, 8.3145)
    import ctypes
    assert ctypes.sizeof(_lib.Config) == 1912   # sizeof(fvh_config) in C


def test_mesh_entry_points_without_gpu():
    L = _lib.load()
    m = meshgen.sod_tube(nx=6, ny=2)
    from fvhyper_amd import Mesh
    pm = Mesh.from_arrays(m)
    counts = (C.c_int64 * 6)()
    assert L.fvh_mesh_counts(pm._h, counts) == 0
    assert list(counts) == [pm.n_nodes, 24, 24 + m.n_bnd, pm.n_edges, m.n_bnd, 2]
    assert L.fvh_mesh_export(pm._h, b"noSuchField", None, 0) < 0
    assert b"unknown mesh field" in L.fvh_last_error()
    buf = np.zeros(3)
    assert L.fvh_mesh_export(pm._h, b"cellsAreas", buf.ctypes.data_as(C.c_void_p), 3) < 0   # too small
    assert L.fvh_mesh_counts(None, counts) < 0
