"""The entry point that allocates nothing (fvh_min_dt_device: min_dt of
explicit.cpp:555-564 on a device array, in place, its per-block minima in a
file-scope device array) against the allocating one (fvh_phase_min_dt) on
the same values: bit for bit, on two calls in a row, the second smaller."""
import ctypes as C

import numpy as np
import pytest

from fvhyper_amd import FvhError, Mesh, Solver, _lib
from helpers import CONFIGS, MESHES, bits_equal

pytestmark = pytest.mark.gpu


def _device_min(torch, values, stream):
    t = torch.tensor(values, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().fvh_min_dt_device(C.c_void_p(t.data_ptr()), t.numel(), C.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return t.cpu().numpy()


def test_min_dt_device_matches_the_allocating_entry_point():
    import torch
    torch.cuda.set_device(0)
    m = MESHES["naca"]()
    s = Solver(Mesh.from_arrays(m), CONFIGS["second_michalak_local"])
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(77)
    n_max = 4 * (m.cells.shape[0] + m.bnd.shape[0])
    for n in (n_max, 1000):   # the second call smaller: stale partials must not leak in
        d = rng.uniform(1e-4, 1e-2, n)
        d[rng.integers(1, n, 5)] = np.nan   # ignored, as std::min from dt[0] ignores them
        got = _device_min(torch, d, stream)
        want = s.phase_min_dt(d)
        print("n", n, "min", got[0], flush=True)
        assert bits_equal(got, want)
        assert got[0] == np.nanmin(d)
    d = np.full(300, 2.0)
    d[0] = np.nan
    assert np.all(np.isnan(_device_min(torch, d, stream)))
    with pytest.raises(FvhError, match="at most"):
        _lib.check(_lib.load().fvh_min_dt_device(C.c_void_p(1), (1 << 31) + 1, None))
