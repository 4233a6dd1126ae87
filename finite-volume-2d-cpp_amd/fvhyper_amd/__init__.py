"""fvhyper_amd -- MI355X-native explicit finite-volume path of fvhyper.

The compute path is libfvh.so (hand-written HIP kernels for gfx950 behind a
C ABI, include/fvh.h); this package is its Python face plus the seeded mesh
generators and workload definitions.
"""
from . import meshgen, workloads
from ._lib import FvhError, load
from .solver import Batch, Mesh, Solver, autotune, make_config, stream_copy_gbs

__all__ = ["meshgen", "workloads", "Mesh", "Solver", "Batch", "autotune", "make_config", "stream_copy_gbs", "load",
           "FvhError"]
