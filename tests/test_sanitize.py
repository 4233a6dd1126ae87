"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize): the mesh reader and builder (fvh_mesh.cpp) on good and
malformed MSH 4.1 and SU2 files, every exported member, both device orders
and the geometry reverse; and the drop-in's host code (mesh reading, plugin
probing, VTU writing).  CPU only."""
import os
import shutil
import subprocess

import pytest

from fvhyper_amd import meshgen
from helpers import su2_fixture_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


@pytest.fixture(scope="module")
def binaries():
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return os.path.join(SAN, "mesh_asan"), os.path.join(SAN, "dropin_asan")


def _env(leaks):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=%d:halt_on_error=1" % leaks
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return env


def test_mesh_code_under_sanitizers(binaries, tmp_path):
    d = str(tmp_path)
    meshgen.write_msh(os.path.join(d, "mixed.msh"), meshgen.mixed_channel(nx=12, ny=6))
    meshgen.write_msh(os.path.join(d, "step.msh"), meshgen.forward_step(n_per_unit=20))
    meshgen.write_msh(os.path.join(d, "naca.msh"), meshgen.naca0012_sized(3000))
    meshgen.write_su2(os.path.join(d, "naca_m.su2"), su2_fixture_mesh())
    txt = open(os.path.join(d, "mixed.msh")).read()
    open(os.path.join(d, "trunc.msh"), "w").write(txt[:len(txt) // 2])
    open(os.path.join(d, "empty.msh"), "w").write("")
    open(os.path.join(d, "garbage.msh"), "w").write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\nx y z\n")
    open(os.path.join(d, "bad.su2"), "w").write("NDIME= 2\nNELEM= 1\n7 0 1 2 3 4 0\nNPOIN= 3\n0 0 0\n1 0 1\n"
                                                "0 1 2\nNMARK= 0\n")
    open(os.path.join(d, "trunc.su2"), "w").write("NDIME= 2\nNELEM= 4\n5 0 1 2 0\n")
    r = subprocess.run([binaries[0], "mixed.msh", "step.msh", "naca.msh", "naca_m.su2", "--", "trunc.msh",
                        "empty.msh", "garbage.msh", "bad.su2", "trunc.su2", "missing.msh"],
                       cwd=d, capture_output=True, text=True, timeout=600, env=_env(1))
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.count(" ok") == 4


def test_dropin_host_code_under_sanitizers(binaries, tmp_path):
    d = str(tmp_path)
    meshgen.write_msh(os.path.join(d, "naca.msh"), meshgen.naca0012_sized(3000))
    r = subprocess.run([binaries[1], "naca"], cwd=d, capture_output=True, text=True, timeout=600, env=_env(0))
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "drop-in host ok" in r.stdout
    assert os.path.getsize(os.path.join(d, "out.vtu")) > 0 and os.path.getsize(os.path.join(d, "out_full.vtu")) > 0
