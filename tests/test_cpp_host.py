"""C++ host classes (include/bcs): bcs::Sesdf / bcs::Ssesdf over the C ABI, driven by a compiled
program (tests/cpp/host_api.cpp) the way a reference caller uses the classes."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "host_api")
SRC = os.path.join(ROOT, "tests", "cpp", "host_api.cpp")


def test_host_headers_compile_cleanly(tmp_path):
    """The headers compile with a host compiler, warnings as errors, and link against the library."""
    out = tmp_path / "host_api"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), SRC, "-o", str(out), "-L", os.path.join(ROOT, "sesdf_amd"), "-lsesdf_hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_host_classes_match_python_path(gpu, tmp_path):
    """Same counts through bcs::Sesdf / bcs::Ssesdf as through the Python mirror on the same atoms."""
    from sesdf_amd import Sesdf, Ssesdf, synth

    assert os.path.exists(BIN), "tests/cpp/host_api is not built (sesdf_amd/build.py host_api)"
    atoms = synth.globule(2000, synth.SEED_BASE)
    path = tmp_path / "atoms.f32"
    atoms.astype(np.float32).tofile(path)
    print("host_api on", atoms.shape[0], "atoms", flush=True)
    r = subprocess.run([BIN, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    s = Sesdf(atoms)
    c = s.counts()
    e = Ssesdf(atoms)
    ce = e.counts()
    n = atoms.shape[0]
    assert got["atoms"] == n and got["perm_sum"] == n * (n - 1) // 2
    assert got["sectors"] == c["sector_nb"] and got["sector_bytes"] == 16 * c["sector_nb"]
    assert got["segments"] == c["segment_nb"] and got["full_circles"] == c["full_circle_nb"]
    assert got["concave"] == c["concave_patch_nb"]
    assert got["exterior_torus"] == ce["torus_nb"] and got["exterior_concave"] == ce["concave_patch_nb"]
    assert got["max_neighbors"] == 128 and got["moved"] is True and got["status"] == 0
