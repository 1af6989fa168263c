"""C++ host classes (include/bcs): bcs::Sesdf / bcs::Ssesdf over the C ABI, driven by a compiled
program (tests/cpp/host_api.cpp) the way a reference caller uses the classes."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from sesdf_amd import synth

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


def _run_cpp(name, atoms, tmp_path, *args):
    path = tmp_path / "atoms.f32"
    np.ascontiguousarray(atoms, np.float32).tofile(path)
    exe = os.path.join(ROOT, "tests", "cpp", name)
    assert os.path.exists(exe), f"tests/cpp/{name} is not built (sesdf_amd/build.py host_api)"
    print(name, "on", atoms.shape[0], "atoms", *args, flush=True)
    r = subprocess.run([exe, str(path), *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def oracle_census(res):
    """Circle census of an FP64 oracle build (the counts app/circles/main.cu:13-87 makes), with the
    number of circles whose class the epsilon band leaves open."""
    K, n = res.K, len(res["nb_count"])
    nb, st = res["nb_ids"].reshape(n, K), res["status"].reshape(n, K)
    amb = res["amb_circle"].reshape(n, K)
    seg_circles = {(int(i), int(j)) for i, j in res["segments"][:, :2]}
    t2g = res["trimmed_to_global"]
    amb_seg = {(int(t2g[t]) // K, int(nb.reshape(-1)[t2g[t]])) for t in np.nonzero(res["amb_seg_circle"])[0]}
    out = dict(buried=0, full=0, without_intersections=0, with_intersections=0, total=0)
    band = 0
    for i in range(n):
        for l in range(int(res["nb_count"][i])):
            j = int(nb[i, l])
            if j < i:
                continue
            out["total"] += 1
            s = int(st[i, l])
            if s == 1 and (i, j) in seg_circles:
                s = 3
            out[{0: "buried", 1: "without_intersections", 2: "full", 3: "with_intersections"}[s]] += 1
            band += bool(amb[i, l]) or (i, j) in amb_seg
    return out, band


@pytest.mark.gpu
def test_circle_census_matches_oracle(gpu, tmp_path):
    """The census of app/circles/main.cu over the C++ host class (getData + toHost getters) on the
    100k-atom globule, against the FP64 oracle's; every segment's circle has status 1."""
    import oracle as O

    atoms, _ = synth.globule_voids(100000, 10, synth.SEED_BASE + 1)
    got = _run_cpp("circles_census", atoms, tmp_path)
    res = O.build(atoms, mode=0, precision=64, track=True, threads=8)
    want, band = oracle_census(res)
    print("census gpu", got, "oracle", want, "band circles", band, flush=True)
    assert got["status"] == 0 and got["segment_not_intersected"] == 0 and got["unknown_neighbor"] == 0
    assert got["total"] == want["total"]
    for k in ("buried", "full", "without_intersections", "with_intersections"):
        assert abs(got[k] - want[k]) <= band, k
    assert got["with_intersections"] > 0 and got["buried"] > 0
    # globalToTrimmed: N * K + 1 prefix counts, the last one the number of intersected circles
    assert got["g2t_size"] == 100000 * 128 + 1 and got["g2t_last"] == got["intersected"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["complete", "exterior"])
def test_stage_entry_points_give_the_build(gpu, tmp_path, mode):
    """Stages run one by one (app/detailed shape) give build()'s counts; out-of-order or wrong-surface
    stage calls are refused."""
    atoms = synth.globule(20000, synth.SEED_BASE + 4)
    got = _run_cpp("detailed_stages", atoms, tmp_path, *(["exterior"] if mode == "exterior" else []))
    print(got, flush=True)
    assert got["same_counts"] is True and min(got["ms"]) >= 0
    assert got["concave"] > 0 and (got["segments"] > 0 if mode == "complete" else got["torus"] > 0)
    assert got["out_of_order_status"] == (1 if mode == "complete" else 3)  # SESDF_ERR_INVALID / _UNSUPPORTED


def test_contour_buildup_header_compiles_cleanly(tmp_path):
    """bcs::ContourBuildup compiles with a host compiler, warnings as errors, and links."""
    out = tmp_path / "contour_buildup"
    src = os.path.join(ROOT, "tests", "cpp", "contour_buildup.cpp")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), src, "-o", str(out), "-L", os.path.join(ROOT, "sesdf_amd"), "-lsesdf_hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_contour_buildup_class_matches_python_path(gpu, tmp_path):
    """bcs::ContourBuildup (constructed as app/benchmarks/main.cu:168 does, built twice, moved) gives
    the counts of the Python mirror on the same atoms."""
    from sesdf_amd import ContourBuildup

    atoms = synth.globule(2000, synth.SEED_BASE)
    got = _run_cpp("contour_buildup", atoms, tmp_path)
    c = ContourBuildup(atoms).counts()
    assert got["atoms"] == atoms.shape[0] and got["status"] == 0 and got["vao"] == 0
    assert got["probes"] == got["data_probes"] == got["again_probes"] == c["probe_nb"]
    assert got["tori"] == got["data_tori"] == got["again_tori"] == c["torus_nb"]
    assert got["probe_neighbors"] == 32
