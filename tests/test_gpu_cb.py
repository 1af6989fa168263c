"""Contour-Buildup comparator on the GPU (sesdf_cb_* through sesdf_amd.ContourBuildup) against the
oracle (oracle/cb_oracle.cpp): neighbour lists bit-exact, probes (triples and sides) and tori exact
outside the epsilon band, positions and torus values within the FP32 bounds of cb_parity, and the
singularity texture bit-exact against the oracle's FP32 pass over the GPU's own probes."""
import os

import numpy as np
import pytest

import cb_parity
import oracle as O
from sesdf_amd import ContourBuildup, load, synth

pytestmark = pytest.mark.gpu

MMTF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mmtf")
CODES = ["1AGA", "101M", "1VIS", "7SC0", "3EAM", "7DBB", "1A8R", "7O0U", "1AON", "7RGD", "3JC8", "7CGO"]


def _say(*a):
    print("[cb]", *a, flush=True)


def gpu_dict(cb):
    d, it = cb.get_data(), cb.get_internal()
    return {"sorted_to_initial": it["sorted_to_initial"], "neighbor_count": it["neighbor_count"],
            "neighbor_ids": it["neighbor_ids"], "probe_atoms": it["probe_atoms"],
            "probe_positions": d["probe_positions"], "torus_atoms": it["torus_atoms"],
            "torus_positions": d["torus_positions"], "torus_axes": d["torus_axes"],
            "torus_visibility_spheres": d["torus_visibility_spheres"], "data": d, "internal": it}


def check_triangles(g):
    """Spherical-triangle vectors are the atom centres minus the probe (exact in FP32; a probe of a
    degenerate circle is NaN in the reference's arithmetic, and stays NaN here)."""
    d, it = g["data"], g["internal"]
    A, P = d["atoms"], d["probe_positions"]
    ids = it["probe_atoms"]
    for vec, col in ((d["sphere_triangle_vec1"], 0), (d["sphere_triangle_vec2"], 1), (d["sphere_triangle_vec3"], 2)):
        assert np.array_equal(vec[:, :3], A[ids[:, col], :3] - P[:, :3], equal_nan=True)
    assert (d["sphere_triangle_vec1"][:, 3] == 1).all() and (d["sphere_triangle_vec2"][:, 3] == 1).all()
    assert (d["sphere_triangle_vec3"][:, 3] == np.float32(d["probe_radius"]) * np.float32(d["probe_radius"])).all()
    assert (P[:, 3] == np.float32(d["probe_radius"])).all()
    assert (ids[:, 0] < ids[:, 1]).all() and (ids[:, 1] < ids[:, 2]).all()


def check_singularities(g, orc):
    d, it = g["data"], g["internal"]
    tc, st, order = orc.singularities(d["probe_positions"])
    assert np.array_equal(it["probe_sorted"], order)
    assert np.array_equal(d["tex_coords"], tc)
    cnt = tc[order, 0].astype(np.int64)
    rows = d["singularity_texture"]
    for s in np.nonzero(cnt)[0]:
        assert np.array_equal(rows[s, :3 * cnt[s]], st[s, :3 * cnt[s]]), s
    return int(cnt.sum())


def run_case(atoms, label, threads=16, **kw):
    cb = ContourBuildup(atoms, **kw)
    c = cb.counts()
    g = gpu_dict(cb)
    args = dict(probe_radius=kw.get("probe_radius", 1.4), grid_dim=kw.get("grid_dimensions", 16), threads=threads)
    orc = O.cb_build(atoms, precision=64, **args)
    rep = cb_parity.compare(g, orc, atoms)
    # the FP32 restatement's own error against FP64 on the same input sets the scale of the bound
    own = cb_parity.compare(cb_parity.as_gpu(O.cb_build(atoms, precision=32, **args)), orc, atoms)
    _say(label, {k: c[k] for k in ("probe_nb", "torus_nb", "neighbor_overflow_atoms", "arc_capacity", "build_attempts")},
         {k: rep[k] for k in ("probes_excluded", "excluded_circles", "probe_pos_max_err", "probe_pos_p99_err",
                              "torus_max_err")}, "FP32 oracle p99 %.2e max %.2e" % (own["probe_pos_p99_err"], own["probe_pos_max_err"]))
    cb_parity.assert_parity(rep, pos_tol=2 * own["probe_pos_max_err"] + 1e-5, pos_p99_tol=2 * own["probe_pos_p99_err"] + 5e-6)
    check_triangles(g)
    pairs = check_singularities(g, orc)
    assert c["probe_nb"] == len(g["probe_atoms"]) and c["torus_nb"] == len(g["torus_atoms"])
    return cb, c, rep, pairs


@pytest.mark.parametrize("n", [2000, 20000])
def test_globule_matches_oracle(gpu, n):
    atoms = synth.globule(n, synth.SEED_BASE)
    _, c, rep, _ = run_case(atoms, f"globule {n}")
    assert c["probe_nb"] > n // 2 and rep["probes_excluded"][1] < 0.05 * c["probe_nb"]


def test_headline_globule_matches_oracle(gpu):
    atoms, _ = synth.globule_voids(100_000)
    _, c, rep, pairs = run_case(atoms, "globule 100k with voids")
    assert c["probe_nb"] > 50_000 and pairs > 0


@pytest.mark.parametrize("code", CODES)
def test_sample_protein_matches_oracle(gpu, code):
    atoms = load(os.path.join(MMTF, code + ".mmtf"))
    run_case(atoms, f"{code} ({len(atoms)} atoms)")


@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_tiny_inputs(gpu, n):
    a = np.array([[0, 0, 0, 1.6], [3, 0, 0, 1.6], [1.5, 2.6, 0, 1.6]], np.float32)[:n]
    cb = ContourBuildup(a)
    c = cb.counts()
    assert c["probe_nb"] == (2 if n == 3 else 0)
    assert c["torus_nb"] == {0: 0, 1: 0, 2: 1, 3: 3}[n]
    if n:
        run_case(a, f"{n} atoms", threads=1)


@pytest.mark.parametrize("dim", [1, 4, 32])
def test_grid_dimensions(gpu, dim):
    atoms = synth.globule(2000, synth.SEED_BASE + 7)
    run_case(atoms, f"grid {dim}^3", grid_dimensions=dim)


@pytest.mark.parametrize("rp", [0.0, 2.4])
def test_probe_radius(gpu, rp):
    atoms = synth.globule(2000, synth.SEED_BASE + 8)
    run_case(atoms, f"probe {rp}", probe_radius=rp)


def test_wide_arc_lists_and_capacity_growth_match(gpu):
    """The 80-arc kernel and builds whose outputs outgrow small capacities give the same results."""
    atoms = synth.globule(20000, synth.SEED_BASE)
    ref = gpu_dict(ContourBuildup(atoms))
    cb = ContourBuildup(atoms, build_surface=False)
    cb.set_tuning(arc_capacity=80, probe_capacity=64, torus_capacity=64)
    cb.build()
    c = cb.counts()
    assert c["arc_capacity"] == 80 and c["build_attempts"] == 2
    got = gpu_dict(cb)
    for k in ("probe_atoms", "probe_positions", "torus_atoms", "torus_positions", "torus_axes",
              "torus_visibility_spheres", "neighbor_ids"):
        assert np.array_equal(got[k], ref[k], equal_nan=True), k
    assert np.array_equal(got["data"]["tex_coords"], ref["data"]["tex_coords"])


def test_rebuild_after_moving_atoms(gpu):
    """A second build reads the new atom span on the grid of the first box (as the reference, whose
    grid parameters are fixed at construction); atoms that leave the box wrap in the hash."""
    atoms = synth.globule(5000, synth.SEED_BASE + 9)
    cb = ContourBuildup(atoms)
    first = len(cb.get_internal()["probe_atoms"])
    moved = atoms.copy()
    moved[:, :3] += np.float32(0.37)
    cb.set_atoms(moved)
    cb.build()
    got = gpu_dict(cb)
    orc = O.cb_build(moved, aabb=O.aabb_of(atoms), precision=64, threads=16)
    cb_parity.assert_parity(cb_parity.compare(got, orc, moved))
    assert first > 0 and len(got["probe_atoms"]) > 0


def test_replay_matches_build(gpu):
    atoms = synth.globule(20000, synth.SEED_BASE)
    cb = ContourBuildup(atoms)
    first = gpu_dict(cb)
    for _ in range(3):
        cb.enqueue()
    cb.sync()
    again = gpu_dict(cb)
    for k in ("probe_atoms", "probe_positions", "torus_atoms", "torus_positions", "neighbor_ids"):
        assert np.array_equal(first[k], again[k], equal_nan=True), k
    assert np.array_equal(first["data"]["tex_coords"], again["data"]["tex_coords"])


@pytest.mark.parametrize("case,auto", [("globule_100k", 0), ("3JC8", 1), ("7CGO", 1), ("1A8R", 1), ("1AGA", 0), ("globule_2000", 0)])
def test_singularity_search_grid_matches_stencil_scan(gpu, case, auto):
    """The singularity texture from the 2 rp search grid is the reference's stencil scan, row for row:
    the same build with either search gives identical texture coordinates and rows. The search grid
    is chosen by itself where the reference's cells are large."""
    if case == "globule_100k":
        atoms, _ = synth.globule_voids(100_000)
    elif case == "globule_2000":
        atoms = synth.globule(2000, synth.SEED_BASE)
    else:
        atoms = load(os.path.join(MMTF, case + ".mmtf"))
    cb = ContourBuildup(atoms)
    assert cb.counts()["singularity_grid"] == auto
    out = {}
    for mode in (1, 2):
        cb.set_tuning(singularity_search=mode)
        cb.build()
        assert cb.counts()["singularity_grid"] == (1 if mode == 2 else 0)
        d = cb.get_data()
        out[mode] = (d["tex_coords"].copy(), d["singularity_texture"].copy())
    tc, order = out[1][0], cb.get_internal()["probe_sorted"]
    assert np.array_equal(out[2][0], tc)
    cnt = tc[order, 0].astype(np.int64)
    rows1, rows2 = out[1][1], out[2][1]
    for s in np.nonzero(cnt)[0]:
        assert np.array_equal(rows1[s, :3 * cnt[s]], rows2[s, :3 * cnt[s]]), s
    _say(case, "probes", len(order), "neighbour pairs", int(cnt.sum()), "full rows", int((cnt == 32).sum()))
    assert cnt.sum() > 0


def test_singularity_search_value_is_checked(gpu):
    from sesdf_amd import _lib as L

    cb = ContourBuildup(synth.globule(300, synth.SEED_BASE))
    with pytest.raises(L.SesdfError, match="singularity_search"):
        cb.set_tuning(singularity_search=3)
    cb.set_tuning(singularity_search=2)
    cb.build()
    assert cb.counts()["singularity_grid"] == 1
