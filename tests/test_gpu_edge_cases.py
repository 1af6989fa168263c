"""GPU parity at the edges of the input domain, through the C ABI, against the FP64 oracle:
no atoms; neighbour lists truncated at K (max_neighbors 32 / 64); a cubic lattice (exact ties
everywhere); collinear centres (a sphere centred on a circle's centre or axis); near-coincident
atoms (ill-conditioned trilateration, which the kernels solve again in FP64); coordinates far
from the origin; a large probe radius with K above 128."""
import numpy as np
import pytest

import oracle as O
import parity
from sesdf_amd import Sesdf, Ssesdf, synth

pytestmark = pytest.mark.gpu


def _surf(atoms, mode, **kw):
    cls = Sesdf if mode == 0 else Ssesdf
    print(f"[gpu] {cls.__name__} n={len(atoms)} {kw}", flush=True)
    return cls(atoms, **kw)


def _check(atoms, mode, threads=8, same_lists=False, **kw):
    """same_lists: the oracle takes the GPU build's neighbour lists and sorted order (lists truncated
    at K depend on the grid order, and the head atom of a triple bounds its probe tests)."""
    surf = _surf(atoms, mode, **kw)
    extra = {}
    if same_lists:
        extra = dict(neighbors=_input_lists(surf), order=surf.get_atom_indices())
    orc = O.build(atoms, mode=mode, precision=64, track=True, threads=threads,
                  max_neighbors=kw.get("max_neighbors", 128), probe_radius=kw.get("probe_radius", 1.4), **extra)
    rep = parity.compare(parity.gpu_arrays(surf), orc, atoms, mode)
    print({k: rep[k] for k in rep if k != "mismatch"}, rep["mismatch"], surf.counts(), flush=True)
    parity.assert_parity(rep)
    return surf, orc, rep


def lattice(m=6, a=3.0, r=1.6):
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3) * a
    return np.c_[g, np.full(len(g), r)].astype(np.float32)


def chain(n=40, step=2.2, wiggle=0.0, r=1.6, seed=0):
    rng = np.random.default_rng(seed)
    yz = rng.uniform(-wiggle, wiggle, (n, 2))
    return np.c_[np.arange(n) * step, yz, np.full(n, r)].astype(np.float32)


def near_coincident(eps, n=300, twins=40, seed=4):
    """A globule plus copies of some of its atoms moved by eps in random directions."""
    base = synth.globule(n, synth.SEED_BASE + 40)
    rng = np.random.default_rng(seed)
    pick = rng.choice(n, twins, replace=False)
    v = rng.normal(size=(twins, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    dup = base[pick].copy()
    dup[:, :3] += (eps * v).astype(np.float32)
    return np.concatenate([base, dup]).astype(np.float32)


@pytest.mark.parametrize("mode", [0, 1])
def test_empty_input(gpu, mode):
    import torch

    atoms = np.zeros((0, 4), np.float32)
    surf = _surf(atoms, mode)
    c = surf.counts()
    assert all(c[k] == 0 for k in ("atom_nb", "neighbor_nb", "sector_nb", "concave_patch_nb", "segment_nb", "torus_nb"))
    d = surf.get_data()
    assert d["atoms"].shape == (0, 4) and len(d["concave_patches_position"]) == 0
    assert len(surf.get_atom_indices()) == 0
    surf.build()  # rebuild of an empty input
    if mode == 1:
        return
    band = 2.0
    pts = torch.tensor([[0.0, 0.0, 0.0, 0.0], [5.0, -1.0, 2.0, 0.0]], device="cuda")
    out = torch.zeros(2, device="cuda")
    codes = torch.zeros(2, device="cuda", dtype=torch.int64)
    surf.sample_points_ptr(pts.data_ptr(), 2, out.data_ptr(), band=band, d_codes=codes.data_ptr())
    grid = torch.zeros(4 * 4 * 4, device="cuda")
    surf.sample_grid_ptr((0, 0, 0), (0.5, 0.5, 0.5), (4, 4, 4), grid.data_ptr(), band=band)
    grad = torch.full((1, 4), 7.0, device="cuda")  # untouched: there are no atoms
    w = torch.ones(2, device="cuda")
    surf.sdf_backward_ptr(pts.data_ptr(), w.data_ptr(), 2, grad.data_ptr(), grad.data_ptr(), band=band)
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == band).all() and (grid.cpu().numpy() == band).all()
    assert ((codes.cpu().numpy() >> 40) == 5).all()  # clamped
    assert (grad.cpu().numpy() == 7.0).all()
    labels, ext = surf.surface_components()
    assert len(labels) == 0 and ext == 0


def _input_lists(surf):
    """GPU neighbour lists per input atom, in input ids (for the oracle's override)."""
    cnt, ids = surf.get_neighbors()
    s2i = surf.get_atom_indices().astype(np.int64)
    K = ids.shape[1]
    c_in = np.zeros(len(s2i), np.uint32)
    i_in = np.zeros((len(s2i), K), np.uint32)
    c_in[s2i] = cnt
    for i in range(len(s2i)):
        i_in[s2i[i], : cnt[i]] = s2i[ids[i, : cnt[i]]]
    return c_in, i_in


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("K", [32, 64])
def test_truncated_neighbor_lists(gpu, mode, K):
    atoms = synth.globule(3000, synth.SEED_BASE + 50)
    surf, orc, rep = _check(atoms, mode, same_lists=True, max_neighbors=K)
    c = surf.counts()
    # the kept neighbours are the K lowest sorted indices of the full list (cell order), unless a
    # pair within the predicate's epsilon band is involved
    full = O.build(atoms, mode=mode, precision=64, track=True, threads=8, max_neighbors=128)
    s2i_o, s2i_g = full["sorted_to_initial"], surf.get_atom_indices()
    g_of_input = np.empty(len(atoms), np.int64)
    g_of_input[s2i_g] = np.arange(len(atoms))
    amb_atoms = {int(s2i_o[a]) for p in full["amb_pairs"] for a in p}
    cnt_g, ids_g = surf.get_neighbors()
    cnt_o, ids_o = full["nb_count"], full["nb_ids"].reshape(-1, 128)
    truncated = checked = 0
    for io in range(len(atoms)):
        a = int(s2i_o[io])
        if a in amb_atoms:
            continue
        true = np.sort(g_of_input[s2i_o[ids_o[io, : cnt_o[io]]]])
        ig = g_of_input[a]
        got = ids_g[ig, : cnt_g[ig]]
        assert cnt_g[ig] == min(len(true), K)
        np.testing.assert_array_equal(got, true[:K])
        truncated += len(true) > K
        checked += 1
    print("atoms checked", checked, "truncated", truncated, flush=True)
    assert truncated > (100 if K == 32 else 0)
    assert c["neighbor_overflow_atoms"] >= truncated


@pytest.mark.parametrize("mode", [0, 1])
def test_cubic_lattice_ties(gpu, mode):
    _check(lattice(), mode)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("wiggle", [0.0, 0.02])
def test_collinear_chain(gpu, mode, wiggle):
    surf, orc, rep = _check(chain(wiggle=wiggle), mode)
    if wiggle == 0.0:
        # sphere i+1 sits on the centre of circle (i, i+2): that circle is hidden
        assert rep["status_hist"][0] > 0


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("eps", [0.1, 0.02, 0.005, 0.002])
def test_near_coincident_atoms(gpu, mode, eps):
    surf, orc, rep = _check(near_coincident(eps), mode)
    assert rep["concave_nb"] > 400
    assert surf.counts()["fp64_resolve_nb"] > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_far_from_origin(gpu, mode):
    atoms = synth.globule(2000, synth.SEED_BASE + 21)
    atoms[:, :3] += np.float32([300.0, -300.0, 150.0])
    _check(atoms, mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_large_probe_radius_wide_lists(gpu, mode):
    atoms = synth.globule(1500, synth.SEED_BASE + 30)
    surf, orc, rep = _check(atoms, mode, probe_radius=2.4, max_neighbors=288)
    assert orc["nb_count"].max() > 128  # the lists are wider than the 128-slot path holds
    if mode == 1:
        return
    import torch

    band = 2.0
    rng = np.random.default_rng(8)
    lo, hi = atoms[:, :3].min(0) - 4, atoms[:, :3].max(0) + 4
    q = (lo + (hi - lo) * rng.random((6000, 3))).astype(np.float32)
    P = torch.from_numpy(np.c_[q, np.zeros(len(q), np.float32)]).cuda()
    out = torch.empty(len(q), device="cuda")
    surf.sample_points_ptr(P.data_ptr(), len(q), out.data_ptr(), band=band)
    torch.cuda.synchronize()
    ref = orc.sdf(q.astype(np.float64), band=band, threads=8)
    err = np.abs(out.cpu().numpy().astype(np.float64) - ref)
    print("sdf max err", err.max(), flush=True)
    assert err.max() <= 1e-4


@pytest.mark.parametrize("mode", [0, 1])
def test_neighbor_overflow_grows_the_capacity(gpu, mode):
    """With the grow policy a build whose lists exceed K = 32 raises K to the longest list (a
    multiple of 32) and builds again: no list is truncated, and the whole build equals the oracle's
    with untruncated lists (K = 128)."""
    atoms = synth.globule(3000, synth.SEED_BASE + 50)
    surf = _surf(atoms, mode, max_neighbors=32, grow_neighbors=True)
    c = surf.counts()
    print("after growth:", {k: c[k] for k in ("max_neighbors", "neighbor_growth_nb", "neighbor_overflow_atoms")}, flush=True)
    assert c["neighbor_growth_nb"] >= 1 and c["neighbor_overflow_atoms"] == 0
    assert 32 < c["max_neighbors"] <= 128 and c["max_neighbors"] % 32 == 0
    cnt, _ = surf.get_neighbors()
    assert cnt.max() > 32 and c["max_neighbors"] - 32 < cnt.max() <= c["max_neighbors"]
    orc = O.build(atoms, mode=mode, precision=64, track=True, threads=8, max_neighbors=128)
    rep = parity.compare(parity.gpu_arrays(surf), orc, atoms, mode)
    parity.assert_parity(rep)
    # the default policy truncates, as the reference
    plain = _surf(atoms, mode, max_neighbors=32)
    assert plain.counts()["neighbor_overflow_atoms"] > 0 and plain.counts()["max_neighbors"] == 32


@pytest.mark.parametrize("mode", [0, 1])
def test_tangent_spheres_within_epsilon(gpu, mode):
    """Spheres farther apart than Ri + Rj by less than the predicate's slack are neighbours whose
    circle has radius^2 < 0. The reference's sector angle is then acos(clamp(NaN)) = acos(1) = 0
    (helper_math clamp is fmaxf(lo, fminf(v, hi))): a point cap, which hides no part of the face.
    Here such a pair has no circle at all (status 0 on both sides, no sector): the same surface."""
    d = np.float32(np.sqrt(np.float64(6.2) ** 2 + 5e-5))
    atoms = np.array([[0, 0, 0, 1.7], [d, 0, 0, 1.7]], np.float32)
    surf, orc, rep = _check(atoms, mode)
    if mode == 1:
        assert surf.counts()["torus_nb"] == 0
        return
    data = surf.get_data()
    assert surf.counts()["neighbor_nb"] == 2
    assert data["sectors"].shape == (0, 4) and orc["sectors"].size == 0
    assert len(data["circle_patches"]) == 0 and len(data["segment_patches"]) == 0
    import torch

    # the face of each atom stays exposed: inside the SAS ball of atom 0, on its far side, the field
    # is rp - (R - |q - c|) (a hidden face would leave only the degenerate circle to win)
    q = np.array([[-2.5, 0.0, 0.0], [-2.0, 1.0, 0.3]], np.float32)
    P = torch.from_numpy(np.c_[q, np.zeros(len(q), np.float32)]).cuda()
    out = torch.empty(len(q), device="cuda")
    surf.sample_points_ptr(P.data_ptr(), len(q), out.data_ptr(), band=2.0)
    torch.cuda.synchronize()
    ref = orc.sdf(q.astype(np.float64), band=2.0)
    face = 1.4 - (3.1 - np.linalg.norm(q.astype(np.float64), axis=1))
    np.testing.assert_allclose(ref, face, atol=1e-6)
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-4)


def random_cluster(seed):
    """5-60 atoms, radii 1.2-2.0 A, in a box sized for a liquid-like density (overlaps allowed)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 61))
    side = (n / 0.05) ** (1 / 3)
    xyz = rng.uniform(0, side, (n, 3))
    r = rng.uniform(1.2, 2.0, n)
    return np.c_[xyz, r].astype(np.float32)


@pytest.mark.parametrize("seed", range(12))
def test_random_small_clusters(gpu, seed):
    """Random small clusters with random probe radii (1.0-2.0 A), both paths, against the oracle."""
    atoms = random_cluster(synth.SEED_BASE + 100 + seed)
    rp = float(np.random.default_rng(seed).uniform(1.0, 2.0))
    for mode in (0, 1):
        _check(atoms, mode, probe_radius=rp)


@pytest.mark.parametrize("rp", [0.0, 0.3])
@pytest.mark.parametrize("mode", [0, 1])
def test_small_probe_radius(gpu, mode, rp):
    """Probe radius 0 (the SES is the van der Waals surface: probes sit on triple points of the
    spheres, none has neighbours) and 0.3 A; the complete-surface SDF too."""
    atoms = synth.globule(1500, synth.SEED_BASE + 77)
    surf, orc, rep = _check(atoms, mode, probe_radius=rp)
    assert rep["concave_nb"] > 1000
    if mode == 0:
        from test_gpu_sdf import gpu_sdf, near_surface_points

        pts = near_surface_points(atoms, orc, 20000, 2.0, 11)
        g = gpu_sdf(surf, pts, 2.0)
        o = orc.sdf(pts.astype(np.float64), band=2.0, threads=8)
        assert np.abs(g - o).max() <= 1e-4


@pytest.mark.parametrize("mode", [0, 1])
def test_clusters_far_apart(gpu, mode):
    """Two globules 20,000 A apart: the dense cell brick would exceed its cap, so its cells are
    enlarged (to about 60 A); the topology must not change. Probe positions are FP32 outputs: at
    |x| = 20,000 A one ulp is 2e-3 A, so they are held to the FP32 rounding of the FP64 position."""
    a = synth.globule(800, synth.SEED_BASE + 81)
    b = synth.globule(800, synth.SEED_BASE + 82) + np.float32([20000.0, 0.0, 0.0, 0.0])
    atoms = np.concatenate([a, b])
    surf = _surf(atoms, mode, cell_size=2.5)
    orc = O.build(atoms, mode=mode, precision=64, track=True, threads=8)
    rep = parity.compare(parity.gpu_arrays(surf), orc, atoms, mode)
    print({k: rep[k] for k in rep if k != "mismatch"}, rep["mismatch"], flush=True)
    assert not any(rep["mismatch"].values()), rep
    assert rep["concave_nb"] > 1000
    assert rep["probe_pos_max_err"] <= 1e-4 + float(np.spacing(np.float32(20000.0)))
