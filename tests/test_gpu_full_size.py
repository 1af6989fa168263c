"""Parity at the full sizes of the benchmark workloads.

* 100,000-atom globule with 10 voids, probe 1.4 A (complete and exterior surface): the whole build
  against the FP64 oracle (vectorised comparator), and the SDF on the 256^3 grid at 500,000 of its
  vertices against the FP64 SDF oracle (outside the few decided by band elements), plus the
  1-Lipschitz property along every grid axis.
* 2,500,000 atoms (500 globules of 5,000 on a 360 A Fibonacci sphere), exterior surface: balls of
  20 A cut around surface atoms of the large system, each built by the oracle on its own; the
  neighbour lists and circles of atoms 12 A inside a ball and the patches of atoms 11 A inside do
  not see the cut (checked on the CPU in test_parity_fast.py), so there they must match the GPU
  build of the whole system; and the whole system against the FP64 oracle run on all of it. The
  topology must also not depend on the cell edge at this size. The complete surface of the system
  against the FP64 oracle on all atoms, and its exterior SDF at 1M points against the FP64 SDF
  oracle.
"""
import numpy as np
import pytest

import oracle as O
import parity
import parity_fast as PF
from sesdf_amd import Sesdf, Ssesdf, synth

pytestmark = pytest.mark.gpu

SDF_TOL = 1e-4


def _say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def headline():
    atoms, _ = synth.globule_voids(100_000, 10, synth.SEED_BASE + 1)
    return atoms


@pytest.mark.parametrize("mode", [0, 1])
def test_headline_build_matches_fp64(gpu, headline, mode):
    atoms = headline
    cls = Sesdf if mode == 0 else Ssesdf
    _say(f"{cls.__name__} 100k globule with voids, whole build")
    surf = cls(atoms)
    g = parity.gpu_arrays(surf)
    orc = O.build(atoms, mode=mode, precision=64, track=True, threads=16)
    rep = PF.compare_fast(g, orc, atoms, mode)
    _say({k: rep[k] for k in rep if k not in ("mismatch", "band", "band_atoms", "segment_band_circles")}, rep["mismatch"], rep["band"])
    PF.assert_parity(rep)
    assert rep["neighbor_pairs"] > 4_000_000 and rep["concave_nb"] > 50_000
    if mode == 1:
        # torus record values: the oracles take the GPU's sorted order, so that the same head atom
        # builds each record on both sides
        order = surf.get_atom_indices()
        o64 = O.build(atoms, mode=1, precision=64, track=False, threads=16, order=order)
        o32 = O.build(atoms, mode=1, precision=32, track=False, threads=16, order=order)
        compared, outside = PF.torus_values_within_fp32(g, o64, o32, atoms)
        _say(f"torus records: {rep['torus_nb']}, values compared {compared}, outside the FP32 bound {outside}")
        assert rep["torus_nb"] > 50_000 and compared > 0.99 * rep["torus_nb"] and outside == 0
    surf.close()


@pytest.mark.parametrize("exterior", [False, True])
def test_headline_sdf_grid_256(gpu, headline, exterior):
    """Within 1e-4 A of the FP64 oracle wherever neither side's winning element belongs to an atom
    of a topology difference put down to the epsilon band (where the two sides sample different
    surfaces: reported and counted, as the topology differences themselves)."""
    import torch

    from test_gpu_sdf import gpu_sdf

    atoms = headline
    _say(f"SDF 256^3 of the 100k globule, exterior={exterior}")
    surf = Sesdf(atoms)
    lo = atoms[:, :3].min(0) - atoms[:, 3].max() - 3
    hi = atoms[:, :3].max(0) + atoms[:, 3].max() + 3
    dim, band = 256, 2.0
    sp = ((hi - lo) / (dim - 1)).astype(np.float32)
    out = torch.empty(dim ** 3, device="cuda")
    surf.sample_grid_ptr(lo, sp, (dim,) * 3, out.data_ptr(), band=band, exterior=exterior)
    torch.cuda.synchronize()
    f = out.cpu().numpy()
    # 1-Lipschitz along each axis, over the whole grid (both fields are signed distances, shifted
    # by rp and clamped)
    grid = f.reshape(dim, dim, dim)
    for ax, h in ((2, sp[0]), (1, sp[1]), (0, sp[2])):
        jumps = np.abs(np.diff(grid, axis=ax)) > h * (1 + 1e-4) + 1e-5
        _say(f"axis {ax}: {int(jumps.sum())} steps above the grid spacing")
        assert not jumps.any()
    # 500,000 vertices, at the points exactly as the kernel forms them; the point path must give
    # the grid's values, and its winner codes name the elements
    rng = np.random.default_rng(17)
    ii = rng.choice(dim ** 3, 500_000, replace=False)
    lo32 = lo.astype(np.float32)
    pts = np.stack([lo32[0] + sp[0] * (ii % dim).astype(np.float32),
                    lo32[1] + sp[1] * ((ii // dim) % dim).astype(np.float32),
                    lo32[2] + sp[2] * (ii // dim ** 2).astype(np.float32)], 1)
    pv, pc = gpu_sdf(surf, pts, band, exterior, codes=True)
    np.testing.assert_array_equal(pv, f[ii])
    orc = O.build(atoms, mode=0, precision=64, track=True, threads=16)
    o, oc = orc.sdf(pts.astype(np.float64), band=band, exterior=exterior, threads=16, codes=True)
    g = parity.gpu_arrays(surf)
    rep = PF.compare_fast(g, orc, atoms, 0)
    PF.assert_parity(rep)
    s2i_g, s2i_o = g["s2i"].astype(np.int64), orc["sorted_to_initial"].astype(np.int64)
    wa_g = PF.winner_atoms(pc, s2i_g, g["full_circles"], g["segments"], g["concave_ids"])
    wa_o = PF.winner_atoms(oc, s2i_o, orc["full_circles"], orc["segments"], orc["concave_ids"])
    excused = (np.isin(wa_g, rep["band_atoms"]) | np.isin(wa_o, rep["band_atoms"])).any(axis=1)
    err = np.abs(pv - o)
    _say(f"max |GPU - FP64| {err[~excused].max():.2e} over {int((~excused).sum())} vertices; "
         f"{int(excused.sum())} decided by band elements ({len(rep['band_atoms'])} atoms), max there {err[excused].max() if excused.any() else 0:.2e}; "
         f"inside {np.mean(o < 0):.3f}, in band {np.mean(np.abs(o) < band):.3f}")
    if err[~excused].max() > SDF_TOL:
        for q in np.argsort(-np.where(excused, 0, err))[:8]:
            _say(f"vertex {int(ii[q])} at {pts[q]}: GPU {pv[q]:.6f} (code {int(pc[q]):#x}) oracle {o[q]:.6f} (code {int(oc[q]):#x})")
    assert err[~excused].max() <= SDF_TOL
    assert excused.mean() < 1e-3
    # (the exterior field fills the voids and the packing's small cavities: fewer band vertices)
    assert (o < 0).mean() > 0.2 and (np.abs(o) < band).mean() > (0.05 if exterior else 0.1)
    surf.close()


@pytest.fixture(scope="module")
def large():
    return synth.fibonacci_globules(500, 5000, 360.0, synth.SEED_BASE + 2)


def test_large_exterior_balls_match_fp64(gpu, large):
    atoms = large
    _say("Ssesdf 2.5M atoms, whole build")
    surf = Ssesdf(atoms)
    g = parity.gpu_arrays(surf)
    counts = surf.counts()
    _say("counts", {k: counts[k] for k in ("concave_patch_nb", "torus_nb", "neighbor_nb") if k in counts})
    rng = np.random.default_rng(23)
    per = 5000
    total_pairs = total_patches = 0
    for b in range(6):
        m = int(rng.integers(0, len(atoms) // per))
        mol = atoms[m * per:(m + 1) * per, :3]
        dc = np.linalg.norm(mol - mol.mean(0), axis=1)
        centre = mol[rng.choice(np.nonzero(dc > dc.max() - 5)[0])]
        d = np.linalg.norm(atoms[:, :3] - centre, axis=1)
        ball = np.nonzero(d < 20.0)[0]
        _say(f"ball {b}: globule {m}, {len(ball)} atoms")
        sub = O.build(atoms[ball], mode=1, precision=64, track=True, threads=16)
        rep = PF.compare_fast(g, sub, atoms, 1, orc_to_input=ball, keep=np.nonzero(d < 12.0)[0],
                              keep_patches=np.nonzero(d < 11.0)[0])
        _say(rep["mismatch"], rep["band"], rep["neighbor_pairs"], rep["concave_nb"], f"{rep['probe_pos_max_err']:.2e}")
        PF.assert_parity(rep)
        total_pairs += rep["neighbor_pairs"]
        total_patches += rep["concave_nb"]
    assert total_pairs > 5000 and total_patches > 300
    surf.close()


def test_large_exterior_whole_system_matches_fp64(gpu, large):
    """The whole 2.5M-atom exterior build against the FP64 oracle on all atoms (threaded): lists,
    statuses, concave patches, probe positions, probe-neighbour sets and torus records."""
    atoms = large
    _say("Ssesdf 2.5M atoms, whole build against the FP64 oracle on the whole system")
    surf = Ssesdf(atoms)
    g = parity.gpu_arrays(surf)
    surf.close()
    orc = O.build(atoms, mode=1, precision=64, track=True, threads=16)
    _say("oracle built")
    rep = PF.compare_fast(g, orc, atoms, 1, log=_say)
    _say({k: rep[k] for k in ("neighbor_pairs", "concave_nb", "probe_pos_max_err", "torus_nb", "probe_sets_compared")},
         rep["mismatch"], rep["band"])
    PF.assert_parity(rep)
    assert rep["neighbor_pairs"] > 100_000_000 and rep["concave_nb"] > 2_000_000 and rep["torus_nb"] > 2_500_000
    assert rep["probe_sets_compared"] > 500_000


def test_large_complete_whole_system_and_exterior_sdf_at_1m_points(gpu, large):
    """The complete surface of the 2.5M-atom system (500 separate globules, whose outer surfaces
    all bound the unbounded region) against the FP64 oracle on all atoms, and its exterior SDF at
    1M points within 2 A of atom spheres against the FP64 SDF oracle."""
    import torch
    from scipy.spatial import cKDTree

    atoms = large
    _say("Sesdf 2.5M atoms, whole complete build and exterior SDF at 1M points")
    rng = np.random.default_rng(29)
    a = rng.integers(0, len(atoms), 1_000_000)
    u = rng.standard_normal((1_000_000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = atoms[a, 3] + 1.4 + rng.uniform(-2.0, 2.0, 1_000_000)
    pts = (atoms[a, :3] + u * r[:, None]).astype(np.float32)
    surf = Sesdf(atoms)
    P = torch.from_numpy(np.c_[pts, np.zeros(len(pts), np.float32)]).cuda()
    out = torch.empty(len(pts), device="cuda")
    codes = torch.empty(len(pts), device="cuda", dtype=torch.int64)
    surf.sample_points_ptr(P.data_ptr(), len(pts), out.data_ptr(), band=2.0, exterior=True, d_codes=codes.data_ptr())
    torch.cuda.synchronize()
    got, pc = out.cpu().numpy(), codes.cpu().numpy()
    g = parity.gpu_arrays(surf)
    overflow = surf.counts()["component_overflow_atoms"]
    surf.close()
    orc = O.build(atoms, mode=0, precision=64, track=True, threads=16)
    _say("oracle built")
    rep = PF.compare_fast(g, orc, atoms, 0, log=_say)
    _say({k: rep[k] for k in ("neighbor_pairs", "concave_nb", "probe_pos_max_err", "sector_max_err", "sector_excess",
                              "probe_sets_compared")}, rep["mismatch"], rep["band"])
    PF.assert_parity(rep)
    assert rep["neighbor_pairs"] > 100_000_000 and rep["concave_nb"] > 2_000_000
    o, oc = orc.sdf(pts.astype(np.float64), band=2.0, exterior=True, threads=16, codes=True)
    _say("oracle field evaluated")
    # A point decided by an element of an atom whose topology the two builds settle differently
    # inside the epsilon band is excused. The exterior test also reads the surface components, which
    # such an element can join or split; a differing point decided elsewhere must lie within 10 A
    # of such an atom, and such points must be rare.
    band_atoms = np.asarray(rep["band_atoms"], np.int64)
    s2i_g, s2i_o = g["s2i"].astype(np.int64), orc["sorted_to_initial"].astype(np.int64)
    wa_g = PF.winner_atoms(pc, s2i_g, g["full_circles"], g["segments"], g["concave_ids"])
    wa_o = PF.winner_atoms(oc, s2i_o, orc["full_circles"], orc["segments"], orc["concave_ids"])
    excused = (np.isin(wa_g, band_atoms) | np.isin(wa_o, band_atoms)).any(axis=1)
    err = np.abs(got - o)
    bad = (err > SDF_TOL) & ~excused
    dist = cKDTree(atoms[band_atoms, :3].astype(np.float64)).query(pts[bad].astype(np.float64))[0] if len(band_atoms) else \
        np.full(int(bad.sum()), np.inf)
    _say(f"{len(band_atoms)} band atoms; {int(excused.sum())} points decided by their elements; "
         f"{int(bad.sum())} other points differ, farthest {dist.max() if len(dist) else 0.0:.2f} A from a band atom; "
         f"component overflow atoms {overflow}; max |GPU - FP64| elsewhere {err[~excused & ~bad].max():.2e}; "
         f"inside {np.mean(o < 0):.3f}, in band {np.mean(np.abs(o) < 2.0):.3f}")
    assert bad.sum() <= 1e-4 * len(pts) and (dist <= 10.0).all()
    assert excused.mean() < 1e-3 and (o < 0).mean() > 0.2 and (np.abs(o) < 2.0).mean() > 0.2


def test_large_exterior_topology_independent_of_cell(gpu, large):
    """Another cell edge changes the grid order and so which atom of a circle or triple evaluates
    it: the two builds may differ only in near-degenerate records, each confirmed as such by the
    oracle on a ball around it."""
    atoms = large
    keys, torus = [], []
    for cell in (3.2, 5.0):
        _say(f"Ssesdf 2.5M atoms, cell {cell}")
        surf = Ssesdf(atoms, cell_size=cell)
        d = surf.get_data()
        keys.append(np.unique(PF.concave_keys(d["concave_patches_id"], d["concave_patches_position"], surf.get_atom_indices(), atoms)))
        torus.append(surf.counts()["torus_nb"])
        surf.close()
    diff = np.setxor1d(keys[0], keys[1])
    _say(f"{len(keys[0])} / {len(keys[1])} patches, {len(diff)} differ; torus records {torus}")
    assert len(keys[0]) > 1_000_000
    assert len(diff) <= 1e-5 * len(keys[0])
    assert abs(torus[0] - torus[1]) <= 2 * len(diff) + 10
    for k in diff:
        tri = np.array([k["a"], k["b"], k["c"]])
        centre = atoms[tri, :3].mean(0)
        ball = np.nonzero(np.linalg.norm(atoms[:, :3] - centre, axis=1) < 14.0)[0]
        sub = O.build(atoms[ball], mode=1, precision=64, track=True, threads=16)
        s2i = ball[sub["sorted_to_initial"].astype(np.int64)]
        amb = set()
        ck = PF.concave_keys(sub["concave_ids"], sub["concave_pos64"], s2i, atoms)
        for q in np.nonzero(sub["amb_concave"])[0]:
            amb.add((int(ck["a"][q]), int(ck["b"][q]), int(ck["c"][q])))
        for r in sub["amb_rejected"].reshape(-1, 4):
            amb.add(tuple(sorted(int(s2i[v]) for v in r[:3])))
        _say(f"patch {tuple(int(v) for v in tri)} side {int(k['s'])}: near-degenerate {tuple(int(v) for v in tri) in amb}")
        assert tuple(int(v) for v in tri) in amb


def test_batch_4096_molecules_match_fp64(gpu):
    """The batch workload (4,096 molecules of 20 to 5,000 atoms in one launch sequence, a 32^3
    grid each): 40 molecules across the size range against the FP64 oracle, build and grid. Grid
    vertices within 8 A of an atom of a band difference are left out (reported)."""
    import torch

    from sesdf_amd.batch import SesdfBatch, molecule_slice

    sizes = synth.batch_sizes(4096, 20, 5000, synth.SEED_BASE + 3)
    atoms, offs = synth.batch(sizes, synth.SEED_BASE + 3)
    _say(f"batch of {len(sizes)} molecules, {len(atoms)} atoms")
    batch = SesdfBatch(atoms, offs)
    dim, margin, band = 32, 3.0, 2.0
    grids = torch.empty(len(sizes) * dim ** 3, device="cuda")
    batch.sample_grids_ptr(dim, grids.data_ptr(), margin=margin, band=band)
    torch.cuda.synchronize()
    grids = grids.cpu().numpy().reshape(len(sizes), -1)
    counts, ids = batch.get_neighbors()
    status = batch.get_circle_status()
    s2i = batch.get_atom_indices().astype(np.int64)
    data, o = batch.get_data(), batch.offsets()
    K = batch.get_max_neighbor_per_atom()
    order = np.argsort(np.diff(offs.astype(np.int64)), kind="stable")
    chosen = order[np.linspace(0, len(order) - 1, 40).astype(int)]
    worst, excused_total = 0.0, 0
    for m in chosen:
        m = int(m)
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        a0, a1 = int(o["atoms"][m]), int(o["atoms"][m + 1])
        bm = molecule_slice(data, o, m)
        gd = {"K": K, "s2i": s2i[a0:a1] - int(offs[m]), "nb_count": counts[a0:a1],
              "nb_ids": ids[a0:a1].astype(np.int64) - a0, "status": status[a0:a1],
              "concave_ids": bm["concave_patches_id"], "concave_pos": bm["concave_patches_position"],
              "concave_pos_w": bm["concave_patches_position"][:, 3], "full_circles": bm["circle_patches"],
              "sectors": bm["sectors"], "convex_patches": bm["convex_patches"], "segments": bm["segment_patches"],
              "concave_nb": bm["concave_patches_neighbors"], "max_nb": int(data["max_concave_neighbors"]),
              "counts": {"concave_with_neighbors_nb": bm["concave_with_neighbors_nb"]}}
        orc = O.build(mol, mode=0, precision=64, track=True, threads=16)
        rep = PF.compare_fast(gd, orc, mol, 0)
        PF.assert_parity(rep)
        lo = (mol[:, :3] - mol[:, 3:4]).min(0) - margin
        hi = (mol[:, :3] + mol[:, 3:4]).max(0) + margin
        lo32, sp = lo.astype(np.float32), ((hi - lo) / (dim - 1)).astype(np.float32)
        ii = np.arange(dim ** 3)
        pts = np.stack([lo32[0] + sp[0] * (ii % dim).astype(np.float32),
                        lo32[1] + sp[1] * ((ii // dim) % dim).astype(np.float32),
                        lo32[2] + sp[2] * (ii // dim ** 2).astype(np.float32)], 1)
        ref = orc.sdf(pts.astype(np.float64), band=band, threads=16)
        near = np.zeros(len(pts), bool)
        for a in rep["band_atoms"]:
            near |= np.linalg.norm(pts - mol[a, :3], axis=1) < 8.0
        err = np.abs(grids[m] - ref)[~near]
        worst = max(worst, float(err.max()) if len(err) else 0.0)
        excused_total += int(near.sum())
        assert err.max() <= SDF_TOL, (m, len(mol), float(err.max()))
    _say(f"40 molecules of {len(atoms[int(offs[chosen[0]]):int(offs[chosen[0] + 1])])} to "
         f"{len(atoms[int(offs[chosen[-1]]):int(offs[chosen[-1] + 1])])} atoms: grid max |GPU - FP64| {worst:.2e}, "
         f"{excused_total} of {40 * dim ** 3} vertices near band differences")
    batch.close()


def test_differentiation_workload(gpu):
    """20,000-atom globule, 1M queries drawn uniformly within 2 A of the surface, w ~ N(0, 1): the
    SDF values, the backward with stored and with recomputed winners against the FP64 analytic
    gradient, and against FP64 central differences (h = 1e-4 A, topology held fixed) on 256 random
    atoms of the loss over 50,000 queries, per atom."""
    import torch

    atoms = synth.globule(20000, synth.SEED_BASE + 4)
    _say("20k globule, 1M queries: values and gradients")
    surf = Sesdf(atoms)
    orc = O.build(atoms, mode=0, precision=64, track=True, threads=16)
    rng = np.random.default_rng(29)
    band = 2.0
    lo, hi = atoms[:, :3].min(0) - 4, atoms[:, :3].max(0) + 4
    cand = (lo + (hi - lo) * rng.random((4_600_000, 3))).astype(np.float32)  # about 25% fall in the band
    s = orc.sdf(cand.astype(np.float64), band=band, threads=16)
    pts = cand[np.abs(s) < band][:1_000_000]
    assert len(pts) == 1_000_000
    p64 = pts.astype(np.float64)
    w = rng.standard_normal(len(pts)).astype(np.float32)
    P = torch.from_numpy(np.c_[pts, np.zeros(len(pts), np.float32)]).cuda()
    W = torch.from_numpy(w).cuda()
    val = torch.empty(len(pts), device="cuda")
    codes = torch.empty(len(pts), device="cuda", dtype=torch.int64)
    surf.sample_points_ptr(P.data_ptr(), len(pts), val.data_ptr(), band=band, d_codes=codes.data_ptr())
    grads = {}
    for store in (True, False):
        grad = torch.zeros(len(atoms), 4, device="cuda")
        scratch = torch.empty(len(atoms), 4, device="cuda")
        surf.sdf_backward_ptr(P.data_ptr(), W.data_ptr(), len(pts), grad.data_ptr(), scratch.data_ptr(), band=band,
                              d_codes=codes.data_ptr() if store else None)
        torch.cuda.synchronize()
        grads[store] = grad.cpu().numpy().astype(np.float64)
    v, c = val.cpu().numpy(), codes.cpu().numpy()

    # values, outside the points decided by elements of band differences
    o, oc = orc.sdf(p64, band=band, threads=16, codes=True)
    g = parity.gpu_arrays(surf)
    rep = PF.compare_fast(g, orc, atoms, 0)
    PF.assert_parity(rep)
    wa_g = PF.winner_atoms(c, g["s2i"].astype(np.int64), g["full_circles"], g["segments"], g["concave_ids"])
    wa_o = PF.winner_atoms(oc, orc["sorted_to_initial"].astype(np.int64), orc["full_circles"], orc["segments"], orc["concave_ids"])
    excused = (np.isin(wa_g, rep["band_atoms"]) | np.isin(wa_o, rep["band_atoms"])).any(axis=1)
    verr = np.abs(v - o)[~excused]
    _say(f"values: max |GPU - FP64| {verr.max():.2e}, {int(excused.sum())} decided by band elements")
    assert verr.max() <= SDF_TOL

    # gradients against the FP64 analytic gradient: per-atom error over the gradient scale
    ref = orc.sdf_grad(p64, w.astype(np.float64), band=band, threads=16)
    scale = np.abs(ref).max()
    for store, gg in grads.items():
        err = np.abs(gg - ref).max(1) / scale
        _say(f"backward, stored winners {store}: max {err.max():.2e}, p99.9 {np.quantile(err, 0.999):.2e}")
        assert np.quantile(err, 0.999) <= 1e-3
        assert np.mean(err > 1e-3) < 1e-3

    # and against FP64 central differences (h = 1e-4 A, topology held fixed) on 256 random atoms
    # of the loss over the first 50,000 queries. An atom is stable when no query near it changes
    # its winning element under the +-h moves (a change puts a kink inside the difference). Per
    # atom: ||g - fd|| / max(||fd||, floor), floor = 1% of the median ||fd|| of the 256 atoms.
    sub = 50_000
    grad = torch.zeros(len(atoms), 4, device="cuda")
    scratch = torch.empty(len(atoms), 4, device="cuda")
    surf.sdf_backward_ptr(P.data_ptr(), W.data_ptr(), sub, grad.data_ptr(), scratch.data_ptr(), band=band,
                          d_codes=codes.data_ptr())
    torch.cuda.synchronize()
    g_sub = grad.cpu().numpy().astype(np.float64)
    ref_sub = orc.sdf_grad(p64[:sub], w[:sub].astype(np.float64), band=band, threads=16)
    norm_ref = np.linalg.norm(ref_sub, axis=1)
    touched = np.nonzero(norm_ref > 1e-3 * norm_ref.max())[0]
    ids = np.sort(rng.choice(touched, 256, replace=False)).astype(np.uint32)
    fd, stable = orc.sdf_fd(p64[:sub], w[:sub].astype(np.float64), ids, h=1e-4, band=band, threads=16)
    nfd = np.linalg.norm(fd, axis=1)
    floor = 1e-2 * np.median(nfd)
    rel = np.linalg.norm(g_sub[ids] - fd, axis=1) / np.maximum(nfd, floor)
    _say(f"finite differences: {int(stable.sum())} of 256 atoms stable; per-atom ||g - fd|| / max(||fd||, {floor:.3g}): "
         f"max {rel[stable].max():.2e}, median {np.median(rel[stable]):.2e}; unstable atoms max {rel[~stable].max() if (~stable).any() else 0:.2e}")
    assert stable.mean() >= 0.9
    assert np.all(rel[stable] <= 1e-3)
    surf.close()
