"""GPU SDF sampling and its gradient vs the FP64 oracle (values within 1e-4 A; gradients within 1e-3
relative of the FP64 analytic gradient, which itself matches FP64 central differences)."""
import numpy as np
import pytest

import oracle as O
from sesdf_amd import Sesdf, synth

pytestmark = pytest.mark.gpu
SDF_TOL = 1e-4
GRAD_TOL = 1e-3


def _torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def near_surface_points(atoms, orc, count, band, seed):
    rng = np.random.default_rng(seed)
    lo = atoms[:, :3].min(0) - 4
    hi = atoms[:, :3].max(0) + 4
    q = (lo + (hi - lo) * rng.random((count * 6, 3))).astype(np.float32)
    s = orc.sdf(q.astype(np.float64), band=band, threads=8)
    keep = np.abs(s) < band * 0.999
    sel = np.concatenate([q[keep][: count // 2], q[~keep][: count // 2]])
    return sel


def gpu_sdf(surf, pts, band, exterior=False, codes=False):
    torch = _torch()
    P = torch.from_numpy(np.c_[pts, np.zeros(len(pts), np.float32)]).cuda()
    out = torch.empty(len(pts), device="cuda")
    cd = torch.empty(len(pts), device="cuda", dtype=torch.int64) if codes else None
    surf.sample_points_ptr(P.data_ptr(), len(pts), out.data_ptr(), band=band, exterior=exterior,
                           d_codes=cd.data_ptr() if codes else None)
    torch.cuda.synchronize()
    return (out.cpu().numpy(), cd.cpu().numpy() if codes else None) if codes else out.cpu().numpy()


def test_surface_components_match_oracle(gpu):
    """Same partition of atoms into boundary components (labels up to renaming; atoms whose faces
    lie in several components and buried atoms compared as such) and the same exterior component."""
    atoms, _ = synth.globule_voids(3000, 3, synth.SEED_BASE + 106)
    surf = Sesdf(atoms)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    lg, eg = surf.surface_components()
    lo, eo = orc.components()
    s2i_g, s2i_o = surf.get_atom_indices(), orc["sorted_to_initial"]
    MIXED, NONE = 0xFFFFFFFE, 0xFFFFFFFF

    def canon(labels, s2i, ext):
        # per input atom: the smallest input index with its label; -1 mixed, -2 buried, -3 exterior
        first = {}
        for a, l in zip(s2i.tolist(), labels.tolist()):
            first[l] = min(first.get(l, a), a)
        out = np.empty(len(labels), np.int64)
        out[s2i] = [-1 if l == MIXED else -2 if l == NONE else -3 if l == ext else first[l] for l in labels.tolist()]
        return out

    cg, co = canon(lg, s2i_g, eg), canon(lo, s2i_o, eo)
    print("components gpu", len(set(lg.tolist())), "oracle", len(set(lo.tolist())), "mixed", int((cg == -1).sum()),
          "exterior", int((cg == -3).sum()), flush=True)
    assert (cg == -1).sum() > 0 and (cg == -3).sum() > 0
    assert np.array_equal(cg, co)


def spatial_order(pts, edge=2.0):
    """Points sorted by a coarse cell key: waves of consecutive points are spatially compact."""
    key = np.floor((pts - pts.min(0)) / edge).astype(np.int64)
    return pts[np.lexsort((key[:, 0], key[:, 1], key[:, 2]))]


@pytest.mark.parametrize("exterior,order", [(False, "random"), (True, "random"), (False, "sorted"), (True, "sorted"),
                                            (False, "few"), (True, "few")])
def test_sdf_points_match_fp64(gpu, exterior, order):
    """40,000 points, random or spatial order: the library sorts them into cell order and waves
    evaluate cooperatively. 6,000 random points (below the sort threshold of 8,192): lanes of a
    wave evaluate alone."""
    atoms = synth.globule(3000, synth.SEED_BASE + 60)
    surf = Sesdf(atoms)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    band = 2.0
    pts = near_surface_points(atoms, orc, 6000 if order == "few" else 40000, band, 3)
    if order == "sorted":
        pts = spatial_order(pts)
    g = gpu_sdf(surf, pts, band, exterior)
    o = orc.sdf(pts.astype(np.float64), band=band, exterior=exterior, threads=8)
    err = np.abs(g.astype(np.float64) - o)
    print("max |gpu - fp64| =", err.max(), "inside band:", np.sum(np.abs(o) < band), flush=True)
    assert err.max() <= SDF_TOL


@pytest.mark.parametrize("exterior", [False, True])
def test_sdf_grid_64_config0(gpu, exterior):
    """The small workload: 2,000 atoms, SDF on a 64^3 vertex grid over the bounding box + 3 A."""
    torch = _torch()
    atoms = synth.globule(2000, synth.SEED_BASE)
    surf = Sesdf(atoms)
    lo = atoms[:, :3].min(0) - atoms[:, 3].max() - 3
    hi = atoms[:, :3].max(0) + atoms[:, 3].max() + 3
    dims = (64, 64, 64)
    spacing = (hi - lo) / 63
    out = torch.empty(64 ** 3, device="cuda")
    band = 3.0
    surf.sample_grid_ptr(lo, spacing, dims, out.data_ptr(), band=band, exterior=exterior)
    torch.cuda.synchronize()
    g = out.cpu().numpy()
    ii = np.arange(64 ** 3)
    # the grid points exactly as the kernel forms them (FP32 origin + spacing * index)
    pts = np.stack([np.float32(lo[0]) + np.float32(spacing[0]) * (ii % 64).astype(np.float32),
                    np.float32(lo[1]) + np.float32(spacing[1]) * ((ii // 64) % 64).astype(np.float32),
                    np.float32(lo[2]) + np.float32(spacing[2]) * (ii // 4096).astype(np.float32)], 1)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    o = orc.sdf(pts.astype(np.float64), band=band, exterior=exterior, threads=8)
    assert np.abs(g - o).max() <= SDF_TOL
    # invariants of the field: 1-Lipschitz along the grid axes, vdW lower bound
    f = g.reshape(64, 64, 64)
    assert np.abs(np.diff(f, axis=2)).max() <= spacing[0] * (1 + 1e-4) + 1e-5
    assert (g < 0).any() and (g > 0).any()


@pytest.mark.parametrize("store,count", [(True, 60000), (False, 60000), (False, 5000)])
def test_backward_matches_fp64_gradient(gpu, store, count):
    """Stored winners; recomputed ones (a cooperative forward pass into scratch codes: in cell order
    from 8,192 queries on, one wave per query below)."""
    torch = _torch()
    atoms = synth.globule(3000, synth.SEED_BASE + 61)
    surf = Sesdf(atoms)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    band = 2.0
    pts = near_surface_points(atoms, orc, count, band, 5)
    w = np.random.default_rng(7).standard_normal(len(pts)).astype(np.float32)
    P = torch.from_numpy(np.c_[pts, np.zeros(len(pts), np.float32)]).cuda()
    W = torch.from_numpy(w).cuda()
    val = torch.empty(len(pts), device="cuda")
    codes = torch.empty(len(pts), device="cuda", dtype=torch.int64)
    surf.sample_points_ptr(P.data_ptr(), len(pts), val.data_ptr(), band=band, d_codes=codes.data_ptr())
    grad = torch.zeros(len(atoms), 4, device="cuda")
    scratch = torch.empty(len(atoms), 4, device="cuda")
    surf.sdf_backward_ptr(P.data_ptr(), W.data_ptr(), len(pts), grad.data_ptr(), scratch.data_ptr(), band=band,
                          d_codes=codes.data_ptr() if store else None)
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    ref = orc.sdf_grad(pts.astype(np.float64), w.astype(np.float64), band=band)
    # per-atom relative error against the gradient scale; atoms whose winners flip between FP32 and
    # FP64 (boundary of an element's region) are counted
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max(1) / scale
    print("max rel err", err.max(), "p99", np.quantile(err, 0.99), flush=True)
    assert np.quantile(err, 0.999) <= GRAD_TOL
    assert np.mean(err > GRAD_TOL) < 1e-3


def test_torch_autograd_function(gpu):
    torch = _torch()
    from sesdf_amd.torch_sdf import SurfaceSDF

    atoms = synth.globule(1500, synth.SEED_BASE + 62)
    surf = Sesdf(atoms, build_surface=False)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    pts = near_surface_points(atoms, orc, 20000, 2.0, 9)
    A = torch.from_numpy(atoms).cuda().requires_grad_(True)
    P = torch.from_numpy(pts).cuda()
    w = torch.randn(len(pts), device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    out = SurfaceSDF.apply(A, P, surf, 2.0, True, False)
    (out * w).sum().backward()
    g_store = A.grad.detach().cpu().numpy()
    A.grad = None
    out2 = SurfaceSDF.apply(A, P, surf, 2.0, False, False)
    (out2 * w).sum().backward()
    g_rec = A.grad.detach().cpu().numpy()
    assert torch.equal(out, out2)
    np.testing.assert_allclose(g_store, g_rec, rtol=0, atol=1e-4 * np.abs(g_store).max())
    # release: the forward frees the whole build, the backward builds it again
    A.grad = None
    held = surf.counts()["device_bytes"]
    out3 = SurfaceSDF.apply(A, P, surf, 2.0, "release", False)
    assert surf.counts()["device_bytes"] == 0 < held
    (out3 * w).sum().backward()
    g_rel = A.grad.detach().cpu().numpy()
    assert torch.equal(out, out3)
    assert surf.counts()["device_bytes"] > 0
    # the rebuilt topology is the same, and the accumulation exact: the recompute gradient, bit for bit
    assert np.array_equal(g_rel, g_rec)
    ref = orc.sdf_grad(pts.astype(np.float64), w.cpu().numpy().astype(np.float64), band=2.0)
    assert np.abs(g_store - ref).max() <= GRAD_TOL * np.abs(ref).max() * 10
    # finite differences of the FP64 oracle with the topology held fixed, on a few atoms
    ids = np.arange(0, 1500, 150)
    fd, stable = orc.sdf_fd(pts.astype(np.float64), w.cpu().numpy().astype(np.float64), ids, h=1e-4, band=2.0)
    rel = np.abs(fd - g_store[ids]).max(1) / np.abs(ref).max()
    assert np.all(rel[stable] <= GRAD_TOL)


def test_sampling_after_enqueue_uses_the_new_build(gpu):
    """A build launched without readback (sesdf_enqueue) on new atoms in the input buffer: the next
    sampling sees the new surface, not the per-build data prepared for the previous one."""
    import ctypes

    import sesdf_amd._lib as L

    torch = _torch()
    atoms = synth.globule(1500, synth.SEED_BASE + 62)
    moved = atoms.copy()
    moved[:, :3] += np.array([0.37, -0.21, 0.55], np.float32)
    surf = Sesdf(atoms)
    lo = atoms[:, :3].min(0) - 6
    hi = atoms[:, :3].max(0) + 6
    dims = (40, 40, 40)
    sp = (hi - lo) / 39
    out = torch.empty(40 ** 3, device="cuda")
    surf.sample_grid_ptr(lo, sp, dims, out.data_ptr(), band=2.0)  # prepares data for the first build
    buf = ctypes.c_void_p()
    L.check(L.lib().sesdf_input_buffer(surf._h, ctypes.byref(buf)), "input_buffer")
    L.memcpy(buf, moved.ctypes.data, moved.nbytes, 1)  # host to device
    L.check(L.lib().sesdf_enqueue(surf._h), "enqueue")
    surf.sample_grid_ptr(lo, sp, dims, out.data_ptr(), band=2.0)
    torch.cuda.synchronize()
    L.check(L.lib().sesdf_sync(surf._h), "sync")
    fresh = Sesdf(moved)
    ref = torch.empty(40 ** 3, device="cuda")
    fresh.sample_grid_ptr(lo, sp, dims, ref.data_ptr(), band=2.0)
    torch.cuda.synchronize()
    # same surface; the two handles differ only in grid layout and sorted order (rounding)
    assert (out - ref).abs().max().item() <= 1e-5


def test_grid_of_more_than_2_31_bricks_is_refused(gpu):
    """A grid whose 4^3-vertex bricks number more than 2^31 (its output alone would need over 500 GB)
    is refused on the host before any launch; the output pointer is never touched."""
    from sesdf_amd import _lib as L

    atoms = synth.globule(500, synth.SEED_BASE)
    surf = Sesdf(atoms, probe_radius=1.4)
    dims = (65536, 65536, 4096)  # 16384 x 16384 x 1024 bricks = 2^38
    with pytest.raises(L.SesdfError, match="bricks"):
        surf.sample_grid_ptr((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), dims, 16, band=2.0)
    # a grid under the limit still samples
    torch = _torch()
    out = torch.empty(8 ** 3, device="cuda")
    surf.sample_grid_ptr(atoms[:, :3].min(0), (1.0, 1.0, 1.0), (8, 8, 8), out.data_ptr(), band=2.0)
    torch.cuda.synchronize()
    assert np.isfinite(out.cpu().numpy()).all()
