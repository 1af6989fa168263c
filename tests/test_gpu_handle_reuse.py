"""One handle, several builds: atoms moved far from the grid laid out at creation, and two forward
passes of the differentiable SDF on one handle before a single backward pass."""
import ctypes

import numpy as np
import pytest

import oracle as O
import parity
import sesdf_amd._lib as L
from sesdf_amd import Sesdf, synth

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def test_moved_atoms_lay_the_grid_out_again(gpu):
    """Atoms translated by (15, -12, 9) A and spread by 1.25 about their centre: most now lie
    outside the cell brick of the first build. The same handle must build the same surface as a
    fresh one (the grid is laid out again from the current atoms)."""
    atoms = synth.globule(2000, synth.SEED_BASE + 70)
    surf = Sesdf(atoms)
    c0 = surf.counts()
    assert c0["grid_relayout_nb"] == 0
    moved = atoms.copy()
    centre = moved[:, :3].mean(0)
    moved[:, :3] = (moved[:, :3] - centre) * 1.25 + centre + np.array([15.0, -12.0, 9.0], np.float32)
    moved = np.ascontiguousarray(moved, np.float32)
    print("moved-atoms case: 2000 atoms, shift (15, -12, 9), spread 1.25", flush=True)
    L.check(L.lib().sesdf_set_atoms(surf._h, ctypes.c_void_p(moved.ctypes.data)), "set_atoms")
    surf.build()
    c1 = surf.counts()
    assert c1["grid_relayout_nb"] == 1
    fresh = Sesdf(moved)
    cf = fresh.counts()
    for k in ("neighbor_nb", "sector_nb", "full_circle_nb", "concave_patch_nb", "segment_nb"):
        assert c1[k] == cf[k], k
    orc = O.build(moved, mode=0, precision=64, track=True, threads=8)
    parity.assert_parity(parity.compare(parity.gpu_arrays(surf), orc, moved, 0))
    # a second build on the same (moved) atoms keeps the new grid
    surf.build()
    assert surf.counts()["grid_relayout_nb"] == 0


def test_backward_after_two_forwards_on_one_handle(gpu):
    """Two conformers through one handle, one backward: each gradient equals the one computed alone
    on a handle of the same grid layout (the backward builds the forward's atoms again when the
    handle holds another build)."""
    torch = _torch()
    from sesdf_amd.torch_sdf import SurfaceSDF

    atoms = synth.globule(1500, synth.SEED_BASE + 71)
    other = atoms.copy()
    other[:, :3] += np.array([0.31, -0.17, 0.23], np.float32)
    rng = np.random.default_rng(5)
    lo, hi = atoms[:, :3].min(0) - 3, atoms[:, :3].max(0) + 3
    pts = torch.from_numpy((lo + (hi - lo) * rng.random((20000, 3))).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.standard_normal(20000).astype(np.float32)).cuda()
    for store in (True, False):
        surf = Sesdf(atoms, build_surface=False)
        A = torch.from_numpy(atoms).cuda().requires_grad_(True)
        B = torch.from_numpy(other).cuda().requires_grad_(True)
        loss = (SurfaceSDF.apply(A, pts, surf, 2.0, store, False) * w).sum() + \
               (SurfaceSDF.apply(B, pts, surf, 2.0, store, False) * w).sum()
        loss.backward()
        for X, ref_atoms in ((A, atoms), (B, other)):
            # a handle laid out from the first conformer, as the shared one is (the second lies in
            # its cell bricks): the same atoms head each circle and triple, so the per-point values
            # round alike, and the fixed-point accumulation makes the sums bit-identical
            alone = Sesdf(atoms, build_surface=False)
            Y = torch.from_numpy(ref_atoms).cuda().requires_grad_(True)
            (SurfaceSDF.apply(Y, pts, alone, 2.0, store, False) * w).sum().backward()
            g, want = X.grad.cpu().numpy(), Y.grad.cpu().numpy()
            print("store", store, "max |diff|", np.abs(g - want).max(), flush=True)
            # the accumulation is exact (fixed point), so the gradients are bit-identical
            assert np.array_equal(g, want)
