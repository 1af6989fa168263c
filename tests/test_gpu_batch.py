"""Batched builds: every molecule's slice of a batch equals its standalone build exactly; launch
tuning (cell edge, block shapes) does not change any output."""
import numpy as np
import pytest

import oracle as O
import parity
from sesdf_amd import Sesdf, Ssesdf, synth
from sesdf_amd.batch import SesdfBatch, SsesdfBatch, molecule_slice, set_tuning

pytestmark = pytest.mark.gpu


def small_batch(count=24, seed=synth.SEED_BASE + 3):
    sizes = synth.batch_sizes(count, 20, 1500, seed)
    atoms, offs = synth.batch(sizes, seed)
    return atoms, offs


def _eq(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    np.testing.assert_array_equal(a, b, err_msg=what)


def _slices_equal_standalone(atoms, offs, mode):
    cls_b, cls_s = (SesdfBatch, Sesdf) if mode == 0 else (SsesdfBatch, Ssesdf)
    print(f"[gpu] batch of {len(offs) - 1} molecules, {len(atoms)} atoms, mode {mode}", flush=True)
    batch = cls_b(atoms, offs)
    data, o = batch.get_data(), batch.offsets()
    s2i = batch.get_atom_indices()
    for m in range(len(offs) - 1):
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        single = cls_s(mol)
        sd = single.get_data()
        bm = molecule_slice(data, o, m)
        _eq(bm["atoms"], sd["atoms"], "atoms")
        _eq(s2i[int(o["atoms"][m]):int(o["atoms"][m + 1])] - int(offs[m]), single.get_atom_indices().astype(np.int64), "sorted_to_initial")
        _eq(bm["concave_patches_id"][:, :3], sd["concave_patches_id"][:, :3], "concave ids")
        _eq(bm["concave_patches_position"][:, :3], sd["concave_patches_position"][:, :3], "concave positions")
        if mode == 0:
            _eq(bm["convex_patches"].astype(np.int64), sd["convex_patches"].astype(np.int64), "convex patches")
            _eq(bm["sectors"], sd["sectors"], "sectors")
            _eq(bm["circle_patches"], sd["circle_patches"].astype(np.int64), "full circles")
            _eq(bm["segment_patches"], sd["segment_patches"].astype(np.int64), "segments")
        else:
            for k in ("circle_position", "circle_axis", "circle_vs"):
                _eq(bm[k], sd[k], k)
        single.close()
    return batch


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_slices_equal_standalone_builds(gpu, mode):
    atoms, offs = small_batch()
    _slices_equal_standalone(atoms, offs, mode)


def empty_and_tiny_batch():
    """Molecules of 60, 0, 1, 2, 0, 200 and 0 atoms."""
    g60 = synth.globule(60, synth.SEED_BASE + 35)
    g200 = synth.globule(200, synth.SEED_BASE + 36) + np.float32([40.0, 0.0, 0.0, 0.0])
    one = np.array([[5.0, 5.0, 5.0, 1.7]], np.float32)
    two = np.array([[0.0, -9.0, 0.0, 1.7], [3.0, -9.0, 0.5, 1.55]], np.float32)
    parts = [g60, np.zeros((0, 4), np.float32), one, two, np.zeros((0, 4), np.float32), g200, np.zeros((0, 4), np.float32)]
    offs = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)
    return np.concatenate(parts).astype(np.float32), offs


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_with_empty_and_tiny_molecules(gpu, mode):
    atoms, offs = empty_and_tiny_batch()
    batch = _slices_equal_standalone(atoms, offs, mode)
    if mode == 1:
        return
    import torch

    M, dim, margin, band = len(offs) - 1, 8, 3.0, 2.0
    out = torch.empty(M * dim ** 3, device="cuda")
    batch.sample_grids_ptr(dim, out.data_ptr(), margin=margin, band=band)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(M, -1)
    for m in range(M):
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        if len(mol) == 0:
            assert (got[m] == band).all()  # no surface: +band everywhere
            continue
        single = Sesdf(mol)
        lo = (mol[:, :3] - mol[:, 3:4]).min(0) - margin
        hi = (mol[:, :3] + mol[:, 3:4]).max(0) + margin
        sp = ((hi - lo) / (dim - 1)).astype(np.float32)
        ref = torch.empty(dim ** 3, device="cuda")
        single.sample_grid_ptr(lo.astype(np.float32), sp, (dim, dim, dim), ref.data_ptr(), band=band)
        torch.cuda.synchronize()
        np.testing.assert_allclose(got[m], ref.cpu().numpy(), atol=1e-5)
        single.close()


def test_batch_matches_oracle_per_molecule(gpu):
    """Each molecule of a small batch also matches the FP64 oracle (canonical comparison)."""
    atoms, offs = small_batch(6, synth.SEED_BASE + 31)
    batch = SesdfBatch(atoms, offs)
    data, o = batch.get_data(), batch.offsets()
    for m in range(len(offs) - 1):
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        single = Sesdf(mol)
        orc = O.build(mol, mode=0, precision=64, track=True, threads=8)
        rep = parity.compare(parity.gpu_arrays(single), orc, mol, 0)
        parity.assert_parity(rep)
        bm = molecule_slice(data, o, m)
        _eq(bm["concave_patches_position"], single.get_data()["concave_patches_position"], "batch vs single")


@pytest.mark.parametrize("cell,cw,iw", [(2.0, 2, 8), (4.0, 8, 2), (6.5, 4, 4)])
def test_tuning_does_not_change_outputs(gpu, cell, cw, iw):
    atoms, offs = small_batch(12, synth.SEED_BASE + 32)
    ref = SesdfBatch(atoms, offs)
    tuned = SesdfBatch(atoms, offs, build_surface=False)
    set_tuning(tuned, cell, cw, iw)
    tuned.build()
    a, b = ref.get_data(), tuned.get_data()
    # cell edges change the sorted order inside a molecule: compare canonically per molecule
    oa, ob = ref.offsets(), tuned.offsets()
    for m in range(len(offs) - 1):
        ma, mb = molecule_slice(a, oa, m), molecule_slice(b, ob, m)
        s2i_a = ref.get_atom_indices()[int(oa["atoms"][m]):int(oa["atoms"][m + 1])] - int(offs[m])
        s2i_b = tuned.get_atom_indices()[int(ob["atoms"][m]):int(ob["atoms"][m + 1])] - int(offs[m])
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        ka = {parity.concave_key(ma["concave_patches_id"][q], ma["concave_patches_position"][q], s2i_a, mol)
              for q in range(len(ma["concave_patches_id"]))}
        kb = {parity.concave_key(mb["concave_patches_id"][q], mb["concave_patches_position"][q], s2i_b, mol)
              for q in range(len(mb["concave_patches_id"]))}
        assert ka == kb
        assert len(ma["segment_patches"]) == len(mb["segment_patches"])
        assert len(ma["sectors"]) == len(mb["sectors"])


def test_batch_sdf_grids_equal_standalone(gpu):
    import torch

    atoms, offs = small_batch(8, synth.SEED_BASE + 33)
    batch = SesdfBatch(atoms, offs)
    dim, margin, band = 16, 3.0, 2.0
    out = torch.empty(8 * dim ** 3, device="cuda")
    batch.sample_grids_ptr(dim, out.data_ptr(), margin=margin, band=band)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(8, -1)
    for m in range(8):
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        single = Sesdf(mol)
        lo = (mol[:, :3] - mol[:, 3:4]).min(0) - margin
        hi = (mol[:, :3] + mol[:, 3:4]).max(0) + margin
        sp = ((hi - lo) / (dim - 1)).astype(np.float32)
        ref = torch.empty(dim ** 3, device="cuda")
        single.sample_grid_ptr(lo.astype(np.float32), sp, (dim, dim, dim), ref.data_ptr(), band=band)
        torch.cuda.synchronize()
        np.testing.assert_allclose(got[m], ref.cpu().numpy(), atol=1e-5)


def test_autotune_leaves_the_handle_built_and_outputs_unchanged(gpu, tmp_path, monkeypatch):
    """A tuned handle (fresh search, then from the cache) samples the same grid as an untuned one."""
    import torch

    import sesdf_amd.autotune as T

    monkeypatch.setattr(T, "_CACHE", str(tmp_path / "tuning.json"))
    atoms = synth.globule(1500, synth.SEED_BASE + 37)
    ref_s = Sesdf(atoms)
    lo = atoms[:, :3].min(0) - 4
    sp = ((atoms[:, :3].max(0) + 4 - lo) / 15).astype(np.float32)
    ref = torch.empty(16 ** 3, device="cuda")
    ref_s.sample_grid_ptr(lo.astype(np.float32), sp, (16, 16, 16), ref.data_ptr(), band=2.0)
    for cached in (False, True):
        s = Sesdf(atoms)
        t = T.autotune(s, reps=1)
        assert t["cached"] == cached
        got = torch.empty(16 ** 3, device="cuda")
        s.sample_grid_ptr(lo.astype(np.float32), sp, (16, 16, 16), got.data_ptr(), band=2.0)
        torch.cuda.synchronize()
        # another cell edge may settle near-ties between elements differently: last-bit differences
        np.testing.assert_allclose(got.cpu().numpy(), ref.cpu().numpy(), atol=1e-5)
        # the same probes, in input indices (the sorted order follows the cell edge)
        def triples(h):
            ids = h.get_data()["concave_patches_id"][:, :3].astype(np.int64)
            t = np.sort(h.get_atom_indices().astype(np.int64)[ids], axis=1)
            return t[np.lexsort(t.T[::-1])]
        _eq(triples(s), triples(ref_s), "concave triples")
        s.close()
