"""Boundary pieces beyond build/getData: the reference's globalToTrimmed map and the render-facing
buffer layouts of getGraphics (sesdf/src/sesdf/sesdf.cu:266-299, sesdf/src/ssesdf/ssesdf.cu:189-210)."""
import ctypes

import numpy as np
import pytest

import sesdf_amd._lib as L
from sesdf_amd import Sesdf, Ssesdf, synth

pytestmark = pytest.mark.gpu


class Span(ctypes.Structure):
    _fields_ = [("buffer", ctypes.c_uint32), ("offset", ctypes.c_uint64), ("size", ctypes.c_uint64)]


class Layout(ctypes.Structure):
    _fields_ = [("buffer_nb", ctypes.c_uint32), ("buffer_bytes", ctypes.c_uint64 * 5), ("spans", Span * 8)]


def _layout(surf):
    lay = Layout()
    L.check(L.lib().sesdf_get_graphics_layout(surf._h, ctypes.byref(lay)), "layout")
    return lay


def test_global_to_trimmed(gpu):
    atoms = synth.globule(3000, synth.SEED_BASE + 80)
    surf = Sesdf(atoms)
    d = ctypes.c_void_p()
    cnt = ctypes.c_uint64()
    L.check(L.lib().sesdf_get_global_to_trimmed(surf._h, ctypes.byref(d), ctypes.byref(cnt)), "g2t")
    K = surf.get_max_neighbor_per_atom()
    assert cnt.value == 3000 * K + 1
    g2t = np.empty(cnt.value, np.uint32)
    L.memcpy(g2t.ctypes.data, d.value, g2t.nbytes, L.D2H)
    st = surf.get_circle_status().reshape(-1)
    cnt_nb, ids = surf.get_neighbors()
    slots = np.arange(3000 * K)
    i, l = slots // K, slots % K
    mask = (l < cnt_nb[i]) & (st == 1) & (ids.reshape(-1) > i)
    want = np.concatenate([[0], np.cumsum(mask)])  # exclusive prefix over N * K + 1 slots
    assert np.array_equal(g2t, want)
    # exterior handles have no trimmed ids
    ext = Ssesdf(atoms)
    assert L.lib().sesdf_get_global_to_trimmed(ext._h, ctypes.byref(d), ctypes.byref(cnt)) == 3  # UNSUPPORTED


@pytest.mark.parametrize("cls", [Sesdf, Ssesdf])
def test_graphics_layout_and_pack(gpu, cls):
    import torch

    atoms = synth.globule(2001, synth.SEED_BASE + 81)
    surf = cls(atoms)
    c = surf.counts()
    lay = _layout(surf)
    N, I, Iw = c["atom_nb"], c["concave_patch_nb"], c["concave_with_neighbors_nb"]
    nb = Iw * 32 * 16
    bufs = [torch.full((int(lay.buffer_bytes[b]) + 16,), 0xAB, dtype=torch.uint8, device="cuda") for b in range(lay.buffer_nb)]
    ptrs = (ctypes.c_void_p * 5)(*[b.data_ptr() for b in bufs])
    L.check(L.lib().sesdf_pack_graphics(surf._h, ptrs), "pack")
    torch.cuda.synchronize()
    host = [b.cpu().numpy() for b in bufs]
    data = surf.get_data()

    def field(q, dtype, cols):
        sp = lay.spans[q]
        raw = host[sp.buffer][sp.offset: sp.offset + sp.size]
        return raw.view(dtype).reshape(-1, cols)

    if cls is Sesdf:
        Cf, V, G = c["full_circle_nb"], c["sector_nb"], c["segment_nb"]
        assert lay.buffer_nb == 5
        sector_offset = Cf * 8 + (Cf * 8) % 16  # sesdf.cu:276-279
        assert [(s.buffer, s.offset, s.size) for s in lay.spans] == [
            (0, 0, N * 16), (1, 0, N * 8), (2, sector_offset, V * 16), (3, 0, G * 16), (2, 0, Cf * 8),
            (4, I * 16, I * 16), (4, 0, I * 16), (4, I * 32, nb)]
        assert np.array_equal(field(2, np.float32, 4), data["sectors"])
        assert np.array_equal(field(3, np.uint32, 4), data["segment_patches"])
        assert np.array_equal(field(4, np.uint32, 2), data["circle_patches"])
        assert np.array_equal(field(1, np.uint32, 2), data["convex_patches"])
        assert Cf % 2 == 0 or (host[2][Cf * 8: sector_offset] == 0).all()
    else:
        T = c["torus_nb"]
        assert lay.buffer_nb == 3
        assert [(s.buffer, s.offset, s.size) for s in lay.spans[:7]] == [
            (0, 0, N * 16), (1, 0, T * 16), (1, T * 16, T * 16), (1, T * 32, T * 16), (2, 0, I * 16), (2, I * 16, I * 16),
            (2, I * 32, nb)]
        assert np.array_equal(field(1, np.float32, 4), data["circle_position"])
        assert np.array_equal(field(3, np.float32, 4), data["circle_vs"])
    pos_q, id_q = (5, 6) if cls is Sesdf else (4, 5)
    assert np.array_equal(field(0, np.float32, 4), data["atoms"])
    assert np.array_equal(field(pos_q, np.float32, 4), data["concave_patches_position"])
    assert np.array_equal(field(id_q, np.int32, 4), data["concave_patches_id"])
