"""The allocation-free sampling entry point (sesdf_sample_points_noalloc) against
sesdf_sample_points_ex, through the C-ABI: two calls in a row, the second with fewer points, on the
complete and the exterior field; sizes beyond its maximum and a handle whose sampling state is not
in place are refused."""
import ctypes

import numpy as np
import pytest

import sesdf_amd._lib as L
from sesdf_amd import Sesdf, synth

pytestmark = pytest.mark.gpu

MAX_POINTS = 1048576  # SESDF_NOALLOC_MAX_POINTS
ERR_CAPACITY = 4      # SESDF_ERR_CAPACITY


def _torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def _points(torch, atoms, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = atoms[:, :3].min(0) - 3, atoms[:, :3].max(0) + 3
    q = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    return torch.from_numpy(np.c_[q, np.zeros(n, np.float32)]).cuda()


def _sample(fn, surf, P, n, flags, torch):
    out = torch.full((n,), float("nan"), device="cuda")
    codes = torch.full((n,), -1, device="cuda", dtype=torch.int64)
    st = fn(surf._h, ctypes.c_void_p(P.data_ptr()), n, 2.0, flags, ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(codes.data_ptr()))
    return st, out, codes


@pytest.mark.parametrize("flags", [0, 1])
def test_noalloc_matches_sample_points_ex(gpu, flags):
    torch = _torch()
    lib = L.lib()
    atoms = synth.globule(20000, synth.SEED_BASE + 80)
    surf = Sesdf(atoms)
    # two inputs, the second smaller; both sorted in the scratch array (8,192 points and more)
    sizes = (300_000, 50_000)
    inputs = [_points(torch, atoms, n, 40 + k) for k, n in enumerate(sizes)]
    refs = []
    for P, n in zip(inputs, sizes):
        st, out, codes = _sample(lib.sesdf_sample_points_ex, surf, P, n, flags, torch)
        L.check(st, "sesdf_sample_points_ex")
        refs.append((out, codes))
    torch.cuda.synchronize()
    held = surf.counts()["device_bytes"]
    for (P, n), (ref, ref_codes) in zip(zip(inputs, sizes), refs):
        print(f"noalloc call: flags {flags}, {n} points", flush=True)
        st, out, codes = _sample(lib.sesdf_sample_points_noalloc, surf, P, n, flags, torch)
        L.check(st, "sesdf_sample_points_noalloc")
        torch.cuda.synchronize()
        assert surf.counts()["device_bytes"] == held  # the handle allocated nothing
        assert not torch.isnan(out).any()
        # the same kernels on the same cell order: bit-identical values and winning elements
        assert torch.equal(out, ref) and torch.equal(codes, ref_codes)
    surf.close()


def test_noalloc_refusals(gpu):
    torch = _torch()
    lib = L.lib()
    atoms = synth.globule(3000, synth.SEED_BASE + 81)
    surf = Sesdf(atoms)
    P = _points(torch, atoms, 20_000, 50)
    held = surf.counts()["device_bytes"]
    # no sampling state yet for this build: refused, nothing allocated
    st, _, _ = _sample(lib.sesdf_sample_points_noalloc, surf, P, 20_000, 0, torch)
    assert st == ERR_CAPACITY and b"not in place" in lib.sesdf_last_error()
    assert surf.counts()["device_bytes"] == held
    st, ref, _ = _sample(lib.sesdf_sample_points_ex, surf, P, 20_000, 0, torch)
    L.check(st, "sesdf_sample_points_ex")
    # the exterior field needs its own state (the surface components)
    st, _, _ = _sample(lib.sesdf_sample_points_noalloc, surf, P, 20_000, 1, torch)
    assert st == ERR_CAPACITY
    # beyond the maximum: refused before any device work
    st = lib.sesdf_sample_points_noalloc(surf._h, ctypes.c_void_p(P.data_ptr()), MAX_POINTS + 1, 2.0, 0,
                                         ctypes.c_void_p(ref.data_ptr()), None)
    assert st == L.ERR_INVALID and b"SESDF_NOALLOC_MAX_POINTS" in lib.sesdf_last_error()
    st, out, _ = _sample(lib.sesdf_sample_points_noalloc, surf, P, 20_000, 0, torch)
    L.check(st, "sesdf_sample_points_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    surf.close()
