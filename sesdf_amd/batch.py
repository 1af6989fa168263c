"""Batched builds: many independent molecules in one launch sequence (no reference counterpart;
the reference builds one molecule per handle, app/benchmarks/main.cu:174-186).

Molecules are concatenated molecule-major. Every molecule gets its own grid brick (cells are
indexed per molecule, so neighbour searches never cross molecules) and the whole stage sequence
runs once for the batch, replayed as one hipGraph. Outputs are the concatenation; ``offsets()``
returns per-molecule ranges into every output array. Atom indices in the outputs are global
(grid-sorted) indices: molecule m owns sorted atoms [offsets['atoms'][m], offsets['atoms'][m+1]).
"""
import ctypes

import numpy as np

from . import _lib as L
from .surface import Sesdf, Ssesdf

OFFSET_KEYS = ("atoms", "sectors", "full_circles", "concave_with_neighbors", "concave_without_neighbors", "segments",
               "torus")


def _create(self, atoms, offsets, probe_radius, mode, max_neighbors, max_intersection_neighbors, cell_size, device,
            stream, use_graph):
    lib = L.lib()
    self._atoms = np.ascontiguousarray(np.asarray(atoms, np.float32).reshape(-1, 4))
    self._offsets = np.ascontiguousarray(np.asarray(offsets, np.uint64))
    self.molecule_nb = len(self._offsets) - 1
    self.probe_radius = float(probe_radius)
    info = L.BatchInfo()
    info.xyzr = self._atoms.ctypes.data if self._atoms.size else None
    info.atom_offsets = self._offsets.ctypes.data
    info.molecule_nb = self.molecule_nb
    info.probe_radius = self.probe_radius
    info.mode = mode
    info.max_neighbors = max_neighbors
    info.max_intersection_neighbors = max_intersection_neighbors
    info.cell_size = cell_size
    info.device = device
    info.stream = stream
    info.use_graph = 1 if use_graph else 0
    h = ctypes.c_void_p()
    L.check(lib.sesdf_batch_create(ctypes.byref(info), ctypes.byref(h)), "sesdf_batch_create")
    self._h = h


class _BatchMixin:
    def offsets(self):
        """Per-molecule prefix offsets (molecule_nb + 1) into each concatenated output."""
        out = np.zeros((len(OFFSET_KEYS), self.molecule_nb + 1), np.uint64)
        L.check(L.lib().sesdf_batch_offsets(self._h, out.ctypes.data), "sesdf_batch_offsets")
        return dict(zip(OFFSET_KEYS, out))

    def set_tuning(self, cell_size=0.0, circle_waves=0, intersection_waves=0):
        set_tuning(self, cell_size, circle_waves, intersection_waves)


def set_tuning(surf, cell_size=0.0, circle_waves=0, intersection_waves=0):
    """Apply launch tuning (grid cell edge, waves per workgroup) to any handle."""
    t = Tuning(cell_size, circle_waves, intersection_waves)
    L.check(L.lib().sesdf_set_tuning(surf._h, ctypes.cast(ctypes.pointer(t), ctypes.c_void_p)), "sesdf_set_tuning")


class Tuning(ctypes.Structure):
    _fields_ = [("cell_size", ctypes.c_float), ("circle_waves", ctypes.c_uint32), ("intersection_waves", ctypes.c_uint32)]


class SesdfBatch(_BatchMixin, Sesdf):
    """Complete surfaces of a batch of molecules."""

    def __init__(self, atoms, offsets, probe_radius=1.4, build_surface=True, *, max_neighbors=128,
                 max_intersection_neighbors=32, cell_size=0.0, device=0, stream=None, use_graph=True):
        _create(self, atoms, offsets, probe_radius, L.MODE_COMPLETE, max_neighbors, max_intersection_neighbors,
                cell_size, device, stream, use_graph)
        if build_surface:
            self.build()

    def sample_grids_ptr(self, dim, d_out, margin=3.0, band=2.0):
        """One dim^3 SDF grid per molecule (bounding box + margin), molecule-major, into d_out."""
        L.check(L.lib().sesdf_batch_sample_grids(self._h, dim, margin, band, ctypes.c_void_p(d_out)),
                "sesdf_batch_sample_grids")


class SsesdfBatch(_BatchMixin, Ssesdf):
    """Exterior-path surfaces (torus records + concave patches) of a batch of molecules."""

    def __init__(self, atoms, offsets, probe_radius=1.4, build_surface=True, *, max_neighbors=128,
                 max_intersection_neighbors=32, cell_size=0.0, device=0, stream=None, use_graph=True):
        _create(self, atoms, offsets, probe_radius, L.MODE_EXTERIOR, max_neighbors, max_intersection_neighbors,
                cell_size, device, stream, use_graph)
        if build_surface:
            self.build()


def molecule_slice(data, offs, m):
    """The outputs of molecule m of a complete-surface batch, with atom indices made molecule-local."""
    a0 = int(offs["atoms"][m])
    iw = int(offs["concave_with_neighbors"][-1])
    w0, w1 = int(offs["concave_with_neighbors"][m]), int(offs["concave_with_neighbors"][m + 1])
    n0, n1 = int(offs["concave_without_neighbors"][m]), int(offs["concave_without_neighbors"][m + 1])
    conc = np.r_[w0:w1, iw + n0:iw + n1]
    remap = -np.ones(len(data["concave_patches_id"]) + 1, np.int64)
    remap[conc] = np.arange(len(conc))
    out = {"atoms": data["atoms"][a0:int(offs["atoms"][m + 1])]}
    ids = data["concave_patches_id"][conc].copy()
    ids[:, :3] -= a0
    out["concave_patches_id"] = ids
    out["concave_patches_position"] = data["concave_patches_position"][conc]
    maxnb = int(data["max_concave_neighbors"])
    out["concave_patches_neighbors"] = data["concave_patches_neighbors"][w0 * maxnb:w1 * maxnb]
    out["concave_with_neighbors_nb"] = w1 - w0
    if "sectors" in data:
        out["convex_patches"] = data["convex_patches"][a0:int(offs["atoms"][m + 1])] - int(offs["sectors"][m])
        out["sectors"] = data["sectors"][int(offs["sectors"][m]):int(offs["sectors"][m + 1])]
        fc = data["circle_patches"][int(offs["full_circles"][m]):int(offs["full_circles"][m + 1])].astype(np.int64)
        out["circle_patches"] = fc - a0
        seg = data["segment_patches"][int(offs["segments"][m]):int(offs["segments"][m + 1])].astype(np.int64)
        seg[:, :2] -= a0
        seg[:, 2] = remap[seg[:, 2]]
        seg[:, 3] = remap[seg[:, 3]]
        out["segment_patches"] = seg
    else:
        t0, t1 = int(offs["torus"][m]), int(offs["torus"][m + 1])
        for k in ("circle_position", "circle_axis", "circle_vs"):
            out[k] = data[k][t0:t1]
    return out
