"""Python mirror of the reference host classes bcs::Sesdf and bcs::Ssesdf.

Constructor, ``build()``, ``get_data()`` and the analysis getters follow
sesdf/include/bcs/sesdf/sesdf.hpp:15-88 and sesdf/include/bcs/ssesdf/ssesdf.hpp:14-77 (same names in
snake_case, same argument meaning). Results live on the GPU; ``get_data()`` copies them to numpy
arrays (the reference returns device pointers; ``get_device_data()`` does that here).
Indices in the results are in grid-sorted atom order, as in the reference; ``get_atom_indices()``
maps sorted -> input order.
"""
import ctypes

import numpy as np

from . import _lib as L

MAX_INTERSECTIONS_PER_CIRCLE = 16  # sesdf/src/sesdf/sesdf.cu:17


def _np(ptr, count, dtype, cols=None):
    dt = np.dtype(dtype)
    out = np.empty(int(count) * (cols or 1), dt)
    if count:
        L.memcpy(out.ctypes.data, ptr, out.nbytes, L.D2H)
    return out.reshape(-1, cols) if cols else out


def get_aabb(atoms):
    """bcs::getAabb (sesdf/src/core/molecule.cpp:132-143): min/max of centre -/+ vdW radius."""
    a = np.asarray(atoms, np.float32).reshape(-1, 4)
    if a.shape[0] == 0:
        return np.zeros(3, np.float32), np.zeros(3, np.float32)
    return (a[:, :3] - a[:, 3:4]).min(0).astype(np.float32), (a[:, :3] + a[:, 3:4]).max(0).astype(np.float32)


class _Surface:
    MODE = L.MODE_COMPLETE

    def __init__(self, molecule, aabb=None, probe_radius=1.4, build_surface=True, graphics=False, *,
                 max_neighbors=128, max_intersection_neighbors=32, cell_size=0.0, device=0, stream=None,
                 use_graph=True, grow_neighbors=False):
        lib = L.lib()
        self._atoms = np.ascontiguousarray(np.asarray(molecule, np.float32).reshape(-1, 4))
        self.probe_radius = float(probe_radius)
        info = L.CreateInfo()
        info.xyzr = self._atoms.ctypes.data if self._atoms.size else None
        info.atom_nb = self._atoms.shape[0]
        info.probe_radius = self.probe_radius
        if aabb is not None:
            lo, hi = aabb
            for d in range(3):
                info.aabb[d] = float(lo[d])
                info.aabb[3 + d] = float(hi[d])
        info.mode = self.MODE
        info.max_neighbors = max_neighbors
        info.max_intersection_neighbors = max_intersection_neighbors
        info.cell_size = cell_size
        info.device = device
        info.stream = stream
        info.graphics = 1 if graphics else 0
        info.use_graph = 1 if use_graph else 0
        info.build_surface = 0
        h = ctypes.c_void_p()
        L.check(lib.sesdf_create(ctypes.byref(info), ctypes.byref(h)), "sesdf_create")
        self._h = h
        if grow_neighbors:
            # lists longer than max_neighbors: grow K and build again instead of truncating
            L.check(lib.sesdf_set_neighbor_policy(h, 1), "sesdf_set_neighbor_policy")
        if build_surface:
            self.build()

    # -- reference API --------------------------------------------------------------------------
    def build(self):
        """bcs::Sesdf::build (sesdf/src/sesdf/sesdf.cu:172-264)."""
        L.check(L.lib().sesdf_build(self._h), "sesdf_build")

    def build_device(self, d_xyzr):
        """Build from atoms already resident in device memory (pointer to float4 array)."""
        L.check(L.lib().sesdf_build_device(self._h, ctypes.c_void_p(d_xyzr)), "sesdf_build_device")

    def release(self):
        """Free every device buffer of the build (topology and sampling state); the next build
        allocates them again."""
        L.check(L.lib().sesdf_release(self._h), "sesdf_release")

    def get_max_neighbor_per_atom(self):
        return self.internal().max_neighbors

    def get_max_intersections_per_circles(self):
        return MAX_INTERSECTIONS_PER_CIRCLE

    def get_max_intersection_neighbors(self):
        return self.internal().max_intersection_neighbors

    def get_atom_indices(self):
        """bcs::Sesdf::getAtomIndices: sorted -> input permutation."""
        out = np.empty(self.atom_nb, np.uint32)
        if self.atom_nb:
            L.check(L.lib().sesdf_get_atom_indices(self._h, out.ctypes.data), "sesdf_get_atom_indices")
        return out

    # -- counts / internals ---------------------------------------------------------------------
    @property
    def atom_nb(self):
        return self._atoms.shape[0]

    def counts(self):
        c = L.Counts()
        L.check(L.lib().sesdf_get_counts(self._h, ctypes.byref(c)), "sesdf_get_counts")
        return c.as_dict()

    def internal(self):
        it = L.Internal()
        L.check(L.lib().sesdf_get_internal(self._h, ctypes.byref(it)), "sesdf_get_internal")
        return it

    def get_neighbors(self):
        """(counts[N], ids[N, K]) -- getDAtomNeighborsCount / getDAtomNeighborsIndices."""
        it = self.internal()
        n, K = self.atom_nb, it.max_neighbors
        return _np(it.neighbor_count, n, np.uint32), _np(it.neighbor_ids, n * K, np.uint32).reshape(n, K)

    def get_circle_status(self):
        """getDCircleVisibilityStatus: status[N, K] (0 hidden, 1 intersected, 2 full)."""
        it = self.internal()
        n, K = self.atom_nb, it.max_neighbors
        return _np(it.circle_status, n * K, np.uint8).reshape(n, K)

    def stream(self):
        s = ctypes.c_void_p()
        L.check(L.lib().sesdf_get_stream(self._h, ctypes.byref(s)), "sesdf_get_stream")
        return s.value

    def close(self):
        if getattr(self, "_h", None):
            try:
                L.lib().sesdf_destroy(self._h)
            except (TypeError, AttributeError):  # interpreter shutdown: the library is already gone
                pass
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Sesdf(_Surface):
    """Complete surface (bcs::Sesdf): convex sectors, segments, full circles, concave patches."""
    MODE = L.MODE_COMPLETE

    def get_device_data(self):
        d = L.SesdfData()
        L.check(L.lib().sesdf_get_data(self._h, ctypes.byref(d)), "sesdf_get_data")
        return d

    def get_data(self):
        """bcs::Sesdf::getData as host numpy arrays."""
        d = self.get_device_data()
        iw = self.counts()["concave_with_neighbors_nb"]
        return {
            "atoms": _np(d.atoms, d.atom_nb, np.float32, 4),
            "convex_patches": _np(d.convex_patches, d.convex_patch_nb, np.uint32, 2),
            "sectors": _np(d.sectors, d.sector_nb, np.float32, 4),
            "segment_patches": _np(d.segment_patches, d.segment_patch_nb, np.uint32, 4),
            "circle_patches": _np(d.circle_patches, d.circle_patch_nb, np.uint32, 2),
            "concave_patches_position": _np(d.concave_patches_position, d.concave_patch_nb, np.float32, 4),
            "concave_patches_id": _np(d.concave_patches_id, d.concave_patch_nb, np.int32, 4),
            "concave_patches_neighbors": _np(d.concave_patches_neighbors, iw * d.max_concave_neighbors, np.float32, 4),
            "probe_radius": d.probe_radius,
            "max_concave_neighbors": d.max_concave_neighbors,
        }

    # -- SDF (no reference counterpart; see include/sesdf_c.h) --------------------------------------
    def sample_points_ptr(self, d_points, n, d_out, band=2.0, exterior=False, d_codes=None):
        """SDF at n device points (float4 xyz_) into d_out (float); optional winner codes (uint64)."""
        L.check(L.lib().sesdf_sample_points_ex(self._h, ctypes.c_void_p(d_points), n, band, 1 if exterior else 0,
                                               ctypes.c_void_p(d_out), ctypes.c_void_p(d_codes) if d_codes else None),
                "sesdf_sample_points")

    def sample_grid_ptr(self, origin, spacing, dims, d_out, band=2.0, exterior=False):
        g = L.GridDesc()
        for d in range(3):
            g.origin[d], g.spacing[d], g.dims[d] = float(origin[d]), float(spacing[d]), int(dims[d])
        L.check(L.lib().sesdf_sample_grid_ex(self._h, ctypes.byref(g), band, 1 if exterior else 0, ctypes.c_void_p(d_out)),
                "sesdf_sample_grid")

    def sdf_backward_ptr(self, d_points, d_weights, n, d_grad, d_scratch, band=2.0, exterior=False, d_codes=None):
        """Gradient of sum w_i SDF(q_i) per input atom (float4) into d_grad; d_scratch: float4 * atoms."""
        L.check(L.lib().sesdf_sdf_backward_ex(self._h, ctypes.c_void_p(d_points), ctypes.c_void_p(d_weights), n, band,
                                              1 if exterior else 0, ctypes.c_void_p(d_codes) if d_codes else None,
                                              ctypes.c_void_p(d_scratch), ctypes.c_void_p(d_grad)), "sesdf_sdf_backward")

    def surface_components(self):
        """(labels per sorted atom, exterior label) of the surface components used by the exterior SDF."""
        labels = np.empty(self.atom_nb, np.uint32)
        ext = ctypes.c_uint32()
        L.check(L.lib().sesdf_surface_components(self._h, labels.ctypes.data, ctypes.byref(ext)), "sesdf_surface_components")
        return labels, int(ext.value)

    def get_trimmed(self):
        it = self.internal()
        ci = self.counts()["intersected_circle_nb"]
        return _np(it.trimmed_to_global, ci, np.uint32), _np(it.circle_intersection_nb, ci, np.uint32)


class Ssesdf(_Surface):
    """Exterior path (bcs::Ssesdf): torus records and concave patches."""
    MODE = L.MODE_EXTERIOR

    def get_device_data(self):
        d = L.SsesdfData()
        L.check(L.lib().sesdf_get_exterior_data(self._h, ctypes.byref(d)), "sesdf_get_exterior_data")
        return d

    def get_data(self):
        """bcs::Ssesdf::getData as host numpy arrays."""
        d = self.get_device_data()
        iw = self.counts()["concave_with_neighbors_nb"]
        return {
            "atoms": _np(d.atoms, d.atom_nb, np.float32, 4),
            "circle_position": _np(d.circle_position, d.circle_patch_nb, np.float32, 4),
            "circle_axis": _np(d.circle_axis, d.circle_patch_nb, np.float32, 4),
            "circle_vs": _np(d.circle_vs, d.circle_patch_nb, np.float32, 4),
            "concave_patches_position": _np(d.concave_patches_position, d.concave_patch_nb, np.float32, 4),
            "concave_patches_id": _np(d.concave_patches_id, d.concave_patch_nb, np.int32, 4),
            "concave_patches_neighbors": _np(d.concave_patches_neighbors, iw * d.max_concave_neighbors, np.float32, 4),
            "probe_radius": d.probe_radius,
            "max_concave_neighbors": d.max_concave_neighbors,
        }
