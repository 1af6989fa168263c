"""Python mirror of the reference's Contour-Buildup comparator, bcs::ContourBuildup.

Constructor, ``build()`` and ``get_graphics()`` follow
sesdf/include/bcs/contour_buildup/contour_buildup.hpp:24-43 (same names in snake_case, same
argument meaning: molecule, aabb, probe radius, grid dimensions, build surface). The OpenGL buffers
of the reference are device buffers of the handle; ``get_data()`` copies them to numpy arrays,
``get_device_data()`` returns the device pointers. Atom indices are in grid-sorted order, as in the
reference; ``get_internal()["sorted_to_initial"]`` maps them to input order.
"""
import ctypes

import numpy as np

from . import _lib as L
from .surface import _np


class ContourBuildup:
    def __init__(self, molecule, aabb=None, probe_radius=1.4, grid_dimensions=16, build_surface=True, *,
                 tex_size=32768, device=0, stream=None, use_graph=True):
        lib = L.lib()
        self._atoms = np.ascontiguousarray(np.asarray(molecule, np.float32).reshape(-1, 4))
        info = L.CbCreateInfo()
        info.xyzr = self._atoms.ctypes.data if self._atoms.size else None
        info.atom_nb = self._atoms.shape[0]
        info.probe_radius = float(probe_radius)
        if aabb is not None:
            lo, hi = aabb
            for d in range(3):
                info.aabb[d] = float(lo[d])
                info.aabb[3 + d] = float(hi[d])
        info.grid_dimensions = int(grid_dimensions)
        info.tex_size = int(tex_size)
        info.device = device
        info.stream = stream
        info.use_graph = 1 if use_graph else 0
        info.build_surface = 0
        h = ctypes.c_void_p()
        L.check(lib.sesdf_cb_create(ctypes.byref(info), ctypes.byref(h)), "sesdf_cb_create")
        self._h = h
        if build_surface:
            self.build()

    @property
    def atom_nb(self):
        return self._atoms.shape[0]

    def build(self):
        """bcs::ContourBuildup::build (sesdf/src/contour_buildup/contour_buildup.cu:400-660)."""
        L.check(L.lib().sesdf_cb_build(self._h), "sesdf_cb_build")

    def set_atoms(self, molecule):
        """Replace the atom span read by the next build (same atom count); the grid keeps the box of
        construction, as the reference's does."""
        a = np.ascontiguousarray(np.asarray(molecule, np.float32).reshape(-1, 4))
        if a.shape != self._atoms.shape:
            raise ValueError("set_atoms needs the same atom count")
        self._atoms = a
        L.check(L.lib().sesdf_cb_set_atoms(self._h, a.ctypes.data if a.size else None), "sesdf_cb_set_atoms")

    def build_device(self, d_xyzr=None):
        """Build from atoms already resident in device memory (float4 array; None: the input buffer)."""
        L.check(L.lib().sesdf_cb_build_device(self._h, ctypes.c_void_p(d_xyzr) if d_xyzr else None),
                "sesdf_cb_build_device")

    def enqueue(self):
        L.check(L.lib().sesdf_cb_enqueue(self._h), "sesdf_cb_enqueue")

    def sync(self):
        L.check(L.lib().sesdf_cb_sync(self._h), "sesdf_cb_sync")

    def input_buffer(self):
        p = ctypes.c_void_p()
        L.check(L.lib().sesdf_cb_input_buffer(self._h, ctypes.byref(p)), "sesdf_cb_input_buffer")
        return p.value

    def stream(self):
        s = ctypes.c_void_p()
        L.check(L.lib().sesdf_cb_get_stream(self._h, ctypes.byref(s)), "sesdf_cb_get_stream")
        return s.value

    def set_tuning(self, arc_capacity=0, probe_capacity=0, torus_capacity=0, singularity_search=0):
        """Capacities of the next build (arcs per circle on chip: 8 or 80; output slots) and the
        singularity texture's neighbour search (0 by cell size, 1 the reference's stencil scan,
        2 a grid of 2 rp cells; the texture is the same)."""
        t = L.CbTuning(arc_capacity, probe_capacity, torus_capacity, singularity_search)
        L.check(L.lib().sesdf_cb_set_tuning(self._h, ctypes.byref(t)), "sesdf_cb_set_tuning")

    def counts(self):
        c = L.CbCounts()
        L.check(L.lib().sesdf_cb_get_counts(self._h, ctypes.byref(c)), "sesdf_cb_get_counts")
        return c.as_dict()

    def get_graphics(self):
        """bcs::ContourBuildup::getGraphics (contour_buildup.cu:662-665): the counts of
        ContourBuildupGraphics; the vertex arrays and the texture are 0 (no OpenGL here)."""
        c = self.counts()
        return {"atomNb": c["atom_nb"], "atomVao": 0, "sphericalTriangleNb": c["probe_nb"],
                "sphericalTriangleVao": 0, "torusNb": c["torus_nb"], "torusVao": 0, "singularityTexture": 0}

    def get_device_data(self):
        d = L.CbData()
        L.check(L.lib().sesdf_cb_get_data(self._h, ctypes.byref(d)), "sesdf_cb_get_data")
        return d

    def get_data(self):
        """The buffers behind getGraphics() as host numpy arrays."""
        d = self.get_device_data()
        return {
            "atoms": _np(d.atoms, d.atom_nb, np.float32, 4),
            "probe_positions": _np(d.probe_positions, d.probe_nb, np.float32, 4),
            "sphere_triangle_vec1": _np(d.sphere_triangle_vec1, d.probe_nb, np.float32, 4),
            "sphere_triangle_vec2": _np(d.sphere_triangle_vec2, d.probe_nb, np.float32, 4),
            "sphere_triangle_vec3": _np(d.sphere_triangle_vec3, d.probe_nb, np.float32, 4),
            "torus_positions": _np(d.torus_positions, d.torus_nb, np.float32, 4),
            "torus_axes": _np(d.torus_axes, d.torus_nb, np.float32, 4),
            "torus_visibility_spheres": _np(d.torus_visibility_spheres, d.torus_nb, np.float32, 4),
            "tex_coords": _np(d.tex_coords, d.probe_nb, np.float32, 3),
            "singularity_texture": _np(d.singularity_texture, d.probe_nb, np.float32, d.probe_neighbor_nb * 3),
            "probe_radius": d.probe_radius,
            "tex_size": d.tex_size,
        }

    def get_internal(self):
        """Neighbour lists, sorted order and the atom triples / pairs behind probes and tori."""
        it = L.CbInternal()
        L.check(L.lib().sesdf_cb_get_internal(self._h, ctypes.byref(it)), "sesdf_cb_get_internal")
        c = self.counts()
        n, K = self.atom_nb, it.max_neighbors
        return {
            "sorted_to_initial": _np(it.sorted_to_initial, n, np.uint32),
            "neighbor_count": _np(it.neighbor_count, n, np.uint32),
            "neighbor_ids": _np(it.neighbor_ids, n * K, np.uint32).reshape(n, K),
            "probe_atoms": _np(it.probe_atoms, c["probe_nb"], np.uint32, 4),
            "torus_atoms": _np(it.torus_atoms, c["torus_nb"], np.uint32, 2),
            "probe_sorted": _np(it.probe_sorted, c["probe_nb"], np.uint32),
        }

    def close(self):
        if getattr(self, "_h", None):
            try:
                L.lib().sesdf_cb_destroy(self._h)
            except (TypeError, AttributeError):  # interpreter shutdown
                pass
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
