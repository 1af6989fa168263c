"""ctypes declarations of the C ABI in include/sesdf_c.h (libsesdf_hip.so).

The HIP library is mandatory: importing the product API on a machine where it is not built raises
immediately; there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsesdf_hip.so")

u8, u16, u32, u64, i32, f32 = (ctypes.c_uint8, ctypes.c_uint16, ctypes.c_uint32, ctypes.c_uint64,
                               ctypes.c_int32, ctypes.c_float)
vp = ctypes.c_void_p
fp = ctypes.POINTER(ctypes.c_float)

SESDF_OK = 0
ERR_INVALID = 1  # SESDF_ERR_INVALID
MODE_COMPLETE = 0
MODE_EXTERIOR = 1


class CreateInfo(ctypes.Structure):
    _fields_ = [("xyzr", vp), ("atom_nb", u32), ("probe_radius", f32), ("aabb", f32 * 6), ("mode", i32),
                ("max_neighbors", u32), ("max_intersection_neighbors", u32), ("cell_size", f32), ("device", i32),
                ("stream", vp), ("graphics", i32), ("use_graph", i32), ("build_surface", i32)]


class SesdfData(ctypes.Structure):
    _fields_ = [("atom_nb", u32), ("atoms", vp), ("convex_patch_nb", u32), ("convex_patches", vp),
                ("sector_nb", u32), ("sectors", vp), ("segment_patch_nb", u32), ("segment_patches", vp),
                ("circle_patch_nb", u32), ("circle_patches", vp), ("concave_patch_nb", u32),
                ("concave_patches_position", vp), ("concave_patches_id", vp), ("concave_patches_neighbors", vp),
                ("probe_radius", f32), ("max_concave_neighbors", u16)]


class SsesdfData(ctypes.Structure):
    _fields_ = [("atom_nb", u32), ("atoms", vp), ("circle_patch_nb", u32), ("circle_position", vp),
                ("circle_axis", vp), ("circle_vs", vp), ("concave_patch_nb", u32), ("concave_patches_position", vp),
                ("concave_patches_id", vp), ("concave_patches_neighbors", vp), ("probe_radius", f32),
                ("max_concave_neighbors", u16)]


class Counts(ctypes.Structure):
    _fields_ = [("atom_nb", u32), ("neighbor_nb", u32), ("sector_nb", u32), ("intersected_circle_nb", u32),
                ("full_circle_nb", u32), ("visible_circle_nb", u32), ("concave_patch_nb", u32),
                ("concave_with_neighbors_nb", u32), ("segment_nb", u32), ("torus_nb", u32),
                ("neighbor_overflow_atoms", u32), ("probe_neighbor_overflow", u32),
                ("circle_intersection_overflow", u32), ("jk_anomaly_nb", u32), ("fp64_resolve_nb", u32),
                ("rebuild_nb", u32), ("max_atom_records", u32), ("device_bytes", u64), ("build_serial", u64),
                ("grid_relayout_nb", u32), ("component_overflow_atoms", u32), ("max_neighbors", u32),
                ("neighbor_growth_nb", u32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Internal(ctypes.Structure):
    _fields_ = [("sorted_to_initial", vp), ("neighbor_count", vp), ("neighbor_ids", vp), ("circle_status", vp),
                ("trimmed_to_global", vp), ("circle_intersection_nb", vp), ("max_neighbors", u32),
                ("max_intersections_per_circle", u16), ("max_intersection_neighbors", u16)]


class GridDesc(ctypes.Structure):
    _fields_ = [("origin", f32 * 3), ("spacing", f32 * 3), ("dims", u32 * 3)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("xyzr", vp), ("atom_offsets", vp), ("molecule_nb", u32), ("probe_radius", f32), ("mode", i32),
                ("max_neighbors", u32), ("max_intersection_neighbors", u32), ("cell_size", f32), ("device", i32),
                ("stream", vp), ("use_graph", i32)]


class CbCreateInfo(ctypes.Structure):
    _fields_ = [("xyzr", vp), ("atom_nb", u32), ("probe_radius", f32), ("aabb", f32 * 6), ("grid_dimensions", u32),
                ("tex_size", u32), ("device", i32), ("stream", vp), ("use_graph", i32), ("build_surface", i32)]


class CbData(ctypes.Structure):
    _fields_ = [("atom_nb", u32), ("atoms", vp), ("probe_nb", u32), ("probe_positions", vp),
                ("sphere_triangle_vec1", vp), ("sphere_triangle_vec2", vp), ("sphere_triangle_vec3", vp),
                ("torus_nb", u32), ("torus_positions", vp), ("torus_axes", vp), ("torus_visibility_spheres", vp),
                ("tex_coords", vp), ("singularity_texture", vp), ("probe_neighbor_nb", u32), ("tex_size", u32),
                ("probe_radius", f32)]


class CbCounts(ctypes.Structure):
    _fields_ = [("atom_nb", u32), ("probe_nb", u32), ("torus_nb", u32), ("neighbor_overflow_atoms", u32),
                ("no_contour_atoms", u32), ("arc_capacity", u32), ("build_attempts", u32), ("grid_dimensions", u32),
                ("singularity_grid", u32), ("device_bytes", u64), ("build_serial", u64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class CbInternal(ctypes.Structure):
    _fields_ = [("sorted_to_initial", vp), ("neighbor_count", vp), ("neighbor_ids", vp), ("max_neighbors", u32),
                ("probe_atoms", vp), ("torus_atoms", vp), ("probe_sorted", vp)]


class CbTuning(ctypes.Structure):
    _fields_ = [("arc_capacity", u32), ("probe_capacity", u32), ("torus_capacity", u32), ("singularity_search", u32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "sesdf_cb_create": (i32, [ctypes.POINTER(CbCreateInfo), ctypes.POINTER(vp)]),
    "sesdf_cb_destroy": (i32, [vp]),
    "sesdf_cb_build": (i32, [vp]),
    "sesdf_cb_build_device": (i32, [vp, vp]),
    "sesdf_cb_set_atoms": (i32, [vp, vp]),
    "sesdf_cb_input_buffer": (i32, [vp, ctypes.POINTER(vp)]),
    "sesdf_cb_enqueue": (i32, [vp]),
    "sesdf_cb_sync": (i32, [vp]),
    "sesdf_cb_get_stream": (i32, [vp, ctypes.POINTER(vp)]),
    "sesdf_cb_get_data": (i32, [vp, ctypes.POINTER(CbData)]),
    "sesdf_cb_get_counts": (i32, [vp, ctypes.POINTER(CbCounts)]),
    "sesdf_cb_get_internal": (i32, [vp, ctypes.POINTER(CbInternal)]),
    "sesdf_cb_set_tuning": (i32, [vp, ctypes.POINTER(CbTuning)]),
    "sesdf_create": (i32, [ctypes.POINTER(CreateInfo), ctypes.POINTER(vp)]),
    "sesdf_destroy": (i32, [vp]),
    "sesdf_build": (i32, [vp]),
    "sesdf_release": (i32, [vp]),
    "sesdf_set_neighbor_policy": (i32, [vp, ctypes.c_int]),
    "sesdf_build_device": (i32, [vp, vp]),
    "sesdf_set_atoms": (i32, [vp, vp]),
    "sesdf_input_buffer": (i32, [vp, ctypes.POINTER(vp)]),
    "sesdf_enqueue": (i32, [vp]),
    "sesdf_sync": (i32, [vp]),
    "sesdf_get_data": (i32, [vp, ctypes.POINTER(SesdfData)]),
    "sesdf_get_exterior_data": (i32, [vp, ctypes.POINTER(SsesdfData)]),
    "sesdf_get_counts": (i32, [vp, ctypes.POINTER(Counts)]),
    "sesdf_get_internal": (i32, [vp, ctypes.POINTER(Internal)]),
    "sesdf_get_atom_indices": (i32, [vp, vp]),
    "sesdf_get_stream": (i32, [vp, ctypes.POINTER(vp)]),
    "sesdf_build_timed": (i32, [vp, vp, ctypes.c_int]),
    "sesdf_sample_points": (i32, [vp, vp, u64, f32, vp]),
    "sesdf_sample_grid": (i32, [vp, ctypes.POINTER(GridDesc), f32, vp]),
    "sesdf_sdf_backward": (i32, [vp, vp, vp, u64, f32, ctypes.c_int, vp]),
    "sesdf_sample_points_ex": (i32, [vp, vp, u64, f32, ctypes.c_int, vp, vp]),
    "sesdf_sample_points_noalloc": (i32, [vp, vp, u64, f32, ctypes.c_int, vp, vp]),
    "sesdf_load_mmtf": (i32, [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)), ctypes.POINTER(u64)]),
    "sesdf_free_atoms": (None, [ctypes.POINTER(ctypes.c_float)]),
    "sesdf_element_radius": (f32, [ctypes.c_char_p]),
    "sesdf_sample_grid_ex": (i32, [vp, ctypes.POINTER(GridDesc), f32, ctypes.c_int, vp]),
    "sesdf_sdf_backward_ex": (i32, [vp, vp, vp, u64, f32, ctypes.c_int, vp, vp, vp]),
    "sesdf_surface_components": (i32, [vp, vp, vp]),
    "sesdf_batch_sample_grids": (i32, [vp, u32, f32, f32, vp]),
    "sesdf_set_tuning": (i32, [vp, ctypes.c_void_p]),
    "sesdf_batch_create": (i32, [ctypes.POINTER(BatchInfo), ctypes.POINTER(vp)]),
    "sesdf_batch_offsets": (i32, [vp, vp]),
    "sesdf_upload": (i32, [vp]),
    "sesdf_find_circles": (i32, [vp]),
    "sesdf_find_intersections": (i32, [vp]),
    "sesdf_build_conic_patches": (i32, [vp]),
    "sesdf_handle_intersection_singularities": (i32, [vp]),
    "sesdf_fill_circles": (i32, [vp]),
    "sesdf_get_global_to_trimmed": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(u64)]),
    "sesdf_get_graphics_layout": (i32, [vp, vp]),
    "sesdf_pack_graphics": (i32, [vp, vp]),
    "sesdf_device_to_host": (i32, [vp, vp, ctypes.c_size_t]),
    "sesdf_status_string": (ctypes.c_char_p, [i32]),
    "sesdf_last_error": (ctypes.c_char_p, []),
}

_LIB = None


def lib():
    """Load libsesdf_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _LIB = L
    return _LIB


class SesdfError(RuntimeError):
    pass


def check(status, what=""):
    if status != SESDF_OK:
        L = lib()
        msg = L.sesdf_status_string(status).decode()
        detail = L.sesdf_last_error().decode()
        raise SesdfError(f"{what}: {msg} ({detail})")


# -- minimal HIP runtime access for device <-> host copies (no torch needed) ------------------------
_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("libamdhip64.so not found")
        _HIP.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipMemcpy.restype = ctypes.c_int
        _HIP.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        _HIP.hipMalloc.restype = ctypes.c_int
        _HIP.hipFree.argtypes = [vp]
        _HIP.hipFree.restype = ctypes.c_int
        _HIP.hipDeviceSynchronize.restype = ctypes.c_int
        _HIP.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _HIP.hipGetDeviceCount.restype = ctypes.c_int
        _HIP.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _HIP.hipEventCreate.restype = ctypes.c_int
        _HIP.hipEventRecord.argtypes = [vp, vp]
        _HIP.hipEventRecord.restype = ctypes.c_int
        _HIP.hipEventSynchronize.argtypes = [vp]
        _HIP.hipEventSynchronize.restype = ctypes.c_int
        _HIP.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _HIP.hipEventElapsedTime.restype = ctypes.c_int
        _HIP.hipEventDestroy.argtypes = [vp]
        _HIP.hipEventDestroy.restype = ctypes.c_int
        _HIP.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        _HIP.hipMemGetInfo.restype = ctypes.c_int
    return _HIP


class Event:
    """hipEvent recorded on an explicit stream (the stream the kernels are launched on)."""

    def __init__(self):
        self.e = vp()
        if hip().hipEventCreate(ctypes.byref(self.e)) != 0:
            raise SesdfError("hipEventCreate failed")

    def record(self, stream):
        hip().hipEventRecord(self.e, stream)

    def sync(self):
        hip().hipEventSynchronize(self.e)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        hip().hipEventElapsedTime(ctypes.byref(ms), self.e, later.e)
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.e)
        except Exception:
            pass


def device_synchronize():
    hip().hipDeviceSynchronize()


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total))
    return free.value, total.value


H2D, D2H, D2D = 1, 2, 3


def memcpy(dst, src, nbytes, kind):
    if nbytes == 0:
        return
    rc = hip().hipMemcpy(dst, src, nbytes, kind)
    if rc != 0:
        raise SesdfError(f"hipMemcpy failed ({rc})")
