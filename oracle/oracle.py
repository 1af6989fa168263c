"""ORACLE front-end (test infrastructure only): ctypes over oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
It runs the CPU restatement of the reference's stage sequence and the FP64 SDF / gradient oracle
(both in oracle/sesdf_oracle.cpp), and the restatement of the reference's Contour-Buildup
comparator (oracle/cb_oracle.cpp), and returns numpy arrays.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ARRAYS = {
    "atoms": np.float32, "sorted_to_initial": np.uint32, "nb_count": np.uint32, "nb_ids": np.uint32,
    "status": np.uint8, "trimmed_to_global": np.uint32, "convex_patches": np.uint32, "sectors": np.float32,
    "full_circles": np.uint32, "concave_ids": np.int32, "concave_pos": np.float32, "concave_pos64": np.float64,
    "concave_nb": np.float32, "circle_int_nb": np.uint32, "segments": np.uint32, "torus": np.float32,
    "torus_slots": np.uint32, "amb_pairs": np.uint32, "amb_circle": np.uint8, "amb_concave": np.uint8,
    "amb_hasnb": np.uint8, "amb_rejected": np.uint32, "amb_seg_circle": np.uint8, "amb_probe_nb": np.uint8,
    "stats": np.uint64, "grid": np.float32, "counters": np.uint64,
}
SHAPES = {"atoms": 4, "convex_patches": 2, "sectors": 4, "full_circles": 2, "concave_ids": 4, "concave_pos": 4,
          "concave_pos64": 3, "concave_nb": 4, "segments": 4, "amb_pairs": 2, "amb_rejected": 4}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.oracle_build.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_uint32, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oracle_build.restype = ctypes.c_void_p
        L.oracle_build_nb.argtypes = L.oracle_build.argtypes + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_build_nb.restype = ctypes.c_void_p
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_get.restype = ctypes.c_uint64
        L.oracle_sdf.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_sdf.restype = ctypes.c_int
        L.oracle_components.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_components.restype = ctypes.c_uint32
        L.oracle_sdf_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
        L.oracle_sdf_grad.restype = ctypes.c_int
        L.oracle_sdf_fd.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int]
        L.oracle_sdf_fd.restype = ctypes.c_int
        _LIB = L
    return _LIB


class OracleResult:
    """Arrays of one CPU build (reference grid-sorted index space, see sorted_to_initial)."""

    def __init__(self, handle, mode, n, K, rp):
        self._h = handle
        self.mode, self.n, self.K, self.rp = mode, n, K, rp
        self._cache = {}

    def __getitem__(self, name):
        if name not in self._cache:
            L = lib()
            es = ctypes.c_uint32(0)
            cnt = L.oracle_get(self._h, name.encode(), None, ctypes.byref(es))
            arr = np.zeros(cnt, ARRAYS[name])
            assert cnt == 0 or arr.itemsize == es.value, (name, arr.itemsize, es.value)
            if cnt:
                L.oracle_get(self._h, name.encode(), ctypes.c_void_p(arr.ctypes.data), None)
            if name in SHAPES:
                arr = arr.reshape(-1, SHAPES[name])
            self._cache[name] = arr
        return self._cache[name]

    def sdf(self, points, band=4.0, exterior=False, threads=1, codes=False):
        """FP64 SDF at points [n, 3] (complete-surface builds only); optionally the winner codes."""
        p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 3))
        out = np.zeros(p.shape[0], np.float64)
        cd = np.zeros(p.shape[0], np.int64) if codes else None
        rc = lib().oracle_sdf(self._h, ctypes.c_void_p(p.ctypes.data), p.shape[0], 1 if exterior else 0, band,
                              ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(cd.ctypes.data) if codes else None, threads)
        if rc:
            raise ValueError("oracle SDF needs a complete-surface build")
        return (out, cd) if codes else out

    def sdf_grad(self, points, weights, band=2.0, exterior=False, threads=8):
        """FP64 analytic gradient of sum_q w_q SDF(q) w.r.t. (x, y, z, r) per atom, input order [n, 4]."""
        p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 3))
        w = np.ascontiguousarray(np.asarray(weights, np.float64).reshape(-1))
        g = np.zeros((self.n, 4), np.float64)
        rc = lib().oracle_sdf_grad(self._h, ctypes.c_void_p(p.ctypes.data), ctypes.c_void_p(w.ctypes.data), p.shape[0],
                                   1 if exterior else 0, band, ctypes.c_void_p(g.ctypes.data), threads)
        if rc:
            raise ValueError("oracle SDF needs a complete-surface build")
        return g

    def sdf_fd(self, points, weights, atom_ids, h=1e-4, band=2.0, exterior=False, threads=8):
        """Central finite differences (FP64, topology fixed) for the listed input atoms: ([k, 4], stable[k])."""
        p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 3))
        w = np.ascontiguousarray(np.asarray(weights, np.float64).reshape(-1))
        ids = np.ascontiguousarray(np.asarray(atom_ids, np.uint32))
        out = np.zeros((len(ids), 4), np.float64)
        st = np.zeros(len(ids), np.uint8)
        rc = lib().oracle_sdf_fd(self._h, ctypes.c_void_p(p.ctypes.data), ctypes.c_void_p(w.ctypes.data), p.shape[0],
                                 1 if exterior else 0, band, ctypes.c_void_p(ids.ctypes.data), len(ids), h,
                                 ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(st.ctypes.data), threads)
        if rc:
            raise ValueError("oracle SDF needs a complete-surface build")
        return out, st.astype(bool)

    def components(self):
        comp = np.zeros(self.n, np.uint32)
        ext = lib().oracle_components(self._h, ctypes.c_void_p(comp.ctypes.data))
        return comp, int(ext)

    @property
    def iw(self):
        return int(self["counters"][0])

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().oracle_free(self._h)
            except (TypeError, AttributeError):  # interpreter shutdown
                pass
            self._h = None


CB_ARRAYS = {
    "atoms": (np.float32, 4), "sorted_to_initial": (np.uint32, 0), "nb_count0": (np.uint32, 0),
    "nb_count": (np.uint32, 0), "nb_ids": (np.uint32, 64), "removed": (np.uint8, 64), "arc_count": (np.uint8, 64),
    "visible": (np.uint8, 64), "probe_pos": (np.float32, 4), "probe_pos64": (np.float64, 3),
    "probe_ids": (np.uint32, 4), "tria1": (np.float32, 4), "tria2": (np.float32, 4), "tria3": (np.float32, 4),
    "torus_pos": (np.float32, 4), "torus_axis": (np.float32, 4), "torus_vs": (np.float32, 4),
    "torus_ids": (np.uint32, 2), "probe_sorted": (np.uint32, 0), "tex_coord": (np.float32, 3),
    "sing_tex": (np.float32, 96), "amb_atom": (np.uint8, 0), "amb_cover": (np.uint8, 64), "amb_circle": (np.uint8, 64),
    "amb_probe": (np.uint8, 0), "stats": (np.uint64, 0),
}


def _cb_lib():
    L = lib()
    if not getattr(L, "_cb_ready", False):
        L.oracle_cb_build.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int]
        L.oracle_cb_build.restype = ctypes.c_void_p
        L.oracle_cb_singularities.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_cb_singularities.restype = ctypes.c_int
        L.oracle_cb_free.argtypes = [ctypes.c_void_p]
        L.oracle_cb_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_cb_get.restype = ctypes.c_uint64
        L.oracle_cb_stat_names.restype = ctypes.c_char_p
        L._cb_ready = True
    return L


class CbResult:
    """Arrays of one Contour-Buildup oracle build (grid-sorted atom index space, see sorted_to_initial)."""

    def __init__(self, handle, n):
        self._h, self.n = handle, n
        self._cache = {}

    def __getitem__(self, name):
        if name not in self._cache:
            L = _cb_lib()
            dt, width = CB_ARRAYS[name]
            es = ctypes.c_uint32(0)
            cnt = L.oracle_cb_get(self._h, name.encode(), None, ctypes.byref(es))
            arr = np.zeros(cnt, dt)
            assert cnt == 0 or arr.itemsize == es.value, (name, arr.itemsize, es.value)
            if cnt:
                L.oracle_cb_get(self._h, name.encode(), ctypes.c_void_p(arr.ctypes.data), None)
            if width:
                arr = arr.reshape(-1, width)
            self._cache[name] = arr
        return self._cache[name]

    def stats(self):
        names = _cb_lib().oracle_cb_stat_names().decode().split()
        return dict(zip(names, (int(v) for v in self["stats"])))

    def singularities(self, probes):
        """FP32 singularity texture of the given probe array [P, 4] on this build's grid:
        (tex_coord [P, 3] in probe order, sing_tex [P, 96] in sorted probe order, sorted probe -> probe)."""
        p = np.ascontiguousarray(probes, np.float32).reshape(-1, 4)
        P = p.shape[0]
        tc = np.zeros((P, 3), np.float32)
        st = np.zeros((P, 96), np.float32)
        order = np.zeros(P, np.uint32)
        _cb_lib().oracle_cb_singularities(self._h, ctypes.c_void_p(p.ctypes.data), P, ctypes.c_void_p(tc.ctypes.data),
                                          ctypes.c_void_p(st.ctypes.data), ctypes.c_void_p(order.ctypes.data))
        return tc, st, order

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _cb_lib().oracle_cb_free(self._h)
            except (TypeError, AttributeError):  # interpreter shutdown
                pass
            self._h = None


def aabb_of(atoms):
    """The reference's getAabb: min / max of centre -/+ vdW radius (FP32)."""
    a = np.asarray(atoms, np.float32).reshape(-1, 4)
    lo = (a[:, :3] - a[:, 3:4]).min(axis=0)
    hi = (a[:, :3] + a[:, 3:4]).max(axis=0)
    return np.concatenate([lo, hi]).astype(np.float32)


def cb_build(atoms, probe_radius=1.4, aabb=None, grid_dim=16, tex_size=32768, precision=64, track=True, delta=5e-5,
             delta_theta=1e-5, threads=1):
    """Contour-Buildup oracle build (bcs::ContourBuildup::build()). precision 32 mirrors the reference's
    FP32 arithmetic; 64 is the accuracy oracle (with the epsilon-band tracking when track).

    delta (A) bands the distance-like predicates, about ten times the FP32 rounding of coordinates of
    tens of A; angle comparisons on a circle of radius w are banded by delta_theta + delta / w."""
    a = np.ascontiguousarray(atoms, np.float32).reshape(-1, 4)
    box = np.ascontiguousarray(aabb_of(a) if aabb is None else aabb, np.float32)
    h = _cb_lib().oracle_cb_build(ctypes.c_void_p(a.ctypes.data), a.shape[0], probe_radius,
                                  ctypes.c_void_p(box.ctypes.data), grid_dim, tex_size, precision, int(track), delta,
                                  delta_theta, threads)
    if not h:
        raise ValueError("grid_dim must be a power of two and tex_size positive")
    return CbResult(h, a.shape[0])


def build(atoms, probe_radius=1.4, mode=0, precision=64, max_neighbors=128, track=True, delta=1e-3,
          delta_theta=1e-4, threads=1, neighbors=None, order=None):
    """Run the oracle build. mode 0 = complete surface (Sesdf), 1 = exterior path (Ssesdf).

    neighbors: optional (count[n], ids[n, max_neighbors]) per input atom, in input ids, used in place
    of the grid query's lists; order: optional sorted order (sorted index -> input index) used in place
    of the grid's. Together they check a build whose lists were truncated at K in another grid order."""
    a = np.ascontiguousarray(atoms, np.float32)
    if order is not None:
        order = np.ascontiguousarray(order, np.uint32)
        assert order.shape == (a.shape[0],) and np.array_equal(np.sort(order), np.arange(a.shape[0]))
    if neighbors is None and order is None:
        h = lib().oracle_build(ctypes.c_void_p(a.ctypes.data), a.shape[0], probe_radius, mode, precision,
                               max_neighbors, int(track), delta, delta_theta, threads)
    else:
        cnt = ids = None
        if neighbors is not None:
            cnt = np.ascontiguousarray(neighbors[0], np.uint32)
            ids = np.ascontiguousarray(neighbors[1], np.uint32).reshape(-1)
            assert cnt.shape == (a.shape[0],) and ids.size == a.shape[0] * max_neighbors
            assert (cnt <= max_neighbors).all() and (ids < max(a.shape[0], 1)).all()
        ptr = lambda v: ctypes.c_void_p(v.ctypes.data) if v is not None else None  # noqa: E731
        h = lib().oracle_build_nb(ctypes.c_void_p(a.ctypes.data), a.shape[0], probe_radius, mode, precision,
                                  max_neighbors, int(track), delta, delta_theta, threads, ptr(cnt), ptr(ids), ptr(order))
    return OracleResult(h, mode, a.shape[0], max_neighbors, probe_radius)
