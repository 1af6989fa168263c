"""Synthetic atom clouds (float32 xyzr, shape [N, 4]) for tests and benchmarks.

Thin ctypes front-end over ``libsesdf_synth.so`` (sesdf_amd/csrc/synth.cpp). Seeds follow one
rule: workload c uses ``0x5E5DF000 + c``; molecules / globules / query streams use sub-streams.
"""
import ctypes
import math
import os

import numpy as np

_LIB = None
SEED_BASE = 0x5E5DF000


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsesdf_synth.so")
        if not os.path.exists(path):
            raise RuntimeError("libsesdf_synth.so is not built: run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        u32, u64, fp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
        lib.sesdf_synth_globule.argtypes = [u32, u64, fp]
        lib.sesdf_synth_globule.restype = u32
        lib.sesdf_synth_globule_voids.argtypes = [u32, u32, u64, fp, fp]
        lib.sesdf_synth_globule_voids.restype = u32
        lib.sesdf_synth_fibonacci.argtypes = [u32, u32, ctypes.c_float, u64, fp]
        lib.sesdf_synth_fibonacci.restype = u32
        lib.sesdf_synth_batch_sizes.argtypes = [u32, u32, u32, u64, fp]
        lib.sesdf_synth_batch_sizes.restype = None
        lib.sesdf_synth_batch.argtypes = [u32, fp, u64, fp]
        lib.sesdf_synth_batch.restype = u64
        lib.sesdf_synth_uniform.argtypes = [u64, u64, fp]
        lib.sesdf_synth_uniform.restype = None
        lib.sesdf_synth_normal.argtypes = [u64, u64, fp]
        lib.sesdf_synth_normal.restype = None
        _LIB = lib
    return _LIB


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def globule_radius(n):
    return (3.0 * n / (4.0 * math.pi * 0.05)) ** (1.0 / 3.0)


def globule(n, seed=SEED_BASE):
    """n atoms uniform in a ball at density 0.05 / A^3, min pair distance 1.2 A, C/N/O/S mix."""
    out = np.zeros((n, 4), np.float32)
    got = _lib().sesdf_synth_globule(n, seed, _ptr(out))
    return out[:got]


def globule_voids(n, n_voids=10, seed=SEED_BASE + 1):
    """Globule with n_voids empty spherical voids (radius U(4, 8) A); returns (atoms, voids[cx,cy,cz,r])."""
    out = np.zeros((n, 4), np.float32)
    voids = np.zeros((n_voids, 4), np.float32)
    got = _lib().sesdf_synth_globule_voids(n, n_voids, seed, _ptr(out), _ptr(voids))
    return out[:got], voids


def fibonacci_globules(n_globules=500, atoms_per=5000, shell_radius=360.0, seed=SEED_BASE + 2):
    """Globules of atoms_per atoms centred on a Fibonacci sphere, each randomly rotated."""
    out = np.zeros((n_globules * atoms_per, 4), np.float32)
    got = _lib().sesdf_synth_fibonacci(n_globules, atoms_per, shell_radius, seed, _ptr(out))
    return out[:got]


def batch_sizes(n_molecules=4096, lo=20, hi=5000, seed=SEED_BASE + 3):
    sizes = np.zeros(n_molecules, np.uint32)
    _lib().sesdf_synth_batch_sizes(n_molecules, lo, hi, seed, _ptr(sizes))
    return sizes


def batch(sizes, seed=SEED_BASE + 3):
    """Concatenated globules (molecule-major) and the per-molecule atom offsets [M+1]."""
    sizes = np.ascontiguousarray(sizes, np.uint32)
    out = np.zeros((int(sizes.sum()), 4), np.float32)
    got = _lib().sesdf_synth_batch(len(sizes), _ptr(sizes), seed ^ 0xBA7C4, _ptr(out))
    assert got == out.shape[0]
    offsets = np.zeros(len(sizes) + 1, np.uint64)
    offsets[1:] = np.cumsum(sizes, dtype=np.uint64)
    return out, offsets


def uniform(n, seed):
    out = np.zeros(n, np.float64)
    _lib().sesdf_synth_uniform(n, seed, _ptr(out))
    return out


def normal(n, seed):
    out = np.zeros(n, np.float64)
    _lib().sesdf_synth_normal(n, seed, _ptr(out))
    return out


def aabb(atoms):
    """bcs::getAabb semantics (sesdf/src/core/molecule.cpp:132-143): centre -/+ vdW radius."""
    a = np.asarray(atoms, np.float32)
    lo = (a[:, :3] - a[:, 3:4]).min(axis=0)
    hi = (a[:, :3] + a[:, 3:4]).max(axis=0)
    return lo.astype(np.float32), hi.astype(np.float32)
