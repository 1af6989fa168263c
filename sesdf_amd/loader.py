"""Molecule loading, as the reference's bcs::load (sesdf/src/core/loader.cpp:14-71) for MMTF files.

``load(path)`` returns the atoms of the first model in file order as a (n, 4) float32 array of
x, y, z and the van der Waals radius of each atom's element (the table of
sesdf/src/core/molecule.cpp:9-131; unknown elements 1.20 A). The decoding runs in the C library
(``sesdf_load_mmtf``), the same code C++ callers reach through ``bcs::load``.
"""
import ctypes
import os

import numpy as np

from . import _lib as L


def load(path):
    lib = L.lib()
    ptr = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_uint64()
    L.check(lib.sesdf_load_mmtf(os.fsencode(path), ctypes.byref(ptr), ctypes.byref(n)), f"load {path}")
    try:
        out = np.ctypeslib.as_array(ptr, shape=(n.value * 4,)).reshape(-1, 4).copy() if n.value else \
            np.zeros((0, 4), np.float32)
    finally:
        lib.sesdf_free_atoms(ptr)
    return out


def element_radius(symbol):
    """van der Waals radius of an element symbol (any case), as bcs::getRadius."""
    return float(L.lib().sesdf_element_radius(symbol.encode()))
