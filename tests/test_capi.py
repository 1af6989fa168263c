"""The C-ABI library loads and exports every entry point declared in include/sesdf_c.h (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "sesdf_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sesdf_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    import sesdf_amd._lib as L

    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python signature table covers every declared entry point
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def test_status_strings_without_gpu():
    import sesdf_amd._lib as L

    lib = L.lib()
    assert lib.sesdf_status_string(0) == b"ok"
    assert lib.sesdf_status_string(3) == b"unsupported"


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors agree with the C struct sizes the C compiler gives the header's structs."""
    import subprocess

    import sesdf_amd._lib as L

    names = {"sesdf_create_info": L.CreateInfo, "sesdf_data": L.SesdfData, "ssesdf_data": L.SsesdfData,
             "sesdf_grid_desc": L.GridDesc, "sesdf_counts": L.Counts, "sesdf_internal": L.Internal,
             "sesdf_batch_info": L.BatchInfo, "sesdf_cb_create_info": L.CbCreateInfo, "sesdf_cb_data": L.CbData,
             "sesdf_cb_counts": L.CbCounts, "sesdf_cb_internal": L.CbInternal, "sesdf_cb_tuning": L.CbTuning}
    prog = tmp_path / "sizes.c"
    prog.write_text('#include <stdio.h>\n#include "sesdf_c.h"\nint main(void) {\n' +
                    "".join(f'    printf("%zu\\n", sizeof({n}));\n' for n in names) + "    return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert sizes == [ctypes.sizeof(t) for t in names.values()], dict(zip(names, sizes))


def test_synthetic_generators_deterministic():
    from sesdf_amd import synth

    a = synth.globule(1000, 7)
    b = synth.globule(1000, 7)
    np.testing.assert_array_equal(a, b)
    assert a.shape == (1000, 4)
    d = np.sqrt(((a[:, None, :3] - a[None, :, :3]) ** 2).sum(-1))
    np.fill_diagonal(d, 10)
    assert d.min() >= 1.2 - 1e-5
    R = synth.globule_radius(1000)
    assert np.linalg.norm(a[:, :3], axis=1).max() <= R + 1e-4
    radii = np.unique(a[:, 3]).astype(np.float64)
    assert all(np.min(np.abs(r - np.array([1.52, 1.55, 1.70, 1.80]))) < 1e-6 for r in radii)
    sizes = synth.batch_sizes(4096)
    assert sizes.min() >= 20 and sizes.max() <= 5000


def test_batch_offsets_validated_before_device_work():
    """Offsets must start at 0 and never decrease; both are rejected without touching a GPU."""
    import ctypes

    import numpy as np

    import sesdf_amd._lib as L

    lib = L.lib()
    atoms = np.zeros((4, 4), np.float32)
    for offs in ([1, 2, 4], [0, 3, 2], [0, 4, 3]):
        o = np.array(offs, np.uint64)
        info = L.BatchInfo()
        info.xyzr = atoms.ctypes.data
        info.atom_offsets = o.ctypes.data
        info.molecule_nb = len(offs) - 1
        info.probe_radius = 1.4
        h = ctypes.c_void_p()
        assert lib.sesdf_batch_create(ctypes.byref(info), ctypes.byref(h)) == L.ERR_INVALID, offs
        assert not h.value
        assert b"atom_offsets" in lib.sesdf_last_error()


def test_non_finite_input_and_bad_aabb_rejected_before_device_work():
    import ctypes

    import numpy as np

    import sesdf_amd._lib as L

    lib = L.lib()
    good = np.array([[0, 0, 0, 1.7], [3, 0, 0, 1.7]], np.float32)
    cases = []
    for bad in (np.nan, np.inf):
        a = good.copy()
        a[1, 2] = bad
        cases.append((a, [0.0] * 6))
    cases.append((good, [0, 0, 0, -1, 5, 5]))        # min > max
    cases.append((good, [0, 0, 0, np.inf, 5, 5]))    # not finite
    for atoms, aabb in cases:
        info = L.CreateInfo()
        info.xyzr = atoms.ctypes.data
        info.atom_nb = len(atoms)
        info.probe_radius = 1.4
        for d in range(6):
            info.aabb[d] = aabb[d]
        h = ctypes.c_void_p()
        assert lib.sesdf_create(ctypes.byref(info), ctypes.byref(h)) == L.ERR_INVALID
        assert not h.value
    a = np.concatenate([good, good]).copy()
    a[3, 3] = np.inf
    o = np.array([0, 2, 4], np.uint64)
    info = L.BatchInfo()
    info.xyzr = a.ctypes.data
    info.atom_offsets = o.ctypes.data
    info.molecule_nb = 2
    info.probe_radius = 1.4
    h = ctypes.c_void_p()
    assert lib.sesdf_batch_create(ctypes.byref(info), ctypes.byref(h)) == L.ERR_INVALID
    assert b"finite" in lib.sesdf_last_error()
