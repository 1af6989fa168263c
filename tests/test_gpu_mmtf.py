"""The reference's 12 sample proteins (samples/*.mmtf, 126 to 335,722 atoms, the benchmark cases of
app/benchmarks/main.cu:97-113), loaded with sesdf_amd.load and built on the GPU in both modes,
against the FP64 oracle: the whole topology (neighbour lists, circle statuses, sectors, concave
patches and their positions, probe-neighbour sets, segments or torus records), the torus values
within the FP32 bound, and the SDF (complete and exterior) at 20,000 points near each surface."""
import os

import numpy as np
import pytest

import oracle as O
import parity
import parity_fast as PF
from sesdf_amd import Sesdf, Ssesdf, load
from test_gpu_sdf import gpu_sdf

pytestmark = pytest.mark.gpu

MMTF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mmtf")
CODES = ["1AGA", "101M", "1VIS", "7SC0", "3EAM", "7DBB", "1A8R", "7O0U", "1AON", "7RGD", "3JC8", "7CGO"]
SDF_TOL = 1e-4


def _say(*a):
    print("[mmtf]", *a, flush=True)


@pytest.mark.parametrize("code", CODES)
@pytest.mark.parametrize("mode", [0, 1])
def test_sample_protein_matches_fp64(gpu, code, mode):
    atoms = load(os.path.join(MMTF, code + ".mmtf"))
    cls = Sesdf if mode == 0 else Ssesdf
    _say(f"{code}: {len(atoms)} atoms, {cls.__name__}")
    surf = cls(atoms)
    counts = surf.counts()
    assert counts["neighbor_overflow_atoms"] == 0
    g = parity.gpu_arrays(surf)
    orc = O.build(atoms, mode=mode, precision=64, track=True, threads=16)
    rep = PF.compare_fast(g, orc, atoms, mode)
    _say({k: rep[k] for k in ("neighbor_pairs", "concave_nb", "probe_pos_max_err", "probe_sets_compared") if k in rep},
         rep["mismatch"], rep["band"])
    PF.assert_parity(rep)
    if mode == 1:
        order = surf.get_atom_indices()
        o64 = O.build(atoms, mode=1, precision=64, track=False, threads=16, order=order)
        o32 = O.build(atoms, mode=1, precision=32, track=False, threads=16, order=order)
        compared, outside = PF.torus_values_within_fp32(g, o64, o32, atoms)
        _say(f"torus records {rep['torus_nb']}, values compared {compared}, outside the FP32 bound {outside}")
        assert compared > 0.99 * rep["torus_nb"] and outside == 0
        surf.close()
        return
    # SDF at points within 2 A of the SAS spheres, both fields
    rng = np.random.default_rng(61)
    n = 20_000
    a = rng.integers(0, len(atoms), n)
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = atoms[a, 3] + 1.4 + rng.uniform(-2.0, 2.0, n)
    pts = (atoms[a, :3] + u * r[:, None]).astype(np.float32)
    s2i_g, s2i_o = g["s2i"].astype(np.int64), orc["sorted_to_initial"].astype(np.int64)
    for exterior in (False, True):
        pv, pc = gpu_sdf(surf, pts, 2.0, exterior, codes=True)
        o, oc = orc.sdf(pts.astype(np.float64), band=2.0, exterior=exterior, threads=16, codes=True)
        wa_g = PF.winner_atoms(pc, s2i_g, g["full_circles"], g["segments"], g["concave_ids"])
        wa_o = PF.winner_atoms(oc, s2i_o, orc["full_circles"], orc["segments"], orc["concave_ids"])
        excused = (np.isin(wa_g, rep["band_atoms"]) | np.isin(wa_o, rep["band_atoms"])).any(axis=1)
        err = np.abs(pv - o)
        _say(f"exterior={exterior}: max |GPU - FP64| {err[~excused].max():.2e}, {int(excused.sum())} points decided by "
             f"band elements; inside {np.mean(o < 0):.3f}")
        assert err[~excused].max() <= SDF_TOL
        assert excused.mean() < 2e-3
    surf.close()
