"""GPU exterior SDF: the classification of points outside the SAS union into the unbounded region
and probe-accessible cavities, against a voxel flood fill of the complement and the FP64 oracle, and
the voids of the 100,000-atom globule (complete field positive, exterior field negative)."""
import numpy as np
import pytest

import oracle as O
from sesdf_amd import Sesdf, synth
from test_gpu_sdf import SDF_TOL, gpu_sdf
from test_oracle_components import FLOOD_CASES, RP, exterior_class, flood_case, sample_complement

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,n_voids,seed,band", FLOOD_CASES)
def test_exterior_rule_matches_flood_fill(gpu, n, n_voids, seed, band):
    atoms = flood_case(n, n_voids, seed)
    pts, want = sample_complement(atoms, seed)
    print("flood-fill case", n, n_voids, seed, band, "points", len(pts), "cavity", int((~want).sum()), flush=True)
    surf = Sesdf(atoms)
    got = gpu_sdf(surf, pts.astype(np.float32), band=band, exterior=True)
    assert np.count_nonzero(exterior_class(got, band) != want) == 0
    # values against the FP64 oracle at the same (FP32) points
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    ref = orc.sdf(pts.astype(np.float32).astype(np.float64), band=band, exterior=True, threads=8)
    assert np.abs(got - ref).max() <= SDF_TOL


def test_void_centres_of_the_100k_globule(gpu):
    atoms, voids = synth.globule_voids(100000, 10, synth.SEED_BASE + 1)
    surf = Sesdf(atoms)
    c = voids[:, :3].astype(np.float32)
    complete = gpu_sdf(surf, c, band=4.0, exterior=False)
    exterior = gpu_sdf(surf, c, band=4.0, exterior=True)
    print("void centres complete", complete, "exterior", exterior, flush=True)
    assert (complete > 0).all() and (exterior < 0).all()
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    for ext, got in ((False, complete), (True, exterior)):
        ref = orc.sdf(c.astype(np.float64), band=4.0, exterior=ext, threads=8)
        assert np.abs(got - ref).max() <= SDF_TOL
