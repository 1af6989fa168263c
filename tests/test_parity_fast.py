"""The vectorised full-size comparator against the reference comparator, and the locality argument
behind comparing a ball cut out of a large system (CPU only: oracle builds on both sides)."""
import numpy as np
import pytest

import oracle as O
import parity
import parity_fast as PF
from oracle_util import as_build_arrays
from sesdf_amd import synth

COVERED = ("neighbors", "status", "concave", "has_neighbors", "full_circles", "segments", "probe_neighbors")


@pytest.mark.parametrize("mode", [0, 1])
def test_agrees_with_reference_comparator(mode):
    atoms = synth.globule(2000, synth.SEED_BASE)
    f32 = as_build_arrays(O.build(atoms, mode=mode, precision=32, track=False))
    f64 = O.build(atoms, mode=mode, precision=64, track=True)
    ref = parity.compare(f32, f64, atoms, mode)
    fast = PF.compare_fast(f32, f64, atoms, mode)
    for k in COVERED:
        if k in ref["mismatch"]:
            assert fast["mismatch"][k] == ref["mismatch"][k], k
            assert fast["band"][k] == ref["band"][k], k
    assert fast["concave_nb"] == ref["concave_nb"]
    assert fast["probe_pos_max_err"] == pytest.approx(ref["probe_pos_max_err"], abs=1e-12)
    if mode == 0:
        assert fast["sector_max_err"] == pytest.approx(ref["sector_max_err"], abs=1e-12)
    PF.assert_parity(fast)


def test_planted_mistakes_are_reported():
    atoms = synth.globule(600, synth.SEED_BASE + 9)
    f64 = O.build(atoms, mode=0, precision=64, track=True)
    base = as_build_arrays(O.build(atoms, mode=0, precision=32, track=False))
    assert not any(PF.compare_fast(base, f64, atoms, 0)["mismatch"].values())

    arr = dict(base)  # a dropped concave patch
    keep = np.ones(len(arr["concave_ids"]), bool)
    keep[len(keep) // 2] = False
    arr["concave_ids"], arr["concave_pos"], arr["concave_pos_w"] = (arr[k][keep] for k in ("concave_ids", "concave_pos", "concave_pos_w"))
    arr["segments"] = arr["segments"][:0]  # their endpoint indices refer to the full list
    rep = PF.compare_fast(arr, f64, atoms, 0)
    assert rep["mismatch"]["concave"] >= 1 and rep["mismatch"]["segments"] >= 1

    arr = dict(base)  # an intersected circle reported full (slot away from the band)
    st = arr["status"].copy()
    amb = f64["amb_circle"].astype(bool)
    K = base["K"]
    slot = next(i * K + l for i in range(len(atoms)) for l in range(int(base["nb_count"][i]))
                if st[i * K + l] == 1 and not amb[i * K + l])
    st[slot] = 2
    arr["status"] = st
    assert PF.compare_fast(arr, f64, atoms, 0)["mismatch"]["status"] >= 1

    arr = dict(base)  # a probe-neighbour list one entry short (a probe away from the band)
    ids = arr["concave_ids"].copy()
    amb_p = f64["amb_probe_nb"].astype(bool)
    q = next(q for q in range(f64.iw) if 1 <= ids[q, 3] < 32 and not amb_p[q])
    ids[q, 3] -= 1
    arr["concave_ids"] = ids
    rep = PF.compare_fast(arr, f64, atoms, 0)
    assert rep["mismatch"]["probe_neighbors"] == 1 and rep["probe_sets_compared"] > 20

    arr = dict(base)  # a probe moved by 1e-3 A
    pos = arr["concave_pos"].copy()
    pos[0, 0] += 1e-3
    arr["concave_pos"] = pos
    with pytest.raises(AssertionError):
        PF.assert_parity(PF.compare_fast(arr, f64, atoms, 0))


@pytest.mark.parametrize("mode", [0, 1])
def test_ball_cut_out_is_exact_inside(mode):
    """Circles and neighbour lists of atoms 8 A inside a 20 A ball, and patches of atoms 9 A
    inside, do not see the cut: the ball's build agrees with the whole build there; atoms at the
    rim do not (the comparison has teeth)."""
    atoms = synth.globule(6000, synth.SEED_BASE + 21)
    whole = as_build_arrays(O.build(atoms, mode=mode, precision=32, track=False))
    rng = np.random.default_rng(5)
    d_c = np.linalg.norm(atoms[:, :3] - atoms[:, :3].mean(0), axis=1)
    centre = atoms[rng.choice(np.nonzero(d_c > d_c.max() - 6)[0]), :3]  # on the surface
    d = np.linalg.norm(atoms[:, :3] - centre, axis=1)
    ball = np.nonzero(d < 20.0)[0]
    sub = O.build(atoms[ball], mode=mode, precision=64, track=True)
    rep = PF.compare_fast(whole, sub, atoms, mode, orc_to_input=ball, keep=np.nonzero(d < 12.0)[0],
                          keep_patches=np.nonzero(d < 11.0)[0])
    assert rep["neighbor_pairs"] > 1000 and rep["concave_nb"] > 50
    PF.assert_parity(rep)
    rim = PF.compare_fast(whole, sub, atoms, mode, orc_to_input=ball, keep=np.nonzero(d < 19.5)[0],
                          keep_patches=np.nonzero(d < 19.5)[0])
    assert rim["mismatch"]["neighbors"] > 0


def test_torus_records_compared_by_circle():
    """Exterior path: torus records keyed by circle, their values held to the FP32 restatement's
    own error; a record moved by 1e-3 A is reported."""
    atoms = synth.globule(800, synth.SEED_BASE + 19)
    f64 = O.build(atoms, mode=1, precision=64, track=True)
    arr = as_build_arrays(O.build(atoms, mode=1, precision=32, track=False))
    t = O.build(atoms, mode=1, precision=32, track=False)["torus"].reshape(3, -1, 4)
    arr.update(torus_pos=t[0].copy(), torus_axis=t[1].copy(), torus_vs=t[2].copy())
    f32 = O.build(atoms, mode=1, precision=32, track=False)
    rep = PF.compare_fast(arr, f64, atoms, 1)
    assert rep["torus_nb"] > 100 and rep["torus_same_head"] == rep["torus_nb"]
    PF.assert_parity(rep)
    assert PF.torus_values_within_fp32(arr, f64, f32, atoms) == (rep["torus_nb"], 0)
    arr["torus_pos"][len(t[0]) // 2, 1] += 1e-3
    assert PF.torus_values_within_fp32(arr, f64, f32, atoms)[1] == 1
