"""More pins of the CPU oracle (no GPU): the stages whose outputs the first pins left open.

* Probe neighbours (the reference's findIntersectionsNeighbors, sesdf/src/sesdf/operations.cu:271-338;
  ssesdf/src/ssesdf/operations.cu:331-399): every with-neighbour probe lists the other with-neighbour
  probes closer than 2 rp. Checked against an O(Iw^2) enumeration of all pairs, which does not use
  the grid the oracle (and the reference) searches.
* Torus records of the exterior path (fillCircles, ssesdf/src/ssesdf/operations.cu:246-285): checked
  through what the geometry fixes, not through the formula: the centre lies on both SAS spheres and
  on the line of centres, the axis is the unit vector from atom i towards that centre, and the
  visibility sphere is centred on the point dividing c_i c_j in the ratio R_i : R_j with the radius
  the law of cosines gives for any point of the circle.

Each check also runs on a copy of the oracle's output with one planted mistake, which it must report.
"""
import numpy as np
import pytest

import oracle as O
from sesdf_amd import synth

RP = float(np.float32(1.4))


def probe_neighbor_report(res):
    """(checked, mismatches) of the oracle's probe-neighbour lists against all-pairs enumeration.

    Lists shorter than the capacity must equal the set {q : |p - q| < 2 rp}; full lists must be a
    subset of it. Probes the oracle marks ambiguous (a distance within the epsilon band of 2 rp, or
    a neighbour that is itself ambiguous) are skipped."""
    iw, M = res.iw, 32
    ids, pos32, pos64 = res["concave_ids"], res["concave_pos"], res["concave_pos64"]
    nb = res["concave_nb"].reshape(-1, 4)
    amb = res["amb_probe_nb"]
    if iw == 0:
        return 0, 0
    P = pos64[:iw]
    d = np.sqrt(((P[:, None, :] - P[None, :, :]) ** 2).sum(-1))
    want = d < 2.0 * RP
    np.fill_diagonal(want, False)
    index = {tuple(pos32[q, :3].tolist()): q for q in range(iw)}
    checked = bad = 0
    for q in range(iw):
        if amb[q]:
            continue
        cnt = int(ids[q, 3])
        entries = nb[q * M: q * M + cnt]
        got = set()
        for e in entries:
            r = index.get(tuple(e[:3].tolist()))
            got.add(-1 if r is None else r)
            if res.mode == 0:
                px, py, pz = (np.float32(v) for v in pos32[q, :3])
                if e[3] != np.float32(px * py * pz):
                    bad += 1
            elif e[3] != 0.0:
                bad += 1
        brute = set(np.nonzero(want[q])[0].tolist())
        if cnt < M:
            bad += got != brute
        else:
            bad += not (got <= brute and len(brute) >= M)
        if res.mode == 1 and pos32[q, 3] != np.float32(q * M):
            bad += 1
        checked += 1
    return checked, bad


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,seed", [(150, 11), (200, 12), (2000, 0)])
def test_probe_neighbors_brute_force(mode, n, seed):
    atoms = synth.globule(n, synth.SEED_BASE + (0 if n == 2000 else 2000 + seed))
    res = O.build(atoms, mode=mode, precision=64, track=True)
    checked, bad = probe_neighbor_report(res)
    assert bad == 0
    assert checked >= (400 if n == 2000 else 10)


def test_probe_neighbor_check_detects_planted_mistake():
    """Drop one entry of one list (and shorten its count) in a copy of the oracle's output."""
    atoms = synth.globule(2000, synth.SEED_BASE)
    res = O.build(atoms, mode=0, precision=64, track=True)

    class Planted:
        def __init__(self, base):
            self.base, self.mode, self.iw = base, base.mode, base.iw
            self.ids = base["concave_ids"].copy()
            q = next(q for q in range(base.iw) if 2 <= self.ids[q, 3] < 32 and not base["amb_probe_nb"][q])
            self.ids[q, 3] -= 1

        def __getitem__(self, k):
            return self.ids if k == "concave_ids" else self.base[k]

    assert probe_neighbor_report(res)[1] == 0
    assert probe_neighbor_report(Planted(res))[1] >= 1


def torus_report(res, tol=1e-4):
    """Largest violation of each torus-record invariant (exterior builds; FP64 from the inputs)."""
    A = res["atoms"].astype(np.float64)
    K = res.K
    slots = res["torus_slots"].astype(np.int64)
    T = len(slots)
    tor = res["torus"].reshape(3, T, 4).astype(np.float64)
    i = slots // K
    j = res["nb_ids"][slots].astype(np.int64)
    ci, cj = A[i, :3], A[j, :3]
    ri = A[i, 3]
    Ri, Rj = ri + RP, A[j, 3] + RP
    dvec = cj - ci
    d = np.linalg.norm(dvec, axis=1)
    n = dvec / d[:, None]
    # pairs farther apart than R_i + R_j (within the neighbour test's 1e-4 slack) have no circle: the
    # reference takes sqrt of a negative number there (NaN radius), and so does the oracle
    real = d < Ri + Rj
    ci, cj, ri, Ri, Rj, dvec, d, n, tor = ci[real], cj[real], ri[real], Ri[real], Rj[real], dvec[real], d[real], n[real], tor[:, real]
    tc, R = tor[0, :, :3], tor[0, :, 3]
    ax, axw = tor[1, :, :3], tor[1, :, 3]
    vs, vsr = tor[2, :, :3], tor[2, :, 3]
    out = {"tangent": int((~real).sum())}
    # centre and radius: on both SAS spheres, on the line of centres
    out["sphere_i"] = np.abs(((tc - ci) ** 2).sum(1) + R * R - Ri * Ri).max() / (Ri * Ri).max()
    out["sphere_j"] = np.abs(((tc - cj) ** 2).sum(1) + R * R - Rj * Rj).max() / (Rj * Rj).max()
    out["on_line"] = np.linalg.norm(np.cross(tc - ci, n), axis=1).max()
    # axis: unit vector from c_i towards the circle centre; w = probe radius
    off = np.linalg.norm(tc - ci, axis=1)
    ok = off > 1e-2  # the direction of a centre within 0.01 A of c_i is not defined to FP32 precision
    out["axis_unit"] = np.abs(np.linalg.norm(ax, axis=1) - 1).max()
    out["axis_dir"] = np.abs((ax[ok] * (tc[ok] - ci[ok])).sum(1) - off[ok]).max() if ok.any() else 0.0
    out["axis_w"] = np.abs(axw - np.float32(RP)).max()
    # visibility sphere: centre at c_i + R_i / (R_i + R_j) (c_j - c_i); radius by the law of cosines
    want_c = ci + (Ri / (Ri + Rj))[:, None] * dvec
    out["vs_centre"] = np.linalg.norm(vs + tc - want_c, axis=1).max()
    lam = Ri / (Ri + Rj)
    r2 = ri * ri + (lam * d) ** 2 - (ri / Ri) * lam * (Ri * Ri - Rj * Rj + d * d)
    out["vs_radius"] = np.abs(vsr - np.sqrt(np.maximum(r2, 0))).max()
    out["records"] = T
    return out


def _assert_torus(rep, tol=1e-4):
    for k, v in rep.items():
        if k not in ("records", "tangent"):
            assert v <= tol, (k, v, rep)


@pytest.mark.parametrize("n", [2000, 100000])
def test_torus_records_invariants(n):
    if n == 2000:
        atoms = synth.globule(n, synth.SEED_BASE)
    else:
        atoms, _ = synth.globule_voids(n, 10, synth.SEED_BASE + 1)
    res = O.build(atoms, mode=1, precision=64, track=False, threads=8)
    rep = torus_report(res)
    _assert_torus(rep)
    # pairs without a real circle (tangent within the neighbour slack, or contained) have status 0:
    # no record
    assert rep["records"] > n // 3 and rep["tangent"] == 0


def test_torus_check_detects_planted_mistakes():
    """Swap the roles of i and j in the visibility-sphere ratio, and flip one axis."""
    atoms = synth.globule(2000, synth.SEED_BASE)
    res = O.build(atoms, mode=1, precision=64, track=False)
    _assert_torus(torus_report(res))
    T = len(res["torus_slots"])
    K = res.K

    class Planted:
        def __init__(self, base, which):
            self.base, self.K = base, base.K
            tor = base["torus"].reshape(3, T, 4).copy()
            q = T // 2
            if which == "ratio":
                A = base["atoms"].astype(np.float64)
                g = int(base["torus_slots"][q])
                i, j = g // K, int(base["nb_ids"][g])
                Ri, Rj = A[i, 3] + RP, A[j, 3] + RP
                wrong = A[i, :3] + Rj / (Ri + Rj) * (A[j, :3] - A[i, :3])
                tor[2, q, :3] = wrong - tor[0, q, :3]
            else:
                tor[1, q, :3] = -tor[1, q, :3]
            self.tor = tor.reshape(-1)

        def __getitem__(self, k):
            return self.tor if k == "torus" else self.base[k]

    for which in ("ratio", "axis"):
        rep = torus_report(Planted(res, which))
        with pytest.raises(AssertionError):
            _assert_torus(rep)
