"""Pins of the CPU oracle (no GPU): brute force on small inputs, closed forms, invariants.

The reference ships no tests or golden outputs, so the oracle is pinned by mathematics that does
not restate its own formulas: O(N^2) neighbour enumeration, O(N^3) three-sphere trilateration with
visibility against every atom, dense sampling of circles, arc-midpoint exposure of segments, and
the SAS-complex invariants (sum of per-circle counts = 3 I, even counts, G = 3 I / 2).
"""
import itertools

import numpy as np
import pytest

import oracle as O
import parity
from oracle_util import as_build_arrays, tetrahedron
from sesdf_amd import synth

RP = float(np.float32(1.4))  # the probe radius crosses the ABI as FP32


def small_cluster(n, seed):
    """n atoms packed at protein-like density (min distance 1.2 A)."""
    return synth.globule(n, synth.SEED_BASE + 1000 + seed)


def trilaterate(c1, r1, c2, r2, c3, r3):
    """Both intersection points of three spheres (FP64, textbook construction)."""
    ex = (c2 - c1) / np.linalg.norm(c2 - c1)
    i = np.dot(ex, c3 - c1)
    ey = c3 - c1 - i * ex
    if np.linalg.norm(ey) < 1e-12:
        return []
    ey /= np.linalg.norm(ey)
    ez = np.cross(ex, ey)
    d = np.linalg.norm(c2 - c1)
    j = np.dot(ey, c3 - c1)
    x = (r1 * r1 - r2 * r2 + d * d) / (2 * d)
    y = (r1 * r1 - r3 * r3 + i * i + j * j) / (2 * j) - i / j * x
    z2 = r1 * r1 - x * x - y * y
    if z2 < 0:
        return []
    z = np.sqrt(z2)
    base = c1 + x * ex + y * ey
    return [base + z * ez, base - z * ez]


def oracle_sets(res, atoms):
    s2i = res["sorted_to_initial"]
    conc = {parity.concave_key(res["concave_ids"][q], res["concave_pos64"][q], s2i, atoms): q
            for q in range(len(res["concave_ids"]))}
    return conc


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbors_brute_force(seed):
    atoms = small_cluster(250, seed)
    res = O.build(atoms, mode=0, precision=64, track=True)
    s2i, K = res["sorted_to_initial"], res.K
    got = set()
    for i in range(len(atoms)):
        for l in range(res["nb_count"][i]):
            got.add((int(s2i[i]), int(s2i[res["nb_ids"][i * K + l]])))
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    m = d2 - (R[:, None] + R[None, :]) ** 2 <= float(np.float32(1e-4))
    np.fill_diagonal(m, False)
    want = {(int(a), int(b)) for a, b in zip(*np.nonzero(m))}
    assert got == want


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_concave_patches_brute_force(seed):
    atoms = small_cluster(120, seed)
    res = O.build(atoms, mode=0, precision=64, track=True)
    conc = oracle_sets(res, atoms)
    amb = {k for k, q in conc.items() if res["amb_concave"][q]}
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    n = len(atoms)
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    nb = d2 < (R[:, None] + R[None, :]) ** 2
    want, near = set(), set()
    for i, j, k in itertools.combinations(range(n), 3):
        if not (nb[i, j] and nb[i, k] and nb[j, k]):
            continue
        for x in trilaterate(c[i], R[i], c[j], R[j], c[k], R[k]):
            m = ((c - x) ** 2).sum(-1) - R * R
            m[[i, j, k]] = np.inf
            key = parity.concave_key(np.array([i, j, k]), x, np.arange(n), atoms)
            if np.min(m) > 1e-3:
                want.add(key)
            elif np.min(m) > -1e-3:
                near.add(key)
    assert set(conc) - amb - near == want - near - amb
    assert len(want) > 0
    # probe positions lie on the three SAS spheres
    s2i = res["sorted_to_initial"]
    for q in range(len(res["concave_ids"])):
        x = res["concave_pos64"][q]
        for a in res["concave_ids"][q][:3]:
            t = int(s2i[a])
            assert abs(np.linalg.norm(x - c[t]) - R[t]) < 1e-9


def test_circle_status_by_sampling():
    atoms = small_cluster(150, 7)
    res = O.build(atoms, mode=0, precision=64, track=True)
    s2i, K = res["sorted_to_initial"], res.K
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    theta = np.linspace(0, 2 * np.pi, 2048, endpoint=False)
    checked = 0
    for i in range(len(atoms)):
        ni = res["nb_count"][i]
        for l in range(ni):
            if res["amb_circle"][i * K + l]:
                continue
            a, b = int(s2i[i]), int(s2i[res["nb_ids"][i * K + l]])
            d = c[b] - c[a]
            dd = np.linalg.norm(d)
            n = d / dd
            t = (R[a] ** 2 - R[b] ** 2 + dd * dd) / (2 * dd)
            ctr = c[a] + n * t
            rad = np.sqrt(max(R[a] ** 2 - t * t, 0.0))
            u = np.cross(n, [1.0, 0, 0] if abs(n[0]) < 0.9 else [0, 1.0, 0])
            u /= np.linalg.norm(u)
            v = np.cross(n, u)
            pts = ctr + rad * (np.cos(theta)[:, None] * u + np.sin(theta)[:, None] * v)
            hidden = complete_broken = False
            for kk in range(ni):
                if kk == l:
                    continue
                k = int(s2i[res["nb_ids"][i * K + kk]])
                inside = ((pts - c[k]) ** 2).sum(-1) - R[k] ** 2 < 1e-4
                hidden |= bool(inside.all())
                complete_broken |= bool(inside.any())
            want = 0 if hidden else (1 if complete_broken else 2)
            assert res["status"][i * K + l] == want, (a, b)
            checked += 1
    assert checked > 1000


@pytest.mark.parametrize("seed", [0, 5])
def test_segments_arc_midpoints_exposed(seed):
    """Each segment (A -> B, counter-clockwise about n(i->j)) covers an exposed arc of circle ij."""
    atoms = small_cluster(200, seed)
    res = O.build(atoms, mode=0, precision=64, track=True)
    s2i = res["sorted_to_initial"]
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    amb_circ = set()
    t2g = res["trimmed_to_global"]
    for t in np.nonzero(res["amb_seg_circle"])[0]:
        g = int(t2g[t])
        amb_circ.add((g // res.K, int(res["nb_ids"][g])))
    checked = 0
    for i, j, A, B in res["segments"]:
        if (int(i), int(j)) in amb_circ:
            continue
        a, b = int(s2i[i]), int(s2i[j])
        d = c[b] - c[a]
        dd = np.linalg.norm(d)
        n = d / dd
        t = (R[a] ** 2 - R[b] ** 2 + dd * dd) / (2 * dd)
        ctr = c[a] + n * t
        pa = res["concave_pos64"][A] - ctr
        pb = res["concave_pos64"][B] - ctr
        ang = np.arctan2(np.dot(np.cross(pa, pb), n), np.dot(pa, pb)) % (2 * np.pi)
        mid_angle = ang / 2
        # rotate pa by mid_angle about n (Rodrigues)
        mid = pa * np.cos(mid_angle) + np.cross(n, pa) * np.sin(mid_angle) + n * np.dot(n, pa) * (1 - np.cos(mid_angle))
        x = ctr + mid
        m = ((c - x) ** 2).sum(-1) - R * R
        m[[a, b]] = np.inf
        assert m.min() > 0, (a, b, m.min())
        checked += 1
    assert checked > 50


@pytest.mark.parametrize("mode", [0, 1])
def test_invariants_globule(mode):
    atoms = synth.globule(2000, synth.SEED_BASE)
    res = O.build(atoms, mode=mode, precision=64, track=True)
    ids = res["concave_ids"]
    assert np.all(ids[:, 0] < ids[:, 1]) and np.all(ids[:, 1] < ids[:, 2])
    w = np.abs(res["concave_pos"][:, 3]).round().astype(int)
    iw = res.iw
    if mode == 0:
        assert np.all(w[:iw] == 2) and np.all(w[iw:] == 1)
        cnt = res["circle_int_nb"]
        assert cnt.sum() == 3 * len(ids)
        assert np.all(cnt % 2 == 0)
        assert len(res["segments"]) == cnt.sum() // 2 == 3 * len(ids) // 2
        # status symmetry (complete surface evaluates both orientations)
        s2i, K = res["sorted_to_initial"], res.K
        st = {}
        for i in range(len(atoms)):
            for l in range(res["nb_count"][i]):
                st[(i, int(res["nb_ids"][i * K + l]))] = (int(res["status"][i * K + l]), bool(res["amb_circle"][i * K + l]))
        for (i, j), (s, amb) in st.items():
            s2, amb2 = st[(j, i)]
            assert s == s2 or amb or amb2
        # sector angles in [0, pi], convex ranges contiguous
        assert np.all((res["sectors"][:, 3] >= 0) & (res["sectors"][:, 3] <= np.pi + 1e-6))
        cp = res["convex_patches"]
        assert np.all(cp[1:, 0] == cp[:-1, 1]) and cp[-1, 1] == len(res["sectors"])
    else:
        # every torus record is a status-2 circle, i < j
        K = res.K
        for g in res["torus_slots"]:
            i, l = divmod(int(g), K)
            assert res["status"][g] == 2 and res["nb_ids"][g] > i


def test_two_atoms_closed_form():
    atoms = np.array([[0, 0, 0, 1.7], [3.0, 0.0, 0.0, 1.5]], np.float32)
    res = O.build(atoms, mode=0, precision=64, track=True)
    assert len(res["full_circles"]) == 1 and len(res["concave_ids"]) == 0 and len(res["sectors"]) == 2
    Ri, Rj, d = 1.7 + RP, 1.5 + RP, 3.0
    t = (Ri * Ri - Rj * Rj + d * d) / (2 * d)
    # sector half-angle of atom 0: cos = t / Ri
    s2i = res["sorted_to_initial"]
    first = int(np.nonzero(s2i == 0)[0][0])
    q = res["convex_patches"][first, 0]
    assert abs(res["sectors"][q, 3] - np.arccos(t / Ri)) < 1e-6
    ext = O.build(atoms, mode=1, precision=64, track=True)
    tor = ext["torus"].reshape(3, -1, 4)
    assert tor.shape[1] == 1
    assert abs(tor[0, 0, 3] - np.sqrt(Ri * Ri - t * t)) < 1e-5  # torus radius = circle radius


def test_three_atoms_closed_form():
    atoms = np.array([[0, 0, 0, 1.7], [3.2, 0, 0, 1.7], [1.6, 2.8, 0, 1.52]], np.float32)
    res = O.build(atoms, mode=0, precision=64, track=True)
    assert len(res["concave_ids"]) == 2
    assert len(res["segments"]) == 3
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    want = trilaterate(c[0], R[0], c[1], R[1], c[2], R[2])
    got = res["concave_pos64"]
    for x in want:
        assert np.min(np.linalg.norm(got - x, axis=1)) < 1e-9
    # hasNeighbors iff the probe is closer than rp to the plane of the centres (here z = 0)
    for q in range(2):
        has = abs(res["concave_pos"][q, 3]) == 2
        assert has == (abs(got[q, 2]) < RP)


def test_tetrahedron_fourth_sphere_hides():
    atoms = tetrahedron(3.4, 1.6)
    res = O.build(atoms, mode=0, precision=64, track=True)
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    want = set()
    for i, j, k in itertools.combinations(range(4), 3):
        for x in trilaterate(c[i], R[i], c[j], R[j], c[k], R[k]):
            l = ({0, 1, 2, 3} - {i, j, k}).pop()
            if np.sum((c[l] - x) ** 2) - R[l] ** 2 > 0:
                want.add(parity.concave_key(np.array([i, j, k]), x, np.arange(4), atoms))
    assert set(oracle_sets(res, atoms)) == want
    assert len(want) == 4  # the inner probe of each face is buried by the fourth atom


@pytest.mark.parametrize("mode", [0, 1])
def test_fp32_oracle_matches_fp64_outside_band(mode):
    atoms = synth.globule(2000, synth.SEED_BASE)
    f32 = O.build(atoms, mode=mode, precision=32, track=False)
    f64 = O.build(atoms, mode=mode, precision=64, track=True)
    rep = parity.compare(as_build_arrays(f32), f64, atoms, mode)
    parity.assert_parity(rep)


def test_comparator_detects_planted_mistake():
    """A planted oracle-side mistake (a dropped concave patch) must be reported."""
    atoms = synth.globule(600, synth.SEED_BASE + 9)
    f64 = O.build(atoms, mode=0, precision=64, track=True)
    arr = as_build_arrays(O.build(atoms, mode=0, precision=32, track=False))
    keep = np.ones(len(arr["concave_ids"]), bool)
    keep[len(keep) // 2] = False
    arr["concave_ids"] = arr["concave_ids"][keep]
    arr["concave_pos"] = arr["concave_pos"][keep]
    arr["concave_pos_w"] = arr["concave_pos_w"][keep]
    arr["segments"] = arr["segments"][:0]
    rep = parity.compare(arr, f64, atoms, 0)
    assert rep["mismatch"]["concave"] >= 1


@pytest.mark.parametrize("mode", [0, 1])
def test_neighbor_lists_given_back_reproduce_the_query(mode):
    """The oracle's neighbour-list override (used to check builds truncated at K in another grid
    order) reproduces the grid query's build when given its own lists, and has teeth: lists that
    lost an entry change the result."""
    atoms = synth.globule(800, synth.SEED_BASE + 5)
    a = O.build(atoms, mode=mode, precision=64, track=False)
    s2i, K = a["sorted_to_initial"].astype(np.int64), a.K
    nb = a["nb_ids"].reshape(-1, K)
    cnt = np.zeros(len(atoms), np.uint32)
    ids = np.zeros((len(atoms), K), np.uint32)
    cnt[s2i] = a["nb_count"]
    for i in range(len(atoms)):
        ids[s2i[i], : a["nb_count"][i]] = s2i[nb[i, : a["nb_count"][i]]]
    b = O.build(atoms, mode=mode, precision=64, track=False, neighbors=(cnt, ids))
    for name in ("nb_count", "nb_ids", "status", "concave_ids", "concave_pos"):
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)
    short = cnt.copy()
    short[short > 0] -= 1
    c = O.build(atoms, mode=mode, precision=64, track=False, neighbors=(short, ids))
    assert not np.array_equal(a["concave_ids"], c["concave_ids"])
    # another sorted order (untruncated lists): the same surface, in other indices
    order = np.random.default_rng(2).permutation(len(atoms))
    d = O.build(atoms, mode=mode, precision=64, track=True, order=order)
    assert np.array_equal(d["sorted_to_initial"], order)
    rep = parity.compare(as_build_arrays(d), O.build(atoms, mode=mode, precision=64, track=True), atoms, mode)
    parity.assert_parity(rep)
    assert sum(rep["band"].values()) == 0
