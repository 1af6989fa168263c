"""Pins of the Contour-Buildup oracle (oracle/cb_oracle.cpp), CPU only.

What fixes each part, independently of the restatement itself:
  * closed forms: two atoms (one circle, no probe), three atoms (two trilateration points);
  * brute force on small clusters: every triple of extended spheres, both trilateration points,
    kept when no other extended sphere contains them; circles kept when a dense sample of their
    points has one outside every other sphere;
  * brute-force neighbour lists with the reference's stencil of cells [g - c, g + c);
  * the independent SES oracle (sesdf_oracle.cpp), which reaches the same vertices by a different
    algorithm, on inputs whose neighbour lists are complete;
  * brute-force singularity lists.
Each comparison also runs on a copy of the oracle's output with one planted mistake.
"""
import itertools
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import cb_parity  # noqa: E402
import oracle as O  # noqa: E402
from sesdf_amd import synth  # noqa: E402

RP = 1.4


def cluster(n, seed, spread=4.0):
    """n atoms, centres uniform in a ball of radius spread * n^(1/3), at least 1.2 A apart."""
    rng = np.random.default_rng(seed)
    out = []
    R = spread * n ** (1 / 3)
    while len(out) < n:
        p = rng.uniform(-R, R, 3)
        if np.linalg.norm(p) > R or any(np.linalg.norm(p - q[:3]) < 1.2 for q in out):
            continue
        out.append(np.array([*p, rng.choice([1.52, 1.55, 1.7, 1.8])]))
    return np.array(out, np.float32)


def oracle_probe_keys(res, atoms, pos_name="probe_pos64"):
    tab, _ = cb_parity.probe_table(np.asarray(atoms, np.float64), res["sorted_to_initial"], res["probe_ids"],
                                   res[pos_name])
    return tab


def oracle_tori(res):
    s2i = res["sorted_to_initial"]
    return {tuple(sorted((int(s2i[i]), int(s2i[j])))) for i, j in res["torus_ids"]}


# ---------------------------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("precision", [32, 64])
def test_two_atoms_one_torus_no_probe(precision):
    a = np.array([[0, 0, 0, 1.7], [3, 0, 0, 1.5]], np.float32)
    r = O.cb_build(a, precision=precision)
    assert len(r["probe_pos"]) == 0
    assert len(r["torus_pos"]) == 1
    Ri, Rj, d = 1.7 + RP, 1.5 + RP, 3.0
    xc = (Ri * Ri - Rj * Rj + d * d) / (2 * d)
    pos, axis = r["torus_pos"][0], r["torus_axis"][0]
    np.testing.assert_allclose(pos[:3], [xc, 0, 0], atol=1e-5)
    np.testing.assert_allclose(pos[3], np.sqrt(Ri * Ri - xc * xc), atol=1e-5)
    np.testing.assert_allclose(np.abs(axis[:3]), [1, 0, 0], atol=1e-6)
    assert axis[3] == np.float32(RP)
    # the visibility sphere's centre lies on the axis, between the two atom centres
    vs = r["torus_vs"][0]
    c = pos[:3] + vs[:3]
    assert abs(c[1]) < 1e-5 and abs(c[2]) < 1e-5 and 0.0 < c[0] < 3.0 and vs[3] > 0


@pytest.mark.parametrize("precision", [32, 64])
def test_three_atoms_two_probes(precision):
    s = 3.0
    tri = np.array([[0, 0, 0], [s, 0, 0], [s / 2, s * np.sqrt(3) / 2, 0]])
    a = np.concatenate([tri, np.full((3, 1), 1.6)], axis=1).astype(np.float32)
    r = O.cb_build(a, precision=precision)
    R = 1.6 + RP
    centroid = tri.mean(axis=0)
    h = np.sqrt(R * R - (s / np.sqrt(3)) ** 2)
    got = np.sort(r["probe_pos64"][:, 2])
    assert len(r["probe_pos64"]) == 2 and len(r["torus_pos"]) == 3
    tol = 1e-9 if precision == 64 else 2e-5
    np.testing.assert_allclose(got, [-h, h], atol=tol)
    np.testing.assert_allclose(r["probe_pos64"][:, :2], np.tile(centroid[:2], (2, 1)), atol=tol)
    assert sorted(map(tuple, r["probe_ids"][:, :3].tolist())) == [(0, 1, 2), (0, 1, 2)]
    # spherical-triangle vectors: atom centre minus probe
    for q in range(2):
        p = r["probe_pos"][q]
        i, j, k = r["probe_ids"][q, :3]
        for vec, at in ((r["tria1"][q], i), (r["tria2"][q], j), (r["tria3"][q], k)):
            np.testing.assert_allclose(vec[:3], r["atoms"][at, :3] - p[:3], atol=1e-5)
        assert r["tria3"][q, 3] == np.float32(RP) * np.float32(RP) and r["tria1"][q, 3] == 1.0


# ---------------------------------------------------------------------------------------------
# brute force on small clusters
# ---------------------------------------------------------------------------------------------
def brute_vertices(atoms, rp=RP, margin=1e-6):
    """Exposed SAS vertices: (sorted triple, side) -> position, and the ambiguous keys."""
    A = np.asarray(atoms, np.float64)
    c, R = A[:, :3], A[:, 3] + rp
    out, amb = {}, set()
    for t in itertools.combinations(range(len(A)), 3):
        i, j, k = t
        if any(np.linalg.norm(c[x] - c[y]) >= R[x] + R[y] for x, y in ((i, j), (i, k), (j, k))):
            continue
        # planes |x - c_m|^2 = R_m^2 minus the first: two linear equations, then the line / sphere
        ex = c[j] - c[i]
        ey = c[k] - c[i]
        n = np.cross(ex, ey)
        if np.dot(n, n) < 1e-12:
            continue
        b1 = (R[i] ** 2 - R[j] ** 2 + np.dot(c[j], c[j]) - np.dot(c[i], c[i])) / 2
        b2 = (R[i] ** 2 - R[k] ** 2 + np.dot(c[k], c[k]) - np.dot(c[i], c[i])) / 2
        # x . (cj - ci) = b1, x . (ck - ci) = b2 (differences of the sphere equations), and x in
        # the plane of the three centres
        M = np.array([ex, ey, n])
        x0 = np.linalg.solve(M, [b1, b2, np.dot(n, c[i])])
        nh = n / np.linalg.norm(n)
        t2 = R[i] ** 2 - np.dot(x0 - c[i], x0 - c[i])
        if t2 <= 0:
            continue
        for sgn in (1, -1):
            x = x0 + sgn * np.sqrt(t2) * nh
            others = [m for m in range(len(A)) if m not in t]
            gap = min((np.dot(x - c[m], x - c[m]) - R[m] ** 2 for m in others), default=np.inf)
            key = t + (cb_parity.side(A, t, x),)
            if abs(gap) <= margin or t2 < 1e-6:
                amb.add(key)
            if gap > 0:
                out[key] = x
    return out, amb


def brute_visible_circles(atoms, rp=RP, samples=2048, margin=1e-5):
    A = np.asarray(atoms, np.float64)
    c, R = A[:, :3], A[:, 3] + rp
    vis, amb = set(), set()
    for i, j in itertools.combinations(range(len(A)), 2):
        d = np.linalg.norm(c[j] - c[i])
        if d >= R[i] + R[j] or d <= abs(R[i] - R[j]):
            continue
        n = (c[j] - c[i]) / d
        xc = (R[i] ** 2 - R[j] ** 2 + d * d) / (2 * d)
        r = np.sqrt(R[i] ** 2 - xc * xc)
        u = np.cross(n, [1, 0, 0] if abs(n[0]) < 0.9 else [0, 1, 0])
        u /= np.linalg.norm(u)
        v = np.cross(n, u)
        th = np.linspace(0, 2 * np.pi, samples, endpoint=False)
        pts = c[i] + xc * n + r * (np.cos(th)[:, None] * u + np.sin(th)[:, None] * v)
        others = [m for m in range(len(A)) if m not in (i, j)]
        if not others:
            vis.add((i, j))
            continue
        g = np.min([((pts - c[m]) ** 2).sum(1) - R[m] ** 2 for m in others], axis=0)
        best = g.max()
        if abs(best) <= margin * 100 or (best > 0 and best < 1e-3):
            amb.add((i, j))
        if best > 0:
            vis.add((i, j))
    return vis, amb


@pytest.mark.parametrize("n,seed", [(4, 1), (8, 2), (12, 3), (20, 4), (30, 5), (45, 6)])
def test_probes_and_tori_match_brute_force(n, seed):
    a = cluster(n, 900 + seed, spread=1.6)
    r = O.cb_build(a, grid_dim=1, precision=64)
    assert r["nb_count0"].max() < 64
    got = oracle_probe_keys(r, a)
    exp, amb = brute_vertices(a)
    keys = (set(got) | set(exp)) - amb
    assert {k for k in keys if k in got} == {k for k in keys if k in exp}
    assert all(len(got[k]) == 1 for k in got)
    for k in keys & set(got):
        np.testing.assert_allclose(got[k][0], exp[k], atol=1e-6)
    vis, vamb = brute_visible_circles(a)
    tori = oracle_tori(r)
    assert (tori - vamb) == (vis - vamb)
    assert len(keys) >= (2 if n > 4 else 0)


def test_brute_force_comparison_detects_planted_mistake():
    a = cluster(20, 904, spread=1.6)
    r = O.cb_build(a, grid_dim=1, precision=64)
    got = oracle_probe_keys(r, a)
    exp, amb = brute_vertices(a)
    victim = next(k for k in got if k not in amb)
    got.pop(victim)
    keys = (set(got) | set(exp)) - amb
    assert {k for k in keys if k in got} != {k for k in keys if k in exp}


# ---------------------------------------------------------------------------------------------
# neighbour lists: brute force with the reference's stencil
# ---------------------------------------------------------------------------------------------
def stencil_lists(atoms, rp=RP, dim=16):
    """Per sorted atom, the first 64 atoms j != i with |ci - cj| < ri + rj + 2 rp among the cells
    [g - c, g + c) (clamped), in (z, y, x, sorted index) order. FP32 as the reference."""
    a = np.asarray(atoms, np.float32)
    box = O.aabb_of(a)
    lo, hi = box[:3], box[3:]
    cell = ((hi - lo) / np.float32(dim)).astype(np.float32)
    g = np.floor((a[:, :3] - lo) / cell).astype(np.int64)
    m = dim - 1
    h = ((g[:, 2] & m) * dim + (g[:, 1] & m)) * dim + (g[:, 0] & m)
    order = np.argsort(h, kind="stable")
    S, gs = a[order], g[order]
    lists = []
    for i in range(len(S)):
        rng = np.float32(S[i, 3] + np.float32(3.0) + np.float32(2.0) * np.float32(rp))
        c = np.ceil(rng / cell).astype(np.int64)
        lo_c, hi_c = np.maximum(gs[i] - c, 0), np.minimum(gs[i] + c, dim)
        inb = np.all((gs >= lo_c) & (gs < hi_c), axis=1)
        rel = (S[:, :3] - S[i, :3]).astype(np.float32)
        d = np.sqrt(((rel[:, 0] * rel[:, 0] + rel[:, 1] * rel[:, 1]) + rel[:, 2] * rel[:, 2]).astype(np.float32))
        nd = (S[i, 3] + S[:, 3] + np.float32(2.0) * np.float32(rp)).astype(np.float32)
        ok = inb & (d < nd)
        ok[i] = False
        js = np.nonzero(ok)[0]
        js = js[np.lexsort((js, gs[js, 0], gs[js, 1], gs[js, 2]))]
        lists.append(js[:64])
    return order, lists


@pytest.mark.parametrize("n,seed", [(300, 11), (2000, 0)])
def test_neighbour_lists_brute_force(n, seed):
    a = synth.globule(n, synth.SEED_BASE + (0 if n == 2000 else 3000 + seed))
    r = O.cb_build(a, precision=32)
    order, lists = stencil_lists(a)
    assert np.array_equal(order, r["sorted_to_initial"])
    cnt, ids = r["nb_count0"], r["nb_ids"]
    bad = sum(not np.array_equal(ids[i, :cnt[i]], lists[i]) for i in range(n))
    assert bad == 0
    # the stencil misses the upper row of cells: with cells wider than a third of the search range
    # (the 2,000-atom globule: 2.8 A cells) some pairs closer than ri + rj + 2 rp are absent
    A = r["atoms"]
    full = sum(int(((np.linalg.norm(A[:, :3] - A[i, :3], axis=1) < A[i, 3] + A[:, 3] + 2 * RP).sum() - 1)) for i in range(n))
    assert cnt.sum() < full if n == 2000 else cnt.sum() <= full
    # planted mistake: one list loses an entry
    planted = [l.copy() for l in lists]
    q = next(i for i in range(n) if len(planted[i]) > 3)
    planted[q] = planted[q][:-1]
    assert sum(not np.array_equal(ids[i, :cnt[i]], planted[i]) for i in range(n)) == 1


# ---------------------------------------------------------------------------------------------
# agreement with the independent SES oracle where the neighbour lists are complete
# ---------------------------------------------------------------------------------------------
def ses_keys(s, atoms):
    A = np.asarray(atoms, np.float64)
    s2i = s["sorted_to_initial"]
    tab = {}
    for (i, j, k, _), x in zip(s["concave_ids"], s["concave_pos64"]):
        t = tuple(sorted((int(s2i[i]), int(s2i[j]), int(s2i[k]))))
        tab.setdefault(t + (cb_parity.side(A, t, x),), []).append(x)
    return tab


@pytest.mark.parametrize("n,seed", [(60, 0), (120, 1), (200, 2)])
def test_matches_ses_oracle_with_complete_lists(n, seed):
    a = synth.globule(n, synth.SEED_BASE + 100 + seed)
    r = O.cb_build(a, grid_dim=1, precision=64)
    s = O.build(a, mode=0, precision=64, max_neighbors=128)
    assert r["nb_count0"].max() < 64
    cp, sp = oracle_probe_keys(r, a), ses_keys(s, a)
    assert set(cp) == set(sp)
    for k in cp:
        np.testing.assert_allclose(cp[k][0], sp[k][0], atol=1e-9)
    # tori: full circles and the circles of visible vertices
    s2i = s["sorted_to_initial"]
    exp = {tuple(sorted((int(s2i[i]), int(s2i[j])))) for i, j in s["full_circles"]}
    for (i, j, k, _sd) in sp:
        exp |= {(i, j), (i, k), (j, k)}
    assert oracle_tori(r) == exp


# ---------------------------------------------------------------------------------------------
# FP32 restatement against FP64 outside the band; the comparator itself
# ---------------------------------------------------------------------------------------------
as_gpu = cb_parity.as_gpu


@pytest.mark.parametrize("n", [2000, 20000])
def test_fp32_matches_fp64_outside_band(n):
    a = synth.globule(n, synth.SEED_BASE)
    r32 = O.cb_build(a, precision=32, threads=8)
    r64 = O.cb_build(a, precision=64, threads=8)
    rep = cb_parity.compare(as_gpu(r32), r64, a)
    cb_parity.assert_parity(rep)
    assert rep["probes"][1] > n // 2
    assert rep["probes_excluded"][1] < 0.05 * rep["probes"][1]


def test_comparator_detects_planted_mistakes():
    a = synth.globule(2000, synth.SEED_BASE)
    r32 = O.cb_build(a, precision=32)
    r64 = O.cb_build(a, precision=64)
    g = as_gpu(r32)
    skip = cb_parity.excluded_circles(r64)
    q = next(q for q in range(len(g["probe_atoms"])) if (int(g["probe_atoms"][q, 0]), int(g["probe_atoms"][q, 1])) not in skip)
    dropped = dict(g, probe_atoms=np.delete(g["probe_atoms"], q, 0), probe_positions=np.delete(g["probe_positions"], q, 0))
    assert cb_parity.compare(dropped, r64, a)["probe_mismatch"] == 1
    moved = dict(g, probe_positions=g["probe_positions"].copy())
    moved["probe_positions"][q, 0] += 1e-3
    assert cb_parity.compare(moved, r64, a)["probe_pos_max_err"] > 1e-4
    t = next(t for t in range(len(g["torus_atoms"])) if (int(g["torus_atoms"][t, 0]), int(g["torus_atoms"][t, 1])) not in skip)
    bent = dict(g, torus_axes=g["torus_axes"].copy())
    bent["torus_axes"][t, 1] += 1e-3
    assert cb_parity.compare(bent, r64, a)["torus_max_err"] > 2e-4
    ids = g["neighbor_ids"].copy()
    ids[0, 0] += 1
    assert cb_parity.compare(dict(g, neighbor_ids=ids), r64, a)["neighbor_mismatch"] == 1


# ---------------------------------------------------------------------------------------------
# singularity texture: brute force
# ---------------------------------------------------------------------------------------------
def test_singularities_brute_force():
    a = synth.globule(2000, synth.SEED_BASE)
    r = O.cb_build(a, precision=32)
    P = r["probe_pos"]
    box = O.aabb_of(a)
    lo, cell = box[:3], ((box[3:] - box[:3]) / np.float32(16)).astype(np.float32)
    g = np.floor((P[:, :3] - lo) / cell).astype(np.int64)
    order = r["probe_sorted"]
    rank = np.empty(len(P), np.int64)
    rank[order] = np.arange(len(P))
    nd = np.float32(RP) + np.float32(RP)
    c = np.ceil(nd / cell).astype(np.int64)
    tc, st = r["tex_coord"], r["sing_tex"]
    checked = 0
    for q in range(len(P)):
        lo_c, hi_c = np.maximum(g[q] - c, 0), np.minimum(g[q] + c, 16)
        inb = np.all((g >= lo_c) & (g < hi_c), axis=1)
        rel = (P[:, :3] - P[q, :3]).astype(np.float32)
        d = np.sqrt(((rel[:, 0] * rel[:, 0] + rel[:, 1] * rel[:, 1]) + rel[:, 2] * rel[:, 2]).astype(np.float32))
        ok = inb & (d < nd)
        ok[q] = False
        cnt = int(tc[q, 0])
        assert tc[q, 1] == rank[q] // 32768 and tc[q, 2] == rank[q] % 32768
        assert cnt == min(int(ok.sum()), 32)
        if cnt < 32:
            got = {tuple(v) for v in st[rank[q], :3 * cnt].reshape(-1, 3)}
            assert got == {tuple(v) for v in P[ok, :3]}
            checked += 1
    assert checked > 100
