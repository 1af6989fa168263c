"""Canonical comparison of Contour-Buildup builds (GPU or oracle) against the oracle.

Every build is mapped to input atom ids. Probes are keyed by (sorted input triple, side), side =
sign det(cj - ci, ck - ci, x - ci); tori by the input pair. Probes and tori written from a circle
whose decisions fall in the oracle's epsilon band (amb_cover / amb_circle), or from an atom whose
neighbour list or contour decision does (amb_atom), are excluded from the exact comparison and
counted.
"""
from collections import defaultdict

import numpy as np


def side(A, t, x):
    ci, cj, ck = A[t[0], :3], A[t[1], :3], A[t[2], :3]
    v = np.sign(np.dot(np.cross(cj - ci, ck - ci), np.asarray(x, np.float64)[:3] - ci))
    return 0 if np.isnan(v) else int(v)  # NaN probes (degenerate circles) keep a key of their own


def excluded_circles(orc):
    """(i, j) sorted-index pairs whose outputs are not compared exactly. A covered-circle decision
    in the band excludes the whole atom: a circle that is kept in one build and removed in the other
    still cuts the atom's other circles, so their arc ends can change the atom they name."""
    ids, cnt = orc["nb_ids"], orc["nb_count0"]
    amb = orc["amb_circle"].astype(bool)
    cover = orc["amb_cover"].astype(bool)
    atom = orc["amb_atom"].astype(bool) | np.array([cover[i, :cnt[i]].any() for i in range(len(cnt))], bool)
    out = set()
    for i in range(len(cnt)):
        n = int(cnt[i])
        if atom[i]:
            out.update((i, int(j)) for j in ids[i, :n])
            out.add((i, -1))
        else:
            out.update((i, int(ids[i, q])) for q in np.nonzero(amb[i, :n])[0])
    return out


def probe_table(A, s2i, ids, pos, skip=()):
    """key -> sorted list of positions; probes from circles in skip are left out (and counted)."""
    tab = defaultdict(list)
    dropped = 0
    for (i, j, k, _), x in zip(ids, pos):
        if (int(i), int(j)) in skip or (int(i), -1) in skip:
            dropped += 1
            continue
        t = tuple(sorted((int(s2i[i]), int(s2i[j]), int(s2i[k]))))
        tab[t + (side(A, t, x),)].append(np.asarray(x[:3], np.float64))
    for v in tab.values():
        v.sort(key=lambda p: tuple(p))
    return tab, dropped


def torus_table(s2i, ids, pos, axis, vs, skip=()):
    tab = {}
    dropped = 0
    for (i, j), p, a, v in zip(ids, pos, axis, vs):
        if (int(i), int(j)) in skip or (int(i), -1) in skip:
            dropped += 1
            continue
        tab[(int(s2i[i]), int(s2i[j]))] = np.concatenate([p, a, v]).astype(np.float64)
    return tab, dropped


def compare(gpu, orc, atoms):
    """gpu: dict with sorted_to_initial, neighbor_count, neighbor_ids, probe_atoms, probe_positions,
    torus_atoms, torus_positions, torus_axes, torus_visibility_spheres (host arrays);
    orc: FP64 oracle build with tracking. Returns a report dict."""
    A = np.asarray(atoms, np.float64).reshape(-1, 4)
    rep = {}
    s2i_g, s2i_o = gpu["sorted_to_initial"], orc["sorted_to_initial"]
    rep["order_equal"] = bool(np.array_equal(s2i_g, s2i_o))
    # neighbour lists in input ids, atoms outside the band
    amb_atom = orc["amb_atom"].astype(bool)
    cnt_o, ids_o = orc["nb_count"], orc["nb_ids"]
    cnt_g, ids_g = gpu["neighbor_count"], gpu["neighbor_ids"]
    bad_nb = 0
    for s in range(len(cnt_o)):
        if amb_atom[s]:
            continue
        if cnt_o[s] != cnt_g[s] or not np.array_equal(ids_o[s, :cnt_o[s]], ids_g[s, :cnt_g[s]]):
            bad_nb += 1
    rep["neighbor_mismatch"] = bad_nb
    rep["amb_atoms"] = int(amb_atom.sum())
    skip = excluded_circles(orc) if rep["order_equal"] else set()
    rep["excluded_circles"] = len(skip)
    tg, dg = probe_table(A, s2i_g, gpu["probe_atoms"], gpu["probe_positions"], skip)
    to, do = probe_table(A, s2i_o, orc["probe_ids"], orc["probe_pos64"], skip)
    rep["probes"] = (len(gpu["probe_atoms"]), len(orc["probe_ids"]))
    rep["probes_excluded"] = (dg, do)
    miss = 0
    errs = []
    for key in set(tg) | set(to):
        a, b = tg.get(key, []), to.get(key, [])
        if len(a) != len(b):
            miss += abs(len(a) - len(b))
            continue
        errs.extend(float(np.abs(p - q).max()) for p, q in zip(a, b))
    rep["probe_mismatch"] = miss
    rep["probe_pos_max_err"] = max(errs, default=0.0)
    rep["probe_pos_p99_err"] = float(np.percentile(errs, 99)) if errs else 0.0
    ug, _ = torus_table(s2i_g, gpu["torus_atoms"], gpu["torus_positions"], gpu["torus_axes"],
                        gpu["torus_visibility_spheres"], skip)
    uo, _ = torus_table(s2i_o, orc["torus_ids"], orc["torus_pos"], orc["torus_axis"], orc["torus_vs"], skip)
    rep["tori"] = (len(gpu["torus_atoms"]), len(orc["torus_ids"]))
    rep["torus_mismatch"] = len(set(ug) ^ set(uo))
    rep["torus_max_err"] = max((float(np.abs(ug[k] - uo[k]).max()) for k in set(ug) & set(uo)), default=0.0)
    return rep


def as_gpu(res):
    """An oracle build in the form compare() takes for the GPU side."""
    return {"sorted_to_initial": res["sorted_to_initial"], "neighbor_count": res["nb_count"],
            "neighbor_ids": res["nb_ids"], "probe_atoms": res["probe_ids"], "probe_positions": res["probe_pos"],
            "torus_atoms": res["torus_ids"], "torus_positions": res["torus_pos"], "torus_axes": res["torus_axis"],
            "torus_visibility_spheres": res["torus_vs"]}


def assert_parity(rep, pos_tol=1e-3, pos_p99_tol=2e-5, torus_tol=2e-4):
    """Topology exact outside the band. Probe positions: the reference places a probe at the angle of
    its arc end on the small circle (atan2, then cos / sin), which in FP32 loses up to ~1e-4 A on
    nearly tangent cuts (the FP32 restatement shows the same against FP64); the 99th percentile is
    held to pos_p99_tol (callers scale it by the FP32 restatement's own error on the input)."""
    assert rep["order_equal"], rep
    assert rep["neighbor_mismatch"] == 0, rep
    assert rep["probe_mismatch"] == 0, rep
    assert rep["torus_mismatch"] == 0, rep
    assert rep["probe_pos_max_err"] <= pos_tol, rep
    assert rep["probe_pos_p99_err"] <= pos_p99_tol, rep
    assert rep["torus_max_err"] <= torus_tol, rep
