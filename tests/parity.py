"""Canonical comparison of a GPU build against the FP64 oracle (test helper).

Both sides are mapped to INPUT atom indices (sorted -> input through sorted_to_initial); sets are
compared as sets, and every difference must be attributed to an element the oracle classified as
near-degenerate (epsilon band). Returns a report dict; ``assert_parity`` enforces the contract:
  * neighbour pairs, circle statuses, full circles, concave patches (keyed by sorted input triple and
    side), hasNeighbors flags, segments (keyed by circle and endpoint keys, oriented), probe
    neighbour sets (count < max): bit-exact outside the band;
  * probe positions within 1e-4 A of the FP64 oracle; sector axis/angle within 2e-5 outside the band
    (circles of near-tangent or near-degenerate pairs: reported in sector_band_max_err).
"""
import numpy as np

POS_TOL = 1e-4
SECTOR_TOL = 2e-5


def _pairs(nb_count, nb_ids, s2i, K):
    out = {}
    for i in range(len(nb_count)):
        n = int(nb_count[i])
        for l in range(n):
            out[(int(s2i[i]), int(s2i[nb_ids[i * K + l] if nb_ids.ndim == 1 else nb_ids[i, l]]))] = (i, l)
    return out


def concave_key(ids_sorted, pos, s2i, atoms_in):
    """(a, b, c, side) with a < b < c input indices, side = sign det(cb - ca, cc - ca, x - ca)."""
    tri = sorted(int(s2i[int(v)]) for v in ids_sorted[:3])
    ca, cb, cc = (atoms_in[t, :3].astype(np.float64) for t in tri)
    x = np.asarray(pos[:3], np.float64)
    side = 1 if np.dot(np.cross(cb - ca, cc - ca), x - ca) > 0 else -1
    return (tri[0], tri[1], tri[2], side)


def compare(gpu, orc, atoms_in, mode):
    """gpu: dict of arrays from the GPU build (see gpu_arrays); orc: oracle.OracleResult (FP64, tracked)."""
    rep = {"mismatch": {}, "band": {}}
    K_g, K_o = gpu["K"], orc.K
    s2i_g, s2i_o = gpu["s2i"], orc["sorted_to_initial"]

    # -- neighbour pairs ------------------------------------------------------------------------
    pg = _pairs(gpu["nb_count"], gpu["nb_ids"].reshape(-1), s2i_g, K_g)
    po = _pairs(orc["nb_count"], orc["nb_ids"], s2i_o, K_o)
    amb_pairs = {tuple(sorted((int(s2i_o[a]), int(s2i_o[b])))) for a, b in orc["amb_pairs"]}
    diff = set(pg) ^ set(po)
    bad = [d for d in diff if tuple(sorted(d)) not in amb_pairs]
    rep["mismatch"]["neighbors"] = len(bad)
    rep["band"]["neighbors"] = len(diff) - len(bad)
    rep["neighbor_pairs"] = len(po)

    # -- circle status ----------------------------------------------------------------------------
    st_g, st_o = gpu["status"].reshape(-1), orc["status"]
    amb_c = orc["amb_circle"]
    amb_circles = set()
    for (a, b), (i, l) in po.items():
        if amb_c[i * K_o + l]:
            amb_circles.add(tuple(sorted((a, b))))
    for p in amb_pairs:
        amb_circles.add(p)

    def status_map(pairs, st, K):
        m = {}
        for (a, b), (i, l) in pairs.items():
            m[(a, b)] = int(st[i * K + l])
        return m

    sg, so = status_map(pg, st_g, K_g), status_map(po, st_o, K_o)
    if mode == 1:
        # exterior path: each unordered circle is evaluated once (from its lower sorted index)
        def und(m):
            u = {}
            for (a, b), v in m.items():
                k = (min(a, b), max(a, b))
                u[k] = max(u.get(k, 0), v)
            return u
        sg, so = und(sg), und(so)
    # exterior path: a partially visible circle that holds a concave patch is promoted to status 2
    # (ssesdf operations.cu:301-303), so a 1/2 difference on a circle of an ambiguous patch is in band
    promo_amb = set()
    if mode == 1:
        co_ids, co_pos = orc["concave_ids"], orc["concave_pos64"]
        amb_q = np.nonzero(orc["amb_concave"])[0]
        tris = [tuple(concave_key(co_ids[q], co_pos[q], s2i_o, atoms_in)[:3]) for q in amb_q]
        tris += [tuple(sorted((int(s2i_o[i]), int(s2i_o[j]), int(s2i_o[k])))) for i, j, k, _ in orc["amb_rejected"]]
        for t in tris:
            promo_amb.update({(t[0], t[1]), (t[0], t[2]), (t[1], t[2])})
    nbad = nband = 0
    for k in set(sg) & set(so):
        if sg[k] != so[k]:
            if tuple(sorted(k)) in amb_circles:
                nband += 1
            elif {sg[k], so[k]} == {1, 2} and tuple(sorted(k)) in promo_amb:
                nband += 1
            else:
                nbad += 1
    rep["mismatch"]["status"] = nbad
    rep["band"]["status"] = nband
    rep["status_hist"] = np.bincount(np.array(list(so.values()), np.int64), minlength=3).tolist() if so else [0, 0, 0]

    # -- concave patches --------------------------------------------------------------------------
    def conc(ids, pos, s2i, atoms):
        d = {}
        for q in range(len(ids)):
            d[concave_key(ids[q], pos[q], s2i, atoms_in)] = q
        return d

    cg = conc(gpu["concave_ids"], gpu["concave_pos"], s2i_g, None)
    co = conc(orc["concave_ids"], orc["concave_pos64"], s2i_o, None)
    amb_acc = {k for k, q in co.items() if orc["amb_concave"][q]}
    amb_rej = set()
    for i, j, k, side in orc["amb_rejected"]:
        tri = sorted((int(s2i_o[i]), int(s2i_o[j]), int(s2i_o[k])))
        amb_rej.add(tuple(tri))
    bad_missing = [k for k in set(co) - set(cg) if k not in amb_acc]
    bad_extra = [k for k in set(cg) - set(co) if tuple(k[:3]) not in amb_rej and k not in amb_acc]
    rep["mismatch"]["concave"] = len(bad_missing) + len(bad_extra)
    rep["band"]["concave"] = len(set(co) ^ set(cg)) - rep["mismatch"]["concave"]
    rep["concave_nb"] = len(co)
    # positions and hasNeighbors on the common set
    common = [k for k in set(cg) & set(co)]
    maxerr, hasnb_bad, hasnb_band = 0.0, 0, 0
    for k in common:
        qg, qo = cg[k], co[k]
        err = float(np.max(np.abs(gpu["concave_pos"][qg, :3].astype(np.float64) - orc["concave_pos64"][qo])))
        maxerr = max(maxerr, err)
        hg = abs(int(round(gpu["concave_pos_w"][qg]))) == 2 if "concave_pos_w" in gpu else None
        ho = abs(int(round(orc["concave_pos"][qo, 3]))) == 2
        if hg is not None and hg != ho:
            if orc["amb_hasnb"][qo] or orc["amb_concave"][qo]:
                hasnb_band += 1
            else:
                hasnb_bad += 1
    rep["probe_pos_max_err"] = maxerr
    rep["mismatch"]["has_neighbors"] = hasnb_bad
    rep["band"]["has_neighbors"] = hasnb_band

    # -- complete-surface extras ------------------------------------------------------------------
    if mode == 0:
        fg = {tuple(sorted((int(s2i_g[a]), int(s2i_g[b])))) for a, b in gpu["full_circles"]}
        fo = {tuple(sorted((int(s2i_o[a]), int(s2i_o[b])))) for a, b in orc["full_circles"]}
        d = fg ^ fo
        rep["mismatch"]["full_circles"] = len([x for x in d if x not in amb_circles])
        rep["band"]["full_circles"] = len(d) - rep["mismatch"]["full_circles"]

        # sectors keyed by ordered (i, j)
        def sectors(nb_count, nb_ids, st, cp, sec, s2i, K):
            out = {}
            for i in range(len(nb_count)):
                q = int(cp[i, 0])
                for l in range(int(nb_count[i])):
                    if st[i * K + l] != 0:
                        out[(int(s2i[i]), int(s2i[nb_ids[i * K + l]]))] = sec[q]
                        q += 1
                assert q == int(cp[i, 1])
            return out

        sec_g = sectors(gpu["nb_count"], gpu["nb_ids"].reshape(-1), st_g, gpu["convex_patches"], gpu["sectors"], s2i_g, K_g)
        sec_o = sectors(orc["nb_count"], orc["nb_ids"], st_o, orc["convex_patches"], orc["sectors"], s2i_o, K_o)
        serr = berr = 0.0
        for k in set(sec_g) & set(sec_o):
            e = float(np.max(np.abs(sec_g[k].astype(np.float64) - sec_o[k].astype(np.float64))))
            # near-tangent circles (epsilon band): cap angle ill-conditioned in FP32, reported only
            if tuple(sorted(k)) in amb_circles:
                berr = max(berr, e)
            else:
                serr = max(serr, e)
        rep["sector_max_err"] = serr
        rep["sector_band_max_err"] = berr

        # segments keyed by circle + oriented endpoint keys
        def segs(seg, ids, pos, s2i):
            keys = [concave_key(ids[q], pos[q], s2i, atoms_in) for q in range(len(ids))]
            out = set()
            for i, j, A, B in seg:
                a, b = int(s2i[i]), int(s2i[j])
                kA, kB = keys[A], keys[B]
                if a > b:
                    a, b, kA, kB = b, a, kB, kA
                out.add((a, b, kA, kB))
            return out

        sgg = segs(gpu["segments"], gpu["concave_ids"], gpu["concave_pos"], s2i_g)
        sgo = segs(orc["segments"], orc["concave_ids"], orc["concave_pos64"], s2i_o)
        amb_seg = set()
        t2g = orc["trimmed_to_global"]
        for t in np.nonzero(orc["amb_seg_circle"])[0]:
            g = int(t2g[t])
            i, l = divmod(g, K_o)
            amb_seg.add(tuple(sorted((int(s2i_o[i]), int(s2i_o[orc["nb_ids"][g]])))))
        d = sgg ^ sgo
        rep["mismatch"]["segments"] = len([x for x in d if (x[0], x[1]) not in amb_seg and (x[0], x[1]) not in amb_circles])
        rep["band"]["segments"] = len(d) - rep["mismatch"]["segments"]
        rep["segment_nb"] = len(sgo)

    # -- probe neighbours (with-neighbour probes, count < max) ---------------------------------------
    maxnb = gpu["max_nb"]

    def probe_sets(ids, pos, nb, s2i, posw_is_w):
        keys = [concave_key(ids[q], pos[q], s2i, atoms_in) for q in range(len(ids))]
        P = np.array([p[:3] for p in pos], np.float64) if len(pos) else np.zeros((0, 3))
        out = {}
        iw = sum(1 for q in range(len(ids)) if int(ids[q][3]) >= 0 and q < nb.shape[0] // maxnb)
        for q in range(nb.shape[0] // maxnb):
            cnt = int(ids[q][3])
            if cnt >= maxnb:
                continue
            s = set()
            for e in range(cnt):
                v = nb[q * maxnb + e, :3].astype(np.float64)
                d2 = np.sum((P - v) ** 2, axis=1)
                s.add(keys[int(np.argmin(d2))])
            out[keys[q]] = s
        return out

    ng = probe_sets(gpu["concave_ids"], gpu["concave_pos"], gpu["concave_nb"], s2i_g, True)
    no = probe_sets(orc["concave_ids"], orc["concave_pos64"], orc["concave_nb"], s2i_o, True)
    amb_probe = {concave_key(orc["concave_ids"][q], orc["concave_pos64"][q], s2i_o, atoms_in)
                 for q in np.nonzero(orc["amb_probe_nb"])[0]}
    nbad = nband = 0
    for k in set(ng) & set(no):
        if ng[k] != no[k]:
            if k in amb_probe or k in amb_acc:
                nband += 1
            else:
                nbad += 1
    rep["mismatch"]["probe_neighbors"] = nbad
    rep["band"]["probe_neighbors"] = nband
    return rep


def gpu_arrays(surf):
    """Collect the arrays of a sesdf_amd.Sesdf / Ssesdf build for ``compare``."""
    counts, ids = surf.get_neighbors()
    data = surf.get_data()
    out = {
        "K": surf.get_max_neighbor_per_atom(),
        "s2i": surf.get_atom_indices(),
        "nb_count": counts,
        "nb_ids": ids,
        "status": surf.get_circle_status(),
        "concave_ids": data["concave_patches_id"],
        "concave_pos": data["concave_patches_position"],
        "concave_nb": data["concave_patches_neighbors"],
        "max_nb": int(data["max_concave_neighbors"]),
        "counts": surf.counts(),
    }
    if "sectors" in data:
        out.update(full_circles=data["circle_patches"], sectors=data["sectors"], convex_patches=data["convex_patches"],
                   segments=data["segment_patches"], concave_pos_w=data["concave_patches_position"][:, 3])
    else:
        out.update(torus_pos=data["circle_position"], torus_axis=data["circle_axis"], torus_vs=data["circle_vs"],
                   atoms_sorted=data["atoms"])
    return out


def assert_parity(rep):
    bad = {k: v for k, v in rep["mismatch"].items() if v}
    assert not bad, f"topology mismatches outside the epsilon band: {bad} (report {rep})"
    assert rep["probe_pos_max_err"] <= POS_TOL, rep
    if "sector_max_err" in rep:
        assert rep["sector_max_err"] <= SECTOR_TOL, rep


def torus_reference(ai, aj, rp):
    """FP64 restatement of fillCircles (ssesdf/src/ssesdf/operations.cu:246-285) for one circle."""
    ci, cj = ai[:3].astype(np.float64), aj[:3].astype(np.float64)
    Ri, Rj = float(ai[3]) + rp, float(aj[3]) + rp
    d = np.sum((ci - cj) ** 2)
    t = (Ri * Ri - Rj * Rj + d) / (2 * d)
    centre = ci + (cj - ci) * t
    R = np.sqrt(Ri * Ri - np.sum((ci - centre) ** 2))
    ta = (centre - ci) / np.linalg.norm(centre - ci)
    ortho = np.cross(ta, [0.0, 0.0, 1.0])
    ortho /= np.linalg.norm(ortho)
    P = centre + ortho * R
    X = (P - ci) / np.linalg.norm(P - ci) * float(ai[3])
    C = (np.linalg.norm(P - ci) / (np.linalg.norm(P - cj) + np.linalg.norm(P - ci))) * (cj - ci)
    dist = np.linalg.norm(X - C)
    C = (C + ci) - centre
    return np.concatenate([centre, [R]]), np.concatenate([ta, [rp]]), np.concatenate([C, [dist]])
