"""Vectorised canonical comparison of a GPU build against the FP64 oracle, for full-size workloads
(test helper; the same contract as ``parity.compare`` for the rows it covers).

Covered: neighbour pairs, circle statuses (unordered in the exterior path), full circles, sector
values and segments (complete surface), torus records keyed by circle (exterior path; values where
both sides built a record from the same head atom), concave patches keyed by (sorted input triple,
side) with their probe positions and hasNeighbors flags, and the probe-neighbour set of every
with-neighbour probe whose list is shorter than the capacity (neighbours identified by their
positions, which are copies of the probes'). Every difference must fall in the oracle's epsilon
band; ``band_atoms`` lists the atoms of the differences that do (input indices).

``keep`` (input indices) restricts neighbour lists to kept atoms and circles to those whose two atoms
are kept, and concave patches to those whose three atoms are all in ``keep_patches``: a build on a ball cut out of
a larger system is exact for atoms far enough inside the ball (see test_gpu_full_size.py).
"""
import numpy as np

POS_TOL = 1e-4
SECTOR_TOL = 2e-5


def _slots(nb_count, nb_ids, K, rows=None):
    """(row, slot) of the written neighbour slots of the given rows (all rows when None)."""
    ids = nb_ids.reshape(-1, K)
    if rows is None:
        rows = np.arange(ids.shape[0])
    cnt = nb_count[rows].astype(np.int64)
    r = np.repeat(rows, cnt)
    start = np.cumsum(cnt) - cnt
    l = np.arange(int(cnt.sum())) - np.repeat(start, cnt)
    return r, l, ids[r, l]


def _ukey(a, b, n):
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    return lo * n + hi


_TRI = np.dtype([("a", "i8"), ("b", "i8"), ("c", "i8"), ("s", "i1")])
_SEG = np.dtype([("u", "i8"), ("Aa", "i8"), ("Ab", "i8"), ("Ac", "i8"), ("As", "i1"), ("Ba", "i8"), ("Bb", "i8"), ("Bc", "i8"), ("Bs", "i1")])


def concave_keys(ids, pos, s2i, atoms_in):
    """Structured keys (a < b < c input indices, side = sign det(cb - ca, cc - ca, x - ca))."""
    tri = np.sort(s2i[ids[:, :3].astype(np.int64)].astype(np.int64), axis=1)
    ca, cb, cc = (atoms_in[tri[:, m], :3].astype(np.float64) for m in range(3))
    det = np.einsum("ij,ij->i", np.cross(cb - ca, cc - ca), np.asarray(pos[:, :3], np.float64) - ca)
    k = np.empty(len(tri), _TRI)
    k["a"], k["b"], k["c"], k["s"] = tri[:, 0], tri[:, 1], tri[:, 2], np.where(det > 0, 1, -1)
    return k


def _tri_only(k):
    t = np.empty(len(k), _TRI)
    t["a"], t["b"], t["c"], t["s"] = k["a"], k["b"], k["c"], 0
    return t


def compare_fast(gpu, orc, atoms_in, mode, orc_to_input=None, keep=None, keep_patches=None, log=None):
    """gpu: ``parity.gpu_arrays`` of the build; orc: oracle result (FP64, tracked). orc_to_input
    maps the oracle's input indices to those of ``atoms_in`` (a build on a subset). log: called with
    a line after each part (full-size runs print progress)."""
    say = log or (lambda msg: None)
    n_in = len(atoms_in)
    rep = {"mismatch": {}, "band": {}, "details": {}}
    band_atoms = []  # atoms of every difference put down to the band
    K_g, K_o = gpu["K"], orc.K
    s2i_g = gpu["s2i"].astype(np.int64)
    s2i_o = orc["sorted_to_initial"].astype(np.int64)
    if orc_to_input is not None:
        s2i_o = np.asarray(orc_to_input, np.int64)[s2i_o]
    keep_mask = None
    if keep is not None:
        keep_mask = np.zeros(n_in, bool)
        keep_mask[np.asarray(keep, np.int64)] = True
    rows_g = None
    if keep_mask is not None:
        i2s_g = np.empty(n_in, np.int64)
        i2s_g[s2i_g] = np.arange(n_in)
        rows_g = np.sort(i2s_g[np.nonzero(keep_mask)[0]])
    rows_o = None if keep_mask is None else np.nonzero(keep_mask[s2i_o])[0]

    # -- neighbour pairs (ordered, input indices) ----------------------------------------------------
    rg, lg, jg = _slots(gpu["nb_count"], gpu["nb_ids"], K_g, rows_g)
    ro, lo_, jo = _slots(orc["nb_count"], orc["nb_ids"], K_o, rows_o)
    kg = s2i_g[rg] * n_in + s2i_g[jg.astype(np.int64)]
    ko = s2i_o[ro] * n_in + s2i_o[jo.astype(np.int64)]
    ap = orc["amb_pairs"].reshape(-1, 2).astype(np.int64)
    amb_pairs = _ukey(s2i_o[ap[:, 0]], s2i_o[ap[:, 1]], n_in) if len(ap) else np.zeros(0, np.int64)
    diff = np.setxor1d(kg, ko)
    in_band = np.isin(_ukey(diff // n_in, diff % n_in, n_in), amb_pairs)
    band_atoms += [diff[in_band] // n_in, diff[in_band] % n_in]
    rep["mismatch"]["neighbors"] = int((~in_band).sum())
    rep["band"]["neighbors"] = int(in_band.sum())
    rep["neighbor_pairs"] = int(len(ko))

    say("neighbour pairs compared")
    # -- circle statuses ------------------------------------------------------------------------------
    st_g = gpu["status"].reshape(-1, K_g)[rg, lg].astype(np.int64)
    st_o = orc["status"].reshape(-1, K_o)[ro, lo_].astype(np.int64)
    amb_c = orc["amb_circle"].reshape(-1, K_o)[ro, lo_].astype(bool)
    if keep_mask is not None:
        # a circle is compared when both of its atoms are kept (either side may evaluate it)
        fg_, fo_ = keep_mask[kg % n_in], keep_mask[ko % n_in]
        kg, st_g = kg[fg_], st_g[fg_]
        ko, st_o, amb_c = ko[fo_], st_o[fo_], amb_c[fo_]
    amb_circles = np.union1d(_ukey(ko[amb_c] // n_in, ko[amb_c] % n_in, n_in), amb_pairs)

    def reduce_unordered(k, st):
        u = _ukey(k // n_in, k % n_in, n_in)
        order = np.argsort(u, kind="stable")
        u, st = u[order], st[order]
        first = np.r_[True, u[1:] != u[:-1]]
        idx = np.nonzero(first)[0]
        return u[idx], np.maximum.reduceat(st, idx) if len(st) else st

    if mode == 1:
        # the exterior path evaluates each circle once (from its lower sorted index)
        kg_s, sg_v = reduce_unordered(kg, st_g)
        ko_s, so_v = reduce_unordered(ko, st_o)
    else:
        kg_s, sg_v, ko_s, so_v = kg, st_g, ko, st_o
    common, ig, io = np.intersect1d(kg_s, ko_s, return_indices=True)
    neq = sg_v[ig] != so_v[io]
    uc = _ukey(common // n_in, common % n_in, n_in) if mode == 0 else common
    band = np.isin(uc, amb_circles)
    if mode == 1:
        # a partially visible circle that holds an ambiguous patch may be promoted to status 2
        co_k = concave_keys(orc["concave_ids"], orc["concave_pos64"], s2i_o, atoms_in)
        amb_q = orc["amb_concave"].astype(bool)
        tris = [np.stack([co_k["a"][amb_q], co_k["b"][amb_q], co_k["c"][amb_q]], 1)]
        ar = orc["amb_rejected"].reshape(-1, 4).astype(np.int64)
        if len(ar):
            tris.append(np.sort(s2i_o[ar[:, :3]], axis=1))
        t = np.concatenate(tris) if tris else np.zeros((0, 3), np.int64)
        promo = np.concatenate([_ukey(t[:, 0], t[:, 1], n_in), _ukey(t[:, 0], t[:, 2], n_in), _ukey(t[:, 1], t[:, 2], n_in)])
        one_two = np.minimum(sg_v[ig], so_v[io]) == 1
        one_two &= np.maximum(sg_v[ig], so_v[io]) == 2
        band |= one_two & np.isin(uc, promo)
    band_atoms += [uc[neq & band] // n_in, uc[neq & band] % n_in]
    rep["mismatch"]["status"] = int((neq & ~band).sum())
    bad_st = neq & ~band
    rep["details"]["status"] = np.stack([uc[bad_st] // n_in, uc[bad_st] % n_in, sg_v[ig][bad_st], so_v[io][bad_st]], 1)
    rep["band"]["status"] = int((neq & band).sum())

    say("circle statuses compared")
    # -- concave patches ------------------------------------------------------------------------------
    cg = concave_keys(gpu["concave_ids"], gpu["concave_pos"], s2i_g, atoms_in)
    co = concave_keys(orc["concave_ids"], orc["concave_pos64"], s2i_o, atoms_in)
    sel_g = sel_o = None
    if keep_patches is not None:
        kp = np.zeros(n_in, bool)
        kp[np.asarray(keep_patches, np.int64)] = True
        sel_g = kp[cg["a"]] & kp[cg["b"]] & kp[cg["c"]]
        sel_o = kp[co["a"]] & kp[co["b"]] & kp[co["c"]]
    else:
        sel_g = np.ones(len(cg), bool)
        sel_o = np.ones(len(co), bool)
    qg, qo = np.nonzero(sel_g)[0], np.nonzero(sel_o)[0]
    cg_s, co_s = cg[qg], co[qo]
    amb_acc = co_s[orc["amb_concave"][qo].astype(bool)]
    ar = orc["amb_rejected"].reshape(-1, 4).astype(np.int64)
    amb_rej = np.empty(len(ar), _TRI)
    if len(ar):
        t = np.sort(s2i_o[ar[:, :3]], axis=1)
        amb_rej["a"], amb_rej["b"], amb_rej["c"], amb_rej["s"] = t[:, 0], t[:, 1], t[:, 2], 0
    missing = np.setdiff1d(co_s, cg_s)
    extra = np.setdiff1d(cg_s, co_s)
    bad_missing = missing[~np.isin(missing, amb_acc)]
    bad_extra = extra[~(np.isin(_tri_only(extra), amb_rej) | np.isin(extra, amb_acc))]
    for dd in (missing[np.isin(missing, amb_acc)], extra[np.isin(_tri_only(extra), amb_rej) | np.isin(extra, amb_acc)]):
        band_atoms += [dd["a"], dd["b"], dd["c"]]
    rep["mismatch"]["concave"] = int(len(bad_missing) + len(bad_extra))
    rep["band"]["concave"] = int(len(missing) + len(extra)) - rep["mismatch"]["concave"]
    rep["concave_nb"] = int(len(co_s))
    both, jg_, jo_ = np.intersect1d(cg_s, co_s, return_indices=True)
    pg = gpu["concave_pos"][qg[jg_], :3].astype(np.float64)
    po = orc["concave_pos64"][qo[jo_]]
    rep["probe_pos_max_err"] = float(np.max(np.abs(pg - po))) if len(both) else 0.0
    if len(both):
        perr = np.abs(pg - po).max(1)
        worst = np.argsort(-perr)[:8]
        rep["details"]["probe_pos"] = [(tuple(int(both[w][f]) for f in ("a", "b", "c", "s")), float(perr[w]), po[w].tolist())
                                       for w in worst]
    if "concave_pos_w" in gpu:
        hg = np.abs(np.rint(gpu["concave_pos_w"][qg[jg_]])) == 2
        ho = np.abs(np.rint(orc["concave_pos"][qo[jo_], 3])) == 2
        amb_h = (orc["amb_hasnb"][qo[jo_]] | orc["amb_concave"][qo[jo_]]).astype(bool)
        rep["mismatch"]["has_neighbors"] = int(((hg != ho) & ~amb_h).sum())
        rep["band"]["has_neighbors"] = int(((hg != ho) & amb_h).sum())
    else:  # exterior path: the flag is not part of its records
        rep["mismatch"]["has_neighbors"] = rep["band"]["has_neighbors"] = 0

    say("concave patches compared")
    # -- probe neighbours: the set of each with-neighbour probe (lists shorter than the capacity) ---------
    maxnb = int(gpu["max_nb"])
    iw_g = int(gpu["counts"]["concave_with_neighbors_nb"])
    iw_o = int(orc.iw)
    allk = np.unique(np.concatenate([cg, co]))

    def probe_pairs(ids, pos32, nb, keys, iw, sel):
        cnt = ids[:iw, 3].astype(np.int64)
        q = np.repeat(np.arange(iw), cnt)
        e = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        ent = np.ascontiguousarray(nb.reshape(-1, 4)[q * maxnb + e, :3], np.float32)
        pv = np.ascontiguousarray(pos32[:iw, :3], np.float32).view(np.dtype((np.void, 12))).ravel()
        ev = ent.view(np.dtype((np.void, 12))).ravel()
        order = np.argsort(pv)
        loc = np.minimum(np.searchsorted(pv[order], ev), max(iw - 1, 0))
        found = pv[order][loc] == ev if iw else np.zeros(0, bool)
        r = order[loc]
        kid = np.searchsorted(allk, keys[:iw])
        code = kid[q] * len(allk) + np.where(found, kid[r], len(allk) - 1)
        comparable = (cnt < maxnb) & sel[:iw]
        return code, kid, comparable

    if iw_g and iw_o:
        pc_g, kid_g, cmp_g = probe_pairs(gpu["concave_ids"], gpu["concave_pos"], gpu["concave_nb"], cg, iw_g, sel_g)
        pc_o, kid_o, cmp_o = probe_pairs(orc["concave_ids"], orc["concave_pos"], orc["concave_nb"], co, iw_o, sel_o)
        amb_p = (orc["amb_probe_nb"][:iw_o] | orc["amb_concave"][:iw_o]).astype(bool)
        probes = np.intersect1d(kid_g[cmp_g], kid_o[cmp_o & ~amb_p])
        band_probes = np.intersect1d(kid_g[cmp_g], kid_o[cmp_o & amb_p])
        pg_ = pc_g[np.isin(pc_g // len(allk), probes)]
        po_ = pc_o[np.isin(pc_o // len(allk), probes)]
        # a differing neighbour that is a with-neighbour probe on one side only (a patch, or its
        # hasNeighbors flag, that differs in the band) puts the difference in the band
        both_w = np.intersect1d(kid_g, kid_o)
        xor = np.setxor1d(pg_, po_)
        one_side = ~np.isin(xor % len(allk), both_w)
        bad_p = np.unique(xor[~one_side] // len(allk))
        band_extra = np.setdiff1d(np.unique(xor[one_side] // len(allk)), bad_p)
        rep["mismatch"]["probe_neighbors"] = int(len(bad_p))
        rep["details"]["probe_neighbors"] = [(tuple(int(allk[k][f]) for f in ("a", "b", "c", "s")),
                                              sorted(tuple(int(allk[x][f]) for f in ("a", "b", "c", "s")) for x in pg_[pg_ // len(allk) == k] % len(allk)),
                                              sorted(tuple(int(allk[x][f]) for f in ("a", "b", "c", "s")) for x in po_[po_ // len(allk) == k] % len(allk)))
                                             for k in bad_p[:6]]
        bg_ = pc_g[np.isin(pc_g // len(allk), band_probes)]
        bo_ = pc_o[np.isin(pc_o // len(allk), band_probes)]
        rep["band"]["probe_neighbors"] = int(len(np.unique(np.setxor1d(bg_, bo_) // len(allk)))) + int(len(band_extra))
        rep["probe_sets_compared"] = int(len(probes))
    else:
        rep["mismatch"]["probe_neighbors"] = int(iw_g != iw_o)
        rep["band"]["probe_neighbors"] = 0
        rep["probe_sets_compared"] = 0

    say("probe-neighbour sets compared")
    # -- complete surface: full circles and sectors ------------------------------------------------------
    if mode == 0:
        fg, fo = gpu["full_circles"].astype(np.int64), orc["full_circles"].reshape(-1, 2).astype(np.int64)
        ufg = _ukey(s2i_g[fg[:, 0]], s2i_g[fg[:, 1]], n_in)
        ufo = _ukey(s2i_o[fo[:, 0]], s2i_o[fo[:, 1]], n_in)
        if keep_mask is not None:
            ufg = ufg[keep_mask[ufg // n_in] & keep_mask[ufg % n_in]]
            ufo = ufo[keep_mask[ufo // n_in] & keep_mask[ufo % n_in]]
        d = np.setxor1d(ufg, ufo)
        b = np.isin(d, amb_circles)
        band_atoms += [d[b] // n_in, d[b] % n_in]
        rep["mismatch"]["full_circles"] = int((~b).sum())
        rep["band"]["full_circles"] = int(b.sum())

        def sector_table(nb_count, nb_ids, st, cp, sec, s2i, K):
            # the visible slots of each atom in slot order are its sectors [cp[i, 0], cp[i, 1])
            r, l, j = _slots(nb_count, nb_ids, K)
            vis = st.reshape(-1, K)[r, l] != 0
            r, j = r[vis], j[vis]
            cnt = np.bincount(r, minlength=len(nb_count))
            assert np.array_equal(cp[:, 1].astype(np.int64) - cp[:, 0].astype(np.int64), cnt)
            start = np.repeat(cp[:, 0].astype(np.int64), cnt)
            rank = np.arange(len(r)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            return s2i[r] * n_in + s2i[j.astype(np.int64)], sec[start + rank]

        tg, vg = sector_table(gpu["nb_count"], gpu["nb_ids"], gpu["status"], gpu["convex_patches"], gpu["sectors"], s2i_g, K_g)
        to, vo = sector_table(orc["nb_count"], orc["nb_ids"], orc["status"], orc["convex_patches"], orc["sectors"], s2i_o, K_o)
        if keep_mask is not None:
            fg_, fo_ = keep_mask[tg // n_in] & keep_mask[tg % n_in], keep_mask[to // n_in] & keep_mask[to % n_in]
            tg, vg, to, vo = tg[fg_], vg[fg_], to[fo_], vo[fo_]
        both, a, c = np.intersect1d(tg, to, return_indices=True)
        err = np.abs(vg[a].astype(np.float64) - vo[c].astype(np.float64)).max(axis=1) if len(a) else np.zeros(0)
        # a circle of a near-tangent pair has a cap angle that FP32 inputs fix only to ~1e-4 rad
        # (acos near 1 amplifies the rounding of the centre distance): reported, not bounded
        nb = np.isin(_ukey(both // n_in, both % n_in, n_in), amb_circles)
        # Outside the band the bound is SECTOR_TOL plus what FP32 inputs leave open: r^2 of circle
        # (i, j) is known to about 2e-6 A^2 (the rounding of r + rp and of |c_j - c_i|^2), which
        # moves the half angle by 2e-6 / (2 r R_i) (r the circle radius, R_i the SAS radius).
        ii, jj = both // n_in, both % n_in
        dvec = atoms_in[jj, :3].astype(np.float64) - atoms_in[ii, :3].astype(np.float64)
        d = np.linalg.norm(dvec, axis=1)
        Ri_ = atoms_in[ii, 3].astype(np.float64) + float(np.float32(orc.rp))
        Rj_ = atoms_in[jj, 3].astype(np.float64) + float(np.float32(orc.rp))
        tt = (Ri_ ** 2 - Rj_ ** 2 + d ** 2) / (2 * d)
        rr = np.sqrt(np.maximum(Ri_ ** 2 - tt ** 2, 1e-30))
        tol = SECTOR_TOL + 2e-6 / (2 * rr * Ri_)
        rep["sector_excess"] = float((err - tol)[~nb].max()) if (~nb).any() else -1.0
        rep["sector_max_err"] = float(err[~nb].max()) if (~nb).any() else 0.0
        if (~nb).any():
            wq = np.argsort(-np.where(nb, -1.0, err))[:4]
            rep["details"]["sectors"] = [(int(both[q] // n_in), int(both[q] % n_in), float(err[q])) for q in wq]
        rep["sector_band_max_err"] = float(err[nb].max()) if nb.any() else 0.0
        # segments keyed by circle (input pair, lower first) and oriented endpoint patch keys
        def seg_keys(seg, ids, pos, s2i):
            seg = seg.reshape(-1, 4).astype(np.int64)
            ck = concave_keys(ids, pos, s2i, atoms_in)
            a, b = s2i[seg[:, 0]], s2i[seg[:, 1]]
            kA, kB = ck[seg[:, 2]], ck[seg[:, 3]]
            sw = a > b
            a2, b2 = np.where(sw, b, a), np.where(sw, a, b)
            kA2, kB2 = np.where(sw, kB, kA), np.where(sw, kA, kB)
            k = np.empty(len(seg), _SEG)
            k["u"] = a2 * n_in + b2
            for f in ("a", "b", "c", "s"):
                k["A" + f], k["B" + f] = kA2[f], kB2[f]
            return k

        sgg = seg_keys(gpu["segments"], gpu["concave_ids"], gpu["concave_pos"], s2i_g)
        sgo = seg_keys(orc["segments"], orc["concave_ids"], orc["concave_pos64"], s2i_o)
        t2g = orc["trimmed_to_global"].astype(np.int64)
        amb_t = np.nonzero(orc["amb_seg_circle"])[0]
        gsl = t2g[amb_t]
        amb_seg = _ukey(s2i_o[gsl // K_o], s2i_o[orc["nb_ids"][gsl].astype(np.int64)], n_in)
        if keep_mask is not None:
            sgg = sgg[keep_mask[sgg["u"] // n_in] & keep_mask[sgg["u"] % n_in]]
            sgo = sgo[keep_mask[sgo["u"] // n_in] & keep_mask[sgo["u"] % n_in]]
        d = np.setxor1d(sgg, sgo)
        b = np.isin(d["u"], amb_seg) | np.isin(d["u"], amb_circles)
        rep["mismatch"]["segments"] = int((~b).sum())
        rep["details"]["segments"] = [(int(x["u"] // n_in), int(x["u"] % n_in), (int(x["Aa"]), int(x["Ab"]), int(x["Ac"]), int(x["As"])),
                                       (int(x["Ba"]), int(x["Bb"]), int(x["Bc"]), int(x["Bs"])), bool(np.isin(x, sgg)))
                                      for x in d[~b][:12]]
        rep["band"]["segments"] = int(b.sum())
        rep["segment_nb"] = int(len(sgo))
        rep["segment_band_circles"] = np.unique(d["u"][b])
        band_atoms += [d["u"][b] // n_in, d["u"][b] % n_in]
    # -- exterior path: torus records, keyed by circle (head atom, partner) --------------------------------
    if mode == 1 and "torus_pos" in gpu:
        def torus_keys(r, j, s2i):
            return s2i[r] * n_in + s2i[j.astype(np.int64)]

        # the GPU emits an atom's records in slot order, atoms in sorted order
        rr, ll, jj = _slots(gpu["nb_count"], gpu["nb_ids"], K_g)
        two = gpu["status"].reshape(-1, K_g)[rr, ll] == 2
        tg = torus_keys(rr[two], jj[two], s2i_g)
        vg = np.concatenate([gpu["torus_pos"], gpu["torus_axis"], gpu["torus_vs"]], axis=1).astype(np.float64)
        assert len(tg) == len(vg), (len(tg), len(vg))
        slots = orc["torus_slots"].astype(np.int64)
        to = torus_keys(slots // K_o, orc["nb_ids"][slots], s2i_o)
        vo = orc["torus"].reshape(3, -1, 4).transpose(1, 0, 2).reshape(-1, 12).astype(np.float64)
        if keep_mask is not None:
            fg_, fo_ = keep_mask[tg // n_in] & keep_mask[tg % n_in], keep_mask[to // n_in] & keep_mask[to % n_in]
            tg, vg, to, vo = tg[fg_], vg[fg_], to[fo_], vo[fo_]
        ug, uo = _ukey(tg // n_in, tg % n_in, n_in), _ukey(to // n_in, to % n_in, n_in)
        d = np.setxor1d(ug, uo)
        b = np.isin(d, amb_circles) | np.isin(d, promo)
        band_atoms += [d[b] // n_in, d[b] % n_in]
        rep["mismatch"]["torus"] = int((~b).sum())
        rep["details"]["torus"] = np.stack([d[~b] // n_in, d[~b] % n_in], 1)
        rep["band"]["torus"] = int(b.sum())
        rep["torus_nb"] = int(len(uo))
        # values where both sides built the record from the same head atom (same sorted order);
        # reported here, bounded by torus_values_within_fp32 (some fields are ill-conditioned)
        _, a, c = np.intersect1d(tg, to, return_indices=True)
        rep["torus_same_head"] = int(len(a))
        rep["torus_max_err"] = float(np.nanmax(np.abs(vg[a] - vo[c]))) if len(a) else 0.0
    rep["band_atoms"] = np.unique(np.concatenate([np.asarray(x, np.int64) for x in band_atoms])) if band_atoms else np.zeros(0, np.int64)
    return rep


def assert_parity(rep):
    bad = {k: v for k, v in rep["mismatch"].items() if v}
    assert not bad, f"topology mismatches outside the epsilon band: {bad} (report {rep})"
    assert rep["probe_pos_max_err"] <= POS_TOL, rep
    if "sector_excess" in rep:
        assert rep["sector_excess"] <= 0, rep



def _torus_table(res_or_gpu, is_gpu, K, s2i, n_in):
    if is_gpu:
        rr, ll, jj = _slots(res_or_gpu["nb_count"], res_or_gpu["nb_ids"], K)
        two = res_or_gpu["status"].reshape(-1, K)[rr, ll] == 2
        keys = s2i[rr[two]] * n_in + s2i[jj[two].astype(np.int64)]
        vals = np.concatenate([res_or_gpu["torus_pos"], res_or_gpu["torus_axis"], res_or_gpu["torus_vs"]], axis=1)
    else:
        slots = res_or_gpu["torus_slots"].astype(np.int64)
        keys = s2i[slots // K] * n_in + s2i[res_or_gpu["nb_ids"][slots].astype(np.int64)]
        vals = res_or_gpu["torus"].reshape(3, -1, 4).transpose(1, 0, 2).reshape(-1, 12)
    return keys, vals.astype(np.float64)


def torus_values_within_fp32(gpu, orc64, orc32, atoms_in, slack=1e-4):
    """Torus record values (centre, radius, axis, visibility sphere) of the GPU against FP64, where
    the FP64 and FP32 oracles ran in the GPU's sorted order (the same head atom builds each record).
    Some fields are ill-conditioned in FP32 (the radius of a near-tangent circle, the axis of a
    circle whose centre lies near its head atom's centre), so each field of each record may differ
    from FP64 by twice what the FP32 restatement of the reference's algorithm differs, plus
    `slack`. A NaN must be a NaN on both sides. Returns (records compared, records outside)."""
    n_in = len(atoms_in)
    kg, vg = _torus_table(gpu, True, gpu["K"], gpu["s2i"].astype(np.int64), n_in)
    k64, v64 = _torus_table(orc64, False, orc64.K, orc64["sorted_to_initial"].astype(np.int64), n_in)
    k32, v32 = _torus_table(orc32, False, orc32.K, orc32["sorted_to_initial"].astype(np.int64), n_in)
    both, a, b = np.intersect1d(kg, k64, return_indices=True)
    _, c, d = np.intersect1d(both, k32, return_indices=True)
    g, r64, r32 = vg[a[c]], v64[b[c]], v32[d]
    nan_ok = np.isnan(g) == np.isnan(r64)
    eg = np.where(np.isnan(g) & np.isnan(r64), 0.0, np.abs(g - r64))
    e32 = np.where(np.isnan(r32) & np.isnan(r64), 0.0, np.abs(r32 - r64))
    bad = ~nan_ok | ~( eg <= 2 * np.nan_to_num(e32, nan=0.0) + slack )
    return int(len(g)), int(bad.any(axis=1).sum())


def winner_atoms(codes, s2i, full_circles, segments, concave_ids):
    """Input indices of the atoms that carry each winning element ([n, 3], -1 where none)."""
    t, idx = codes >> 40, (codes & 0xFFFFFFFFFF).astype(np.int64)
    out = np.full((len(codes), 3), -1, np.int64)
    m = t == 1  # face of an atom
    out[m, 0] = s2i[idx[m]]
    m = t == 2  # full circle
    out[m, :2] = s2i[full_circles.reshape(-1, 2)[idx[m]].astype(np.int64)]
    m = t == 3  # arc of a segment: its circle
    out[m, :2] = s2i[segments.reshape(-1, 4)[idx[m], :2].astype(np.int64)]
    m = t == 4  # probe vertex
    out[m] = s2i[concave_ids.reshape(-1, 4)[idx[m], :3].astype(np.int64)]
    return out
