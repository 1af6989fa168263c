"""Precision of the gradient's nu-solve (sum_m nu_m (p - c_m) = e over the winner's sphere
constraints: 1 for a face, 2 for an arc, 3 for a vertex) in FP32 against FP64.

The backward kernel solves it in FP64 for every query. This test restates the solve with numpy in
both precisions from the same FP32 inputs, pins the FP64 restatement to the oracle's analytic
gradient, and measures what FP32 would lose: the per-query error of the contributions
-nu_m (p - c_m) relative to the query's largest contribution, by winner type. Vertices whose three
constraint vectors are nearly coplanar make the 3 x 3 normal equations ill-conditioned (the Gram
determinant is the squared volume), and FP32 then loses several digits.
"""
import numpy as np

import oracle as O
from sesdf_amd import synth

RP = 1.4


def _winner_geometry(res, pts, codes, dtype):
    """Per query: constraint vectors a_m = p - c_m ([n, 3, 3], zero rows beyond m), m, e (unit) and
    the sorted atoms of the winner, all formed in `dtype` from the FP32 inputs."""
    atoms = res["atoms"].reshape(-1, 4).astype(dtype)
    c, R = atoms[:, :3], atoms[:, 3] + dtype(RP)
    q = pts.astype(dtype)
    t, idx = codes >> 40, (codes & 0xFFFFFFFFFF).astype(np.int64)
    n = len(q)
    at = np.full((n, 3), -1, np.int64)
    p = np.zeros((n, 3), dtype)
    m = np.zeros(n, np.int64)
    face = t == 1
    at[face, 0] = idx[face]
    v = q[face] - c[idx[face]]
    p[face] = c[idx[face]] + v * (R[idx[face]] / np.linalg.norm(v, axis=1))[:, None]
    m[face] = 1
    vert = t == 4
    ids = res["concave_ids"].reshape(-1, 4)[idx[vert], :3].astype(np.int64)
    at[vert] = ids
    pos = res["concave_pos64"].reshape(-1, 3)[idx[vert]]
    p[vert] = pos if dtype == np.float64 else pos.astype(np.float32)
    m[vert] = 3
    arc = (t == 2) | (t == 3)
    ij = np.zeros((n, 2), np.int64)
    fc = res["full_circles"].reshape(-1, 2).astype(np.int64)
    sg = res["segments"].reshape(-1, 4).astype(np.int64)
    ij[t == 2] = fc[idx[t == 2]]
    ij[t == 3] = sg[idx[t == 3], :2]
    i, j = ij[arc, 0], ij[arc, 1]
    at[arc, 0], at[arc, 1] = i, j
    dj = c[j] - c[i]
    d2 = (dj * dj).sum(1)
    nrm = dj / np.sqrt(d2)[:, None]
    ctr = c[i] + dj * ((R[i] ** 2 - R[j] ** 2 + d2) / (2 * d2))[:, None]
    rad = np.sqrt(np.maximum(R[i] ** 2 - ((ctr - c[i]) ** 2).sum(1), 0))
    vv = q[arc] - ctr
    wv = vv - nrm * (vv * nrm).sum(1)[:, None]
    p[arc] = ctr + wv * (rad / np.linalg.norm(wv, axis=1))[:, None]
    m[arc] = 2
    e = p - q
    e = e / np.linalg.norm(e, axis=1)[:, None]
    a = np.zeros((n, 3, 3), dtype)
    for u in range(3):
        ok = at[:, u] >= 0
        a[ok, u] = p[ok] - c[at[ok, u]]
    return a, m, e, at


def _solve(a, m, e):
    """Normal equations G nu = A e (G = A A^T), by the closed forms of the kernel (Cramer)."""
    dt = a.dtype.type
    G = np.einsum("nij,nkj->nik", a, a)
    r = np.einsum("nij,nj->ni", a, e)
    nu = np.zeros_like(r)
    s1 = m == 1
    nu[s1, 0] = r[s1, 0] / G[s1, 0, 0]
    s2 = m == 2
    g = G[s2]
    det = g[:, 0, 0] * g[:, 1, 1] - g[:, 0, 1] * g[:, 1, 0]
    nu[s2, 0] = (r[s2, 0] * g[:, 1, 1] - g[:, 0, 1] * r[s2, 1]) / det
    nu[s2, 1] = (g[:, 0, 0] * r[s2, 1] - g[:, 1, 0] * r[s2, 0]) / det
    s3 = m == 3
    g3 = G[s3]
    d = np.linalg.det(g3.astype(np.float64)).astype(dt) if dt == np.float64 else _det3(g3)
    for col in range(3):
        A = g3.copy()
        A[:, :, col] = r[s3]
        nu[s3, col] = (np.linalg.det(A) if dt == np.float64 else _det3(A)) / d
    return nu


def _det3(A):
    return (A[:, 0, 0] * (A[:, 1, 1] * A[:, 2, 2] - A[:, 1, 2] * A[:, 2, 1])
            - A[:, 0, 1] * (A[:, 1, 0] * A[:, 2, 2] - A[:, 1, 2] * A[:, 2, 0])
            + A[:, 0, 2] * (A[:, 1, 0] * A[:, 2, 1] - A[:, 1, 1] * A[:, 2, 0]))


def test_nu_solve_fp32_against_fp64():
    atoms = synth.globule(3000, synth.SEED_BASE + 61)
    res = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    rng = np.random.default_rng(11)
    lo, hi = atoms[:, :3].min(0) - 4, atoms[:, :3].max(0) + 4
    cand = (lo + (hi - lo) * rng.random((400_000, 3))).astype(np.float32)
    val, codes = res.sdf(cand.astype(np.float64), band=2.0, threads=8, codes=True)
    keep = (codes >> 40 >= 1) & (codes >> 40 <= 4) & (np.abs(val) < 2.0)
    pts, codes = cand[keep][:60_000], codes[keep][:60_000]
    w = rng.standard_normal(len(pts))

    a64, m, e64, at = _winner_geometry(res, pts, codes, np.float64)
    nu64 = _solve(a64, m, e64)
    contrib64 = -nu64[:, :, None] * a64  # [n, 3 constraints, 3]

    # the FP64 restatement is the oracle's gradient (xyz part), accumulated per atom (the two form
    # an arc's nearest point in different orders: ~1e-7 relative where q is near the circle's axis)
    grad = np.zeros((len(atoms), 3))
    s2i = res["sorted_to_initial"].astype(np.int64)
    for u in range(3):
        ok = at[:, u] >= 0
        np.add.at(grad, s2i[at[ok, u]], w[ok, None] * contrib64[ok, u])
    ref = res.sdf_grad(pts.astype(np.float64), w, band=2.0, threads=8)[:, :3]
    assert np.abs(grad - ref).max() <= 1e-6 * np.abs(ref).max()

    a32, _, e32, _ = _winner_geometry(res, pts, codes, np.float32)
    nu32 = _solve(a32, m, e32)
    contrib32 = -nu32[:, :, None].astype(np.float64) * a32.astype(np.float64)
    scale = np.abs(contrib64).max(axis=(1, 2))
    err = np.abs(contrib32 - contrib64).max(axis=(1, 2)) / scale
    t = codes >> 40
    for name, sel in (("face", t == 1), ("arc", (t == 2) | (t == 3)), ("vertex", t == 4)):
        if sel.any():
            print(f"nu-solve FP32 vs FP64, {name}: {int(sel.sum())} queries, relative error max {err[sel].max():.2e}, "
                  f"p99.9 {np.quantile(err[sel], 0.999):.2e}, above 1e-3: {int((err[sel] > 1e-3).sum())}", flush=True)
    # faces: one division, FP32 suffices; the three-sphere solve is where FP32 falls short
    assert err[t == 1].max() < 1e-5
    assert (t == 4).sum() > 1000
