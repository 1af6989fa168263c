"""Pins of the oracle's SDF and gradient by methods independent of its element-wise evaluation.

* one atom: SDF(q) = |q - c| - r everywhere (closed form);
* inside the SAS union A, rp - SDF(q) is the distance to the boundary of A: compared with a brute
  force over dense samples of the exposed parts of every SAS sphere (points outside all other
  spheres), which bounds it from above and, up to the sample spacing, from below;
* outside A: SDF(q) = min_i |q - c_i| - r_i (closed form);
* the field is 1-Lipschitz, never exceeds the vdW bound min_i |q - c_i| - r_i, and the exterior
  variant never exceeds the complete one;
* the analytic gradient matches central finite differences of the FP64 field with the topology held
  fixed (probe vertices re-solved from the moved atoms).
"""
import numpy as np
import pytest

import oracle as O
from sesdf_amd import synth

RP = 1.4


def fibonacci_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = i * np.pi * (3 - np.sqrt(5))
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)


def exposed_samples(atoms, per_sphere, per_circle=4000):
    """Points of the SAS boundary: samples of each sphere |x - c_i| = R_i, and of each pairwise
    intersection circle (which holds the boundary's edges and corners), outside every other sphere."""
    R = atoms[:, 3].astype(np.float64) + RP
    c = atoms[:, :3].astype(np.float64)
    u = fibonacci_sphere(per_sphere)
    out = []

    def keep(p, skip):
        d2 = ((p[:, None, :] - c[None, :, :]) ** 2).sum(-1) - R[None, :] ** 2
        d2[:, skip] = 1.0
        return p[(d2 >= 0).all(1)]

    for i in range(len(atoms)):
        out.append(keep(c[i] + R[i] * u, [i]))
    t = np.linspace(0, 2 * np.pi, per_circle, endpoint=False)
    for i in range(len(atoms)):
        for j in range(i + 1, len(atoms)):
            d = np.linalg.norm(c[j] - c[i])
            if d >= R[i] + R[j] or d <= abs(R[i] - R[j]):
                continue
            n = (c[j] - c[i]) / d
            a = (R[i] ** 2 - R[j] ** 2 + d * d) / (2 * d)  # distance of the circle plane from c_i
            rad = np.sqrt(R[i] ** 2 - a * a)
            e1 = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(n, e1)
            p = c[i] + a * n + rad * (np.cos(t)[:, None] * e1 + np.sin(t)[:, None] * e2)
            out.append(keep(p, [i, j]))
    return np.concatenate(out), 4 * np.pi * R.max() ** 2 / per_sphere


def small_cluster(seed, n=14):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-4.5, 4.5, 3)
        if all(np.linalg.norm(p - q) > 1.9 for q in pts):
            pts.append(p)
    r = rng.choice([1.52, 1.55, 1.70, 1.80], n)
    return np.c_[np.array(pts), r].astype(np.float32)


def test_single_atom_closed_form():
    atoms = np.array([[0.3, -0.2, 0.1, 1.7]], np.float32)
    orc = O.build(atoms, mode=0, precision=64, track=False)
    q = np.random.default_rng(1).uniform(-6, 6, (4000, 3))
    band = 3.0
    got = orc.sdf(q, band=band)
    want = np.clip(np.linalg.norm(q - atoms[0, :3].astype(np.float64), axis=1) - np.float64(atoms[0, 3]), -band, band)
    assert np.abs(got - want).max() < 1e-9


@pytest.mark.parametrize("seed", [3, 4])
def test_inside_distance_matches_boundary_samples(seed):
    atoms = small_cluster(seed)
    orc = O.build(atoms, mode=0, precision=64, track=False)
    band = 12.0  # no clamping: rp - SDF is the raw distance to the boundary
    samples, area_per = exposed_samples(atoms, 40000)
    spacing = np.sqrt(area_per)
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    rng = np.random.default_rng(seed + 10)
    q = rng.uniform(c.min(0) - 2, c.max(0) + 2, (4000, 3))
    g = (np.linalg.norm(q[:, None, :] - c[None], axis=-1) - R[None]).min(1)
    inside = q[g < 0][:600]
    assert len(inside) > 200
    d_oracle = RP - orc.sdf(inside, band=band)
    d_brute = np.array([np.sqrt(((samples - p) ** 2).sum(1).min()) for p in inside])
    # samples lie on the boundary: the true distance is at most d_brute, and at least
    # d_brute minus the sample spacing
    assert (d_oracle <= d_brute + 1e-9).all()
    assert (d_oracle >= d_brute - spacing).all(), (d_brute - d_oracle).max()
    # the bounds have teeth: a field that lost its vertices (or any 0.1 A error) fails them
    codes = orc.sdf(inside, band=band, codes=True)[1]
    lost = d_oracle + np.where((codes >> 40) == 4, 0.1, 0.0)
    assert not ((lost <= d_brute + 1e-9).all() and (lost >= d_brute - spacing).all())


def test_outside_is_nearest_sphere_and_invariants():
    atoms = synth.globule(300, synth.SEED_BASE + 70)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=4)
    band = 3.0
    c = atoms[:, :3].astype(np.float64)
    r = atoms[:, 3].astype(np.float64)
    rng = np.random.default_rng(5)
    q = rng.uniform(c.min(0) - 4, c.max(0) + 4, (6000, 3))
    s = orc.sdf(q, band=band, threads=4)
    dv = (np.linalg.norm(q[:, None, :] - c[None], axis=-1) - r[None]).min(1)  # vdW bound
    g = dv - RP
    out = g >= 0
    assert np.abs(s[out] - np.clip(dv[out], -band, band)).max() < 1e-9
    assert (s <= np.clip(dv, -band, band) + 1e-9).all()
    # 1-Lipschitz along short random segments
    step = rng.normal(size=q.shape)
    step *= (rng.uniform(0.05, 0.6, len(q)) / np.linalg.norm(step, axis=1))[:, None]
    s2 = orc.sdf(q + step, band=band, threads=4)
    assert (np.abs(s2 - s) <= np.linalg.norm(step, axis=1) + 1e-9).all()
    # the exterior variant only adds inside region and removes boundary elements
    se = orc.sdf(q, band=band, exterior=True, threads=4)
    assert (se <= s + 1e-9).all()


def test_gradient_matches_finite_differences():
    atoms = synth.globule(400, synth.SEED_BASE + 71)
    orc = O.build(atoms, mode=0, precision=64, track=False, threads=4)
    band = 2.0
    c = atoms[:, :3].astype(np.float64)
    rng = np.random.default_rng(9)
    q = rng.uniform(c.min(0) - 3, c.max(0) + 3, (20000, 3))
    s = orc.sdf(q, band=band, threads=4)
    q = q[np.abs(s) < band * 0.95][:3000]
    w = rng.standard_normal(len(q))
    g = orc.sdf_grad(q, w, band=band)
    ids = rng.choice(len(atoms), 48, replace=False)
    fd, stable = orc.sdf_fd(q, w, ids, h=1e-4, band=band, threads=4)
    assert stable.sum() >= 24
    ref = g[ids][stable]
    err = np.abs(fd[stable] - ref)
    scale = np.abs(ref).max(1, keepdims=True) + 1e-6
    assert (err / scale).max() < 1e-3, (err / scale).max()
