"""Pins of the exterior-surface rule of the SDF oracle (no GPU).

The exterior SDF counts only the boundary component that bounds the unbounded region of the
complement of the SAS union A; points outside A in other regions (probe-accessible cavities) are
inside. The oracle forms components over arcs, vertices and faces (an atom's exposed sphere may
hold several faces of different components). Pinned here by:

* a voxel flood fill of the complement of A (0.2 A voxels, face connectivity): points at least
  two voxels from A are classified by the flood-fill label and by the oracle's field (outside A the
  exterior field is min(rp + g, band) in the unbounded region and at most rp - g in a cavity, g the
  distance to A), on globules with carved voids and on the many small cavities their packing
  leaves; with a narrow band, so that cavity points farther than rp + band from A occur; and on a
  system of two separate globules with an atom floating in one of the voids (the outer surfaces of
  both globules bound the unbounded region, the floating atom's sphere bounds the void's region);
* the voids of the 100,000-atom globule: the complete field is positive at every void centre and
  the exterior field negative;
* teeth: the rule the first version used (components of atoms joined by arcs, a point taking the
  component of its nearest sphere) fails both checks.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from sesdf_amd import synth

RP = float(np.float32(1.4))


def flood_fill(atoms, h=0.2, pad=4.0):
    """Distance-to-A grid and complement labels; returns (origin, h, dist, labels, unbounded label)."""
    c = atoms[:, :3].astype(np.float64)
    R = atoms[:, 3].astype(np.float64) + RP
    lo = c.min(0) - R.max() - pad
    hi = c.max(0) + R.max() + pad
    dims = np.ceil((hi - lo) / h).astype(int) + 1
    dist = np.full(dims, np.inf, np.float32)
    for a in range(len(c)):
        i0 = np.maximum(np.floor((c[a] - R[a] - 2 * h - lo) / h).astype(int), 0)
        i1 = np.minimum(np.ceil((c[a] + R[a] + 2 * h - lo) / h).astype(int) + 1, dims)
        xs = [lo[d] + h * np.arange(i0[d], i1[d]) for d in range(3)]
        dd = np.sqrt((xs[0][:, None, None] - c[a, 0]) ** 2 + (xs[1][None, :, None] - c[a, 1]) ** 2
                     + (xs[2][None, None, :] - c[a, 2]) ** 2) - R[a]
        sl = tuple(slice(i0[d], i1[d]) for d in range(3))
        np.minimum(dist[sl], dd.astype(np.float32), out=dist[sl])
    lab, _ = ndimage.label(dist > 0)
    return lo, h, dist, lab, lab[0, 0, 0]


def two_globules_and_a_floating_atom(seed):
    """Two globules 40 A apart (the second with voids) and one atom at the centre of its largest void."""
    a, _ = synth.globule_voids(1500, 1, synth.SEED_BASE + 300 + seed)
    b, voids = synth.globule_voids(2500, 2, synth.SEED_BASE + 400 + seed)
    shift = np.array([a[:, 0].max() - b[:, 0].min() + 40.0, 0, 0, 0], np.float32)
    v = voids[np.argmax(voids[:, 3])]
    floating = np.array([[v[0], v[1], v[2], 1.0]], np.float32) + shift
    return np.concatenate([a, b + shift, floating]).astype(np.float32)


def exterior_class(values, band, h=0.2):
    """Unbounded region iff the field reaches min(band, rp + 2h) (cavity values stay below rp - 2h)."""
    return values >= min(band, RP + 2 * h) - 1e-6


def sample_complement(atoms, seed, n_any=3000, n_cavity=1500):
    """Points outside A, two voxels or more from it, with their flood-fill class (True: unbounded)."""
    lo, h, dist, lab, ext = flood_fill(atoms)
    rng = np.random.default_rng(seed)
    near = np.argwhere((dist > 2 * h) & (dist < 3.0))
    pick = near[rng.choice(len(near), min(n_any, len(near)), replace=False)]
    cav = np.argwhere((dist > 2 * h) & (lab != ext) & (lab > 0))
    if len(cav):
        pick = np.concatenate([pick, cav[rng.choice(len(cav), min(n_cavity, len(cav)), replace=False)]])
    return lo + h * pick, lab[tuple(pick.T)] == ext


def atom_level_rule(res, pts):
    """The first version's rule, restated: atoms joined by arcs form components; the exterior one
    holds the atom of largest x + r; a point outside A takes its nearest sphere's component."""
    A = res["atoms"].astype(np.float64)
    n = len(A)
    par = np.arange(n)

    def find(a):
        while par[a] != a:
            par[a] = par[par[a]]
            a = par[a]
        return a

    for u, v in list(res["full_circles"][:, :2]) + list(res["segments"][:, :2]):
        a, b = find(int(u)), find(int(v))
        if a != b:
            par[max(a, b)] = min(a, b)
    comp = np.array([find(a) for a in range(n)])
    ext = comp[int(np.argmax(A[:, 0] + A[:, 3]))]
    d = np.sqrt(((pts[:, None, :] - A[None, :, :3]) ** 2).sum(-1)) - (A[:, 3] + RP)[None, :]
    return comp[np.argmin(d, axis=1)] == ext


def flood_case(n, n_voids, seed):
    if n == "two":
        return two_globules_and_a_floating_atom(seed)
    return synth.globule_voids(n, n_voids, synth.SEED_BASE + 100 + seed)[0]


FLOOD_CASES = [(1500, 2, 5, 10.0), (3000, 3, 6, 10.0), (800, 1, 7, 10.0), (3000, 3, 6, 1.2), ("two", 0, 8, 1.2)]


@pytest.mark.parametrize("n,n_voids,seed,band", FLOOD_CASES)
def test_exterior_rule_matches_flood_fill(n, n_voids, seed, band):
    atoms = flood_case(n, n_voids, seed)
    pts, want = sample_complement(atoms, seed)
    res = O.build(atoms, mode=0, precision=64, track=True)
    sd = res.sdf(pts, band=band, exterior=True, threads=8)
    got = exterior_class(sd, band)
    assert (~want).sum() > 500 and want.sum() > 500
    if band < 2:
        # cavity points beyond rp + band from A are classified too (by their ray)
        g = np.sqrt(((pts[:, None, :] - atoms[None, :, :3]) ** 2).sum(-1)) - (atoms[:, 3] + RP)[None, :]
        assert ((~want) & (g.min(1) > RP + band)).sum() >= 5
    assert np.count_nonzero(got != want) == 0
    # the complete field does not tell cavities apart: positive everywhere outside A
    assert (res.sdf(pts, band=band, exterior=False, threads=8) >= min(band, RP) - 1e-9).all()


def test_void_centres_of_the_100k_globule():
    atoms, voids = synth.globule_voids(100000, 10, synth.SEED_BASE + 1)
    res = O.build(atoms, mode=0, precision=64, track=False, threads=8)
    complete = res.sdf(voids[:, :3], band=4.0, exterior=False, threads=8)
    exterior = res.sdf(voids[:, :3], band=4.0, exterior=True, threads=8)
    assert (complete > 0).all(), complete
    assert (exterior < 0).all(), exterior
    # the atom-level rule puts some of these cavities in the exterior component
    assert not (~atom_level_rule(res, voids[:, :3].astype(np.float64))).all()


def test_atom_level_rule_fails_flood_fill():
    """Teeth of the flood-fill check: the atom-level rule misclassifies cavity points."""
    atoms, _ = synth.globule_voids(3000, 3, synth.SEED_BASE + 106)
    pts, want = sample_complement(atoms, 6)
    res = O.build(atoms, mode=0, precision=64, track=True)
    assert np.count_nonzero(atom_level_rule(res, pts) != want) > 0


def test_component_labels_are_consistent():
    """Atoms whose boundary lies in one component carry its label; mixed atoms span several."""
    atoms, _ = synth.globule_voids(3000, 3, synth.SEED_BASE + 106)
    res = O.build(atoms, mode=0, precision=64, track=True)
    lab, ext = res.components()
    mixed, none = lab == 0xFFFFFFFE, lab == 0xFFFFFFFF
    assert mixed.sum() > 0 and none.sum() > 0
    assert (lab == ext).sum() > 0.2 * len(lab)
    # atoms on an arc have a label; atoms with no arc are buried, or alone (no cap on their sphere)
    on_arc = np.zeros(len(atoms), bool)
    on_arc[res["segments"][:, :2].reshape(-1)] = True
    on_arc[res["full_circles"][:, :2].reshape(-1)] = True
    assert not (on_arc & none).any()
    assert not (~on_arc & mixed).any()
