"""GPU parity: the HIP build through the C ABI vs the FP64 oracle on the same inputs."""
import numpy as np
import pytest

import oracle as O
import parity
from oracle_util import tetrahedron
from sesdf_amd import Sesdf, Ssesdf, synth

pytestmark = pytest.mark.gpu


def _build(cls, atoms, **kw):
    print(f"[gpu] {cls.__name__} n={len(atoms)} {kw}", flush=True)
    return cls(atoms, **kw)


def _valid_neighbors(d):
    """Written part of each probe's neighbour block (slots past the count are left unset, as in the reference)."""
    m = int(d["max_concave_neighbors"])
    nb = d["concave_patches_neighbors"].reshape(-1, m, 4)
    cnt = d["concave_patches_id"][: nb.shape[0], 3]
    return [nb[q, : min(int(cnt[q]), m)] for q in range(nb.shape[0])]


def _assert_same(da, db):
    for k in da:
        if k == "concave_patches_neighbors":
            for x, y in zip(_valid_neighbors(da), _valid_neighbors(db)):
                np.testing.assert_array_equal(x, y, err_msg=k)
        elif isinstance(da[k], np.ndarray):
            np.testing.assert_array_equal(da[k], db[k], err_msg=k)


def _check(atoms, mode, threads=4, **kw):
    cls = Sesdf if mode == 0 else Ssesdf
    surf = _build(cls, atoms, **kw)
    orc = O.build(atoms, mode=mode, precision=64, track=True, threads=threads,
                  max_neighbors=kw.get("max_neighbors", 128))
    rep = parity.compare(parity.gpu_arrays(surf), orc, atoms, mode)
    print({k: rep[k] for k in rep if k != "mismatch"}, rep["mismatch"], flush=True)
    parity.assert_parity(rep)
    return surf, orc, rep


@pytest.mark.parametrize("mode", [0, 1])
def test_single_atom(gpu, mode):
    atoms = np.array([[1.0, 2.0, 3.0, 1.7]], np.float32)
    surf = _build(Sesdf if mode == 0 else Ssesdf, atoms)
    c = surf.counts()
    assert c["concave_patch_nb"] == 0 and c["segment_nb"] == 0 and c["torus_nb"] == 0
    d = surf.get_data()
    if mode == 0:
        assert d["convex_patches"].tolist() == [[0, 0]]


@pytest.mark.parametrize("mode", [0, 1])
def test_two_atoms_full_circle(gpu, mode):
    atoms = np.array([[0, 0, 0, 1.7], [3.0, 0.5, -0.2, 1.55]], np.float32)
    surf, orc, rep = _check(atoms, mode)
    c = surf.counts()
    if mode == 0:
        assert c["full_circle_nb"] == 1 and c["sector_nb"] == 2
    else:
        assert c["torus_nb"] == 1
        d = surf.get_data()
        s2i = surf.get_atom_indices()
        ai, aj = atoms[s2i[0]], atoms[s2i[1]]
        pos, axis, vs = parity.torus_reference(ai, aj, 1.4)
        np.testing.assert_allclose(d["circle_position"][0], pos, atol=1e-4)
        np.testing.assert_allclose(d["circle_axis"][0], axis, atol=1e-5)
        np.testing.assert_allclose(d["circle_vs"][0], vs, atol=1e-4)


@pytest.mark.parametrize("mode", [0, 1])
def test_three_atoms_two_probes(gpu, mode):
    atoms = np.array([[0, 0, 0, 1.7], [3.2, 0, 0, 1.7], [1.6, 2.8, 0, 1.52]], np.float32)
    surf, orc, rep = _check(atoms, mode)
    c = surf.counts()
    assert c["concave_patch_nb"] == 2
    if mode == 0:
        assert c["segment_nb"] == 3


@pytest.mark.parametrize("mode", [0, 1])
def test_tetrahedron(gpu, mode):
    _check(tetrahedron(3.4, 1.6), mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_globule_2000(gpu, mode):
    atoms = synth.globule(2000, synth.SEED_BASE)
    surf, orc, rep = _check(atoms, mode)
    assert rep["concave_nb"] > 500


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("cell", [2.5, 4.0, 7.0])
def test_cell_size_independence(gpu, mode, cell):
    atoms = synth.globule(3000, synth.SEED_BASE + 11)
    _check(atoms, mode, cell_size=cell)


@pytest.mark.parametrize("mode", [0, 1])
def test_graph_and_direct_identical(gpu, mode):
    atoms = synth.globule(5000, synth.SEED_BASE + 12)
    cls = Sesdf if mode == 0 else Ssesdf
    a = _build(cls, atoms, use_graph=True)
    b = _build(cls, atoms, use_graph=False)
    da, db = a.get_data(), b.get_data()
    _assert_same(da, db)
    a.build()  # replay: deterministic outputs
    _assert_same(da, a.get_data())


@pytest.mark.parametrize("mode", [0, 1])
def test_globule_20000(gpu, mode):
    atoms = synth.globule(20000, synth.SEED_BASE + 4)
    _check(atoms, mode, threads=8)
