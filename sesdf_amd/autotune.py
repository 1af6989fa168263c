"""Autotuner for the launch parameters of a build: grid cell edge and the block shapes (waves per
workgroup) of the circle and intersection kernels.

Coordinate search on the live handle, timing hipGraph replays of the whole build with hipEvents on
the build stream: first the cell edge over ``CELLS`` (topology does not depend on it; only the
neighbour-search cost does), then the circle-kernel waves, then the intersection-kernel waves.
Results are cached per workload signature (atom count and molecule count buckets, mode, probe
radius, neighbour capacity) in ``tuning.json`` next to this file.
"""
import json
import math
import os

from . import _lib as L
from .batch import set_tuning

CELLS = (1.5, 2.0, 2.5, 3.2, 3.4, 4.0, 5.0, 6.5)
WAVES = (2, 4, 8)
_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def signature(surf):
    n = max(surf.atom_nb, 1)
    mols = getattr(surf, "molecule_nb", 1)
    return f"n2^{int(math.log2(n))}-m2^{int(math.log2(max(mols, 1)))}-mode{surf.MODE}-rp{surf.probe_radius:.2f}-K{surf.get_max_neighbor_per_atom()}"


def time_build(surf, reps=5, warmup=2, after=None):
    """Mean device time (ms) of one graph-replayed build, plus the work `after()` enqueues on the
    build stream after each build (e.g. SDF sampling) when given."""
    lib = L.lib()
    stream = surf.stream()
    L.check(lib.sesdf_build(surf._h), "build")  # grows capacities if needed, captures the graph
    for _ in range(warmup):
        L.check(lib.sesdf_enqueue(surf._h), "enqueue")
        if after:
            after()
    e0, e1 = L.Event(), L.Event()
    e0.record(stream)
    for _ in range(reps):
        L.check(lib.sesdf_enqueue(surf._h), "enqueue")
        if after:
            after()
    e1.record(stream)
    e1.sync()
    L.check(lib.sesdf_sync(surf._h), "sync")
    return e0.elapsed_ms(e1) / reps


def load_cache():
    try:
        with open(_CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(surf, reps=5, use_cache=True, log=None, after=None, tag=""):
    """Pick (cell_size, circle_waves, intersection_waves) for this handle, apply it, return it.
    With `after`, the objective is the build plus the work it enqueues (`tag` names it in the cache).
    The handle is left built with the chosen tuning (a new cell edge re-lays out the grid)."""
    sig = signature(surf) + tag
    cache = load_cache()
    if use_cache and sig in cache:
        best = cache[sig]
        set_tuning(surf, best["cell_size"], best["circle_waves"], best["intersection_waves"])
        surf.build()
        return dict(best, cached=True)
    cur = {"cell_size": 3.2, "circle_waves": 4, "intersection_waves": 4}
    trials = []

    def trial(cfg):
        set_tuning(surf, cfg["cell_size"], cfg["circle_waves"], cfg["intersection_waves"])
        ms = time_build(surf, reps, after=after)
        trials.append(dict(cfg, ms=ms))
        if log:
            log(f"autotune {cfg} -> {ms:.3f} ms")
        return ms

    best_ms = math.inf
    for key, values in (("cell_size", CELLS), ("circle_waves", WAVES), ("intersection_waves", WAVES)):
        best_v = cur[key]
        for v in values:
            cfg = dict(cur, **{key: v})
            ms = trial(cfg)
            if ms < best_ms:
                best_ms, best_v = ms, v
        cur[key] = best_v
    set_tuning(surf, cur["cell_size"], cur["circle_waves"], cur["intersection_waves"])
    surf.build()
    result = dict(cur, ms=best_ms, signature=sig)
    cache[sig] = result
    try:
        with open(_CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return dict(result, trials=trials, cached=False)
