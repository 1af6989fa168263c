"""Host C++ baseline for the benchmark: the oracle's FP32 restatement of the reference's stage
sequence (grid, neighbours, visibility, sectors, intersections, segments, probe neighbours, torus
records), threaded over atoms, timed on this machine's cores. Test and measurement infrastructure
only: bench.py runs it as a child process per workload, so that each one's peak resident memory is
its own, and reports it next to the GPU numbers as a baseline.

    python oracle/cpu_baseline.py WORKLOAD THREADS   ->  one JSON line on stdout

Workloads (the synthetic inputs of the benchmark, seeded as there):
  c0   2,000-atom globule, complete surface; its SDF on a 64^3 grid (FP64 field evaluator)
  c1   100,000-atom globule with 10 voids, complete surface
  c1x  the same atoms, exterior path
  c2   2,500,000 atoms (500 globules of 5,000 on a 360 A sphere), exterior path: one timed run
  c3   the first 64 molecules of the 4,096-molecule batch, complete surface and an SDF 32^3 grid
       each; molecules spread over the threads (one thread per molecule)
"""
import concurrent.futures
import json
import os
import resource
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.dirname(HERE)]

import oracle as O  # noqa: E402
from sesdf_amd import synth  # noqa: E402


def _median_runs(fn, runs, warmup=1):
    for _ in range(warmup):
        fn()
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return float(np.median(times)), [round(x, 4) for x in times]


def _grid(atoms, dim, margin=3.0):
    lo = atoms[:, :3].min(0) - atoms[:, 3].max() - margin
    hi = atoms[:, :3].max(0) + atoms[:, 3].max() + margin
    axes = [np.linspace(lo[k], hi[k], dim) for k in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)


def run(workload, threads):
    out = {"workload": workload, "threads": threads}
    if workload == "c0":
        atoms = synth.globule(2000, synth.SEED_BASE)
        med, runs = _median_runs(lambda: O.build(atoms, mode=0, precision=32, track=False, threads=threads), 5)
        res = O.build(atoms, mode=0, precision=32, track=False, threads=threads)
        pts = _grid(atoms, 64)
        smed, sruns = _median_runs(lambda: res.sdf(pts, band=2.0, threads=threads), 3)
        out.update(atoms=len(atoms), build_s=med, runs_s=runs, atoms_per_s=len(atoms) / med, sdf_64_s=smed, sdf_runs_s=sruns,
                   sdf_samples_per_s=len(pts) / smed)
    elif workload in ("c1", "c1x"):
        atoms, _ = synth.globule_voids(100_000, 10, synth.SEED_BASE + 1)
        mode = 0 if workload == "c1" else 1
        med, runs = _median_runs(lambda: O.build(atoms, mode=mode, precision=32, track=False, threads=threads), 5)
        stats = O.build(atoms, mode=mode, precision=32, track=False, threads=threads)["stats"]
        out.update(atoms=len(atoms), build_s=med, runs_s=runs, atoms_per_s=len(atoms) / med,
                   predicate_counts={"visibility_tests": int(stats[0]), "triplet_tests": int(stats[1]),
                                     "occluder_tests": int(stats[2])})
    elif workload == "c2":
        atoms = synth.fibonacci_globules(500, 5000, 360.0, synth.SEED_BASE + 2)
        med, runs = _median_runs(lambda: O.build(atoms, mode=1, precision=32, track=False, threads=threads), 1, warmup=0)
        out.update(atoms=len(atoms), build_s=med, runs_s=runs, atoms_per_s=len(atoms) / med)
    elif workload == "c3":
        sizes = synth.batch_sizes(4096, 20, 5000, synth.SEED_BASE + 3)
        atoms, offs = synth.batch(sizes, synth.SEED_BASE + 3)
        P = 64
        mols = [atoms[int(offs[m]):int(offs[m + 1])] for m in range(P)]

        def one(mol):
            res = O.build(mol, mode=0, precision=32, track=False, threads=1)
            res.sdf(_grid(mol, 32), band=2.0, threads=1)

        def all_():
            with concurrent.futures.ThreadPoolExecutor(threads) as ex:
                list(ex.map(one, mols))

        med, runs = _median_runs(all_, 3, warmup=0)
        out.update(molecules=P, atoms=int(sum(len(m) for m in mols)), build_plus_sdf_s=med, runs_s=runs, molecules_per_s=P / med)
    else:
        raise SystemExit(f"unknown workload {workload}")
    out["peak_rss_bytes"] = peak_rss()
    return out


def peak_rss():
    """Peak resident set of this process image (VmHWM; ru_maxrss would carry the parent's over exec)."""
    try:
        for line in open("/proc/self/status"):
            if line.startswith("VmHWM:"):
                return int(line.split()[1]) * 1024
    except OSError:
        pass
    return resource.getrusage(resource.RUSAGE_SELF).ru_maxrss * 1024


if __name__ == "__main__":
    print(json.dumps(run(sys.argv[1], int(sys.argv[2]))), flush=True)
