"""SES build throughput on one MI355X (atoms/s), with roofline and host-CPU baseline.

Workload: a 100,000-atom globule (density 0.05 / A^3, C/N/O/S mix, min pair distance 1.2 A) with
10 empty voids of radius U(4, 8) A, probe 1.4 A, complete surface. A step is one full build()
(grid, circles, intersections, segments, probe neighbours) on atoms already resident in HBM,
replayed as one hipGraph. Also reported: the exterior path on the same atoms and the exterior path
on 2.5M atoms (500 globules of 5,000 on a 360 A Fibonacci sphere).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--no-cpu", action="store_true", help="skip the host-CPU baseline leg")
    p.add_argument("--no-large", action="store_true", help="skip the 2.5M-atom leg")
    p.add_argument("--cell", type=float, default=0.0)
    return p.parse_args()


def time_builds(surf, steps, warmup):
    """Time `steps` graph replays of the whole stage sequence with hipEvents on the build stream."""
    import sesdf_amd._lib as L

    lib = L.lib()
    stream = surf.stream()
    for _ in range(warmup):
        L.check(lib.sesdf_enqueue(surf._h), "enqueue")
    L.check(lib.sesdf_sync(surf._h), "sync")
    L.device_synchronize()
    e0, e1 = L.Event(), L.Event()
    e0.record(stream)
    for _ in range(steps):
        L.check(lib.sesdf_enqueue(surf._h), "enqueue")
    e1.record(stream)
    e1.sync()
    L.check(lib.sesdf_sync(surf._h), "sync")
    L.device_synchronize()
    return e0.elapsed_ms(e1) / steps


def stage_times(surf, reps=10):
    """Per-stage device times (ms) from event pairs around each stage (direct launches)."""
    import ctypes

    import sesdf_amd._lib as L

    acc = np.zeros(8)
    buf = (ctypes.c_float * 8)()
    for _ in range(reps):
        L.check(L.lib().sesdf_build_timed(surf._h, buf, 8), "build_timed")
        acc += np.array(buf[:])
    return acc / reps


def circles_alg_bytes(c, n):
    """Algorithmic HBM bytes of the fused circle kernel: atom read + one gather per atom (32 N),
    neighbour ids + statuses (5 B per ordered circle), sectors (16 V), per-atom ranges (8 N),
    intersected ids (4 Ci), full circles (8 Cf)."""
    return 32 * n + 5 * c["neighbor_nb"] + 16 * c["sector_nb"] + 8 * n + 4 * c["intersected_circle_nb"] + 8 * c["full_circle_nb"]


def cpu_baseline(atoms, budget_s=20.0):
    """Host C++ build of the same stage sequence (the FP32 oracle restatement) on this box's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = min(os.cpu_count() or 1, 16)
    t0 = time.time()
    O.build(atoms, mode=0, precision=32, track=False, threads=threads)  # warm-up
    times = []
    while time.time() - t0 < budget_s and len(times) < 5:
        t = time.time()
        O.build(atoms, mode=0, precision=32, track=False, threads=threads)
        times.append(time.time() - t)
    med = float(np.median(times))
    return {"value": atoms.shape[0] / med, "unit": "atoms/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full complete-surface builds of the {atoms.shape[0]}-atom workload, FP32, median {med:.3f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    from sesdf_amd import Sesdf, Ssesdf, synth
    import sesdf_amd._lib as L

    atoms, voids = synth.globule_voids(100_000, 10, synth.SEED_BASE + 1)
    n = atoms.shape[0]

    res = {}
    surf = Sesdf(atoms, probe_radius=1.4, cell_size=args.cell, use_graph=True)
    ms = time_builds(surf, args.steps, args.warmup)
    counts = surf.counts()
    stages = stage_times(surf)
    ext = Ssesdf(atoms, probe_radius=1.4, cell_size=args.cell, use_graph=True)
    ms_ext = time_builds(ext, args.steps, args.warmup)

    # dominant kernel of the complete build: the fused circle stage (stage index 1)
    circ_ms = float(stages[1])
    alg = circles_alg_bytes(counts, n)
    achieved = alg / (circ_ms * 1e-3) / 1e9

    large = {}
    if not args.no_large:
        big = synth.fibonacci_globules(500, 5000, 360.0, synth.SEED_BASE + 2)
        free0, _ = L.mem_info()
        sbig = Ssesdf(big, probe_radius=1.4, cell_size=args.cell, use_graph=True)
        free1, _ = L.mem_info()
        ms_big = time_builds(sbig, max(5, args.steps // 5), 2)
        cb = sbig.counts()
        large = {"atoms": int(big.shape[0]), "ms_per_build": ms_big, "atoms_per_s": big.shape[0] / (ms_big * 1e-3),
                 "concave": cb["concave_patch_nb"], "torus": cb["torus_nb"],
                 "device_bytes": cb["device_bytes"], "mem_delta_bytes": int(free0 - free1)}
        del sbig

    cpu = None if args.no_cpu else cpu_baseline(atoms)

    out = {
        "metric": "SES build atoms/s (100k-atom globule, complete surface)",
        "value": n / (ms * 1e-3),
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "100k-atom globule with 10 voids, probe 1.4 A, complete surface (hipGraph replay)",
                   "atoms": n, "probe_radius": 1.4, "cell_size": surf_cell(args.cell)},
        "exterior_atoms_per_s": n / (ms_ext * 1e-3),
        "exterior_ms_per_step": ms_ext,
        "large_2p5M_exterior": large,
        "stage_ms": dict(zip(["grid", "circles", "intersections", "circle_scan", "fill", "segments", "probe_nb", "torus"],
                             [float(x) for x in stages])),
        "counts": {k: int(v) for k, v in counts.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "circles_kernel", "alg_bytes_per_launch": alg, "kernel_ms": circ_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def surf_cell(c):
    return c if c > 0 else 3.2


if __name__ == "__main__":
    main()
