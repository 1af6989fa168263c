"""SES build throughput on one MI355X, with roofline and host-CPU baseline.

Headline (value): a 100,000-atom globule (density 0.05 / A^3, C/N/O/S mix, min pair distance 1.2 A)
with 10 empty voids of radius U(4, 8) A, probe 1.4 A, complete surface. A step is one full build()
(grid, circles, intersections, segments, probe neighbours) on atoms already resident in HBM,
replayed as one hipGraph. Further legs (same JSON line):
  exterior   the exterior path on the same atoms;
  sdf        SDF of the complete surface on a 256^3 grid over the bounding box + 3 A, then the
             exterior field (voids excluded) on the same grid, then small queries at 64 and 4,096
             scattered points;
  protein    both surfaces of 7CGO (335,722 atoms, the reference's largest sample present), from MMTF;
  large      exterior path on 2.5M atoms (500 globules of 5,000 on a 360 A Fibonacci sphere) + memory;
  batch      4,096 molecules (N ~ logU[20, 5000]) in one launch sequence + 32^3 SDF grid each, vs
             one launch sequence per molecule (timed on the first 512 molecules); also the builds
             alone both ways;
  grad       20,000-atom globule, 1M queries within |SDF| <= 2 A, loss = sum w_i SDF(q_i), w ~ N(0, 1):
             forward (build + sampling) vs backward (stored / recomputed winners; or the build
             released after the forward pass and built again in the backward pass);
  cb         the reference's Contour-Buildup comparator against both SES paths on the 12 sample
             proteins and the 100k globule (the reference's benchmark comparison);
  cpu        host C++ build of the same stage sequence (the FP32 oracle restatement) on this machine's
             cores for the workloads above (oracle/cpu_baseline.py); its predicate counts give the
             compute side of the roofline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs main,exterior,...]
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
STAGES = ["grid", "circles_kernel", "circles_offsets_emit", "intersections_kernel", "intersections_offsets_emit",
          "circle_list_scan", "fill", "segments", "probe_neighbors", "torus"]
ALL_LEGS = "main,exterior,sdf,protein,large,batch,grad,cb,cpu"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--legs", default=ALL_LEGS)
    p.add_argument("--cell", type=float, default=0.0)
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="run one device copy of this size first (counter calibration under rocprofv3)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_builds(surf, steps, warmup):
    """Mean device time (ms) of `steps` graph replays of the whole stage sequence, hipEvents on the build stream."""
    import sesdf_amd._lib as L

    lib = L.lib()
    stream = surf.stream()
    L.check(lib.sesdf_build(surf._h), "build")
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


def replay_distribution(surf, reps):
    """Device time of each of `reps` graph replays, one hipEvent pair around each: median and spread."""
    import sesdf_amd._lib as L

    lib = L.lib()
    stream = surf.stream()
    ev = [(L.Event(), L.Event()) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        L.check(lib.sesdf_enqueue(surf._h), "enqueue")
        e1.record(stream)
    ev[-1][1].sync()
    L.check(lib.sesdf_sync(surf._h), "sync")
    t = np.array([e0.elapsed_ms(e1) for e0, e1 in ev])
    return {"n": reps, "median": float(np.median(t)), "p10": float(np.percentile(t, 10)),
            "p90": float(np.percentile(t, 90)), "min": float(t.min()), "max": float(t.max())}


def host_input_rate(surf, atoms, reps):
    """The reference's build() as a caller sees it: atoms from host memory (upload), the whole
    build, and the counts read back (sesdf_set_atoms + sesdf_build, blocking), host clock."""
    import ctypes

    import sesdf_amd._lib as L

    lib = L.lib()
    a = np.ascontiguousarray(atoms, np.float32)
    ptr = ctypes.c_void_p(a.ctypes.data)
    for _ in range(3):
        L.check(lib.sesdf_set_atoms(surf._h, ptr), "set_atoms")
        L.check(lib.sesdf_build(surf._h), "build")
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        L.check(lib.sesdf_set_atoms(surf._h, ptr), "set_atoms")
        L.check(lib.sesdf_build(surf._h), "build")
        t.append((time.perf_counter() - t0) * 1e3)
    t = np.array(t)
    med = float(np.median(t))
    return {"n": reps, "median_ms": med, "p10_ms": float(np.percentile(t, 10)), "p90_ms": float(np.percentile(t, 90)),
            "atoms_per_s": len(atoms) / (med * 1e-3), "input_bytes": int(a.nbytes)}


def stage_times(surf, reps=10):
    """Per-stage device times (ms): event pairs around each stage of a directly launched build."""
    import ctypes

    import sesdf_amd._lib as L

    acc = np.zeros(len(STAGES))
    buf = (ctypes.c_float * len(STAGES))()
    for _ in range(reps):
        L.check(L.lib().sesdf_build_timed(surf._h, buf, len(STAGES)), "build_timed")
        acc += np.array(buf[:])
    return dict(zip(STAGES, (acc / reps).tolist()))


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-calibrated), or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "topology_kernels_100k_traffic.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_dispatch {kernel}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def intersections_alg_bytes(c, n):
    """Algorithmic HBM bytes of one launch of the intersection kernel: atom i + one gather of its
    neighbours (32 N), neighbour ids + statuses (5 B per ordered circle), the touch masks of the
    intersected circles (16 Ci), staged records (4 I) and per-atom counts (12 N)."""
    return 32 * n + 5 * c["neighbor_nb"] + 16 * c["intersected_circle_nb"] + 4 * c["concave_patch_nb"] + 12 * n


def circles_alg_bytes(c, n, K=128):
    """Algorithmic HBM bytes of one launch of the neighbour + visibility kernel: atom i + one gather
    of each atom (16 N), neighbour ids, counts and statuses (4 C + 4 N + 1 C), the touch masks of the
    intersected circles (16 Ci), the intersected-circle bit masks (K / 8 per atom) and per-atom
    counts (16 N). Reads of the cell offsets (4 B per stencil row) are left out: a lower bound."""
    return 16 * n + 5 * c["neighbor_nb"] + 4 * n + 16 * c["intersected_circle_nb"] + (K // 8) * n + 16 * n


# Flops per predicate evaluation (F_alg = the oracle's counts in the reference's loop order, early
# exits included, times these): a circle (i, j) against a sphere k, the reference's test
# (operations.cu:126-204); a candidate triplet (two circle frames, the probe pair solve); one
# occluder against one of the two probe positions (squared distance against R^2).
FLOPS_VISIBILITY, FLOPS_TRIPLET, FLOPS_OCCLUDER = 30, 60, 10
FP32_PEAK_TFLOPS = 157.3


def compute_roofline(counts, main_out):
    """Compute side of the two O(n^2)-per-atom kernels: F_alg from the predicate counts, achieved
    TFLOP/s against the FP32 vector peak, and max(B_alg / 8 TB/s, F_alg / P_FP32) / t."""
    k = main_out["roofline_kernels"]
    flops = {"circles_count_kernel": counts["visibility_tests"] * FLOPS_VISIBILITY,
             "intersections_count_kernel": counts["triplet_tests"] * FLOPS_TRIPLET + counts["occluder_tests"] * FLOPS_OCCLUDER}
    res = {}
    for name, f in flops.items():
        t = k[name]["kernel_ms"] * 1e-3
        b = k[name]["alg_bytes_per_launch"]
        bound_s = max(b / (HBM_PEAK_GBS * 1e9), f / (FP32_PEAK_TFLOPS * 1e12))
        res[name] = {"flops_per_launch": int(f), "achieved_tflops": f / t / 1e12, "peak_tflops": FP32_PEAK_TFLOPS,
                     "frac_fp32": f / t / 1e12 / FP32_PEAK_TFLOPS, "frac_roofline": bound_s / t,
                     "binding": "fp32" if f / (FP32_PEAK_TFLOPS * 1e12) > b / (HBM_PEAK_GBS * 1e9) else "hbm"}
    res["predicate_counts"] = counts
    res["flops_per_test"] = {"visibility": FLOPS_VISIBILITY, "triplet": FLOPS_TRIPLET, "occluder": FLOPS_OCCLUDER}
    return res


def kernel_roofline(kernel, alg, kms):
    achieved = alg / (kms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(kernel), "kernel": kernel, "alg_bytes_per_launch": int(alg), "kernel_ms": kms}


def tuned(surf, args):
    """Launch tuning before any timed build: the autotuner's choice (cell edge, block shapes) unless
    --cell fixes the cell edge."""
    if args.cell > 0:
        return {"cell_size": args.cell, "circle_waves": 4, "intersection_waves": 4}
    from sesdf_amd.autotune import autotune

    t = autotune(surf, log=log)
    return {k: t[k] for k in ("cell_size", "circle_waves", "intersection_waves")}


def leg_main(args, atoms, out):
    from sesdf_amd import Sesdf

    n = atoms.shape[0]
    surf = Sesdf(atoms, probe_radius=1.4, cell_size=args.cell, use_graph=True)
    tune = tuned(surf, args)
    ms = time_builds(surf, args.steps, args.warmup)
    counts = surf.counts()
    stages = stage_times(surf)
    replays = replay_distribution(surf, 120)
    host = host_input_rate(surf, atoms, 100)
    # both O(n^2)-per-atom kernels; the headline roofline is the longer one's (the dominant kernel)
    lines = {
        "circles_count_kernel": kernel_roofline("circles_count_kernel", circles_alg_bytes(counts, n), stages["circles_kernel"]),
        "intersections_count_kernel": kernel_roofline("intersections_count_kernel", intersections_alg_bytes(counts, n),
                                                      stages["intersections_kernel"]),
    }
    dominant = max(lines, key=lambda k: lines[k]["kernel_ms"])
    out.update({
        "value": n / (ms * 1e-3), "ms_per_step": ms, "stage_ms": stages,
        "counts": {k: int(v) for k, v in counts.items()},
        "roofline": lines[dominant],
        "roofline_kernels": lines,
        "replay_ms": replays,
        "host_input": host,
        "tuning": tune,
    })
    log(f"main: {ms:.3f} ms/build, {n / (ms * 1e-3) / 1e6:.1f} M atoms/s; median of {replays['n']} replays "
        f"{replays['median']:.3f} ms; from host memory {host['median_ms']:.3f} ms")
    return surf


def leg_exterior(args, atoms, out):
    from sesdf_amd import Ssesdf

    ext = Ssesdf(atoms, probe_radius=1.4, cell_size=args.cell, use_graph=True)
    tune = tuned(ext, args)
    ms = time_builds(ext, args.steps, args.warmup)
    out["exterior"] = {"atoms_per_s": atoms.shape[0] / (ms * 1e-3), "ms_per_build": ms,
                       "torus": int(ext.counts()["torus_nb"]), "concave": int(ext.counts()["concave_patch_nb"]), "tuning": tune}
    log(f"exterior: {ms:.3f} ms/build")


def leg_sdf(args, atoms, surf, out):
    import torch

    import sesdf_amd._lib as L

    lo = atoms[:, :3].min(0) - atoms[:, 3].max() - 3
    hi = atoms[:, :3].max(0) + atoms[:, 3].max() + 3
    dim = 256
    sp = ((hi - lo) / (dim - 1)).astype(np.float32)
    res = torch.empty(dim ** 3, device="cuda")
    tune = None
    if args.cell <= 0:
        # the grid that suits the build does not suit the sampling: tune for build + grid
        from sesdf_amd.autotune import autotune

        t = autotune(surf, log=log, after=lambda: surf.sample_grid_ptr(lo, sp, (dim,) * 3, res.data_ptr(), band=2.0),
                     tag=f"+sdf{dim}")
        tune = {k: t[k] for k in ("cell_size", "circle_waves", "intersection_waves")}
    stream = surf.stream()
    surf.sample_grid_ptr(lo, sp, (dim,) * 3, res.data_ptr(), band=2.0)  # warm-up
    L.device_synchronize()
    e0, e1 = L.Event(), L.Event()
    reps = 5
    e0.record(stream)
    for _ in range(reps):
        surf.sample_grid_ptr(lo, sp, (dim,) * 3, res.data_ptr(), band=2.0)
    e1.record(stream)
    e1.sync()
    ms = e0.elapsed_ms(e1) / reps
    r = res.cpu().numpy()
    # the exterior field (voids excluded) on the same build and grid
    surf.sample_grid_ptr(lo, sp, (dim,) * 3, res.data_ptr(), band=2.0, exterior=True)  # warm-up
    L.device_synchronize()
    e0.record(stream)
    for _ in range(reps):
        surf.sample_grid_ptr(lo, sp, (dim,) * 3, res.data_ptr(), band=2.0, exterior=True)
    e1.record(stream)
    e1.sync()
    ms_ext = e0.elapsed_ms(e1) / reps
    # small queries at scattered points (caller order, under the point-sort threshold): latency
    g = torch.Generator(device="cuda").manual_seed(7)
    small = {}
    for n in (64, 4096):
        P = torch.zeros(n, 4, device="cuda")
        P[:, :3] = torch.tensor(lo, device="cuda") + torch.tensor(hi - lo, device="cuda") * torch.rand(n, 3, device="cuda", generator=g)
        v = torch.empty(n, device="cuda")
        surf.sample_points_ptr(P.data_ptr(), n, v.data_ptr(), band=2.0)  # warm-up
        L.device_synchronize()
        e0.record(stream)
        for _ in range(20):
            surf.sample_points_ptr(P.data_ptr(), n, v.data_ptr(), band=2.0)
        e1.record(stream)
        e1.sync()
        small[f"points_{n}_ms"] = e0.elapsed_ms(e1) / 20
    out["sdf"] = {"grid": "256^3", "band_A": 2.0, "ms": ms, "samples_per_s": dim ** 3 / (ms * 1e-3),
                  "inside_fraction": float((r < 0).mean()), "surface_band_fraction": float((np.abs(r) < 2).mean()),
                  "exterior_ms": ms_ext, "exterior_inside_fraction": float((res.cpu().numpy() < 0).mean()),
                  "scattered": small, "tuning": tune}
    log(f"sdf 256^3: {ms:.3f} ms, exterior field {ms_ext:.3f} ms; scattered points: "
        + ", ".join(f"{k} {v:.3f}" for k, v in small.items()))


def leg_protein(args, out):
    """The largest of the reference's sample proteins present (7CGO, 335,722 atoms), read from its
    MMTF file with the reference's radii: both surfaces, hipGraph replays."""
    from sesdf_amd import Sesdf, Ssesdf, load

    atoms = load(os.path.join(ROOT, "tests", "golden", "mmtf", "7CGO.mmtf"))
    res = {"protein": "7CGO", "atoms": int(atoms.shape[0])}
    for name, cls in (("complete", Sesdf), ("exterior", Ssesdf)):
        surf = cls(atoms, probe_radius=1.4, cell_size=args.cell, use_graph=True)
        tune = tuned(surf, args)
        ms = time_builds(surf, max(10, args.steps // 2), 3)
        c = surf.counts()
        res[name] = {"ms_per_build": ms, "atoms_per_s": atoms.shape[0] / (ms * 1e-3),
                     "concave": int(c["concave_patch_nb"]), "tuning": tune}
        log(f"7CGO {name}: {ms:.3f} ms/build")
        surf.close()
    out["protein_7CGO"] = res


CB_CASES = ["1AGA", "101M", "1VIS", "7SC0", "3EAM", "7DBB", "1A8R", "7O0U", "1AON", "7RGD", "3JC8", "7CGO"]


def time_cb(cb, steps, warmup):
    """Mean device time (ms) of `steps` graph replays of a Contour-Buildup build, hipEvents on its stream."""
    import sesdf_amd._lib as L

    cb.build()
    for _ in range(warmup):
        cb.enqueue()
    cb.sync()
    e0, e1 = L.Event(), L.Event()
    stream = cb.stream()
    e0.record(stream)
    for _ in range(steps):
        cb.enqueue()
    e1.record(stream)
    e1.sync()
    cb.sync()
    return e0.elapsed_ms(e1) / steps


def leg_cb(args, out, globule):
    """The reference's benchmark comparison (app/benchmarks/main.cu:155-202): Contour-Buildup (16^3 grid,
    the reference's constant), the exterior path and the complete surface, on the sample proteins present
    (the first twelve cases of the reference's list) and the 100k globule; default launch parameters for
    all three (no autotuning), each timed as graph replays on atoms resident in HBM."""
    from sesdf_amd import ContourBuildup, Sesdf, Ssesdf, load

    cases = [(c, load(os.path.join(ROOT, "tests", "golden", "mmtf", c + ".mmtf"))) for c in CB_CASES]
    cases.append(("globule_100k", globule))
    res = []
    steps = max(10, args.steps // 2)
    for name, atoms in cases:
        row = {"case": name, "atoms": int(atoms.shape[0])}
        cb = ContourBuildup(atoms, probe_radius=1.4, grid_dimensions=16)
        row["cb_ms"] = time_cb(cb, steps, 3)
        c = cb.counts()
        row["cb_probes"], row["cb_tori"] = int(c["probe_nb"]), int(c["torus_nb"])
        cb.close()
        for key, cls in (("ssesdf_ms", Ssesdf), ("sesdf_ms", Sesdf)):
            surf = cls(atoms, probe_radius=1.4, use_graph=True)
            row[key] = time_builds(surf, steps, 3)
            surf.close()
        row["cb_over_ssesdf"] = row["cb_ms"] / row["ssesdf_ms"]
        row["cb_over_sesdf"] = row["cb_ms"] / row["sesdf_ms"]
        log(f"cb {name}: CB {row['cb_ms']:.3f} ms, SSESDF {row['ssesdf_ms']:.3f} ms, SESDF {row['sesdf_ms']:.3f} ms")
        res.append(row)
    out["contour_buildup"] = {"grid": "16^3", "probe_radius": 1.4, "cases": res}


def leg_large(args, out):
    from sesdf_amd import Ssesdf, synth
    import sesdf_amd._lib as L

    big = synth.fibonacci_globules(500, 5000, 360.0, synth.SEED_BASE + 2)
    L.device_synchronize()
    free0, _ = L.mem_info()
    sbig = Ssesdf(big, probe_radius=1.4, cell_size=args.cell, use_graph=True)
    tune = tuned(sbig, args)
    ms = time_builds(sbig, max(5, args.steps // 5), 2)
    free1, _ = L.mem_info()
    cb = sbig.counts()
    out["large_2p5M_exterior"] = {
        "atoms": int(big.shape[0]), "ms_per_build": ms, "atoms_per_s": big.shape[0] / (ms * 1e-3),
        "concave": int(cb["concave_patch_nb"]), "torus": int(cb["torus_nb"]),
        "handle_device_bytes": int(cb["device_bytes"]), "device_mem_delta_bytes": int(free0 - free1),
        "stage_ms": stage_times(sbig, 3), "tuning": tune}
    log(f"large: {ms:.3f} ms/build")
    sbig.close()


def leg_batch(args, out):
    import torch

    import sesdf_amd._lib as L
    from sesdf_amd import Sesdf, synth
    from sesdf_amd.autotune import autotune
    from sesdf_amd.batch import SesdfBatch, set_tuning

    sizes = synth.batch_sizes(4096, 20, 5000, synth.SEED_BASE + 3)
    atoms, offs = synth.batch(sizes, synth.SEED_BASE + 3)
    M, dim, margin, band = len(sizes), 32, 3.0, 2.0
    batch = SesdfBatch(atoms, offs, build_surface=False)
    grids = torch.empty(M * dim ** 3, device="cuda")
    # tuned for the whole workload: the build and the per-molecule grids that follow it
    tune = autotune(batch, log=log, after=lambda: batch.sample_grids_ptr(dim, grids.data_ptr(), margin=margin, band=band),
                    tag=f"+grids{dim}")
    ms_build = time_builds(batch, max(5, args.steps // 5), 2)
    batch.sample_grids_ptr(dim, grids.data_ptr(), margin=margin, band=band)
    L.device_synchronize()
    # both ways on the host clock, from an idle device to an idle device: median of three
    runs_batched = []
    for _ in range(3):
        t0 = time.perf_counter()
        L.check(L.lib().sesdf_enqueue(batch._h), "enqueue")
        batch.sample_grids_ptr(dim, grids.data_ptr(), margin=margin, band=band)
        L.device_synchronize()
        runs_batched.append((time.perf_counter() - t0) * 1e3)
        L.check(L.lib().sesdf_sync(batch._h), "sync")
    ms_batched = float(np.median(runs_batched))
    # one launch sequence per molecule (own handle, own graph), same work, first P molecules
    P = 512
    singles = []
    for m in range(P):
        s = Sesdf(atoms[int(offs[m]):int(offs[m + 1])], use_graph=True)
        # the same launch tuning as the batch, so that only the launch structure differs
        set_tuning(s, tune["cell_size"], tune["circle_waves"], tune["intersection_waves"])
        L.check(L.lib().sesdf_build(s._h), "build")
        singles.append(s)
    outs = torch.empty(P * dim ** 3, device="cuda")
    descs = []
    for m in range(P):
        mol = atoms[int(offs[m]):int(offs[m + 1])]
        lo = (mol[:, :3] - mol[:, 3:4]).min(0) - margin
        hi = (mol[:, :3] + mol[:, 3:4]).max(0) + margin
        descs.append((lo.astype(np.float32), ((hi - lo) / (dim - 1)).astype(np.float32)))
    L.device_synchronize()
    runs_single = []
    for _ in range(3):
        t0 = time.perf_counter()
        for m, s in enumerate(singles):
            L.check(L.lib().sesdf_enqueue(s._h), "enqueue")
            s.sample_grid_ptr(descs[m][0], descs[m][1], (dim,) * 3, outs.data_ptr() + 4 * m * dim ** 3, band=band)
        L.device_synchronize()
        runs_single.append((time.perf_counter() - t0) * 1e3)
    ms_single = float(np.median(runs_single))
    # the builds alone, both ways (no SDF grids)
    t0 = time.perf_counter()
    for s in singles:
        L.check(L.lib().sesdf_enqueue(s._h), "enqueue")
    L.device_synchronize()
    ms_single_build = (time.perf_counter() - t0) * 1e3
    # same accuracy: the batch's grids equal the standalone grids
    same = bool(torch.allclose(grids[:P * dim ** 3], outs, atol=1e-5))
    batched_rate = M / (ms_batched * 1e-3)
    single_rate = P / (ms_single * 1e-3)
    out["batch"] = {"molecules": M, "atoms": int(atoms.shape[0]), "sdf_grid": f"{dim}^3",
                    "ms_batched_build_plus_sdf": ms_batched, "runs_ms_batched": runs_batched, "ms_batched_build": ms_build,
                    "runs_ms_per_molecule_launches": runs_single, "clock": "host, device idle to idle, median of 3",
                    "molecules_per_s_batched": batched_rate,
                    "per_molecule_launch_molecules": P, "ms_per_molecule_launches": ms_single,
                    "molecules_per_s_per_launch": single_rate, "speedup": batched_rate / single_rate,
                    "ms_per_molecule_launches_build_only": ms_single_build,
                    "speedup_build_only": (M / ms_build) / (P / ms_single_build),
                    "grids_equal_standalone": same,
                    "tuning": {k: tune[k] for k in ("cell_size", "circle_waves", "intersection_waves")}}
    log(f"batch: {batched_rate:.0f} mol/s batched vs {single_rate:.0f} per launch; "
        f"builds alone {M / ms_build * 1e3:.0f} vs {P / ms_single_build * 1e3:.0f}")
    for s in singles:
        s.close()
    batch.close()


def leg_grad(args, out):
    import torch

    import sesdf_amd._lib as L
    from sesdf_amd import Sesdf, synth

    atoms = synth.globule(20000, synth.SEED_BASE + 4)
    surf = Sesdf(atoms)
    band = 2.0
    g = torch.Generator(device="cuda").manual_seed(int(synth.SEED_BASE + 4))
    lo = torch.tensor(atoms[:, :3].min(0) - 4, device="cuda")
    hi = torch.tensor(atoms[:, :3].max(0) + 4, device="cuda")
    keep = []
    total = 0
    while total < 1_000_000:
        cand = torch.zeros(2_000_000, 4, device="cuda")
        cand[:, :3] = lo + (hi - lo) * torch.rand(2_000_000, 3, device="cuda", generator=g)
        s = torch.empty(cand.shape[0], device="cuda")
        surf.sample_points_ptr(cand.data_ptr(), cand.shape[0], s.data_ptr(), band=band)
        L.device_synchronize()
        sel = cand[s.abs() < band]
        keep.append(sel)
        total += sel.shape[0]
    Q = torch.cat(keep)[:1_000_000].contiguous()
    nq = Q.shape[0]
    w = torch.randn(nq, device="cuda", generator=g)
    d_atoms = torch.from_numpy(atoms).cuda()
    val = torch.empty(nq, device="cuda")
    codes = torch.empty(nq, device="cuda", dtype=torch.int64)
    grad = torch.zeros(len(atoms), 4, device="cuda")
    scratch = torch.empty(len(atoms), 4, device="cuda")
    stream = surf.stream()

    def fwd():
        surf.build_device(d_atoms.data_ptr())
        surf.sample_points_ptr(Q.data_ptr(), nq, val.data_ptr(), band=band, d_codes=codes.data_ptr())

    def bwd(store):
        surf.sdf_backward_ptr(Q.data_ptr(), w.data_ptr(), nq, grad.data_ptr(), scratch.data_ptr(), band=band,
                              d_codes=codes.data_ptr() if store else None)

    def timed(fn, reps=5):
        fn()
        L.device_synchronize()
        e0, e1 = L.Event(), L.Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.sync()
        return e0.elapsed_ms(e1) / reps

    t_f = timed(fwd)
    t_bs = timed(lambda: bwd(True))
    t_br = timed(lambda: bwd(False))

    # release: the forward frees the whole build, the backward builds it again (allocation
    # included), on the host clock from an idle device to an idle device
    def wall(fn, reps=5):
        times = []
        for _ in range(reps + 1):
            L.device_synchronize()
            t0 = time.perf_counter()
            fn()
            L.device_synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(times[1:]))

    fwd_wall = wall(fwd)
    held = surf.counts()["device_bytes"]
    fwd_release = wall(lambda: (fwd(), surf.release()))
    bwd_release = []
    for _ in range(5):
        surf.release()
        L.device_synchronize()
        t0 = time.perf_counter()
        surf.build_device(d_atoms.data_ptr())
        bwd(False)
        L.device_synchronize()
        bwd_release.append((time.perf_counter() - t0) * 1e3)
    t_brel = float(np.median(bwd_release))
    out["grad"] = {"atoms": len(atoms), "queries": nq, "band_A": band, "ms_forward_build_plus_sdf": t_f,
                   "ms_backward_stored": t_bs, "ms_backward_recompute": t_br,
                   "bwd_over_fwd_stored": t_bs / t_f, "bwd_over_fwd_recompute": t_br / t_f,
                   "stored_winner_bytes": nq * 8,
                   "release": {"clock": "host, device idle to idle, median of 5", "ms_forward": fwd_wall,
                               "ms_forward_then_release": fwd_release, "ms_backward_rebuild_recompute": t_brel,
                               "bwd_over_fwd": t_brel / fwd_wall, "device_bytes_held_between": 0,
                               "device_bytes_of_the_build": int(held)}}
    log(f"grad: fwd {t_f:.3f} ms, bwd stored {t_bs:.3f} ms, recompute {t_br:.3f} ms; release: fwd {fwd_wall:.3f} ms, "
        f"bwd with rebuild {t_brel:.3f} ms")


def cpu_model():
    try:
        import subprocess

        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def leg_cpu():
    """Host C++ build of the same stage sequence (the oracle's FP32 restatement, threaded over atoms)
    on this machine's cores, one child process per workload (oracle/cpu_baseline.py)."""
    import subprocess

    hw = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = hw
    # at most 16 threads: the share of the host that one GPU's jobs may use on these machines
    threads = max(1, min(allowed, 16))
    res = {}
    for w in ("c0", "c1", "c1x", "c2", "c3"):
        t0 = time.perf_counter()
        p = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), w, str(threads)],
                           capture_output=True, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"cpu baseline {w} failed: {p.stderr[-2000:]}")
        res[w] = json.loads(p.stdout.strip().splitlines()[-1])
        log(f"cpu {w}: {time.perf_counter() - t0:.1f} s wall")
    c1 = res["c1"]
    return {"value": c1["atoms_per_s"], "unit": "atoms/s", "cores": threads, "kind": "port",
            "sample": f"complete-surface builds of the 100,000-atom workload, FP32, {threads} threads, median of "
                      f"{len(c1['runs_s'])} ({c1['build_s']:.3f} s) after one warm-up; the other workloads below",
            "hardware_concurrency": hw, "affinity_cpus": allowed, "cpu_model": cpu_model(),
            "workloads": res}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    legs = set(args.legs.split(","))
    from sesdf_amd import synth

    if args.calibrate_copy_mib:
        import torch

        a = torch.ones(args.calibrate_copy_mib * 262144, device="cuda")
        b = a.clone()
        torch.cuda.synchronize()
        del a, b

    atoms, _ = synth.globule_voids(100_000, 10, synth.SEED_BASE + 1)
    out = {
        "metric": "SES build atoms/s (100k-atom globule, complete surface)", "value": None, "unit": "atoms/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "100k-atom globule with 10 voids, probe 1.4 A, complete surface, hipGraph replay",
                   "atoms": int(atoms.shape[0]), "probe_radius": 1.4,
                   "launch": "autotuned (cell edge, block shapes)" if args.cell <= 0 else f"cell {args.cell} A"},
    }
    surf = leg_main(args, atoms, out) if "main" in legs else None
    if "exterior" in legs:
        leg_exterior(args, atoms, out)
    if "sdf" in legs and surf is not None:
        leg_sdf(args, atoms, surf, out)
    if surf is not None:
        surf.close()
    if "protein" in legs:
        leg_protein(args, out)
    if "large" in legs:
        leg_large(args, out)
    if "batch" in legs:
        leg_batch(args, out)
    if "grad" in legs:
        leg_grad(args, out)
    if "cb" in legs:
        leg_cb(args, out, atoms)
    out["cpu_baseline"] = leg_cpu() if "cpu" in legs else None
    if out["cpu_baseline"] and "roofline_kernels" in out:
        out["roofline_compute"] = compute_roofline(out["cpu_baseline"]["workloads"]["c1"]["predicate_counts"], out)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
