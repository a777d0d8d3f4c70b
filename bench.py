#!/usr/bin/env python3
"""Depth-9 screened Poisson reconstruction throughput on one MI355X.

Workload: 800,000 points on a smooth union of six seeded random ellipsoids
(semi-axes 0.15-0.45), sampling density varying 10x, position noise 0.001,
exact analytic normals; depth 9, full_depth 4, point_weight 2.0, iters 8.
A step is one complete reconstruct() (inputs already resident in HBM) of the
mixed-precision path: float64 inputs, octree stages, vectors, base-level solve
and isovalue, with the finer levels' Laplacian + screening operators stored and
streamed as float32.  The full float64 path and the float32 path are timed the
same way; mixed and float32 fields are compared with the float64 field at
812,000 queries (every input plus 12,000 points straddling refinement
interfaces).

Prints one JSON line (metric/value per the benchmark contract) with a
``roofline`` object for the sliced Gauss-Seidel relaxation of the finer levels
and a ``cpu_baseline`` object (the CPU oracle on the host's cores on the same
input, with the core count and its provenance).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "warp-psr_amd"))

METRIC = "Mpts/s SPSR depth-9 mixed vs FP64 at field err<=1e-6; relax HBM GB/s"
HBM_PEAK_GBS = 8000.0
PARAMS = dict(depth=9, full_depth=4, point_weight=2.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC = ROOT / "profiles" / "relax_traffic_ellipsoid800k_d9.json"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    counter runs (scaled by the copy calibration), or None."""
    if not TRAFFIC.exists():
        return None
    doc = json.loads(TRAFFIC.read_text())
    return doc["bytes_per_launch"] if doc.get("kernel") == kernel else None


def relax_bytes(st, v, s, tiled):
    """Algorithmic bytes of one level's relaxation: per sweep, the operator's
    stored values (nnz * v), the right-hand side, and x read and written once
    per row (3 * R * s), plus the row structure -- one int per row for the
    tiled form (its grid position; neighbours are found by offset), the CSR
    row extents and nnz column indices for the colour-per-launch form."""
    per_sweep = st["nnz"] * v + 3 * st["rows"] * s
    per_sweep += st["rows"] * 4 if tiled else (st["rows"] + 1) * 4 + st["nnz"] * 4
    return per_sweep * st["sweeps"]


def relax_roofline(metadata, value_bytes, vector_bytes, kernel=None):
    """Sliced Gauss-Seidel relaxation of every finer level: algorithmic bytes
    (relax_bytes) over the HIP-event span of its launches on the
    reconstruction stream.  The tiled form (k_relax_tiled) launches once per
    wave, over row records the level's operator pass wrote; the
    colour-per-launch form (k_relax_wave_args) four times per wave.
    """
    stats = metadata["relaxation"]
    v, s = value_bytes, vector_bytes
    tiled = all(st.get("tile", 0) > 0 for st in stats)
    total = sum(relax_bytes(st, v, s, tiled) for st in stats)
    seconds = sum(st["seconds"] for st in stats)
    launches = sum(st["launches"] for st in stats)
    achieved = total / seconds / 1e9
    fine = max(stats, key=lambda st: st["rows"])
    name = "k_relax_tiled" if tiled else "k_relax_wave_args"
    return {
        "kernel": f"{name} (sliced Gauss-Seidel, all finer levels)",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": measured_traffic(kernel) if kernel else None,
        "traffic_unit": "bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
        "levels": [st["depth"] for st in stats],
        "launches": launches,
        "algorithmic_bytes": total,
        "bytes_per_launch": round(total / launches),
        "seconds": seconds,
        "avg_launch_us": round(seconds / launches * 1e6, 3),
        "finest": {"depth": fine["depth"], "rows": fine["rows"], "nnz": fine["nnz"], "launches": fine["launches"],
                   "seconds": fine["seconds"]},
        "per_level": [{"depth": st["depth"], "launches": st["launches"], "us_per_launch": round(st["seconds"] / st["launches"] * 1e6, 3),
                       "GBps": round(relax_bytes(st, v, s, tiled) / st["seconds"] / 1e9, 1), "tile": st.get("tile", 0),
                       "tiles": st.get("tiles", 0), "tile_entries": st.get("tile_entries", 0)} for st in stats],
    }


def mesh_throughput(torch, recon, points, dev, resolution=256, repeats=3):
    """Marching cubes of the float32 depth-9 field on a resolution^3 node grid
    over the input's bounding box (sparse backend, 16-layer slabs): every node
    evaluated once for the classification, the retained cells' corners again
    for the vertices.  Median of `repeats` after one warm-up."""
    from warp_psr_amd.meshing import extract_mesh
    lo, hi = points.min(axis=0).astype(np.float64), points.max(axis=0).astype(np.float64)
    pad = 0.02 * (hi - lo)
    bounds = (tuple(lo - pad), tuple(hi + pad))
    mesh = extract_mesh(recon, resolution=resolution, bounds=bounds)
    times = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t = time.perf_counter()
        mesh = extract_mesh(recon, resolution=resolution, bounds=bounds)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    dt = sorted(times)[len(times) // 2]
    cells = (resolution - 1) ** 3
    log(f"mesh: {resolution}^3 nodes, {mesh.metadata['retained_cells']} retained cells, {len(mesh.vertices)} vertices, "
        f"{dt * 1e3:.1f} ms")
    return {"resolution": resolution, "backend": "sparse", "seconds": round(dt, 4), "Mcells_s": round(cells / dt / 1e6, 2),
            "field_evaluations": mesh.metadata["field_evaluations"], "retained_cells": mesh.metadata["retained_cells"],
            "vertices": int(len(mesh.vertices)), "triangles": int(len(mesh.indices) // 3)}


def batch_throughput(torch, synthetic, dev, concurrency=8):
    """Config 5: 64 independent clouds (5k-200k points, depth 7/8) in one
    batched call on `concurrency` worker streams, in mixed precision (float64
    inputs, float32 finer-level operators) and in float32, against the same
    clouds one by one.  Reports aggregate points/s, the p50/p99 of each
    cloud's completion time within the batch, and the largest field
    difference of any cloud's mixed result from its own float64 run (at the
    cloud's input points)."""
    from warp_psr_amd import reconstruct, reconstruct_batch
    t0 = time.perf_counter()
    work = synthetic.workload_batch()
    log(f"batch workload generated in {time.perf_counter() - t0:.1f}s")
    total = sum(len(p) for p, _, _ in work)
    out = {"clouds": len(work), "points": total, "depths": "7/8", "full_depth": 4, "concurrency": concurrency}
    for label, dtype, extra in (("mixed", np.float64, {"mixed": True}), ("f32", np.float32, {})):
        clouds = [(torch.from_numpy(p.astype(dtype)).to(dev), torch.from_numpy(n.astype(dtype)).to(dev)) for p, n, _ in work]
        opts = [dict(depth=d, full_depth=4, point_weight=2.0, **extra) for _, _, d in work]
        torch.cuda.synchronize()
        reconstruct_batch(clouds, per_cloud=opts, max_concurrency=concurrency)  # warm-up (graph captures)
        torch.cuda.synchronize()
        t = time.perf_counter()
        results = reconstruct_batch(clouds, per_cloud=opts, max_concurrency=concurrency)
        torch.cuda.synchronize()
        batched = time.perf_counter() - t
        done = sorted(r.metadata["batch_completion_seconds"] for r in results)
        leg = {"dtype": label, "seconds": round(batched, 4), "Mpts_s": round(total / batched / 1e6, 3),
               "completion_p50_s": round(float(np.percentile(done, 50)), 4),
               "completion_p99_s": round(float(np.percentile(done, 99)), 4)}
        if label == "mixed":
            worst = 0.0
            for (p, n), o, r in zip(clouds, opts, results):
                o64 = {k: v for k, v in o.items() if k != "mixed"}
                r64 = reconstruct(p, n, **o64)
                v, _, ok = r.evaluate(p)
                v64, _, ok64 = r64.evaluate(p)
                both = (ok & ok64).bool()
                worst = max(worst, float((v - v64).abs()[both].max()))
            leg["max_field_error_vs_fp64"] = worst
            leg["within_1e-6"] = worst <= 1e-6
        del results
        t = time.perf_counter()
        for (p, n), o in zip(clouds, opts):
            r = reconstruct(p, n, **o)
        torch.cuda.synchronize()
        serial = time.perf_counter() - t
        leg.update(one_by_one_seconds=round(serial, 4), one_by_one_Mpts_s=round(total / serial / 1e6, 3))
        log(f"batch {label}: {len(work)} clouds, {total} points, {batched * 1e3:.1f} ms batched "
            f"(p50 {leg['completion_p50_s'] * 1e3:.1f} ms, p99 {leg['completion_p99_s'] * 1e3:.1f} ms), "
            f"{serial * 1e3:.1f} ms one by one")
        out[label] = leg
    return out


def cpu_environment():
    """Logical CPUs, this process's affinity and the visible cgroup-v2 cpu.max
    quotas (the provenance the reference records with its CPU timings,
    tools/cpu_environment.py:7-43), and the thread count they allow."""
    quotas = []
    try:
        group = next(line[3:] for line in Path("/proc/self/cgroup").read_text().splitlines() if line.startswith("0::"))
        root = Path("/sys/fs/cgroup")
        current = root / group.lstrip("/")
        while True:
            limit = current / "cpu.max"
            if limit.exists():
                quota, period = limit.read_text().split()
                quotas.append({"group": str(current.relative_to(root)), "cpu_max": f"{quota} {period}",
                               "cpu_equivalents": None if quota == "max" else int(quota) / int(period)})
            if current == root or root not in current.parents:
                break
            current = current.parent
    except (OSError, StopIteration, ValueError):
        pass
    affinity = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    finite = [q["cpu_equivalents"] for q in quotas if q["cpu_equivalents"] is not None]
    limits = [os.cpu_count() or 1]
    if affinity:
        limits.append(len(affinity))
    if finite:
        limits.append(max(1, int(min(finite))))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        limits.append(int(env))
    return {"logical_cpus": os.cpu_count(), "affinity_cpus": len(affinity) if affinity else None,
            "cgroup_quotas": quotas, "omp_num_threads_env": env, "threads": min(limits)}


def cpu_baseline(points, normals, serial_sample):
    """The CPU oracle on the host cores: its OpenMP build (parallel loops over
    rows, queries and aggregates with unchanged per-row arithmetic; bit-equal
    to the serial build) on the full benchmark input after one warm-up run on
    a 50,000-point sample; and the serial build on that sample for the
    one-core rate."""
    import oracle  # the checker; timed here as the CPU reference path
    env = cpu_environment()
    os.environ["OMP_NUM_THREADS"] = str(env["threads"])
    oracle.reconstruct(points[:serial_sample], normals[:serial_sample], mixed=True, openmp=True, **PARAMS)
    t = time.perf_counter()
    oracle.reconstruct(points, normals, mixed=True, openmp=True, **PARAMS)
    dt = time.perf_counter() - t
    t = time.perf_counter()
    oracle.reconstruct(points[:serial_sample], normals[:serial_sample], mixed=True, **PARAMS)
    ds = time.perf_counter() - t
    log(f"cpu baseline: OpenMP oracle, {env['threads']} threads, {len(points)} points: {dt:.1f} s; serial oracle "
        f"{serial_sample} points: {ds:.1f} s")
    return {
        "value": round(len(points) / dt / 1e6, 6),
        "unit": "Mpts/s",
        "cores": env["threads"],
        "kind": "port",
        "sample": f"the full {len(points):,}-point depth-9 workload in mixed precision, OpenMP C++ oracle on "
                  f"{env['threads']} threads, one run after a warm-up on a {serial_sample:,}-point sample: {dt:.2f} s",
        "environment": env,
        "serial": {"points": serial_sample, "seconds": round(ds, 3), "Mpts_s": round(serial_sample / ds / 1e6, 6),
                   "cores": 1},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=800000)
    ap.add_argument("--cpu-sample", type=int, default=50000, help="points of the CPU warm-up and serial runs")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-mesh", action="store_true")
    ap.add_argument("--skip-float32", action="store_true", help="no float32 leg (implies --skip-mesh)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (read 1 GiB, write 1 GiB) three times, for scaling "
                         "FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: a reconstruction runs on one "
              "MI355X", file=sys.stderr)
        return 2

    import torch
    from warp_psr_amd import reconstruct, synthetic

    if args.copy_calibration:
        n = 1 << 28  # 1 GiB of float32
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "dispatches": 3}), flush=True)
        return 0

    t0 = time.perf_counter()
    p32, n32 = synthetic.workload_ellipsoids(args.points, dtype=np.float32)
    log(f"workload generated in {time.perf_counter() - t0:.1f}s")
    p64, n64 = p32.astype(np.float64), n32.astype(np.float64)
    dev = torch.device("cuda:0")
    tp32, tn32 = torch.from_numpy(p32).to(dev), torch.from_numpy(n32).to(dev)
    tp64, tn64 = torch.from_numpy(p64).to(dev), torch.from_numpy(n64).to(dev)
    torch.cuda.synchronize()

    def timed(tp, tn, steps, warmup, label, **extra):
        for _ in range(warmup):
            r = reconstruct(tp, tn, **PARAMS, **extra)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            r = reconstruct(tp, tn, **PARAMS, **extra)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        log(f"{label}: {dt * 1e3:.1f} ms/step, {len(tp) / dt / 1e6:.3f} Mpts/s; stages {r.metadata['timings_seconds']}")
        return r, dt

    rmx, dtmx = timed(tp64, tn64, args.steps, args.warmup, "mixed", mixed=True)
    out = {
        "metric": METRIC,
        "value": round(args.points / dtmx / 1e6, 4),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dtmx * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed",
        "data": "synthetic",
        "config": {"workload": "six-ellipsoid union, 800,000 points, 10x density variation, noise 0.001, "
                               "depth 9, full_depth 4, point_weight 2.0, iters 8 (SPSR)",
                   "points": args.points, "depth": 9, "full_depth": 4, "point_weight": 2.0,
                   "precision": "mixed: float64 inputs, octree, vectors, base solve and isovalue; "
                                "float32 finer-level operators",
                   "basis_counts": rmx.metadata["basis_counts"]},
    }
    out["roofline"] = relax_roofline(rmx.metadata, 4, 8, "k_relax_tiled<double, float>")
    out["stage_seconds"] = rmx.metadata["timings_seconds"]
    out["normal_average_wait_seconds"] = rmx.metadata["normal_average_wait_seconds"]
    r32 = None
    if not args.skip_float32:
        r32, dt32 = timed(tp32, tn32, max(1, args.steps // 2), 1, "float32")
        out["float32"] = {"Mpts_s": round(args.points / dt32 / 1e6, 4), "ms_per_step": round(dt32 * 1e3, 3),
                          "relax_roofline": relax_roofline(r32.metadata, 4, 4)}
    if not args.skip_fp64:
        r64, dt64 = timed(tp64, tn64, max(1, args.steps // 2), 1, "float64")
        out["float64"] = {"Mpts_s": round(args.points / dt64 / 1e6, 4), "ms_per_step": round(dt64 * 1e3, 3),
                          "relax_roofline": relax_roofline(r64.metadata, 8, 8)}
        keys = r64.diagnostics.array("keys")
        iq = synthetic.interface_queries(keys, 9, r64.diagnostics.array("model_to_cube"),
                                         r64.diagnostics.array("cube_to_model"), p64)
        q64 = np.concatenate([p64, iq])
        v64, _, ok64 = r64.evaluate(torch.from_numpy(q64).to(dev))
        v64, ok64 = v64.cpu().numpy(), ok64.cpu().numpy().astype(bool)

        def error(r, q):
            v, _, ok = r.evaluate(torch.from_numpy(q).to(dev))
            v = v.cpu().numpy().astype(np.float64)
            both = ok.cpu().numpy().astype(bool) & ok64
            err = np.abs(v - v64)[both]
            return {"queries": int(len(q)), "max_abs": float(err.max()), "iso_abs": abs(r.iso_value - r64.iso_value),
                    "same_octree": bool(np.array_equal(r.diagnostics.array("keys"), keys)),
                    "within_1e-6": bool(err.max() <= 1e-6)}

        out["field_error_vs_fp64"] = error(rmx, q64)
        if r32 is not None:
            out["float32"]["field_error_vs_fp64"] = error(r32, q64.astype(np.float32))
    if not args.skip_mesh and r32 is not None:
        out["mesh"] = mesh_throughput(torch, r32, p32, dev)
    if not args.skip_batch:
        out["batch"] = batch_throughput(torch, synthetic, dev)
    if not args.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(p64, n64, args.cpu_sample)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
