"""Inputs and parameters of the reference's threshold fixtures and large torus
runs (docs/results/thresholds.json, torus-100k/1m/1m-d9.json, extracted into
tests/golden/reference_reports.json), shared by the CPU and GPU tests."""
import hashlib
import json
from pathlib import Path

import numpy as np

from warp_psr_amd import synthetic

REPORTS = json.loads((Path(__file__).parent / "golden" / "reference_reports.json").read_text())
THRESHOLDS = REPORTS["thresholds"]
LARGE_TORUS = {r["name"]: r for r in REPORTS["large_torus"]}


def sha(a):
    return hashlib.sha256(np.asarray(a, dtype="<i8").tobytes()).hexdigest()


def threshold_input(f):
    p = f["params"]
    dtype = np.float32 if p["precision"] == "float" else np.float64
    if "ownership" in f["name"]:
        return synthetic.ownership_input(dtype)
    return synthetic.fixture_input(p["shape"], p["count"], p["seed"], p["precision"])


def threshold_params(f):
    p = f["params"]
    return dict(depth=p["depth"], full_depth=p["full_depth"], base_depth=p["base_depth"], solve_depth=p["solve_depth"],
                kernel_depth=p["kernel_depth"], samples_per_node=p["samples_per_node"], scale=p["scale"],
                point_weight=p["weight"], exact_interpolation=p["exact"], iters=p["iters"],
                base_v_cycles=p["base_v_cycles"], cg_accuracy=p["cg_accuracy"])


def selected_splat(diagnostics, f):
    """The raw and top splat depth of the aggregate the reference's bisection
    selected: aggregates in first-input order, cancelled ones left out, as
    in splats.txt (tools/check_thresholds.py:150-156)."""
    p = f["params"]
    first = diagnostics.array("first_indices")
    order = np.argsort(first, kind="stable")
    live = np.linalg.norm(diagnostics.array("normal_sums").reshape(-1, 3)[order], axis=1) > 0
    dw = diagnostics.array("depths_and_weights").reshape(-1, 2)[order][live]
    raw = dw[f["aggregate_index"], 0]
    top = int(np.ceil(np.clip(raw, p["base_depth"], p["depth"])))
    return raw, top


def torus_input(name):
    r = LARGE_TORUS[name]
    count = 100000 if name == "torus-100k" else 1000000
    return synthetic.fixture_input("torus", count, synthetic.SEED, "float")


def torus_params(name):
    return dict(LARGE_TORUS[name]["parameters"])


# Large parity cases whose oracle results are committed as SHA-256s
# (tests/golden/oracle_hashes.json, written by tests/golden/make_oracle_hashes.py).
LARGE_CASES = ("sphere20k-d6", "torus250k-d8-pw0", "torus250k-d8-pw4", "torus-1m-d9", "ellipsoids800k-d9-mixed")


def large_case_inputs(name):
    """(points, normals, queries, params) of one case; shared with the GPU tests."""
    if name == "ellipsoids800k-d9-mixed":
        p, n = synthetic.workload_ellipsoids(800000, dtype=np.float64)
        q = np.concatenate([p[::8], synthetic.cube_queries(p, 12000, np.random.default_rng(5))])
        return p, n, q, dict(depth=9, full_depth=4, point_weight=2.0, mixed=True)
    if name.startswith("torus250k-d8"):
        p, n, q = synthetic.workload_torus(250000)
        return p, n, q, dict(depth=8, full_depth=4, point_weight=float(name.rsplit("pw", 1)[1]))
    if name == "torus-1m-d9":
        p, n = torus_input("torus-1m-d9")
        q = np.concatenate([p[::40], synthetic.cube_queries(p, 5000, np.random.default_rng(6))])
        return p, n, q, torus_params("torus-1m-d9")
    if name == "sphere20k-d6":
        p, n, q = synthetic.workload_sphere(20000)
        return p, n, q, dict(depth=6, full_depth=3, point_weight=2.0)
    raise ValueError(name)


