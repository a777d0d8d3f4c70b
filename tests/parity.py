"""Stage-by-stage comparison of the HIP pipeline against the CPU oracle.

Both sides export the same named arrays (host copies); integer and
floating-point arrays are compared bitwise (NaNs compare equal by position).
"""
import numpy as np

import oracle
from warp_psr_amd import synthetic

# (name, per_level, level_range) -- level_range(params) gives the levels to check
STAGES = [
    ("transform", False), ("bounds", False), ("positions", False), ("unit_normals", False), ("input_status", False),
    ("input_order", False), ("input_group_indices", False), ("sample_keys", False), ("position_sums", False),
    ("normal_sums", False), ("weights", False), ("first_indices", False), ("group_offsets", False),
    ("aggregate_keys", True), ("aggregate_values", True), ("density_keys", "kd"), ("density_values", "kd"),
    ("densities", False), ("depths_and_weights", False), ("density_scale", False),
    ("normal_keys", False), ("normal_values", False), ("normal_average", False), ("normal_offsets", False),
    ("tree_keys", True), ("basis_keys", True),
    ("normal_load", "solve"), ("point_load", "solve"), ("rhs", "solve"),
    ("matrix_offsets", "solve"), ("matrix_columns", "solve"), ("matrix_values", "solve"),
    ("cross_offsets", "fine"), ("cross_columns", "fine"), ("cross_values", "fine"),
    ("prolong_offsets", "prolong"), ("prolong_columns", "prolong"), ("prolong_values", "prolong"),
    ("corrections", "solve"), ("keys", False), ("coefficients", False), ("cumulative", False),
    ("centroid_values", False), ("iso", False), ("cube_to_model", False), ("model_to_cube", False), ("cg_runs", False),
]


def levels_for(kind, p, basis_counts):
    depth = p["depth"]
    if kind is True:
        return range(depth + 1)
    if kind == "kd":
        return range(p["kernel_depth"] + 1)
    solved = [d for d in range(p["base_depth"], p["solve_depth"] + 1)]
    stop = next((d for d in solved if not basis_counts[d]), None)
    if stop is not None:
        solved = [d for d in solved if d < stop]
    if kind == "solve":
        return solved
    if kind == "fine":
        return [d for d in solved if d > p["base_depth"]]
    if kind == "prolong":
        return range(1, depth + 1)
    raise ValueError(kind)


def same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                              b.view(np.uint32 if b.dtype == np.float32 else np.uint64)) or \
            np.array_equal(a, b, equal_nan=True)
    return np.array_equal(a, b)


def describe(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    if a.dtype != b.dtype:
        return f"dtype {a.dtype} vs {b.dtype}"
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b))
    i = bad[0]
    extra = ""
    if a.dtype.kind == "f":
        extra = f" max|diff|={np.nanmax(np.abs(a.astype(np.float64) - b.astype(np.float64))):.3e}"
    return f"{len(bad)} of {a.size} differ, first at {i}: gpu={a.flat[i]!r} oracle={b.flat[i]!r}{extra}"


def compare_stages(gpu_diag, ref, params, stages=STAGES):
    """Return a list of (stage, level, message) mismatches (empty when bit-exact)."""
    counts = [len(ref.array("basis_keys", d)) for d in range(params["depth"] + 1)]
    out = []
    for name, kind in stages:
        levels = [-1] if kind is False else levels_for(kind, params, counts)
        for lv in levels:
            g = gpu_diag.array(name, lv)
            o = ref.array(name, lv)
            if name == "aggregate_values" and g.size != o.size:
                pass
            if not same(g, o):
                out.append((name, lv, describe(g, o)))
    return out


FIXTURES = [
    # name, shape, count, precision, params (reference fixture generator settings)
    ("sphere", "sphere", 512, "float", dict(depth=4, full_depth=2, point_weight=2.0)),
    ("sphere-double", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=2.0)),
    ("sphere-psr", "sphere", 512, "float", dict(depth=4, full_depth=2, point_weight=0.0)),
    ("sparse", "sphere", 32, "float", dict(depth=5, full_depth=1, point_weight=2.0)),
    ("torus", "torus", 8192, "float", dict(depth=6, full_depth=3, point_weight=2.0)),
    ("torus-double", "torus", 8192, "double", dict(depth=6, full_depth=3, point_weight=2.0)),
    ("cluster", "cluster", 1024, "float", dict(depth=5, full_depth=2, point_weight=2.0)),
    ("plane", "plane", 1024, "float", dict(depth=5, full_depth=2, point_weight=2.0)),
    ("sheets-double", "sheets", 1024, "double", dict(depth=5, full_depth=2, point_weight=2.0)),
    ("disconnected", "disconnected", 1024, "float", dict(depth=5, full_depth=2, point_weight=2.0)),
    ("base-zero-psr-double", "sphere", 512, "double", dict(depth=4, full_depth=2, base_depth=0, point_weight=0.0)),
    ("root-solve-float", "sphere", 512, "float", dict(depth=4, full_depth=0, base_depth=0, point_weight=2.0)),
    ("torus-base-zero-float", "torus", 8192, "float", dict(depth=6, full_depth=3, base_depth=0, point_weight=2.0)),
    ("exact", "sphere", 512, "float", dict(depth=4, full_depth=2, point_weight=2.0, exact_interpolation=True)),
    ("cycles", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=2.0, base_v_cycles=2, iters=3)),
    ("torus-mixed", "torus", 8192, "double", dict(depth=6, full_depth=3, point_weight=2.0, mixed=True)),
    ("sphere-psr-mixed", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=0.0, mixed=True)),
]


def fixture_arrays(shape, count, precision, input_scale=1.0, translation=(0., 0., 0.)):
    return synthetic.fixture_input(shape, count, synthetic.SEED, precision, input_scale, translation)


def fixture_queries(points, count=4096, seed=7):
    """Inputs, uniform cube points and exact cube-boundary/grid points."""
    rng = np.random.default_rng(seed)
    q = [points, synthetic.cube_queries(points, count, rng)]
    lo, hi = points.min(axis=0).astype(np.float64), points.max(axis=0).astype(np.float64)
    width = float(np.max(hi - lo)) * 1.1
    center = (hi + lo) / 2
    g = np.linspace(-0.5, 0.5, 9)
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3) * width + center
    q.append(grid.astype(points.dtype))
    q.append(np.array([[np.nan, 0, 0], [1e9, 0, 0]], dtype=points.dtype))
    return np.concatenate(q).astype(points.dtype)
