"""The OpenMP build of the oracle (the CPU baseline's multi-core mode) gives
the serial build's arrays bit for bit: its parallel loops are over rows,
queries or aggregates whose outputs are disjoint, the sliced relaxation runs
the reference's wavefront batching, and sorts break ties by position."""
import os

import numpy as np
import pytest

import oracle
from parity import STAGES, fixture_arrays, levels_for, same
from warp_psr_amd import synthetic


@pytest.mark.parametrize("shape,count,precision,params", [
    ("torus", 8192, "float", dict(depth=6, full_depth=3)),
    ("sphere", 4096, "double", dict(depth=6, full_depth=2, point_weight=0.0)),
    ("ellipsoids", 60000, "double", dict(depth=7, full_depth=3, mixed=True)),
])
def test_openmp_build_equals_serial(shape, count, precision, params):
    if shape == "ellipsoids":
        p, n = synthetic.workload_ellipsoids(count, dtype=np.float64)
    else:
        p, n = fixture_arrays(shape, count, precision)
    serial = oracle.reconstruct(p, n, **params)
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "4"
    try:
        parallel = oracle.reconstruct(p, n, openmp=True, **params)
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    counts = [len(serial.array("basis_keys", d)) for d in range(serial.params["depth"] + 1)]
    for name, kind in STAGES:
        for lv in ([-1] if kind is False else levels_for(kind, serial.params, counts)):
            assert same(parallel.array(name, lv), serial.array(name, lv)), f"{name}[{lv}]"
    q = np.concatenate([p[:2000], synthetic.cube_queries(p, 2000, np.random.default_rng(3))])
    for a, b in zip(parallel.evaluate(q), serial.evaluate(q)):
        assert same(a, b)
