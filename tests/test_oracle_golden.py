"""Pins the CPU oracle (CPU only, no GPU).

* Golden reports of the reference (docs/results/{validation,profiles,
  geometry-profiles,cancelled-normals}.json, extracted into
  tests/golden/reference_reports.json): the regenerated inputs hash to the
  recorded input SHA-256, every level's support and active key arrays hash to
  the recorded SHA-256s (tools/check_tree.py:9-21), cube_to_model is equal and
  the isovalue is within 1e-6 of warp-psr's (3 ulp for screened runs).
* Known answers from the reference's tests (tests/test_samples.py:15-28,
  tests/test_cancelled_normals.py); tests/test_field.py and the other
  known-answer tests are restated in test_oracle_known_answers.py.
* Closed forms: hat and normal integral tables vs brute-force quadrature,
  stiffness row sums, the deterministic log vs math.log, and the wavefront
  relaxation order vs the sequential sliced order (tests/test_relaxation.py:94-107).
"""
import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from warp_psr_amd import synthetic

REPORTS = json.loads((Path(__file__).parent / "golden" / "reference_reports.json").read_text())
FIXTURES = [f for f in REPORTS["fixtures"] if f["params"]["shape"] != "external" and not f["params"].get("variant")]
# base_depth 0 float32 uses the reference's ordered serial normal loads
# (ordered_normal_loads.py:263-337); the isovalue is checked against both
# warp-psr's and PoissonRecon's with the float32 field tolerance.
ORDERED_ROOT = {f["name"] for f in FIXTURES if f["params"]["base_depth"] == 0 and f["params"]["precision"] == "float"}


def run_fixture(f):
    p = f["params"]
    pts, nrm = synthetic.fixture_input(p["shape"], p["count"], p["seed"], p["precision"], p["input_scale"],
                                       p["translation"])
    res = oracle.reconstruct(pts, nrm, depth=p["depth"], full_depth=p["full_depth"], base_depth=p["base_depth"],
                             solve_depth=p["solve_depth"], kernel_depth=p["kernel_depth"],
                             samples_per_node=p["samples_per_node"], scale=p["scale"], point_weight=p["weight"],
                             exact_interpolation=p["exact"], iters=p["iters"], base_v_cycles=p["base_v_cycles"],
                             cg_accuracy=p["cg_accuracy"])
    return pts, nrm, res


def sha(a):
    return hashlib.sha256(np.asarray(a, dtype="<i8").tobytes()).hexdigest()


@pytest.mark.parametrize("f", FIXTURES, ids=[f["name"] for f in FIXTURES])
def test_fixture_tree_iso_transform(f):
    pts, nrm, res = run_fixture(f)
    assert hashlib.sha256(synthetic.fixture_input_text(pts, nrm)).hexdigest() == f["input_sha256"]
    for lv in f["tree"]:
        d = lv["depth"]
        assert len(res.array("tree_keys", d)) == lv["support_count"]
        assert sha(res.array("tree_keys", d)) == lv["support_sha256"], f"support keys differ at depth {d}"
        assert sha(res.array("basis_keys", d)) == lv["active_sha256"], f"active keys differ at depth {d}"
    ctm = res.array("cube_to_model").reshape(4, 4)
    np.testing.assert_array_equal(ctm, np.asarray(f["cube_to_model"], dtype=ctm.dtype))
    dt = np.float32 if f["params"]["precision"] == "float" else np.float64
    iso, warp_iso = dt(res.iso_value), dt(f["iso_value"])
    ulps = abs(float(iso) - float(warp_iso)) / float(np.spacing(abs(warp_iso)))
    screened = f["params"]["weight"] > 0 and not f["params"]["exact"]
    if dt == np.float32:
        # Every float32 isovalue is within the 1e-6 field bound of warp-psr's.
        # Screened runs agree to 3 ulp; unscreened and exact-interpolation
        # runs leave the constant mode to the base-level smoother and move by
        # up to 8.4e-7 (sphere-psr, sphere-exact).  The differing bits come
        # from Warp's FEM quadrature order at the base level and below depth
        # 4 (the oracle sums the same hat products row by row).
        assert abs(float(iso) - float(warp_iso)) <= 1e-6
        if screened:  # (an isovalue near zero, profile-root-solve: 2.7e-9)
            assert ulps <= 3 or abs(float(iso) - float(warp_iso)) <= 1e-8
        if f["name"] in ORDERED_ROOT:
            assert abs(float(iso) - f["reference_iso_value"]) <= 1e-6
    else:
        assert abs(float(iso) - float(warp_iso)) <= 3e-15
        assert abs(res.iso_value - f["reference_iso_value"]) <= 1e-8 + 1e-6 * abs(f["reference_iso_value"])


def test_fixture_isovalues_exact_count():
    """How many float32 isovalues equal warp-psr's published value bit for bit
    (7 of 28; every other one within 3 ulp when screened, 8.4e-7 otherwise)."""
    exact = 0
    floats = [f for f in FIXTURES if f["params"]["precision"] == "float"]
    for f in floats:
        _, _, res = run_fixture(f)
        exact += int(np.float32(res.iso_value) == np.float32(f["iso_value"]))
    assert len(floats) == 28 and exact >= 7


CANCELLED = REPORTS["cancelled_normals"]


def cancelled_input(case):
    points = [[-1, -1, -1], [-1, -1, -1], [1, 1, 1], [1, 1, 1]]
    normals = [[1, 0, 0], [-1, 0, 0], [1, 0, 0], [-1, 0, 0]]
    if "partial" in case:
        points.append([0, 0.1, 0])
        normals.append([0, 0, 1])
    dtype = np.float64 if case.startswith("double") else np.float32
    return np.array(points, dtype=dtype), np.array(normals, dtype=dtype)


@pytest.mark.parametrize("c", CANCELLED, ids=[c["case"] for c in CANCELLED])
def test_cancelled_normals(c):
    """tests/test_cancelled_normals.py:12-48 inputs; trees and averages of cancelled-normals.json."""
    pts, nrm = cancelled_input(c["case"])
    params = dict(c["parameters"])
    res = oracle.reconstruct(pts, nrm, **params)
    for lv in c["tree"]["levels"]:
        assert sha(res.array("tree_keys", lv["depth"])) == lv["support_sha256"]
        assert sha(res.array("basis_keys", lv["depth"])) == lv["active_sha256"]
    np.testing.assert_allclose(res.array("normal_average"), np.asarray(c["normal_average"][0]), rtol=1e-6, atol=1e-12)
    if "partial" not in c["case"]:
        assert res.iso_value == 0
        values, grads, valid = res.evaluate(pts)
        assert valid.all() and (values == 0).all() and (grads == 0).all()
    else:
        assert abs(res.iso_value - c["iso_value"]) <= 1e-5


def test_samples_known_answer():
    """tests/test_samples.py:15-28: multiplicity, rejection and split ownership."""
    points = np.array([[-1, -1, -1], [1, 1, 1], [0, 0, 0], [0, 0, 0], [.1, .1, .1]], dtype=np.float32)
    normals = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 2], [0, 0, -3], [0, 0, 0]], dtype=np.float32)
    res = oracle.reconstruct(points, normals, depth=2, full_depth=1)
    np.testing.assert_array_equal(res.array("weights"), [1, 2, 1])
    np.testing.assert_array_equal(res.array("sample_keys"),
                                  [(2 << 48), (2 << 48) | (1 << 32) | (1 << 16) | 1, (2 << 48) | (3 << 32) | (3 << 16) | 3])
    np.testing.assert_array_equal(res.array("normal_sums").reshape(-1, 3)[1], [0, 0, 0])
    np.testing.assert_array_equal(res.array("input_status"), [0, 0, 0, 0, 1])
    np.testing.assert_array_equal(res.array("input_group_indices"), [0, 2, 1, 1, -1])
    np.testing.assert_array_equal(res.array("input_order"), [0, 2, 3, 1, 4])
    np.testing.assert_array_equal(res.array("group_offsets"), [0, 1, 3, 4])


def test_degenerate_rejected():
    with pytest.raises(ValueError, match="degenerate"):
        oracle.reconstruct(np.zeros((2, 3), np.float32), np.array([[1, 0, 0], [1, 0, 0]], np.float32))
    with pytest.raises(ValueError, match="No accepted samples"):
        oracle.reconstruct(np.array([[0, 0, 0], [1, 1, 1]], np.float32), np.zeros((2, 3), np.float32))


def brute_hat_integrals(depth, index, trial_depth, trial_index, n=20000):
    """Midpoint quadrature of hat*hat and hat'*hat' on [0,1] (Neumann: clipped to the cube)."""
    x = (np.arange(n) + 0.5) / n
    t = x * (1 << depth) - index
    c = x * (1 << trial_depth) - trial_index
    ht = np.maximum(1 - np.abs(t), 0)
    hc = np.maximum(1 - np.abs(c), 0)
    dt = np.where(np.abs(t) < 1, -np.sign(t), 0) * (1 << depth)
    dc = np.where(np.abs(c) < 1, -np.sign(c), 0) * (1 << trial_depth)
    return np.array([np.mean(ht * hc), np.mean(dt * dc)])


@pytest.mark.parametrize("depth,index,trial_depth,trial_index",
                         [(3, 0, 3, 0), (3, 4, 3, 5), (3, 8, 3, 8), (4, 7, 3, 3), (4, 8, 3, 4), (4, 16, 3, 8), (2, 1, 1, 1)])
def test_hat_integrals_closed_form(depth, index, trial_depth, trial_index):
    got = oracle.hat_integrals(depth, index, trial_depth, trial_index)
    ref = brute_hat_integrals(depth, index, trial_depth, trial_index)
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=1e-6)


def bspline2(u):
    """Uniform quadratic B-spline supported on [0, 3)."""
    return np.where((u >= 0) & (u < 1), 0.5 * u * u,
                    np.where((u >= 1) & (u < 2), 0.5 * (-2 * u * u + 6 * u - 3),
                             np.where((u >= 2) & (u < 3), 0.5 * (3 - u) ** 2, 0.0)))


def brute_normal_integral(test_depth, test_index, normal_depth, normal_index, n=40000):
    """Hat (and derivative) against the odd-reflected (Dirichlet) quadratic B-spline."""
    x = (np.arange(n) + 0.5) / n
    nn = 1 << normal_depth
    u = x * nn - normal_index + 1
    b = bspline2(u) - bspline2(x * nn - (-1 - normal_index) + 1) - bspline2(x * nn - (2 * nn - 1 - normal_index) + 1)
    t = x * (1 << test_depth) - test_index
    h = np.maximum(1 - np.abs(t), 0)
    dh = np.where(np.abs(t) < 1, -np.sign(t), 0) * (1 << test_depth)
    return np.array([np.mean(h * b), np.mean(dh * b)])


@pytest.mark.parametrize("td,ti,nd,ni", [(3, 3, 3, 2), (3, 0, 3, 0), (3, 8, 3, 7), (2, 2, 3, 3), (3, 5, 2, 2), (4, 9, 4, 9),
                                         (3, 1, 4, 1)])
def test_normal_integral_closed_form(td, ti, nd, ni):
    got = oracle.normal_integral(td, ti, nd, ni)
    ref = brute_normal_integral(td, ti, nd, ni)
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=2e-6)


def test_stiffness_rows_sum_to_zero():
    """A constant is in the Neumann kernel: every same-level stiffness row sums to 0 (tests/test_uniform.py:48)."""
    pts, nrm = synthetic.fixture_input("sphere", 512)
    res = oracle.reconstruct(pts.astype(np.float64), nrm.astype(np.float64), depth=4, full_depth=2, point_weight=0.0)
    for d in (3, 4):
        off, val = res.array("matrix_offsets", d), res.array("matrix_values", d)
        keys = res.array("basis_keys", d)
        full = [i for i in range(len(keys)) if off[i + 1] - off[i] == 27]
        sums = np.add.reduceat(val, off[:-1])[full]
        assert np.abs(sums).max() < 1e-12


def test_det_log_matches_libm():
    """The shared deterministic log equals the correctly rounded log (checked against math.log within 1 ulp)."""
    import ctypes
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-3, 10, 2000), 10.0 ** rng.uniform(-300, 300, 500), [0.5, 1.0, 2.0, 1e-320]])
    # derived through the density depths of a tiny run would be indirect; compare values directly
    lib = oracle._lib()
    assert lib is not None
    out = np.array([oracle_det_log(x) for x in xs])
    ref = np.array([math.log(x) for x in xs])
    ulp = np.abs(out - ref) / np.spacing(np.abs(ref) + 1e-300)
    assert ulp.max() <= 1.0
    assert np.mean(out == ref) > 0.99


def oracle_det_log(x):
    import ctypes
    f = oracle._lib().oracle_det_log
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    return f(x)


def test_wavefront_order_equals_sequential():
    """relaxation.py:128-143: every coupled (slice, visit) pair keeps its sequential order in the wave batching."""
    for depth in (1, 2, 3, 5):
        for iters in (1, 2, 3, 8):
            n = 1 << depth
            seq = []
            for front in range(n, -iters, -1):
                for z in range(max(front, 0), min(front + iters, n + 1)):
                    seq.append((z, front))
            visit = {}
            order_seq = {}
            for pos, (z, front) in enumerate(seq):
                t = visit.get(z, 0)
                visit[z] = t + 1
                order_seq[(z, t)] = pos
            wave = {}
            for q in range(n + 2 * (iters - 1) + 1):
                first = max(0, (q - n + 1) // 2)
                last = min(iters - 1, q // 2)
                for t in range(first, last + 1):
                    wave[(n - q + 2 * t, t)] = q
            assert set(wave) == set(order_seq)
            for (z, t), pos in order_seq.items():
                for (z2, t2), pos2 in order_seq.items():
                    if abs(z - z2) <= 1 and (z, t) != (z2, t2) and pos < pos2:
                        assert wave[(z, t)] < wave[(z2, t2)], (depth, iters, z, t, z2, t2)


def _brute_relax(keys, offsets, columns, values, b, x, iters, forward, z):
    """solver.py:84-107 (_gs) and :227-238 (relax): full row scans per colour,
    rows in index order, in the values' precision."""
    dt = values.dtype.type
    x = x.copy()
    c = np.stack([(keys >> 32) & 0xFFFF, (keys >> 16) & 0xFFFF, keys & 0xFFFF], axis=1)
    colour = (c[:, 0] & 1) + 2 * (c[:, 1] & 1) + 4 * (c[:, 2] & 1)
    for _ in range(iters):
        for col in (range(8) if forward else range(7, -1, -1)):
            if z >= 0 and (col >> 2) != (z & 1):
                continue
            for i in range(len(keys)):
                if colour[i] != col or (z >= 0 and c[i, 2] != z):
                    continue
                r, diagonal = b[i], dt(0)
                for e in range(offsets[i], offsets[i + 1]):
                    r = dt(r - dt(x[columns[e]] * values[e]))
                    if columns[e] == i:
                        diagonal = values[e]
                if diagonal != dt(0):
                    x[i] = dt(x[i] + dt(r * dt(dt(1) / diagonal)))
    return x


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_relax_scheduled_equals_full_row_scans(dt):
    """tests/test_relaxation.py:9-34 restated: the scheduled 8-colour sweeps
    (forward and reverse) and the one-slice z mode give the bits of the
    full-row-scan relaxation."""
    p, n = synthetic.fixture_input("sphere", 512, synthetic.SEED, "double")
    ref = oracle.reconstruct(p, n, depth=4, full_depth=2)
    depth = 3
    keys = ref.array("basis_keys", depth)
    offsets = ref.array("matrix_offsets", depth).astype(np.int32)
    columns = ref.array("matrix_columns", depth).astype(np.int32)
    rng = np.random.default_rng(5)
    values = rng.uniform(-1.0, 1.0, size=len(columns)).astype(dt)
    rows = np.repeat(np.arange(len(keys)), np.diff(offsets))
    values[columns == rows] = dt(30.0)
    b = rng.normal(size=len(keys)).astype(dt)
    x0 = rng.normal(size=len(keys)).astype(dt)
    for forward in (True, False):
        got = oracle.relax_scheduled(keys, offsets, columns, values, b, x0, depth, 2, forward=forward)
        assert np.array_equal(got, _brute_relax(keys, offsets, columns, values, b, x0, 2, forward, -1))
        assert not np.array_equal(got, x0)
    x = x0.copy()
    want = x0.copy()
    for z in range((1 << depth), -1, -1):
        x = oracle.relax_scheduled(keys, offsets, columns, values, b, x, depth, 1, z=z)
        want = _brute_relax(keys, offsets, columns, values, b, want, 1, True, z)
    assert np.array_equal(x, want)


def test_ordered_normal_loads_match_fem_loads():
    """tests/test_ordered_normal_loads.py:46-106 restated: in float64 the
    ordered path (ordered_normal_loads.py:263-337, taken by float32 runs with
    base depth 0) gives the stencil path's loads to 1e-10 of their scale, on
    every level."""
    p, n = synthetic.fixture_input("sphere", 512, synthetic.SEED, "double")
    ref = oracle.reconstruct(p, n, depth=4, full_depth=0, base_depth=0)
    for d in range(5):
        expected = ref.array("normal_load", d)
        actual = ref.array("ordered_normal_load", d)
        scale = np.max(np.abs(expected)) or 1.0
        np.testing.assert_allclose(actual / scale, expected / scale, atol=1e-10, rtol=1e-9)
    assert np.any(ref.array("normal_load", 4) != 0)


def test_float_base_zero_takes_the_ordered_path():
    """normal_loads.py:84-92: float32 runs with base depth 0 use the ordered
    float32 loads (rounded in float32, widened to float64)."""
    p, n = synthetic.fixture_input("sphere", 512, synthetic.SEED, "float")
    ref = oracle.reconstruct(p, n, depth=4, full_depth=0, base_depth=0)
    for d in range(5):
        load = ref.array("normal_load", d)
        assert np.array_equal(load, ref.array("ordered_normal_load", d))
        assert np.array_equal(load, load.astype(np.float32).astype(np.float64))


def test_mixed_precision_error_bound():
    """Float32 finer-level operators keep the float64 octree exactly and the
    field within 1e-6 of the float64 run (inputs and interface queries)."""
    p, n = synthetic.workload_ellipsoids(30000, dtype=np.float64)
    kw = dict(depth=7, full_depth=4, point_weight=2.0)
    full = oracle.reconstruct(p, n, **kw)
    mixed = oracle.reconstruct(p, n, mixed=True, **kw)
    keys = full.array("keys")
    np.testing.assert_array_equal(mixed.array("keys"), keys)
    # rounding is really applied: finer-level operators are float32 values
    vals = mixed.array("matrix_values", 6)
    assert np.array_equal(vals, vals.astype(np.float32).astype(np.float64))
    assert not np.array_equal(vals, full.array("matrix_values", 6))
    q = np.concatenate([p, synthetic.interface_queries(keys, 7, full.array("model_to_cube"),
                                                       full.array("cube_to_model"), p, count=4000)])
    vf, _, okf = full.evaluate(q)
    vm, _, okm = mixed.evaluate(q)
    both = okf.astype(bool) & okm.astype(bool)
    err = np.abs(vm - vf)[both]
    assert both.sum() > len(p) // 2
    assert 0 < err.max() <= 1e-6
    assert abs(mixed.iso_value - full.iso_value) <= 1e-6


def test_mixed_requires_float64():
    p, n = synthetic.fixture_input("sphere", 64, synthetic.SEED, "float")
    with pytest.raises(ValueError):
        oracle.reconstruct(p, n, mixed=True, depth=4, full_depth=2)
