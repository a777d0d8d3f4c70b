"""Known answers and error paths of the HIP pipeline (reference tests restated).

* tests/test_samples.py:15-28 -- multiplicity, rejection, split ownership;
* tests/test_cancelled_normals.py:12-48 -- the cancelled-normals cases of
  docs/results/cancelled-normals.json (trees by SHA-256, zero field);
* degenerate inputs raise ValueError with the reference's messages;
* psr_relax_sliced (relaxation.py:189-238) on caller CSR operators with
  random values equals the oracle's serial sliced Gauss-Seidel bit for bit
  (tests/test_relaxation.py:110-147 restated: any values on the 27-point
  sparsity of a level give the same bits in every legal order), and
  psr_relax_scheduled (relaxation.py:261-271) equals the oracle's 8-colour
  sweeps and one-slice relaxations (tests/test_relaxation.py:9-34).
"""
import numpy as np
import pytest

import oracle
from parity import compare_stages, fixture_arrays, same
from test_oracle_golden import CANCELLED, cancelled_input, sha

pytestmark = pytest.mark.gpu


def _gpu(cuda, points, normals, **params):
    import torch
    from warp_psr_amd import reconstruct
    return reconstruct(torch.from_numpy(points).to(cuda), torch.from_numpy(normals).to(cuda), **params)


def test_samples_known_answer(cuda):
    points = np.array([[-1, -1, -1], [1, 1, 1], [0, 0, 0], [0, 0, 0], [.1, .1, .1]], dtype=np.float32)
    normals = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 2], [0, 0, -3], [0, 0, 0]], dtype=np.float32)
    d = _gpu(cuda, points, normals, depth=2, full_depth=1).diagnostics
    np.testing.assert_array_equal(d.array("weights"), [1, 2, 1])
    np.testing.assert_array_equal(d.array("sample_keys"),
                                  [(2 << 48), (2 << 48) | (1 << 32) | (1 << 16) | 1, (2 << 48) | (3 << 32) | (3 << 16) | 3])
    np.testing.assert_array_equal(d.array("input_status"), [0, 0, 0, 0, 1])
    np.testing.assert_array_equal(d.array("input_group_indices"), [0, 2, 1, 1, -1])
    np.testing.assert_array_equal(d.array("input_order"), [0, 2, 3, 1, 4])
    np.testing.assert_array_equal(d.array("group_offsets"), [0, 1, 3, 4])


@pytest.mark.parametrize("c", CANCELLED, ids=[c["case"] for c in CANCELLED])
def test_cancelled_normals(cuda, c):
    import torch
    pts, nrm = cancelled_input(c["case"])
    params = dict(c["parameters"])
    res = _gpu(cuda, pts, nrm, **params)
    for lv in c["tree"]["levels"]:
        assert sha(res.diagnostics.array("tree_keys", lv["depth"])) == lv["support_sha256"]
        assert sha(res.diagnostics.array("basis_keys", lv["depth"])) == lv["active_sha256"]
    ref = oracle.reconstruct(pts, nrm, **params)
    assert res.iso_value == ref.iso_value
    if "partial" not in c["case"]:
        assert res.iso_value == 0
        values, grads, valid = res.evaluate(torch.from_numpy(pts).to(cuda))
        assert bool(valid.bool().all()) and bool((values == 0).all()) and bool((grads == 0).all())


def test_degenerate_inputs_rejected(cuda):
    with pytest.raises(ValueError, match="degenerate"):
        _gpu(cuda, np.zeros((2, 3), np.float32), np.array([[1, 0, 0], [1, 0, 0]], np.float32))
    with pytest.raises(ValueError, match="No accepted samples"):
        _gpu(cuda, np.array([[0, 0, 0], [1, 1, 1]], np.float32), np.zeros((2, 3), np.float32))


@pytest.mark.parametrize("precision", ["float", "double"])
def test_relax_sliced_matches_oracle(cuda, precision):
    import torch
    from warp_psr_amd.reconstruction import relax_sliced
    p, n = fixture_arrays("torus", 8192, precision)
    ref = oracle.reconstruct(p, n, depth=6, full_depth=3)
    depth = 6
    keys = ref.array("basis_keys", depth)
    offsets = ref.array("matrix_offsets", depth).astype(np.int32)
    columns = ref.array("matrix_columns", depth).astype(np.int32)
    dt = np.float32 if precision == "float" else np.float64
    rng = np.random.default_rng(11)
    values = rng.uniform(-1.0, 1.0, size=len(columns)).astype(dt)
    rows = np.repeat(np.arange(len(keys)), np.diff(offsets))
    values[columns == rows] = dt(30.0)  # diagonally dominant
    b = rng.normal(size=len(keys)).astype(dt)
    x0 = rng.normal(size=len(keys)).astype(dt)
    expected = oracle.relax_sliced(keys, offsets, columns, values, b, x0, depth, 3)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    x = t(x0)
    relax_sliced(t(keys), t(offsets), t(columns), t(values), t(b), x, depth, 3)
    assert same(x.cpu().numpy(), expected)
    assert not same(x0, expected)


@pytest.mark.parametrize("precision", ["float", "double"])
@pytest.mark.parametrize("shape,count,depth", [("torus", 8192, 6), ("ellipsoids", 60000, 7)])
def test_relax_scheduled_matches_oracle(cuda, precision, shape, count, depth):
    """One workgroup (torus level 6) and one launch per colour group (ellipsoid
    union level 7, more than 65,536 rows) give the oracle's bits."""
    import torch
    from warp_psr_amd import synthetic
    from warp_psr_amd.reconstruction import relax_scheduled
    dt = np.float32 if precision == "float" else np.float64
    if shape == "torus":
        p, n = fixture_arrays(shape, count, precision)
    else:
        p, n = synthetic.workload_ellipsoids(count, dtype=dt)
    ref = oracle.reconstruct(p, n, depth=depth, full_depth=3)
    keys = ref.array("basis_keys", depth)
    offsets = ref.array("matrix_offsets", depth).astype(np.int32)
    columns = ref.array("matrix_columns", depth).astype(np.int32)
    print(f"relax_scheduled {shape} depth {depth} rows {len(keys)} {precision}", flush=True)
    rng = np.random.default_rng(13)
    values = rng.uniform(-1.0, 1.0, size=len(columns)).astype(dt)
    rows = np.repeat(np.arange(len(keys)), np.diff(offsets))
    values[columns == rows] = dt(30.0)
    b = rng.normal(size=len(keys)).astype(dt)
    x0 = rng.normal(size=len(keys)).astype(dt)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    dk, do, dc, dv, db = t(keys), t(offsets), t(columns), t(values), t(b)
    for forward in (True, False):
        expected = oracle.relax_scheduled(keys, offsets, columns, values, b, x0, depth, 3, forward=forward)
        x = t(x0)
        relax_scheduled(dk, do, dc, dv, db, x, depth, 3, forward=forward)
        assert same(x.cpu().numpy(), expected)
    expected = x0.copy()
    x = t(x0)
    for z in ((1 << depth), (1 << depth) // 2 + 1, (1 << depth) // 2, 0):
        expected = oracle.relax_scheduled(keys, offsets, columns, values, b, expected, depth, 2, z=z)
        relax_scheduled(dk, do, dc, dv, db, x, depth, 2, z=z)
    assert same(x.cpu().numpy(), expected)
    assert not same(x0, expected)
