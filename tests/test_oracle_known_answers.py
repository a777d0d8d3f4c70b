"""Known answers of the reference's own tests, restated against the oracle (CPU).

* tests/test_field.py:8-69 -- the affine field x + 2y + 3z under a permuted
  similarity (direct level summation, both precisions), the left derivative
  at a knot [e, e, -e], and the reference evaluation's neighbour/polynomial
  intersection (values .5 -/+ 1e-4, x derivatives [1, -1, 1]) through
  cumulative coefficients;
* tests/test_isovalue_reduction.py:12-34 -- the isovalue reduction against
  math.fsum for sizes 1..1,048,577, with the reference's tolerances, and
  repeatability;
* tests/test_root_solver.py:15-54 -- the constant-mode operator and the
  padded 64-row root system on the rank-seven Laplacian I - 1/8, and the
  base-depth-0 solve of a zero-mean right-hand side;
* tests/test_uniform.py:51-80 -- screening on a uniform level equals the
  brute-force phi^T W phi with the point load sum_p w phi_i, and a uniform
  level's evaluation equals the tensor-product hat sum and its derivatives.

Each check is chosen so that a plausible slip (a transposed linear map, the
right instead of the left knot derivative, a dropped boundary hat, a float
accumulator in the isovalue, a missing mean term) fails it.
"""
import math

import numpy as np
import pytest

import oracle


def _key(d, x, y, z):
    return (d << 48) | (x << 32) | (y << 16) | z


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_affine_gradient_and_domain(dtype):
    keys = np.array([_key(0, x, y, z) for x in range(2) for y in range(2) for z in range(2)], np.int64)
    coefficients = np.array([x + 2 * y + 3 * z for x in range(2) for y in range(2) for z in range(2)], dtype)
    linear = np.array([[0, .25, 0], [0, 0, .25], [.25, 0, 0]])
    translation = np.array([-.2, .3, .1])
    cube = np.array([[.13, .27, .83], [0, 0, 0], [1, 1, 1], [1.01, .5, .5], [np.nan, 0, 0]])
    model = (cube - translation) @ np.linalg.inv(linear).T
    value, gradient, valid = oracle.field_evaluate(keys, coefficients, 0, linear, translation, model.astype(dtype))
    np.testing.assert_array_equal(valid, [1, 1, 1, 0, 0])
    atol = 1e-6 if dtype == np.float32 else 1e-14
    np.testing.assert_allclose(value[:3], cube[:3] @ np.array([1, 2, 3]), atol=atol)
    np.testing.assert_allclose(gradient[:3], np.tile(np.array([1, 2, 3]) @ linear, (3, 1)), atol=atol)
    assert np.isnan(value[3:]).all() and np.isnan(gradient[3:]).all()
    # a transposed linear map moves the gradient: the check above would fail
    assert not np.allclose(np.array([1, 2, 3]) @ linear.T, np.array([1, 2, 3]) @ linear)


def test_knot_uses_reference_left_derivative():
    keys = np.array([_key(1, 1, 1, 1)], np.int64)
    q = np.array([[.5 - 1e-4, .37, .29], [.5, .37, .29], [.5 + 1e-4, .37, .29]], np.float32)
    value, gradient, valid = oracle.field_evaluate(keys, np.array([1.], np.float32), 1, np.eye(3), np.zeros(3), q)
    expected = 2 * (2 * .37) * (2 * .29)
    np.testing.assert_allclose(gradient[:, 0], [expected, expected, -expected], atol=1e-6)
    assert valid.all()
    # the value is continuous across the knot: 1 - |2x - 1| times the y, z hats
    np.testing.assert_allclose(value, [(1 - abs(2 * x - 1)) * (2 * .37) * (2 * .29) for x in q[:, 0]], atol=1e-6)


def test_reference_knot_neighbor_polynomial_intersection():
    keys, coefficients = [], []
    for d in range(2):
        for x in range((1 << d) + 1):
            for y in range((1 << d) + 1):
                for z in range((1 << d) + 1):
                    keys.append(_key(d, x, y, z))
                    coefficients.append(x * .5 if d else 0.)
    keys = np.array(keys, np.int64)
    coefficients = np.array(coefficients, np.float32)
    cumulative = oracle.cumulative(keys, coefficients, 1)
    # level 0 has zero coefficients: the level-1 cumulative values are its own
    np.testing.assert_array_equal(cumulative, coefficients)
    q = np.array([[.5 - 1e-4, .37, .29], [.5, .37, .29], [.5 + 1e-4, .37, .29]], np.float32)
    value, gradient, valid = oracle.field_evaluate(keys, coefficients, 1, np.eye(3), np.zeros(3), q, cumulative=cumulative)
    np.testing.assert_allclose(value, [.5 - 1e-4, .5, .5 + 1e-4], atol=1e-6)
    np.testing.assert_allclose(gradient[:, 0], [1., -1., 1.], atol=1e-6)
    assert valid.all()


def test_cumulative_prolongs_coarse_coefficients():
    """cum_d = coeff_d + sum of trilinear weights of the coarser cumulative:
    a level-0 field x (coefficients at x = 1) prolongs to x/2 at level 1."""
    keys, coefficients = [], []
    for d in range(2):
        for x in range((1 << d) + 1):
            for y in range((1 << d) + 1):
                for z in range((1 << d) + 1):
                    keys.append(_key(d, x, y, z))
                    coefficients.append(float(x) if d == 0 else 0.)
    keys = np.array(keys, np.int64)
    for dtype in (np.float32, np.float64):
        cum = oracle.cumulative(keys, np.array(coefficients, dtype), 1)
        level1 = np.array([(k >> 32) & 65535 for k in keys[8:]], dtype) * dtype(.5)
        np.testing.assert_array_equal(cum[8:], level1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("count", [1, 1023, 1024, 1025, 65537, 1048577])
def test_weighted_mean_matches_accurate_sum(dtype, count):
    rng = np.random.default_rng(235)
    values = rng.normal(size=count).astype(dtype)
    weights = np.exp2(rng.uniform(-20, 20, size=count)).astype(dtype)
    weights[::13] = 0
    weights[0] = 1
    products = values * weights
    expected = dtype(math.fsum(products) / math.fsum(weights))
    first = oracle.isovalue(values, weights)
    second = oracle.isovalue(values, weights)
    np.testing.assert_array_equal(first, second)
    if dtype == np.float64:
        np.testing.assert_allclose(first, expected, atol=3e-15, rtol=3e-14)
    else:
        np.testing.assert_allclose(first, expected, atol=3e-8, rtol=1e-7)
        # a float32 accumulator would not meet the bound at this spread of weights
        if count >= 65537:
            naive = np.float32(0)
            for p in products[:65537].astype(np.float32):
                naive = np.float32(naive + p)
            acc = math.fsum(products[:65537])
            assert abs(float(naive) - acc) > 1e-7 * abs(acc)


def _root_operator(dtype):
    matrix = np.eye(8, dtype=dtype) - dtype(1 / 8)
    offsets = np.arange(0, 72, 8, dtype=np.int32)
    columns = np.tile(np.arange(8, dtype=np.int32), 8)
    keys = np.array([((i // 4) << 32) | ((i // 2 % 2) << 16) | (i % 2) for i in range(8)], np.int64)
    return keys, offsets, columns, matrix.reshape(-1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_root_constant_mode_and_repeated_cycles(dtype):
    keys, offsets, columns, values = _root_operator(dtype)
    vector = np.arange(1, 9, dtype=dtype)
    # the padded system keeps the eight rows at root_row(key) (solver.py:30-33)
    v = np.zeros(64, dtype)
    np.testing.assert_array_equal(oracle.root_padded_apply(keys, offsets, columns, values, v), v)
    # ones at the mapped rows: (I - 1/8) 1 = 0 there and the mean term adds 8/64 to every row
    mapped = []
    for k in keys:
        x, y, z = (k >> 32) & 65535, (k >> 16) & 65535, k & 65535
        x, y, z = x + 2, y + 2, z + 2
        mapped.append(z * 16 + (x & 1) + 2 * (y & 1) + 2 * (x & 2) + 4 * (y & 2))
    assert len(set(mapped)) == 8
    v[mapped] = 1
    out = oracle.root_padded_apply(keys, offsets, columns, values, v)
    np.testing.assert_array_equal(out, np.full(64, 1 / 8, dtype))
    zero_mean = vector - vector.mean()
    for cycles in (1, 3):
        solution = oracle.root_solve(keys, offsets, columns, values, zero_mean, iters=8, cycles=cycles,
                                     accuracy=1e-6, dc=True)
        np.testing.assert_allclose(solution, zero_mean, atol=1e-6 if dtype == np.float32 else 1e-12, rtol=0)
    # without the mean term the same right-hand side is solved only up to the
    # constant mode, which CG from zero does not add: still the zero-mean solution
    solution = oracle.root_solve(keys, offsets, columns, values, zero_mean, iters=8, cycles=1, accuracy=1e-6, dc=False)
    np.testing.assert_allclose(solution, zero_mean, atol=1e-5 if dtype == np.float32 else 1e-12, rtol=0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_point_weights_and_basis(dtype):
    depth = 2
    positions = np.array([[.017, .231, .673], [.991, .823, .401], [.41, .57, .69]], dtype)
    weights = np.array([.31, 2.7, 4.1], dtype)
    n = 1 << depth
    nodes = np.array([[x, y, z] for x in range(n + 1) for y in range(n + 1) for z in range(n + 1)], np.float64) / n
    phi = np.prod(np.maximum(1 - np.abs(positions[:, None].astype(np.float64) - nodes[None]) * n, 0), axis=2)
    dense, load = oracle.screening_dense(positions, weights, depth)
    np.testing.assert_allclose(dense, phi.T @ (weights[:, None] * phi), atol=1e-6, rtol=1e-5)
    np.testing.assert_allclose(load, phi.T @ weights, atol=1e-6, rtol=1e-5)
    # a uniform level's field: value = phi c, gradient of the hat product
    keys = np.array([_key(depth, x, y, z) for x in range(n + 1) for y in range(n + 1) for z in range(n + 1)], np.int64)
    coefficients = np.sin(np.arange(len(nodes))).astype(dtype)
    value, gradient, valid = oracle.field_evaluate(keys, coefficients, depth, np.eye(3), np.zeros(3), positions)
    assert valid.all()
    np.testing.assert_allclose(value, phi @ coefficients, atol=1e-6, rtol=1e-5)
    dx = (positions[:, None].astype(np.float64) - nodes[None]) * n
    hat = np.maximum(1 - np.abs(dx), 0)
    expected = np.zeros((len(positions), 3))
    for a in range(3):
        derivative = np.where(np.abs(dx[:, :, a]) < 1, -n * np.sign(dx[:, :, a]), 0)
        for b in range(3):
            if b != a:
                derivative = derivative * hat[:, :, b]
        expected[:, a] = derivative @ coefficients
    np.testing.assert_allclose(gradient, expected, atol=1e-5, rtol=1e-5)
