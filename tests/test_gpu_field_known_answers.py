"""Reference known answers on the GPU (tests/test_field.py:8-69,
tests/test_isovalue_reduction.py:12-34 restated).

Fields are built by the caller as LevelField(keys, coefficients, linear,
translation, depth, iso_value, cumulative=None) over ROCm tensors and
evaluated by the header-only psr_evaluate_field (psr_field_evaluate): the
direct level summation without cumulative coefficients, the reference
evaluation with them.  Each must meet the reference's known answer and equal
the oracle bit for bit.  The isovalue reduction psr_isovalue must meet the
math.fsum bound and equal the oracle bit for bit.
"""
import math

import numpy as np
import pytest

import oracle
from parity import same

pytestmark = pytest.mark.gpu


def _key(d, x, y, z):
    return (d << 48) | (x << 32) | (y << 16) | z


def _field(cuda, keys, coefficients, linear, translation, depth, cumulative=None):
    import torch
    from warp_psr_amd import LevelField
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    return LevelField(t(keys), t(coefficients), tuple(np.asarray(linear).reshape(-1)), tuple(translation), depth, 0.5,
                      None if cumulative is None else t(cumulative))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_affine_gradient_and_domain(cuda, dtype):
    import torch
    keys = np.array([_key(0, x, y, z) for x in range(2) for y in range(2) for z in range(2)], np.int64)
    coefficients = np.array([x + 2 * y + 3 * z for x in range(2) for y in range(2) for z in range(2)], dtype)
    linear = np.array([[0, .25, 0], [0, 0, .25], [.25, 0, 0]])
    translation = np.array([-.2, .3, .1])
    cube = np.array([[.13, .27, .83], [0, 0, 0], [1, 1, 1], [1.01, .5, .5], [np.nan, 0, 0]])
    model = ((cube - translation) @ np.linalg.inv(linear).T).astype(dtype)
    field = _field(cuda, keys, coefficients, linear, translation, 0)
    value, gradient, valid = (a.cpu().numpy() for a in field.evaluate(torch.from_numpy(model).to(cuda)))
    np.testing.assert_array_equal(valid, [1, 1, 1, 0, 0])
    atol = 1e-6 if dtype == np.float32 else 1e-14
    np.testing.assert_allclose(value[:3], cube[:3] @ np.array([1, 2, 3]), atol=atol)
    np.testing.assert_allclose(gradient[:3], np.tile(np.array([1, 2, 3]) @ linear, (3, 1)), atol=atol)
    assert np.isnan(value[3:]).all() and np.isnan(gradient[3:]).all()
    rv, rg, rvalid = oracle.field_evaluate(keys, coefficients, 0, linear, translation, model)
    assert same(value, rv) and same(gradient, rg)
    np.testing.assert_array_equal(valid, rvalid)


def test_knot_uses_reference_left_derivative(cuda):
    import torch
    keys = np.array([_key(1, 1, 1, 1)], np.int64)
    q = np.array([[.5 - 1e-4, .37, .29], [.5, .37, .29], [.5 + 1e-4, .37, .29]], np.float32)
    field = _field(cuda, keys, np.array([1.], np.float32), np.eye(3), (0., 0., 0.), 1)
    value, gradient, valid = (a.cpu().numpy() for a in field.evaluate(torch.from_numpy(q).to(cuda)))
    expected = 2 * (2 * .37) * (2 * .29)
    np.testing.assert_allclose(gradient[:, 0], [expected, expected, -expected], atol=1e-6)
    assert valid.all()
    rv, rg, _ = oracle.field_evaluate(keys, np.array([1.], np.float32), 1, np.eye(3), np.zeros(3), q)
    assert same(value, rv) and same(gradient, rg)


def test_reference_knot_neighbor_polynomial_intersection(cuda):
    import torch
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
    q = np.array([[.5 - 1e-4, .37, .29], [.5, .37, .29], [.5 + 1e-4, .37, .29]], np.float32)
    field = _field(cuda, keys, coefficients, np.eye(3), (0., 0., 0.), 1, cumulative)
    value, gradient, valid = (a.cpu().numpy() for a in field.evaluate(torch.from_numpy(q).to(cuda)))
    np.testing.assert_allclose(value, [.5 - 1e-4, .5, .5 + 1e-4], atol=1e-6)
    np.testing.assert_allclose(gradient[:, 0], [1., -1., 1.], atol=1e-6)
    rv, rg, _ = oracle.field_evaluate(keys, coefficients, 1, np.eye(3), np.zeros(3), q, cumulative=cumulative)
    assert same(value, rv) and same(gradient, rg)


def test_level_field_validation(cuda):
    import torch
    from warp_psr_amd import LevelField
    k = torch.zeros(2, dtype=torch.int64, device=cuda)
    c = torch.zeros(2, dtype=torch.float32, device=cuda)
    with pytest.raises(ValueError, match="float32/float64 coefficients and int64 keys"):
        LevelField(k.int(), c, (1, 0, 0, 0, 1, 0, 0, 0, 1), (0, 0, 0), 1, 0.)
    with pytest.raises(ValueError, match="depths 0..15"):
        LevelField(k, c, (1, 0, 0, 0, 1, 0, 0, 0, 1), (0, 0, 0), 16, 0.)
    with pytest.raises(ValueError, match="One key per coefficient"):
        LevelField(k[:1], c, (1, 0, 0, 0, 1, 0, 0, 0, 1), (0, 0, 0), 1, 0.)
    field = LevelField(k, c, (1, 0, 0, 0, 1, 0, 0, 0, 1), (0, 0, 0), 1, 0.)
    with pytest.raises(ValueError, match="Query precision/device"):
        field.evaluate(torch.zeros((1, 3), dtype=torch.float64, device=cuda))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("count", [1, 1023, 1024, 1025, 65537, 1048577])
def test_isovalue_matches_accurate_sum(cuda, dtype, count):
    import ctypes
    import torch
    from warp_psr_amd import _lib
    rng = np.random.default_rng(235)
    values = rng.normal(size=count).astype(dtype)
    weights = np.exp2(rng.uniform(-20, 20, size=count)).astype(dtype)
    weights[::13] = 0
    weights[0] = 1
    expected = dtype(math.fsum(values * weights) / math.fsum(weights))
    v, w = torch.from_numpy(values).to(cuda), torch.from_numpy(weights).to(cuda)
    out = torch.empty(2, dtype=v.dtype, device=cuda)
    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    code = 1 if dtype == np.float64 else 0
    for i in range(2):
        _lib.check(lib.psr_isovalue(code, v.data_ptr(), w.data_ptr(), count, out[i:].data_ptr(), stream))
    first, second = out.cpu().numpy()
    np.testing.assert_array_equal(first, second)
    if dtype == np.float64:
        np.testing.assert_allclose(first, expected, atol=3e-15, rtol=3e-14)
    else:
        np.testing.assert_allclose(first, expected, atol=3e-8, rtol=1e-7)
    assert same(np.array([first]), np.array([oracle.isovalue(values, weights)]))
