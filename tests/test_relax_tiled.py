"""The tile-parallel form of the sliced Gauss-Seidel schedule (CPU only).

The HIP library runs each wave of relaxation.py:128-143 (wavefront_groups) as
one launch in which every workgroup updates all four colours of one (x, y)
tile of one slice, recomputing a three-ring halo and keeping loads and stores
of a wave in two value buffers.  The oracle's restatement of that schedule
(relax_sliced_tiled, tiles of each wave in a seeded random order) must give
relax_sliced's bits on the real level operators of reconstructions; planted
mistakes (one halo ring short, one buffer) must not.
"""
import numpy as np
import pytest

import oracle
from parity import fixture_arrays, same


def _operator(shape, count, depth, precision):
    p, n = fixture_arrays(shape, count, precision)
    ref = oracle.reconstruct(p, n, depth=depth, full_depth=3)
    keys = ref.array("basis_keys", depth)
    offsets = ref.array("matrix_offsets", depth).astype(np.int32)
    columns = ref.array("matrix_columns", depth).astype(np.int32)
    values = ref.array("matrix_values", depth)
    rng = np.random.default_rng(count + depth)
    b = rng.normal(size=len(keys)).astype(values.dtype)
    x0 = rng.normal(size=len(keys)).astype(values.dtype)
    return keys, offsets, columns, values, b, x0


@pytest.mark.parametrize("shape,count,depth,precision", [
    ("torus", 8192, 6, "float"), ("sphere", 20000, 6, "double"), ("torus", 30000, 7, "double"),
    ("sphere", 4096, 5, "float")])
@pytest.mark.parametrize("iters", [1, 3, 8])
def test_tiled_schedule_equals_sliced(shape, count, depth, precision, iters):
    keys, offsets, columns, values, b, x0 = _operator(shape, count, depth, precision)
    expected = oracle.relax_sliced(keys, offsets, columns, values, b, x0, depth, iters)
    assert not same(expected, x0)
    for tile in (8, 16, 32, 64):
        got = oracle.relax_sliced_tiled(keys, offsets, columns, values, b, x0, depth, iters, tile, seed=tile * 7 + iters)
        assert same(got, expected), f"tile {tile}"


@pytest.mark.parametrize("plant", [1, 2])
def test_tiled_schedule_planted_mistakes_fail(plant):
    keys, offsets, columns, values, b, x0 = _operator("torus", 30000, 7, "double")
    expected = oracle.relax_sliced(keys, offsets, columns, values, b, x0, 7, 4)
    got = oracle.relax_sliced_tiled(keys, offsets, columns, values, b, x0, 7, 4, 16, seed=3, plant=plant)
    assert not same(got, expected)


def test_tiled_schedule_rejects_wide_couplings():
    """Couplings beyond one cell (not a degree-one operator) are refused."""
    keys = np.array([(3 << 48) | (0 << 32) | (0 << 16) | 0, (3 << 48) | (2 << 32) | (0 << 16) | 0], np.int64)
    offsets = np.array([0, 2, 4], np.int32)
    columns = np.array([0, 1, 0, 1], np.int32)
    values = np.array([4.0, 1.0, 1.0, 4.0])
    with pytest.raises(RuntimeError, match="within one cell"):
        oracle.relax_sliced_tiled(keys, offsets, columns, values, np.ones(2), np.zeros(2), 3, 1, 8)
