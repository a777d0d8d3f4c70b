"""The prolongation structure that k_galerkin's candidate scan relies on,
checked on the oracle's own prolongation matrices.

Trilinear prolongation (transfer_stencils.py:53-66) sends a fine coordinate
f (per axis) to the coarse coordinate f/2 when f is even and to (f - 1)/2 and
(f + 1)/2 when f is odd: every stored entry has |2c - f| <= 1, and 2c = f for
even f.  Hence a coarse slot at offset o (0..2, per axis) from a coarse row
is reached only from the fine offsets 2o - 1 .. 2o + 1 (0..4, around twice
the row), the 27 candidates the kernel scans per slot (it also counts the
contributions at run time and falls back to the full scan if any is missed).
"""
import numpy as np
import pytest

import oracle
from parity import fixture_arrays


def _coords(keys):
    k = np.asarray(keys).astype(np.int64)
    return np.stack([(k >> 32) & 65535, (k >> 16) & 65535, k & 65535], axis=-1)


@pytest.mark.parametrize("shape,count,precision,params", [
    ("torus", 2048, "double", dict(depth=5, full_depth=3)),
    ("sphere", 512, "float", dict(depth=4, full_depth=2, point_weight=0.0)),
    ("disconnected", 1024, "double", dict(depth=5, full_depth=2)),
])
def test_prolongation_parents(shape, count, precision, params):
    p, n = fixture_arrays(shape, count, precision)
    res = oracle.reconstruct(p, n, **params)
    checked = 0
    for d in range(1, params["depth"] + 1):
        fine = _coords(res.array("basis_keys", d))
        coarse = _coords(res.array("basis_keys", d - 1))
        offsets = res.array("prolong_offsets", d).astype(np.int64)
        columns = res.array("prolong_columns", d).astype(np.int64)
        assert len(offsets) == len(fine) + 1
        rows = np.repeat(np.arange(len(fine)), np.diff(offsets))
        f = fine[rows]
        c = coarse[columns]
        gap = 2 * c - f
        assert np.all(np.abs(gap) <= 1), f"depth {d}: a fine node reaches a coarse node beyond its two parents"
        assert np.all(gap[f % 2 == 0] == 0), f"depth {d}: an even fine coordinate reaches a coarse node other than f/2"
        # (for a coarse row r with c at offset o = c - r + 1, the fine offset
        # t = f - 2r + 2 satisfies t - 2o = f - 2c, so |gap| <= 1 is exactly
        # t in 2o - 1 .. 2o + 1)
        checked += len(rows)
    assert checked > 0
