"""The lookup fallbacks give the same bits as the tables they replace.

Levels up to depth 9 use dense z-major indices, levels up to depth 12 (x, y)
column tables, deeper levels searches of the sorted keys.  Runs at the
benchmark sizes never reach the fallbacks, so these tests lower the table
limits (PSR_DENSE_INDEX_MAX_CELLS, PSR_COLUMN_TABLE_MAX_DIM,
PSR_FIELD_TABLE_MAX_DEPTH; read at each use) and check every stage array and
field query against the oracle, bit for bit, on fixtures whose finer levels
would otherwise take the dense path.
"""
import numpy as np
import pytest

import oracle
from parity import FIXTURES, compare_stages, fixture_arrays, fixture_queries, same

pytestmark = pytest.mark.gpu

LIMITS = {
    "column-tables": {"PSR_DENSE_INDEX_MAX_CELLS": "0"},
    "searches": {"PSR_DENSE_INDEX_MAX_CELLS": "0", "PSR_COLUMN_TABLE_MAX_DIM": "0", "PSR_FIELD_TABLE_MAX_DEPTH": "-1"},
}
CASES = [f for f in FIXTURES if f[0] in ("torus", "torus-double", "sparse", "torus-mixed", "torus-base-zero-float")]


@pytest.mark.parametrize("limits", list(LIMITS), ids=list(LIMITS))
@pytest.mark.parametrize("name,shape,count,precision,params", CASES, ids=[c[0] for c in CASES])
def test_fallbacks_bit_exact(cuda, monkeypatch, limits, name, shape, count, precision, params):
    import torch
    from warp_psr_amd import reconstruct
    for k, v in LIMITS[limits].items():
        monkeypatch.setenv(k, v)
    p, n = fixture_arrays(shape, count, precision)
    print(f"lookup fallbacks: {name} with {limits}", flush=True)
    gpu = reconstruct(torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda), **params)
    ref = oracle.reconstruct(p, n, **params)
    mismatches = compare_stages(gpu.diagnostics, ref, ref.params)
    assert not mismatches, "\n".join(f"{s}[{lv}]: {m}" for s, lv, m in mismatches[:20])
    assert gpu.iso_value == ref.iso_value
    q = fixture_queries(p)
    values, grads, valid = gpu.evaluate(torch.from_numpy(q).to(cuda))
    rv, rg, rvalid = ref.evaluate(q)
    np.testing.assert_array_equal(valid.cpu().numpy(), rvalid)
    assert same(values.cpu().numpy(), rv)
    assert same(grads.cpu().numpy(), rg)
