"""HIP pipeline vs the CPU oracle, stage by stage and bit for bit.

Inputs are the reference's fixture shapes (tools/make_fixture.py:29-58,
regenerated bit-exactly; their golden key hashes are checked against the
oracle in test_oracle_golden.py).  Every stage array -- sample groups,
aggregates, density levels, splat depths, normal records, tree and basis keys,
normal/point loads, assembled matrices, right-hand sides, corrections,
cumulative coefficients, centroid values and the isovalue -- must equal the
oracle's bit for bit, and so must field values, gradients and validity at
queries (NaN outside the cube).
"""
import numpy as np
import pytest

import oracle
from parity import FIXTURES, compare_stages, fixture_arrays, fixture_queries, same

pytestmark = pytest.mark.gpu


def _run(cuda, shape, count, precision, params):
    import torch
    from warp_psr_amd import reconstruct
    p, n = fixture_arrays(shape, count, precision)
    tp = torch.from_numpy(p).to(cuda)
    tn = torch.from_numpy(n).to(cuda)
    gpu = reconstruct(tp, tn, **params)
    ref = oracle.reconstruct(p, n, **params)
    return p, gpu, ref


@pytest.mark.parametrize("name,shape,count,precision,params", FIXTURES, ids=[f[0] for f in FIXTURES])
def test_stages_bit_exact(cuda, name, shape, count, precision, params):
    p, gpu, ref = _run(cuda, shape, count, precision, params)
    mismatches = compare_stages(gpu.diagnostics, ref, ref.params)
    assert not mismatches, "\n".join(f"{s}[{lv}]: {m}" for s, lv, m in mismatches[:20])
    assert gpu.iso_value == ref.iso_value


@pytest.mark.parametrize("name,shape,count,precision,params", FIXTURES[:6], ids=[f[0] for f in FIXTURES[:6]])
def test_evaluate_bit_exact(cuda, name, shape, count, precision, params):
    import torch
    p, gpu, ref = _run(cuda, shape, count, precision, params)
    q = fixture_queries(p)
    values, grads, valid = gpu.evaluate(torch.from_numpy(q).to(cuda))
    rv, rg, rvalid = ref.evaluate(q)
    np.testing.assert_array_equal(valid.cpu().numpy(), rvalid)
    assert same(values.cpu().numpy(), rv)
    assert same(grads.cpu().numpy(), rg)
    assert not rvalid[-2:].any()  # NaN and far-away queries are invalid
    assert np.isnan(values.cpu().numpy()[-2:]).all()


def test_repeatable(cuda):
    """Two runs of the same input give identical fields (bitwise)."""
    import torch
    from warp_psr_amd import reconstruct
    p, n = fixture_arrays("torus", 8192, "float")
    tp, tn = torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda)
    a = reconstruct(tp, tn, depth=6, full_depth=3)
    b = reconstruct(tp, tn, depth=6, full_depth=3)
    for name in ("keys", "coefficients", "cumulative"):
        assert same(a.diagnostics.array(name), b.diagnostics.array(name))
    assert a.iso_value == b.iso_value


@pytest.mark.parametrize("name,shape,count,precision,params", [FIXTURES[0], FIXTURES[5], FIXTURES[-2]],
                         ids=[FIXTURES[0][0], FIXTURES[5][0], FIXTURES[-2][0]])
def test_device_field_header_matches_evaluate(cuda, name, shape, count, precision, params):
    """The header-only psr_evaluate_field (via psr_field_evaluate on the
    zero-copy PsrField view) gives the bits of psr_evaluate and the oracle."""
    import torch
    p, gpu, ref = _run(cuda, shape, count, precision, params)
    q = torch.from_numpy(fixture_queries(p)).to(cuda)
    values, grads, valid = gpu.evaluate(q)
    hv, hg, hvalid = gpu.device_field().evaluate(q)
    np.testing.assert_array_equal(hvalid.cpu().numpy(), valid.cpu().numpy())
    assert same(hv.cpu().numpy(), values.cpu().numpy())
    assert same(hg.cpu().numpy(), grads.cpu().numpy())
    rv, _, _ = ref.evaluate(q.cpu().numpy())
    assert same(hv.cpu().numpy(), rv)


def test_batch_matches_single_reconstructions(cuda):
    """psr_reconstruct_batch: every cloud's result equals its own reconstruct() bit for bit."""
    import torch
    from warp_psr_amd import reconstruct, reconstruct_batch
    clouds, opts = [], []
    for i, (shape, count, params) in enumerate([("sphere", 512, dict(depth=4, full_depth=2)),
                                                ("torus", 8192, dict(depth=6, full_depth=3)),
                                                ("cluster", 1024, dict(depth=5, full_depth=2, point_weight=0.0)),
                                                ("plane", 1024, dict(depth=5, full_depth=2)),
                                                ("sphere", 512, dict(depth=5, full_depth=1, iters=3)),
                                                ("disconnected", 1024, dict(depth=5, full_depth=2))]):
        p, n = fixture_arrays(shape, count, "double")
        clouds.append((torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda)))
        opts.append(params)
    batch = reconstruct_batch(clouds, per_cloud=opts, max_concurrency=4)
    assert len(batch) == len(clouds)
    for (p, n), params, b in zip(clouds, opts, batch):
        single = reconstruct(p, n, **params)
        for name in ("keys", "coefficients", "cumulative"):
            assert same(b.diagnostics.array(name), single.diagnostics.array(name)), name
        assert b.iso_value == single.iso_value


@pytest.mark.parametrize("precision", ["double", "float"])
def test_grouped_batch_relaxation_bit_exact(cuda, monkeypatch, precision):
    """With PSR_BATCH_GROUPS=1 batch groups relax each finer level in one
    launch sequence over all members' tiles.  Groups of 4 here mix depths 6
    and 5, first finer levels 4 and 3, and (double inputs) mixed and float64
    operators; every cloud equals its own reconstruct() bit for bit, and the
    batch without groups (PSR_BATCH_GROUPS=0, the default)."""
    import torch
    from warp_psr_amd import reconstruct, reconstruct_batch
    shapes = [("torus", 8192), ("sphere", 2048), ("cluster", 1024), ("plane", 1024), ("disconnected", 1024)]
    clouds, opts = [], []
    for i in range(10):
        shape, count = shapes[i % 5]
        p, n = fixture_arrays(shape, count, precision)
        clouds.append((torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda)))
        depth = 6 if i % 2 == 0 else 5
        params = dict(depth=depth, full_depth=3 if depth == 6 else 2, point_weight=2.0)
        if precision == "double" and i % 3 == 0:
            params["mixed"] = True
        opts.append(params)
    print(f"grouped batch: {len(clouds)} {precision} clouds, concurrency 4", flush=True)
    monkeypatch.setenv("PSR_BATCH_GROUPS", "1")
    batch = reconstruct_batch(clouds, per_cloud=opts, max_concurrency=4)
    monkeypatch.setenv("PSR_BATCH_GROUPS", "0")
    print("grouped batch: the same clouds without groups", flush=True)
    plain = reconstruct_batch(clouds, per_cloud=opts, max_concurrency=4)
    for (p, n), params, b, q in zip(clouds, opts, batch, plain):
        single = reconstruct(p, n, **params)
        for name in ("keys", "coefficients", "cumulative"):
            assert same(b.diagnostics.array(name), single.diagnostics.array(name)), name
            assert same(q.diagnostics.array(name), single.diagnostics.array(name)), name
        assert b.iso_value == single.iso_value == q.iso_value
