"""The HIP pipeline on the reference's threshold fixtures, its large torus
reports, and the benchmark configurations at full size.

* docs/results/thresholds.json: every stage equal to the oracle bit for bit,
  support/active key SHA-256s, cube_to_model, the selected aggregate's splat
  depth at the exact log threshold, isovalue bounds (see
  test_oracle_thresholds.py);
* docs/results/torus-1m.json and torus-1m-d9.json: basis counts and
  cube_to_model exact, isovalue within 1e-6 of warp-psr's;
* the 20,000-point sphere at depth 6 (config 1): every stage and the field at
  its 25,000 queries equal to the oracle bit for bit;
* the 250,000-point noisy torus at depth 8 with point weights 0 and 4
  (config 2), the 1,000,000-point torus at depth 9 and the 800,000-point
  ellipsoid union at depth 9 in mixed precision (config 3): per-level basis
  keys, the field's keys, corrections, cumulative coefficients, isovalue bits
  and values/gradients/validity at a fixed query set equal the oracle's, by
  the SHA-256s committed in tests/golden/oracle_hashes.json (the serial oracle
  takes minutes at these sizes);
* 64 batched clouds (config 5) in mixed precision: every cloud's field within
  1e-6 of its own float64 run at its inputs.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from golden_cases import LARGE_TORUS, THRESHOLDS, large_case_inputs, selected_splat, sha, threshold_input, \
    threshold_params, torus_input, torus_params
from parity import compare_stages, same

pytestmark = pytest.mark.gpu

HASHES = json.loads((Path(__file__).parent / "golden" / "oracle_hashes.json").read_text())


def _gpu(cuda, p, n, **params):
    import torch
    from warp_psr_amd import reconstruct
    return reconstruct(torch.from_numpy(np.ascontiguousarray(p)).to(cuda), torch.from_numpy(np.ascontiguousarray(n)).to(cuda),
                       **params)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("f", THRESHOLDS, ids=[f["name"] for f in THRESHOLDS])
def test_threshold_fixture(cuda, f):
    pts, nrm = threshold_input(f)
    params = threshold_params(f)
    print(f"threshold fixture {f['name']}", flush=True)
    res = _gpu(cuda, pts, nrm, **params)
    ref = oracle.reconstruct(pts, nrm, **params)
    bad = compare_stages(res.diagnostics, ref, ref.params)
    assert not bad, "; ".join(f"{s}[{lv}] {m}" for s, lv, m in bad[:5])
    for lv in f["tree"]:
        assert sha(res.diagnostics.array("tree_keys", lv["depth"])) == lv["support_sha256"]
        assert sha(res.diagnostics.array("basis_keys", lv["depth"])) == lv["active_sha256"]
    ctm = res.diagnostics.array("cube_to_model").reshape(4, 4)
    np.testing.assert_array_equal(ctm, np.asarray(f["cube_to_model"], dtype=ctm.dtype))
    if f.get("selected_splat"):
        raw, top = selected_splat(res.diagnostics, f)
        assert float(raw) == f["selected_splat"]["raw_depth"] and top == f["selected_splat"]["top_depth"]
    assert abs(res.iso_value - f["iso_value"]) <= (1e-6 if f["params"]["precision"] == "float" else 3e-15)


@pytest.mark.parametrize("name", ["torus-1m", "torus-1m-d9"])
def test_large_torus_reports(cuda, name):
    ref = LARGE_TORUS[name]
    p, n = torus_input(name)
    print(f"large torus {name}", flush=True)
    res = _gpu(cuda, p, n, **torus_params(name))
    depth = ref["parameters"]["depth"]
    assert [len(res.diagnostics.array("basis_keys", d)) for d in range(depth + 1)] == ref["basis_counts"]
    ctm = res.diagnostics.array("cube_to_model").reshape(4, 4)
    np.testing.assert_array_equal(ctm, np.asarray(ref["cube_to_model"], dtype=np.float32))
    assert abs(res.iso_value - ref["iso_value"]) <= 1e-6


def test_sphere20k_depth6_bit_exact(cuda):
    import torch
    p, n, q, params = large_case_inputs("sphere20k-d6")
    res = _gpu(cuda, p, n, **params)
    ref = oracle.reconstruct(p, n, **params)
    bad = compare_stages(res.diagnostics, ref, ref.params)
    assert not bad, "; ".join(f"{s}[{lv}] {m}" for s, lv, m in bad[:5])
    values, grads, valid = (a.cpu().numpy() for a in res.evaluate(torch.from_numpy(q).to(cuda)))
    rv, rg, rvalid = ref.evaluate(q)
    assert same(values, rv) and same(grads, rg)
    np.testing.assert_array_equal(valid, rvalid)


@pytest.mark.parametrize("name", ["sphere20k-d6", "torus250k-d8-pw0", "torus250k-d8-pw4", "torus-1m-d9",
                                  "ellipsoids800k-d9-mixed"])
def test_full_size_against_oracle_hashes(cuda, name):
    import torch
    h = HASHES[name]
    p, n, q, params = large_case_inputs(name)
    assert digest(np.column_stack([p, n])) == h["input_sha256"]
    print(f"full-size case {name}: {len(p)} points, {len(q)} queries", flush=True)
    res = _gpu(cuda, p, n, **params)
    d = res.diagnostics
    depth = params["depth"]
    assert [len(d.array("basis_keys", lv)) for lv in range(depth + 1)] == h["basis_counts"]
    assert [digest(d.array("basis_keys", lv)) for lv in range(depth + 1)] == h["basis_keys"]
    assert digest(d.array("keys")) == h["keys"]
    assert digest(d.array("coefficients")) == h["coefficients"], "level corrections differ from the oracle"
    assert digest(d.array("cumulative")) == h["cumulative"]
    dtype = d.array("coefficients").dtype
    assert np.asarray(res.iso_value, dtype=dtype).tobytes().hex() == h["iso_hex"]
    values, grads, valid = res.evaluate(torch.from_numpy(q).to(cuda))
    assert digest(values.cpu().numpy()) == h["values"]
    assert digest(grads.cpu().numpy()) == h["gradients"]
    assert digest(valid.cpu().numpy()) == h["valid"]


def test_batched_mixed_within_1e6_of_float64(cuda):
    """Config 5: 64 clouds (5k-200k points, depth 7/8) reconstructed together in
    mixed precision; each cloud's field at its inputs within 1e-6 of that
    cloud's own float64 reconstruction."""
    import torch
    from warp_psr_amd import reconstruct, reconstruct_batch, synthetic
    work = synthetic.workload_batch(dtype=np.float64)
    clouds = [(torch.from_numpy(p).to(cuda), torch.from_numpy(nn).to(cuda)) for p, nn, _ in work]
    opts = [dict(depth=d, full_depth=4, point_weight=2.0, mixed=True) for _, _, d in work]
    print(f"batch: {len(work)} clouds, {sum(len(p) for p, _, _ in work)} points", flush=True)
    batch = reconstruct_batch(clouds, per_cloud=opts, max_concurrency=8)
    worst = 0.0
    for (p, nn), o, b in zip(clouds, opts, batch):
        full = reconstruct(p, nn, **{k: v for k, v in o.items() if k != "mixed"})
        vb, _, okb = b.evaluate(p)
        vf, _, okf = full.evaluate(p)
        both = (okb.bool() & okf.bool())
        assert bool(both.any())
        err = float((vb - vf).abs()[both].max())
        worst = max(worst, err)
        assert np.array_equal(b.diagnostics.array("keys"), full.diagnostics.array("keys"))
    print(f"batch: worst field difference vs float64 {worst:.3e}", flush=True)
    assert worst <= 1e-6
