"""The reference's threshold fixtures and large torus runs against the oracle (CPU).

* docs/results/thresholds.json (tools/check_thresholds.py): split-plane
  ownership probes (scale 2, screened and unscreened) and the two sphere runs
  whose samples_per_node straddle the exact log threshold of one aggregate's
  top splat depth, in float32 and float64.  The regenerated inputs hash to
  the recorded SHA-256s; every level's support and active keys hash to the
  recorded values; cube_to_model is equal; the selected aggregate's raw splat
  depth and top depth equal the recorded ones (3.0000002 -> 4 and 3.0 -> 3 in
  float32); isovalues are within 1e-6 (float32) and 3e-15 (float64).
* docs/results/torus-100k.json: the 100,000-point torus fixture at depth 8 --
  basis counts and cube_to_model exact, isovalue within one float32 ulp.
  (torus-1m and torus-1m-d9 run on the GPU, against the same reports.)
"""
import hashlib

import numpy as np
import pytest

import oracle
from golden_cases import LARGE_TORUS, THRESHOLDS, selected_splat, sha, threshold_input, threshold_params, torus_input, \
    torus_params
from warp_psr_amd import synthetic


@pytest.mark.parametrize("f", THRESHOLDS, ids=[f["name"] for f in THRESHOLDS])
def test_threshold_fixture(f):
    pts, nrm = threshold_input(f)
    assert hashlib.sha256(synthetic.fixture_input_text(pts, nrm)).hexdigest() == f["input_sha256"]
    res = oracle.reconstruct(pts, nrm, **threshold_params(f))
    for lv in f["tree"]:
        assert sha(res.array("tree_keys", lv["depth"])) == lv["support_sha256"], lv["depth"]
        assert sha(res.array("basis_keys", lv["depth"])) == lv["active_sha256"], lv["depth"]
    ctm = res.array("cube_to_model").reshape(4, 4)
    np.testing.assert_array_equal(ctm, np.asarray(f["cube_to_model"], dtype=ctm.dtype))
    if f.get("selected_splat"):
        raw, top = selected_splat(res, f)
        assert float(raw) == f["selected_splat"]["raw_depth"] and top == f["selected_splat"]["top_depth"]
        assert top == f["integer_depth"] + ("below" in f["name"])
    bound = 1e-6 if f["params"]["precision"] == "float" else 3e-15
    assert abs(res.iso_value - f["iso_value"]) <= bound


def test_torus_100k_depth8():
    name = "torus-100k"
    p, n = torus_input(name)
    res = oracle.reconstruct(p, n, **torus_params(name))
    ref = LARGE_TORUS[name]
    assert [len(res.array("basis_keys", d)) for d in range(9)] == ref["basis_counts"]
    ctm = res.array("cube_to_model").reshape(4, 4)
    np.testing.assert_array_equal(ctm, np.asarray(ref["cube_to_model"], dtype=np.float32))
    assert abs(np.float32(res.iso_value) - np.float32(ref["iso_value"])) <= np.spacing(np.float32(ref["iso_value"]))
