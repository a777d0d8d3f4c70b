"""Size-independent properties at the benchmark size (800,000 points on the
six-ellipsoid union, depth 9, full_depth 4, point_weight 2).

The oracle would take minutes here, so these checks use properties that hold
at any size: repeatability (bitwise), sorted unique keys inside each level's
grid, consistent level offsets, the mixed-precision bound against float64
(|field_mixed - field_f64| <= 1e-6 with the same octree), an isovalue inside
the range of the centroid values, and a field that crosses the isovalue near
the inputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAMS = dict(depth=9, full_depth=4, point_weight=2.0)


@pytest.fixture(scope="module")
def workload(cuda):
    import torch
    from warp_psr_amd import synthetic
    p, n = synthetic.workload_ellipsoids(800000, dtype=np.float64)
    return p, n, torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda)


def test_fullsize_structure_and_repeatability(workload):
    from warp_psr_amd import reconstruct
    p, n, tp, tn = workload
    a = reconstruct(tp.float(), tn.float(), **PARAMS)
    b = reconstruct(tp.float(), tn.float(), **PARAMS)
    for name in ("keys", "coefficients", "cumulative"):
        assert np.array_equal(a.diagnostics.array(name).view(np.uint8), b.diagnostics.array(name).view(np.uint8)), name
    assert a.iso_value == b.iso_value
    counts = a.metadata["basis_counts"]
    keys = a.diagnostics.array("keys")
    assert len(keys) == sum(counts)
    assert np.all(np.diff(keys) > 0)
    off = 0
    for d, c in enumerate(counts):
        level = keys[off:off + c]
        off += c
        assert np.all(level >> 48 == d)
        coords = np.stack([(level >> 32) & 65535, (level >> 16) & 65535, level & 65535])
        assert coords.max() <= (1 << d)
    cv = a.diagnostics.array("centroid_values")
    assert cv.min() <= a.iso_value <= cv.max()


def test_fullsize_mixed_within_1e6_of_float64(workload):
    import torch
    from warp_psr_amd import reconstruct, synthetic
    p, n, tp, tn = workload
    full = reconstruct(tp, tn, **PARAMS)
    mixed = reconstruct(tp, tn, mixed=True, **PARAMS)
    keys = full.diagnostics.array("keys")
    assert np.array_equal(mixed.diagnostics.array("keys"), keys)
    iq = synthetic.interface_queries(keys, 9, full.diagnostics.array("model_to_cube"),
                                     full.diagnostics.array("cube_to_model"), p)
    q = torch.from_numpy(np.concatenate([p, iq])).to(tp.device)
    vf, _, okf = full.evaluate(q)
    vm, _, okm = mixed.evaluate(q)
    both = (okf.bool() & okm.bool()).cpu().numpy()
    err = np.abs(vm.cpu().numpy() - vf.cpu().numpy())[both]
    assert both.sum() > 0.99 * len(q)
    assert err.max() <= 1e-6
    assert abs(mixed.iso_value - full.iso_value) <= 1e-6
    # the surface passes near the inputs: the centered field changes sign
    # between each input and its offset along +-0.01 * normal for most inputs
    sel = np.random.default_rng(3).choice(len(p), 20000, replace=False)
    outer = torch.from_numpy(p[sel] + 0.01 * n[sel]).to(tp.device)
    inner = torch.from_numpy(p[sel] - 0.01 * n[sel]).to(tp.device)
    vo = mixed.evaluate(outer)[0].cpu().numpy() - mixed.iso_value
    vi = mixed.evaluate(inner)[0].cpu().numpy() - mixed.iso_value
    assert np.mean(np.sign(vo) != np.sign(vi)) > 0.9
