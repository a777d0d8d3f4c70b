"""Reverse mode w.r.t. the input normals vs central finite differences.

Mirrors tests/test_reconstruction_adjoint.py:36-70 of the reference: a
centered field loss (values and model-space gradients at queries, random
seeds) is differentiated with reconstruct_differentiable and compared along a
random direction with central differences of the bit-exact forward, after
checking that the perturbed runs keep every discrete assignment (octree keys,
normal-record keys, sample groups, root-solve branch).  The cycles and root
cases run the root conjugate gradient, whose iterations are reversed too.  float64, eps 1e-6,
tolerances as the reference's float64 case (rtol 3e-5, atol 2e-7).
"""
import numpy as np
import pytest

from parity import fixture_arrays

pytestmark = pytest.mark.gpu

CASES = [
    ("screened", "sphere", 512, dict(depth=4, full_depth=2, point_weight=2.0)),
    ("psr", "sphere", 512, dict(depth=4, full_depth=2, point_weight=0.0)),
    ("cycles", "sphere", 512, dict(depth=4, full_depth=2, point_weight=2.0, base_v_cycles=2, iters=3)),
    ("torus", "torus", 8192, dict(depth=6, full_depth=3, point_weight=2.0)),
    ("mixed", "torus", 8192, dict(depth=6, full_depth=3, point_weight=2.0, mixed=True)),
    ("root-padded-psr", "sphere", 512, dict(depth=4, full_depth=2, base_depth=0, point_weight=0.0)),
    ("root-screened", "sphere", 512, dict(depth=4, full_depth=0, base_depth=0, point_weight=2.0)),
]


def _queries(p, rng, count=64):
    lo, hi = p.min(axis=0), p.max(axis=0)
    inside = lo + (hi - lo) * rng.uniform(0.05, 0.95, size=(count, 3))
    return np.concatenate([p[rng.choice(len(p), count, replace=False)], inside])


def _loss(r, q, seed_v, seed_g):
    import torch
    v, g, ok = r.evaluate(q)
    v = v - r.iso_value
    return float((v.double() * seed_v).sum() + (g.double() * seed_g).sum()), ok


def _assignments(r):
    d = r.diagnostics
    return [d.array("keys"), d.array("normal_keys"), d.array("input_group_indices"), r.metadata["root_cg_cycles"]]


@pytest.mark.parametrize("name,shape,count,params", CASES, ids=[c[0] for c in CASES])
def test_normal_gradient_matches_central_differences(cuda, name, shape, count, params):
    import torch
    from warp_psr_amd import reconstruct, reconstruct_differentiable
    p, n = fixture_arrays(shape, count, "double")
    rng = np.random.default_rng(283)
    tp = torch.from_numpy(p).to(cuda)
    tn = torch.from_numpy(n).to(cuda).requires_grad_(True)
    rec = reconstruct_differentiable(tp, tn, **params)
    q = torch.from_numpy(_queries(p, rng)).to(cuda)
    values, grads, valid = rec.evaluate(q, centered=True)
    assert bool(valid.bool().all())
    seed_v = torch.from_numpy(rng.normal(size=len(q))).to(cuda)
    seed_g = torch.from_numpy(rng.normal(size=(len(q), 3)) * 1e-2).to(cuda)
    loss = (values * seed_v).sum() + (grads * seed_g).sum()
    if name.startswith("root") or name == "cycles":
        assert rec.metadata["root_cg_cycles"], "this case exercises the reversed root conjugate gradient"
    loss.backward()
    gradient = tn.grad.cpu().numpy()
    assert np.isfinite(gradient).all() and np.abs(gradient).max() > 0
    direction = rng.normal(size=n.shape)
    eps = 1e-6
    reference = reconstruct(tp, tn.detach(), **params)
    fixed = _assignments(reference)
    losses = []
    for sign in (1, -1):
        tn2 = torch.from_numpy(n + sign * eps * direction).to(cuda)
        r = reconstruct(tp, tn2, **params)
        for a, b in zip(fixed, _assignments(r)):
            np.testing.assert_array_equal(a, b)
        loss, _ = _loss(r, q, seed_v, seed_g)
        losses.append(loss)
    fd = (losses[0] - losses[1]) / (2 * eps)
    np.testing.assert_allclose(np.sum(gradient * direction), fd, rtol=3e-5, atol=2e-7)


def test_query_gradient_matches_central_differences(cuda):
    import torch
    from warp_psr_amd import reconstruct_differentiable
    p, n = fixture_arrays("torus", 8192, "double")
    rng = np.random.default_rng(5)
    tp = torch.from_numpy(p).to(cuda)
    tn = torch.from_numpy(n).to(cuda).requires_grad_(True)
    rec = reconstruct_differentiable(tp, tn, depth=5, full_depth=2)
    q0 = _queries(p, rng, 32)
    q = torch.from_numpy(q0).to(cuda).requires_grad_(True)
    seed_v = torch.from_numpy(rng.normal(size=len(q0))).to(cuda)
    seed_g = torch.from_numpy(rng.normal(size=(len(q0), 3)) * 1e-2).to(cuda)
    values, grads, _ = rec.evaluate(q)
    ((values * seed_v).sum() + (grads * seed_g).sum()).backward()
    gq = q.grad.cpu().numpy()
    direction = rng.normal(size=q0.shape)
    eps = 1e-7
    losses = []
    for sign in (1, -1):
        v, g, _ = rec._primal.evaluate(torch.from_numpy(q0 + sign * eps * direction).to(cuda))
        losses.append(float((v * seed_v).sum() + (g * seed_g).sum()))
    fd = (losses[0] - losses[1]) / (2 * eps)
    np.testing.assert_allclose(np.sum(gq * direction), fd, rtol=1e-5, atol=1e-7)


def test_gradient_accumulates_and_repeats(cuda):
    """Two backward passes through the same evaluation add; repeated calls agree."""
    import torch
    from warp_psr_amd import reconstruct_differentiable
    p, n = fixture_arrays("sphere", 512, "double")
    tp = torch.from_numpy(p).to(cuda)
    tn = torch.from_numpy(n).to(cuda).requires_grad_(True)
    rec = reconstruct_differentiable(tp, tn, depth=4, full_depth=2)
    q = torch.from_numpy(_queries(p, np.random.default_rng(1), 16)).to(cuda)
    values, _, _ = rec.evaluate(q, centered=True)
    values.sum().backward(retain_graph=True)
    first = tn.grad.clone()
    values.sum().backward()
    np.testing.assert_allclose(tn.grad.cpu().numpy(), 2 * first.cpu().numpy(), rtol=1e-9, atol=1e-12)
