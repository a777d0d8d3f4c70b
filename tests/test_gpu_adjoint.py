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


# Positions enter through the sample transform (bounds and their extrema
# owners), the group sums, the density levels and depth/weight, the normal
# splats' B-spline weights and the isovalue centroids
# (reconstruction_adjoint.py:89-177, sample_adjoint.py, splat_adjoint.py);
# screened runs add the constraint points: their screening terms in every
# level's operator, cross operator and point load, the base hierarchy's
# Galerkin products, the root conjugate gradient's matrix and the normal
# average's weight (assembly_adjoint.py:44-91, constraints.py:55-84).
POSITION_CASES = [
    ("psr", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=0.0)),
    ("exact", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=2.0, exact_interpolation=True)),
    ("root-padded-psr", "sphere", 512, "double", dict(depth=4, full_depth=2, base_depth=0, point_weight=0.0)),
    ("torus-psr", "torus", 8192, "double", dict(depth=6, full_depth=3, point_weight=0.0)),
    ("psr-float", "sphere", 512, "float", dict(depth=4, full_depth=2, point_weight=0.0)),
    ("screened", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=2.0)),
    ("cycles", "sphere", 512, "double", dict(depth=4, full_depth=2, point_weight=2.0, base_v_cycles=2, iters=3)),
    ("root-screened", "sphere", 512, "double", dict(depth=4, full_depth=0, base_depth=0, point_weight=2.0)),
    ("torus-screened", "torus", 8192, "double", dict(depth=6, full_depth=3, point_weight=2.0)),
    ("screened-float", "sphere", 512, "float", dict(depth=4, full_depth=2, point_weight=2.0)),
]


def _position_assignments(r, p):
    return _assignments(r) + [np.argmin(p, axis=0), np.argmax(p, axis=0)]


@pytest.mark.parametrize("name,shape,count,precision,params", POSITION_CASES, ids=[c[0] for c in POSITION_CASES])
def test_position_gradient_matches_central_differences(cuda, name, shape, count, precision, params):
    import torch
    from warp_psr_amd import reconstruct, reconstruct_differentiable
    p, n = fixture_arrays(shape, count, precision)
    rng = np.random.default_rng(283)
    tp = torch.from_numpy(p).to(cuda).requires_grad_(True)
    tn = torch.from_numpy(n).to(cuda).requires_grad_(True)
    rec = reconstruct_differentiable(tp, tn, **params)
    q = torch.from_numpy(_queries(p.astype(np.float64), rng).astype(p.dtype)).to(cuda)
    values, grads, valid = rec.evaluate(q, centered=True)
    assert bool(valid.bool().all())
    seed_v = torch.from_numpy(rng.normal(size=len(q))).to(cuda)
    seed_g = torch.from_numpy(rng.normal(size=(len(q), 3)) * 1e-2).to(cuda)
    loss = (values.double() * seed_v).sum() + (grads.double() * seed_g).sum()
    loss.backward()
    eps = 1e-6 if precision == "double" else 1e-4
    rtol, atol = (3e-5, 2e-7) if precision == "double" else (2e-2, 2e-3)
    reference = reconstruct(tp.detach(), tn.detach(), **params)
    fixed = _position_assignments(reference, p)
    for tracked, array, gradient in ((0, p, tp.grad), (1, n, tn.grad)):
        g = gradient.cpu().numpy().astype(np.float64)
        assert np.isfinite(g).all() and np.abs(g).max() > 0
        if tracked == 0:
            # a few random points, plus (float64) the bounds' owners whose runner-up is well clear of
            # the step, so that no cell or owner changes; in float32 an owner's step moves every cell
            owners = []
            if precision == "double":
                for a in range(3):
                    col = np.sort(p[:, a])
                    if col[1] - col[0] > 1e3 * eps:
                        owners.append(np.argmin(p[:, a]))
                    if col[-1] - col[-2] > 1e3 * eps:
                        owners.append(np.argmax(p[:, a]))
            interior = np.setdiff1d(np.arange(len(p)), np.concatenate([np.argmin(p, axis=0), np.argmax(p, axis=0)]))
            chosen = np.unique(np.concatenate([np.asarray(owners, dtype=np.int64),
                                               rng.choice(interior, 24 if precision == "double" else 6, replace=False)]))
            direction = np.zeros(p.shape)
            direction[chosen] = rng.normal(size=(len(chosen), 3))
        else:
            direction = rng.normal(size=array.shape)
        direction = direction.astype(array.dtype)
        losses = []
        for sign in (1, -1):
            moved = (array + sign * eps * direction).astype(array.dtype)
            pp, nn = (moved, n) if tracked == 0 else (p, moved)
            r = reconstruct(torch.from_numpy(pp).to(cuda), torch.from_numpy(nn).to(cuda), **params)
            for a, b in zip(fixed, _position_assignments(r, pp)):
                np.testing.assert_array_equal(a, b)
            loss_value, _ = _loss(r, q, seed_v, seed_g)
            losses.append(loss_value)
        fd = (losses[0] - losses[1]) / (2 * eps)
        np.testing.assert_allclose(np.sum(g * direction.astype(np.float64)), fd, rtol=rtol, atol=atol)
