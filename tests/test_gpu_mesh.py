"""Marching cubes on the GPU vs the CPU restatement (oracle/marching_cubes.py).

Mirrors tests/test_meshing.py and tests/test_reconstruction_mesh_adjoint.py of
the reference: sparse (slab sizes 1 and 7) and dense backends give the same
mesh; all 256 single-cell cases; the exact-threshold policy; nonfinite values
rejected; a reconstruction's mesh; vertex losses differentiated to the input
positions and normals against central differences (float32: eps 1e-4, rtol
3e-2, atol 2e-3 as the reference).  Grid points, vertices and indices are
compared bit for bit with the restatement fed the same node values.
"""
import numpy as np
import pytest

from oracle import marching_cubes as mc
from parity import fixture_arrays

pytestmark = pytest.mark.gpu

BOUNDS = ((-0.9, -0.8, -0.7), (0.95, 0.85, 0.75))
SHAPE = (19, 22, 25)


def _dot(a):
    # elementwise, in a fixed order (a reduction's order may depend on the
    # tensor's shape, and grid slabs and cell corners differ in shape)
    return a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1] + a[:, 2] * a[:, 2]


def analytic(kind, parameter):
    import torch

    def evaluate(points):
        p = points
        if kind == 0:
            return _dot(p) - parameter
        if kind == 1:
            a = p - torch.tensor([-0.48, 0.05, 0.0], device=p.device)
            b = p - torch.tensor([0.42, -0.03, 0.0], device=p.device)
            return torch.minimum(_dot(a), _dot(b)) - parameter
        if kind == 2:
            return p[:, 0] + 0.31 * p[:, 1] - 0.22 * p[:, 2] - parameter
        if kind == 3:
            return torch.ones(p.shape[0], dtype=torch.float32, device=p.device)
        return torch.full((p.shape[0],), float("nan"), dtype=torch.float32, device=p.device)
    return evaluate


def expected_mesh(f, shape, bounds, cuda):
    import torch
    nodes = mc.node_points(shape, *bounds)
    values = f(torch.from_numpy(nodes).to(cuda)).cpu().numpy()
    return mc.extract(values.reshape(shape), shape, *bounds)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_sparse_and_dense_match_restatement(cuda, kind):
    from warp_psr_amd.meshing import extract_isosurface
    f = analytic(kind, 0.0837 if kind == 1 else 0.2941)
    print(f"mesh kind {kind}", flush=True)
    vertices, indices = expected_mesh(f, SHAPE, BOUNDS, cuda)
    for backend, slab in (("sparse", 1), ("sparse", 7), ("dense", 16)):
        mesh = extract_isosurface(f, resolution=SHAPE, bounds=BOUNDS, backend=backend, slab_size=slab, device=cuda)
        np.testing.assert_array_equal(mesh.vertices.cpu().numpy(), vertices)
        np.testing.assert_array_equal(mesh.indices.cpu().numpy(), indices)
        assert not mesh.vertices.requires_grad
        if kind == 3:
            assert len(mesh.vertices) == len(mesh.indices) == mesh.metadata["retained_cells"] == 0
    if kind in (0, 1):
        faces = indices.reshape(-1, 3)
        edges = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]), axis=1)
        _, counts = np.unique(edges, axis=0, return_counts=True)
        assert np.all(counts == 2), "closed analytic surface must have no cracks"


def test_grid_points_match_restatement(cuda):
    import torch
    from warp_psr_amd.meshing import extract_isosurface
    seen = []

    def record(points):
        seen.append(points.cpu().numpy().copy())
        return _dot(points) - 0.2941
    extract_isosurface(record, resolution=SHAPE, bounds=BOUNDS, backend="dense", device=cuda)
    np.testing.assert_array_equal(seen[0], mc.node_points(SHAPE, *BOUNDS))
    assert torch.cuda.is_available()


def test_all_256_cases_and_threshold_policy(cuda):
    import torch
    from warp_psr_amd.meshing import extract_isosurface
    settings = dict(resolution=2, bounds=((0., 0., 0.), (1., 1., 1.)), device=cuda)
    corners = torch.zeros(8, dtype=torch.float32, device=cuda)

    def sample(points):
        index = points[:, 0].round().long() * 4 + points[:, 1].round().long() * 2 + points[:, 2].round().long()
        # corner c of the grid cell sits at mc.CORNERS[c]; map node (x, y, z) to c
        lookup = torch.tensor([0, 4, 3, 7, 1, 5, 2, 6], device=points.device)
        return corners[lookup[index]]
    table = mc.case_table()
    for code in range(256):
        data = np.array([1.0 if code & (1 << j) else -1.0 for j in range(8)], dtype=np.float32)
        corners.copy_(torch.from_numpy(data))
        values = np.empty((2, 2, 2), np.float32)
        for c, (x, y, z) in enumerate(mc.CORNERS):
            values[x, y, z] = data[c]
        v, i = mc.extract(values, (2, 2, 2), (0, 0, 0), (1, 1, 1), table)
        for backend in ("dense", "sparse"):
            mesh = extract_isosurface(sample, backend=backend, slab_size=1, **settings)
            np.testing.assert_array_equal(mesh.vertices.cpu().numpy(), v)
            np.testing.assert_array_equal(mesh.indices.cpu().numpy(), i)
    corners.zero_()
    for backend in ("dense", "sparse"):
        mesh = extract_isosurface(sample, backend=backend, **settings)
        assert len(mesh.indices) == len(mesh.vertices) == 0
    corners.copy_(torch.tensor([0., -1., -1., -1., -1., -1., -1., -1.]))
    for backend in ("dense", "sparse"):
        mesh = extract_isosurface(sample, backend=backend, **settings)
        assert len(mesh.indices) == len(mesh.vertices) == 3
        np.testing.assert_array_equal(mesh.vertices.cpu().numpy(), np.zeros((3, 3)))


def test_nonfinite_field_rejected(cuda):
    from warp_psr_amd.meshing import extract_isosurface
    for backend, extra in (("dense", {}), ("sparse", {}), ("lipschitz", {"lipschitz_bound": 1.0})):
        with pytest.raises(ValueError, match="nonfinite"):
            extract_isosurface(analytic(4, 0.0), resolution=3, bounds=((0., 0., 0.), (1., 1., 1.)), backend=backend,
                               device=cuda, **extra)
    with pytest.raises(ValueError, match="proven"):
        extract_isosurface(analytic(0, 0.2941), resolution=3, bounds=((0., 0., 0.), (1., 1., 1.)), backend="lipschitz",
                           device=cuda)


def _cube_bounds(points, shrink=0.98):
    lo, hi = points.min(axis=0).astype(np.float64), points.max(axis=0).astype(np.float64)
    center, half = (lo + hi) / 2, (hi - lo).max() / 2 * 1.1 * shrink
    return tuple(center - half), tuple(center + half)


def test_reconstruction_mesh_matches_restatement(cuda):
    import torch
    from warp_psr_amd import reconstruct
    from warp_psr_amd.meshing import extract_mesh
    p, n = fixture_arrays("torus", 8192, "float")
    r = reconstruct(torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda), depth=5, full_depth=2)
    bounds = _cube_bounds(p)
    shape = (25, 28, 31)
    print("mesh of torus reconstruction", flush=True)
    values, _, valid = r.evaluate(torch.from_numpy(mc.node_points(shape, *bounds)).to(cuda))
    assert bool(valid.bool().all())
    centered = (values - torch.tensor([r.iso_value], dtype=torch.float32, device=cuda)).cpu().numpy()
    v, i = mc.extract(centered.reshape(shape), shape, *bounds)
    assert len(i) > 0
    for backend, slab in (("dense", 16), ("sparse", 5)):
        mesh = extract_mesh(r, resolution=shape, bounds=bounds, backend=backend, slab_size=slab)
        np.testing.assert_array_equal(mesh.vertices.cpu().numpy(), v)
        np.testing.assert_array_equal(mesh.indices.cpu().numpy(), i)
    with pytest.raises(ValueError, match="backward pass"):
        extract_mesh(r, resolution=shape, bounds=bounds, differentiable=True)
    with pytest.raises(ValueError, match="float32"):
        rd = reconstruct(torch.from_numpy(p.astype(np.float64)).to(cuda), torch.from_numpy(n.astype(np.float64)).to(cuda),
                         depth=4, full_depth=2)
        extract_mesh(rd, resolution=3, bounds=bounds)


def _assignments(r, p):
    d = r.diagnostics
    return [d.array("keys"), d.array("normal_keys"), d.array("input_group_indices"), r.metadata["root_cg_cycles"],
            np.argmin(p, axis=0), np.argmax(p, axis=0)]


@pytest.mark.parametrize("backend", ["dense", "sparse"])
def test_mesh_vertex_loss_to_inputs(cuda, backend):
    """tests/test_reconstruction_mesh_adjoint.py:11-44 of the reference: positions
    move at interior points (the bounds' owners would move every float32 cell)."""
    import torch
    from warp_psr_amd import reconstruct, reconstruct_differentiable
    from warp_psr_amd.meshing import extract_mesh
    p, n = fixture_arrays("sphere", 512, "float")
    options = dict(depth=4, full_depth=2, iters=2)
    bounds = _cube_bounds(p)
    mesh_options = dict(resolution=11, bounds=bounds, backend=backend)
    tp = torch.from_numpy(p).to(cuda).requires_grad_(True)
    tn = torch.from_numpy(n).to(cuda).requires_grad_(True)
    result = reconstruct_differentiable(tp, tn, **options)
    mesh = extract_mesh(result, differentiable=True, **mesh_options)
    reference = reconstruct(tp.detach(), tn.detach(), **options)
    expected = extract_mesh(reference, **mesh_options)
    assert len(mesh.vertices) > 0 and mesh.vertices.requires_grad
    np.testing.assert_array_equal(mesh.indices.cpu().numpy(), expected.indices.cpu().numpy())
    np.testing.assert_array_equal(mesh.vertices.detach().cpu().numpy(), expected.vertices.cpu().numpy())
    rng = np.random.default_rng(372)
    seed = rng.normal(size=tuple(mesh.vertices.shape)) / len(mesh.vertices)
    (mesh.vertices.double() * torch.from_numpy(seed).to(cuda)).sum().backward()
    eps = 1e-4
    fixed = _assignments(reference, p)
    owners = np.concatenate([np.argmin(p, axis=0), np.argmax(p, axis=0)])
    interior = np.setdiff1d(np.arange(len(p)), owners)
    for name, array, tensor in (("normals", n, tn), ("positions", p, tp)):
        gradient = tensor.grad.cpu().numpy().astype(np.float64)
        assert np.isfinite(gradient).all() and np.abs(gradient).max() > 0
        direction = np.zeros(array.shape)
        rows = np.arange(len(array)) if name == "normals" else rng.choice(interior, 24, replace=False)
        direction[rows] = rng.normal(size=(len(rows), 3))
        direction = direction.astype(array.dtype)
        losses = []
        for sign in (1, -1):
            moved = (array + sign * eps * direction).astype(array.dtype)
            pp, nn = (moved, n) if name == "positions" else (p, moved)
            candidate = reconstruct(torch.from_numpy(pp).to(cuda), torch.from_numpy(nn).to(cuda), **options)
            for x, y in zip(fixed, _assignments(candidate, pp)):
                np.testing.assert_array_equal(x, y)
            current = extract_mesh(candidate, **mesh_options)
            np.testing.assert_array_equal(current.indices.cpu().numpy(), mesh.indices.cpu().numpy())
            losses.append(float(np.sum(current.vertices.cpu().numpy().astype(np.float64) * seed)))
        print(name, np.sum(gradient * direction.astype(np.float64)), (losses[0] - losses[1]) / (2 * eps), flush=True)
        np.testing.assert_allclose(np.sum(gradient * direction.astype(np.float64)), (losses[0] - losses[1]) / (2 * eps),
                                   rtol=3e-2, atol=2e-3)


def test_empty_mesh_backward(cuda):
    """tests/test_reconstruction_mesh_adjoint.py:47-60: no vertices, no gradient."""
    import torch
    from warp_psr_amd import reconstruct_differentiable
    from warp_psr_amd.meshing import extract_mesh
    p, n = fixture_arrays("sphere", 512, "float")
    tp = torch.from_numpy(p).to(cuda).requires_grad_(True)
    tn = torch.from_numpy(n).to(cuda).requires_grad_(True)
    result = reconstruct_differentiable(tp, tn, depth=4, full_depth=2, iters=2)
    iso = torch.full((1,), 1e6, dtype=torch.float32, device=cuda, requires_grad=True)
    mesh = extract_mesh(result, differentiable=True, iso_value=iso, resolution=9, bounds=_cube_bounds(p), backend="sparse")
    assert not len(mesh.vertices) and mesh.vertices.requires_grad
    mesh.vertices.sum().backward()
    for t in (tp, tn, iso):
        assert t.grad is None or not bool(t.grad.abs().max() > 0)
