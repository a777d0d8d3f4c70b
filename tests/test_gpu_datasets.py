"""Clouds read by the local loaders feed the HIP path unchanged: a binary PLY
point cloud and a mesh whose normals come from its winding, both written
here, reconstruct bit for bit like the oracle on the same arrays."""
import numpy as np
import pytest

import oracle
from parity import compare_stages
from test_datasets import write_ply

pytestmark = pytest.mark.gpu


def _check(cuda, points, normals, params):
    import torch
    from warp_psr_amd import reconstruct
    gpu = reconstruct(torch.from_numpy(points).to(cuda), torch.from_numpy(normals).to(cuda), **params)
    ref = oracle.reconstruct(points, normals, **params)
    mismatches = compare_stages(gpu.diagnostics, ref, ref.params)
    assert not mismatches, "\n".join(f"{s}[{lv}]: {m}" for s, lv, m in mismatches[:20])
    assert gpu.iso_value == ref.iso_value


def test_ply_point_cloud_reconstructs_like_the_oracle(cuda, tmp_path):
    from warp_psr_amd import datasets, synthetic
    p, n = synthetic.fixture_input("sphere", 3000, precision="float")
    path = tmp_path / "sphere.ply"
    write_ply(path, "binary_little_endian", {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2],
                                             "nx": n[:, 0], "ny": n[:, 1], "nz": n[:, 2]})
    points, normals, meta = datasets.load_input(path, max_points=2000)
    assert points.dtype == np.float32 and meta["selected_count"] == 2000
    print("ply cloud: 2000 of 3000 sphere points, depth 5", flush=True)
    _check(cuda, np.ascontiguousarray(points), np.ascontiguousarray(normals), dict(depth=5, full_depth=2))


def test_mesh_derived_normals_reconstruct_like_the_oracle(cuda, tmp_path, monkeypatch):
    from warp_psr_amd import datasets
    # a UV sphere wound outward
    rings, segs = 24, 48
    v = [[0.0, 0.0, 1.0]]
    for i in range(1, rings):
        t = np.pi * i / rings
        for j in range(segs):
            f = 2 * np.pi * j / segs
            v.append([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)])
    v.append([0.0, 0.0, -1.0])
    v = np.array(v, dtype=np.float32)
    faces = [[0, 1 + j, 1 + (j + 1) % segs] for j in range(segs)]
    for i in range(rings - 2):
        for j in range(segs):
            a, b = 1 + i * segs + j, 1 + i * segs + (j + 1) % segs
            faces += [[a, a + segs, b], [b, a + segs, b + segs]]
    last = len(v) - 1
    faces += [[1 + (rings - 2) * segs + j, last, 1 + (rings - 2) * segs + (j + 1) % segs] for j in range(segs)]
    path = tmp_path / "uv.ply"
    write_ply(path, "binary_big_endian", {"x": v[:, 0], "y": v[:, 1], "z": v[:, 2]}, faces)
    entry = {"file": "uv.ply", "count": len(v), "sha256": datasets.checksum(path), "normal_source": "triangle_winding",
             "faces": len(faces), "description": "test sphere mesh"}
    monkeypatch.setitem(datasets.DATASETS, "uv", entry)
    points, normals, meta = datasets.load_input(path, dataset="uv")
    assert meta["normal_policy"]["zero_vertex_normals"] == 0
    assert np.all(np.sum(normals * points, axis=1) > 0.99)  # outward, near-radial
    print(f"mesh-derived cloud: {len(v)} vertices, depth 5, float64", flush=True)
    _check(cuda, np.ascontiguousarray(points), np.ascontiguousarray(normals), dict(depth=5, full_depth=2))
