"""Marching-cubes case table and CPU restatement (no GPU).

The reference meshes with Warp's fixed case table (warp-lang 1.19.0, absent
here); the library generates its own from the marching-cubes method.  These
tests pin the library's table against the independent restatement in
oracle/marching_cubes.py and pin the restatement by what any correct table
must satisfy (tests/test_meshing.py:74-89, :207-231 of the reference):
closed, consistently oriented surfaces on shared faces, faces toward
increasing values, the exact-threshold policy, and a sphere's topology and
enclosed volume.
"""
import ctypes
from collections import Counter

import numpy as np
import pytest

from oracle import marching_cubes as mc
from warp_psr_amd import _lib
from warp_psr_amd.meshing import _grid


def library_table():
    counts = (ctypes.c_int8 * 256)()
    edges = (ctypes.c_int8 * (256 * 36))()
    _lib.check(_lib.load().psr_mesh_case_table(counts, edges))
    e = np.frombuffer(edges, dtype=np.int8).reshape(256, 36)
    return [[tuple(int(v) for v in e[code][3 * t:3 * t + 3]) for t in range(counts[code])] for code in range(256)]


def test_library_table_equals_restatement():
    assert library_table() == mc.case_table()


def test_every_case_cuts_exactly_its_crossing_edges_into_closed_patches():
    table = mc.case_table()
    for code in range(256):
        pos = [(code >> c) & 1 for c in range(8)]
        crossing = {e for e, (a, b) in enumerate(mc.EDGES) if pos[a] != pos[b]}
        used = {e for tri in table[code] for e in tri}
        assert used == crossing, code
        # every side is used once per direction; a side used in one direction
        # only is on the patch boundary: it joins two edges of one cube face,
        # and the boundary passes through every crossing edge twice
        directed = Counter()
        for a, b, c in table[code]:
            for s, t in ((a, b), (b, c), (c, a)):
                directed[(s, t)] += 1
        face_edges = [{mc._edge(q[k], q[(k + 1) % 4]) for k in range(4)} for q in mc.FACES]
        degree = Counter()
        for (s, t), n in directed.items():
            assert n == 1, (code, s, t)
            if directed[(t, s)] == 0:
                assert any(s in f and t in f for f in face_edges), (code, s, t)
                degree[s] += 1
                degree[t] += 1
        assert all(degree[e] == 2 for e in crossing), code


def _closed_and_oriented(vertices, indices):
    faces = indices.reshape(-1, 3)
    directed = Counter()
    for a, b, c in faces:
        for s, t in ((a, b), (b, c), (c, a)):
            directed[(s, t)] += 1
    return all(n == 1 and directed[(t, s)] == 1 for (s, t), n in directed.items())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_fields_give_closed_oriented_surfaces(seed):
    rng = np.random.default_rng(seed)
    shape = (7, 8, 9)
    values = rng.normal(size=shape).astype(np.float32)
    values[[0, -1], :, :] = -1
    values[:, [0, -1], :] = -1
    values[:, :, [0, -1]] = -1
    vertices, indices = mc.extract(values, shape, (-1, -1, -1), (1, 1, 1))
    assert len(indices) and _closed_and_oriented(vertices, indices)


def test_sphere_topology_volume_and_outward_faces():
    shape = (21, 22, 23)
    lower, upper = (-0.9, -0.8, -0.7), (0.95, 0.85, 0.75)
    p = mc.node_points(shape, lower, upper)
    r2 = np.float32(0.2941)
    values = (p * p).sum(axis=1).astype(np.float32) - r2
    vertices, indices = mc.extract(values.reshape(shape), shape, lower, upper)
    faces = indices.reshape(-1, 3)
    assert _closed_and_oriented(vertices, indices)
    edges = {tuple(sorted(e)) for f in faces for e in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0]))}
    assert len(vertices) - len(edges) + len(faces) == 2
    v = vertices.astype(np.float64)
    volume = np.einsum("ij,ij->i", v[faces[:, 0]], np.cross(v[faces[:, 1]], v[faces[:, 2]])).sum() / 6
    exact = 4 / 3 * np.pi * float(r2) ** 1.5
    assert volume > 0 and abs(volume - exact) / exact < 0.03  # faces toward increasing values: outward
    radius = np.linalg.norm(v, axis=1)
    assert np.abs(radius - np.sqrt(float(r2))).max() < 0.02


def test_exact_threshold_policy():
    """meshing.py docs / tests/test_meshing.py:218-231: zero counts as
    nonnegative; a zero endpoint yields coincident vertices of distinct edges."""
    shape = (2, 2, 2)
    values = np.zeros(shape, np.float32)
    v, i = mc.extract(values, shape, (0, 0, 0), (1, 1, 1))
    assert len(v) == len(i) == 0
    corners = np.array([0., -1., -1., -1., -1., -1., -1., -1.], np.float32)
    values = np.empty(shape, np.float32)
    for c, (x, y, z) in enumerate(mc.CORNERS):
        values[x, y, z] = corners[c]
    v, i = mc.extract(values, shape, (0, 0, 0), (1, 1, 1))
    assert len(v) == len(i) == 3
    np.testing.assert_array_equal(v, np.zeros((3, 3)))


@pytest.mark.parametrize("resolution", [True, 1, (2, 3), (2, 2, 3.5), (2, 2, True), (3000, 3000, 3000), (2, 2, 2**30)])
def test_bad_grid_rejected(resolution):
    """tests/test_meshing.py:193-197: rejected before any sampling."""
    with pytest.raises(ValueError):
        _grid(resolution, ((0., 0., 0.), (1., 1., 1.)))


def test_bad_bounds_rejected():
    for bounds in [((0, 0, 0), (1, 1)), ((0, 0, 0), (1, 1, 0)), ((0, 0, float("nan")), (1, 1, 1)), "ab"]:
        with pytest.raises(ValueError):
            _grid(3, bounds)
