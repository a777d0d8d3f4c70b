"""Local point-cloud loaders (reference tools/datasets.py:97-166): PLY in its
three encodings, ASCII oriented points, mesh-derived winding normals, the
seeded subset and the pinned-dataset checks, on files written here."""
import hashlib

import numpy as np
import pytest

from warp_psr_amd import datasets

_TYPES = {'f4': 'float', 'f8': 'double', 'u1': 'uchar', 'i4': 'int', 'u4': 'uint', 'i2': 'short'}


def write_ply(path, fmt, vertex, faces=None, count_type='u1', index_type='i4'):
    """vertex: {name: 1-D array}; faces: list of index lists."""
    names = list(vertex)
    n = len(vertex[names[0]])
    head = ['ply', f'format {fmt} 1.0', 'comment written by test_datasets', f'element vertex {n}']
    head += [f'property {_TYPES[vertex[k].dtype.str[1:]]} {k}' for k in names]
    if faces is not None:
        head += [f'element face {len(faces)}', f'property list {_TYPES[count_type]} {_TYPES[index_type]} vertex_indices']
    head.append('end_header')
    with open(path, 'wb') as f:
        f.write(('\n'.join(head) + '\n').encode('ascii'))
        if fmt == 'ascii':
            for i in range(n):
                f.write((' '.join(repr(vertex[k][i].item()) for k in names) + '\n').encode('ascii'))
            for face in faces or []:
                f.write((' '.join(str(v) for v in [len(face)] + list(face)) + '\n').encode('ascii'))
        else:
            order = '<' if fmt == 'binary_little_endian' else '>'
            rec = np.zeros(n, dtype=[(k, vertex[k].dtype.newbyteorder(order)) for k in names])
            for k in names:
                rec[k] = vertex[k]
            f.write(rec.tobytes())
            for face in faces or []:
                f.write(np.array([len(face)], dtype=np.dtype(count_type).newbyteorder(order)).tobytes())
                f.write(np.array(face, dtype=np.dtype(index_type).newbyteorder(order)).tobytes())


def cloud(n=257, seed=5):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 3)).astype(np.float32)
    nrm = (p / np.linalg.norm(p, axis=1, keepdims=True)).astype(np.float32)
    return p, nrm


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_point_cloud_roundtrip(tmp_path, fmt):
    p, nrm = cloud()
    rng = np.random.default_rng(1)
    vertex = {'x': p[:, 0], 'y': p[:, 1], 'z': p[:, 2], 'red': rng.integers(0, 256, len(p)).astype(np.uint8),
              'nx': nrm[:, 0], 'ny': nrm[:, 1], 'nz': nrm[:, 2], 'quality': rng.normal(size=len(p))}
    path = tmp_path / f'cloud_{fmt}.ply'
    write_ply(path, fmt, vertex)
    points, normals, meta = datasets.load_input(path)
    assert points.dtype == np.float32 and normals.dtype == np.float32
    assert np.array_equal(points, p) and np.array_equal(normals, nrm)
    assert meta['ignored_properties'] == ['red', 'quality']
    assert meta['original_count'] == meta['selected_count'] == len(p)
    assert meta['source_sha256'] == hashlib.sha256(path.read_bytes()).hexdigest()
    assert meta['selection'] == 'all rows in original file order'
    ply = datasets.read_ply(path)
    assert np.array_equal(ply['vertex']['red'], vertex['red'])
    assert np.array_equal(ply['vertex']['quality'], vertex['quality'])


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_face_lists(tmp_path, fmt):
    p, _ = cloud(8)
    vertex = {'x': p[:, 0], 'y': p[:, 1], 'z': p[:, 2]}
    tris = [[0, 1, 2], [2, 3, 4], [7, 6, 5]]
    path = tmp_path / 'tris.ply'
    write_ply(path, fmt, vertex, tris, count_type='u1', index_type='u4')
    faces = datasets.read_ply(path)['face']['vertex_indices']
    assert faces.shape == (3, 3) and faces.dtype == np.uint32 and np.array_equal(faces, tris)
    mixed = [[0, 1, 2], [0, 1, 2, 3], [4, 5, 6]]  # a quad among triangles
    path = tmp_path / 'mixed.ply'
    write_ply(path, fmt, vertex, mixed, count_type='i4', index_type='i4')
    faces = datasets.read_ply(path)['face']['vertex_indices']
    assert faces.dtype == object and [list(f) for f in faces] == mixed


def test_ply_errors(tmp_path):
    path = tmp_path / 'bad.ply'
    path.write_bytes(b'plx\nformat ascii 1.0\nend_header\n')
    with pytest.raises(ValueError, match='magic'):
        datasets.read_ply(path)
    path.write_bytes(b'ply\nformat ascii 1.0\nelement vertex 1\nproperty quad x\nend_header\n1\n')
    with pytest.raises(ValueError, match='unknown property type'):
        datasets.read_ply(path)
    p, _ = cloud(4)
    write_ply(path, 'binary_little_endian', {'x': p[:, 0], 'y': p[:, 1], 'z': p[:, 2]})
    with pytest.raises(ValueError, match='must supply x, y, z, nx, ny, nz'):
        datasets.load_input(path)
    data = path.read_bytes()
    path.write_bytes(data[:-5])
    with pytest.raises(ValueError, match='ends early'):
        datasets.read_ply(path)
    write_ply(path, 'ascii', {k: np.zeros(0, np.float32) for k in ('x', 'y', 'z', 'nx', 'ny', 'nz')})
    with pytest.raises(ValueError, match='empty'):
        datasets.load_input(path)


def test_ascii_oriented_points(tmp_path):
    rng = np.random.default_rng(3)
    data = rng.normal(size=(100, 6))
    path = tmp_path / 'horse.npts'
    np.savetxt(path, data)
    points, normals, meta = datasets.load_input(path)
    assert points.dtype == np.float64
    assert np.array_equal(points, data[:, :3]) and np.array_equal(normals, data[:, 3:])
    assert meta['ignored_properties'] == []
    np.savetxt(path, data[:, :5])
    with pytest.raises(ValueError, match='exactly six columns'):
        datasets.load_input(path)


def test_seeded_subset(tmp_path):
    rng = np.random.default_rng(4)
    data = rng.normal(size=(1000, 6))
    path = tmp_path / 'c.npts'
    np.savetxt(path, data)
    points, normals, meta = datasets.load_input(path, max_points=100)
    idx = np.sort(np.random.default_rng(2026).choice(1000, 100, replace=False))
    assert np.array_equal(points, data[idx, :3]) and np.array_equal(normals, data[idx, 3:])
    assert meta['selection_seed'] == 2026 and meta['selected_count'] == 100 and meta['original_count'] == 1000
    assert meta['selection_indices_sha256'] == hashlib.sha256(idx.astype('<i8').tobytes()).hexdigest()
    p2, _, m2 = datasets.load_input(path, max_points=100, seed=7)
    assert not np.array_equal(p2, points) and m2['selection_seed'] == 7
    p3, _, m3 = datasets.load_input(path, max_points=5000)
    assert np.array_equal(p3, data[:, :3]) and m3['selection'] == 'all rows in original file order'
    with pytest.raises(ValueError, match='max_points must be positive'):
        datasets.load_input(path, max_points=0)


def tetrahedron():
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64)
    faces = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]])
    return v, faces


def test_mesh_vertex_normals_closed_forms():
    # a regular tetrahedron wound outward: by symmetry each vertex normal is the
    # vertex direction from the centroid
    v, faces = tetrahedron()
    n, info = datasets.mesh_vertex_normals(v, faces)
    assert np.allclose(n, v / np.sqrt(3.0), rtol=0, atol=1e-15)
    assert info['triangles'] == 4 and info['degenerate_triangles'] == 0 and info['zero_vertex_normals'] == 0
    # reversed winding flips every normal
    n_rev, _ = datasets.mesh_vertex_normals(v, faces[:, ::-1])
    assert np.array_equal(n_rev, -n)
    # a flat counter-clockwise grid: every normal is +z exactly
    g = np.array([[x, y, 0.0] for y in range(4) for x in range(4)])
    tris = []
    for y in range(3):
        for x in range(3):
            a = 4 * y + x
            tris += [[a, a + 1, a + 5], [a, a + 5, a + 4]]
    n, _ = datasets.mesh_vertex_normals(g, np.array(tris))
    assert np.array_equal(n, np.tile([0.0, 0.0, 1.0], (16, 1)))
    # an isolated vertex keeps a zero normal; a zero-area triangle is counted
    v2 = np.vstack([v, [[5.0, 5.0, 5.0]], [[0.5, 0.5, 0.5]]])
    f2 = np.vstack([faces, [[0, 0, 1]]])
    n, info = datasets.mesh_vertex_normals(v2, f2)
    assert np.array_equal(n[4], [0.0, 0.0, 0.0]) and np.array_equal(n[5], [0.0, 0.0, 0.0])
    assert info['zero_vertex_normals'] == 2 and info['degenerate_triangles'] == 1


def test_mesh_vertex_normals_sum_order():
    """Vertex 0 receives face areas (2^53, 0, 0) as corner 0 of face 0, (1, 0, 0)
    as corner 1 of face 1 and (-2^53, 0, 0) as corner 0 of face 2.  Corner-major
    order (corner 0 of every face, then corner 1, then corner 2) sums
    2^53 - 2^53 + 1 = 1 exactly: the normal is (1, 0, 0).  Face order would give
    (2^53 + 1) - 2^53 = 0 in float64, a zero normal."""
    v = np.array([[0, 0, 0], [0, 2.0**27, 0], [0, 0, 2.0**26], [0, 0, 1], [0, 1, 0]], dtype=np.float64)
    faces = np.array([[0, 1, 2], [3, 0, 4], [0, 2, 1]])
    n, info = datasets.mesh_vertex_normals(v, faces)
    assert np.array_equal(n[0], [1.0, 0.0, 0.0])
    # vertices 1 and 2 receive the two opposite areas only: exactly zero
    assert np.array_equal(n[1:3], np.zeros((2, 3))) and info['zero_vertex_normals'] == 2


def test_mesh_vertex_normals_errors():
    v, faces = tetrahedron()
    with pytest.raises(ValueError, match='Nx3'):
        datasets.mesh_vertex_normals(v[:, :2], faces)
    with pytest.raises(ValueError, match='triangular integer faces'):
        datasets.mesh_vertex_normals(v, faces.astype(np.float64))
    with pytest.raises(ValueError, match='triangular integer faces'):
        datasets.mesh_vertex_normals(v, np.array([[0, 1, 2, 3]]))
    with pytest.raises(ValueError, match='invalid indices'):
        datasets.mesh_vertex_normals(v, np.array([[0, 1, 4]]))
    with pytest.raises(ValueError, match='invalid indices'):
        datasets.mesh_vertex_normals(v, np.array([[0, -1, 2]]))
    with np.errstate(over='ignore', invalid='ignore'), pytest.raises(ValueError, match='overflowed'):
        datasets.mesh_vertex_normals(v * 1e200, faces)


def test_pinned_mesh_dataset(tmp_path, monkeypatch):
    v, faces = tetrahedron()
    vertex = {'x': v[:, 0].astype(np.float32), 'y': v[:, 1].astype(np.float32), 'z': v[:, 2].astype(np.float32),
              'confidence': np.ones(4, np.float32)}
    path = tmp_path / 'tet.ply'
    write_ply(path, 'binary_little_endian', vertex, faces.tolist())
    sha = hashlib.sha256(path.read_bytes()).hexdigest()
    entry = {'file': 'tet.ply', 'count': 4, 'sha256': sha, 'archive': 'tet.tar.gz', 'archive_sha256': '0' * 64,
             'url': 'local', 'member': 'tet/tet.ply', 'compression': 'tar', 'normal_source': 'triangle_winding',
             'faces': 4, 'source_page': 'local', 'description': 'test tetrahedron'}
    monkeypatch.setitem(datasets.DATASETS, 'tet', entry)
    points, normals, meta = datasets.load_input(path, dataset='tet')
    # float32 positions stacked with float64 derived normals: float64 rows
    assert points.dtype == np.float64 and np.array_equal(points, v)
    assert np.array_equal(normals, datasets.mesh_vertex_normals(v.astype(np.float32), faces)[0])
    assert meta['input_kind'] == 'reconstructed_mesh_vertices' and meta['normal_policy']['triangles'] == 4
    assert meta['dataset'] == 'tet' and meta['archive_member'] == 'tet/tet.ply'
    assert meta['ignored_properties'] == ['confidence']
    assert datasets.dataset_path('tet', tmp_path) == path
    monkeypatch.setitem(datasets.DATASETS, 'tet', dict(entry, faces=5))
    with pytest.raises(ValueError, match='unexpected triangle count'):
        datasets.load_input(path, dataset='tet')
    monkeypatch.setitem(datasets.DATASETS, 'tet', dict(entry, sha256='1' * 64))
    with pytest.raises(ValueError, match='checksum mismatch'):
        datasets.load_input(path, dataset='tet')
    with pytest.raises(ValueError, match='checksum mismatch'):
        datasets.dataset_path('tet', tmp_path)
    with pytest.raises(FileNotFoundError, match='no network'):
        datasets.dataset_path('bunny', tmp_path)
