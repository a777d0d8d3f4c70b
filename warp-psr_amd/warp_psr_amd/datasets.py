"""Oriented point clouds from local files (reference tools/datasets.py:97-166).

``load_input(path, max_points=None, seed=2026, dataset=None)`` reads a PLY
point cloud (x, y, z, nx, ny, nz vertex properties), a mesh whose vertex
normals come from its triangle winding (the pinned mesh-derived datasets), or
an ASCII ``.npts`` file of six columns, and returns ``(points, normals,
metadata)`` as the reference does.  ``mesh_vertex_normals`` is the reference's
area-weighted winding normal.  The PLY reader is this module's own
(``read_ply``: ascii, binary_little_endian and binary_big_endian, scalar and
list properties); the reference uses the ``plyfile`` package.

Nothing here downloads: ``dataset_path`` only finds a pinned file in a local
directory and checks its SHA-256 (the reference's ``download`` fetches it).
"""
import hashlib
from pathlib import Path

import numpy as np

SOURCE = 'https://github.com/mkazhdan/PoissonRecon/blob/262b0f539d404057d1f36e1adc07fc9388678899/README.md'
BASE = 'https://www.cs.jhu.edu/~misha/Code/PoissonRecon/'
STANFORD = 'https://graphics.stanford.edu/data/3Dscanrep/'
STANFORD_FILES = 'https://graphics.stanford.edu/pub/3Dscanrep/'

# The pinned datasets (tools/datasets.py:17-53): file names, vertex and face
# counts and SHA-256s identify the exact inputs the reference benchmarks.
DATASETS = {
    'horse': {'file': 'horse.npts', 'count': 100000,
              'sha256': 'e6fd6bfccccb1aad72c83144acd7543f672f1a59641071ef008ae6055df053ec',
              'description': 'Curvature-weighted, nonuniform samples of a virtual horse; supplied by the PoissonRecon project'},
    'bunny': {'file': 'bunny.points.ply', 'count': 362271,
              'sha256': '859f3a8fda84bd2707abc5ede664b11cde20fef9d5f9ffde36cc52088a486e15',
              'description': 'Merged Stanford Bunny range scans; normals estimated from scan connectivity by the PoissonRecon project'},
    'eagle': {'file': 'eagle.points.ply', 'count': 796825,
              'sha256': '468e41b3679f775b98fe8cc79132cb99e8d6dfc4dcc1bb684e571a90e326e478',
              'description': 'EPFL Scanning 3D Statues from Photos; oriented eagle scan with color'},
    'torso': {'file': 'torso.points.ply', 'count': 3488432,
              'sha256': 'ad7955bb71fe4a042e125b407beed7eeabb79a2062157b28ea955cb621e66524',
              'archive': 'torso.zip', 'archive_sha256': 'a9ee6b6af62921e2223b5cd47f0e80d3345ec93379a7b9b67131d3c1fcf6c25c',
              'description': 'PoissonRecon torso scan; the separate envelope constraint is outside the supported profile'},
    'drill': {'file': 'drill.ply', 'count': 1961,
              'sha256': 'ba5a1fc9fd9d5ec3a7458d5cc3d95989da555e8b03e13ef6120961f604243ba0',
              'archive': 'drill.tar.gz', 'archive_sha256': '2da2acaba36903a9893c920c11506e1bd01c531bdf738d037bc505b313df5e1b',
              'url': STANFORD_FILES + 'drill.tar.gz', 'member': 'drill/reconstruction/drill_shaft_vrip.ply', 'compression': 'tar',
              'normal_source': 'triangle_winding', 'faces': 3855, 'source_page': STANFORD,
              'description': 'Stanford drill reconstruction: thin, elongated geometry; mesh-derived input'},
    'armadillo': {'file': 'Armadillo.ply', 'count': 172974,
                  'sha256': '483563fca686b66168e1e45339c50e71a5dc8a526a04a545a16b9237d5eaa023',
                  'archive': 'Armadillo.ply.gz',
                  'archive_sha256': '8b9b56cc36e66d54429b1e1e75bd89e833645bfe0dc7c1afd1205877a7356a3f',
                  'url': STANFORD_FILES + 'armadillo/Armadillo.ply.gz', 'compression': 'gzip',
                  'normal_source': 'triangle_winding', 'faces': 345944, 'source_page': STANFORD,
                  'description': 'Stanford armadillo reconstruction; mesh-derived input'},
    'dragon': {'file': 'dragon.ply', 'count': 437645,
               'sha256': 'fea87ff48f2aba22fb53e7b67c3ff3f7b8c2a3b3a0653af62c48bba67c6d5744',
               'archive': 'dragon_recon.tar.gz',
               'archive_sha256': '74ac1d90989c9b1732edee82d57e9ce71452144cf4355f108d8c9c616d28d02f',
               'url': STANFORD_FILES + 'dragon/dragon_recon.tar.gz', 'member': 'dragon_recon/dragon_vrip.ply',
               'compression': 'tar', 'normal_source': 'triangle_winding', 'faces': 871414, 'source_page': STANFORD,
               'description': 'Stanford dragon reconstruction; mesh-derived input'},
    'buddha': {'file': 'buddha.ply', 'count': 543652,
               'sha256': '2283371216d748a08376a3c88698e283cc8f18d10ced348d6d133051bcf217ab',
               'archive': 'happy_recon.tar.gz',
               'archive_sha256': '409cd294efbfd8244e15a382b95a9423f153b7776e736c9b09f19ec9d3c10ed0',
               'url': STANFORD_FILES + 'happy/happy_recon.tar.gz', 'member': 'happy_recon/happy_vrip.ply',
               'compression': 'tar', 'normal_source': 'triangle_winding', 'faces': 1087716, 'source_page': STANFORD,
               'description': 'Stanford Happy Buddha reconstruction; mesh-derived input'},
}


def checksum(path):
    """SHA-256 of a file, read in 1 MiB chunks."""
    h = hashlib.sha256()
    with Path(path).open('rb') as stream:
        for chunk in iter(lambda: stream.read(1 << 20), b''):
            h.update(chunk)
    return h.hexdigest()


def dataset_path(name, directory):
    """The pinned file of dataset `name` in `directory`, checked against its
    SHA-256.  There is no download: a missing file raises FileNotFoundError
    naming where the extracted file is expected."""
    entry = DATASETS[name]
    path = Path(directory) / entry['file']
    if not path.exists():
        raise FileNotFoundError(f"{name}: {path} is absent; this build has no network, so place the extracted "
                                f"{entry['file']} there (source: {entry.get('url', BASE + entry.get('archive', entry['file']))})")
    if checksum(path) != entry['sha256']:
        raise ValueError(f'{name}: point-cloud checksum mismatch')
    return path


# ------------------------------------------------------------------ PLY
_PLY_TYPES = {'char': 'i1', 'int8': 'i1', 'uchar': 'u1', 'uint8': 'u1', 'short': 'i2', 'int16': 'i2',
              'ushort': 'u2', 'uint16': 'u2', 'int': 'i4', 'int32': 'i4', 'uint': 'u4', 'uint32': 'u4',
              'float': 'f4', 'float32': 'f4', 'double': 'f8', 'float64': 'f8'}


def _ply_type(name):
    try:
        return np.dtype(_PLY_TYPES[name])
    except KeyError:
        raise ValueError(f'PLY: unknown property type {name!r}') from None


def _ply_header(stream):
    if stream.readline().rstrip(b'\r\n') != b'ply':
        raise ValueError('PLY: missing "ply" magic line')
    fmt, elements = None, []
    while True:
        line = stream.readline()
        if not line:
            raise ValueError('PLY: header has no end_header')
        words = line.decode('ascii', 'replace').split()
        if not words or words[0] in ('comment', 'obj_info'):
            continue
        if words[0] == 'end_header':
            break
        if words[0] == 'format':
            if len(words) != 3 or words[1] not in ('ascii', 'binary_little_endian', 'binary_big_endian'):
                raise ValueError(f'PLY: unsupported format line {line!r}')
            fmt = words[1]
        elif words[0] == 'element':
            if len(words) != 3:
                raise ValueError(f'PLY: bad element line {line!r}')
            elements.append((words[1], int(words[2]), []))
        elif words[0] == 'property':
            if not elements:
                raise ValueError('PLY: property before any element')
            if words[1] == 'list':
                if len(words) != 5:
                    raise ValueError(f'PLY: bad list property line {line!r}')
                elements[-1][2].append((words[4], (_ply_type(words[2]), _ply_type(words[3]))))
            else:
                if len(words) != 3:
                    raise ValueError(f'PLY: bad property line {line!r}')
                elements[-1][2].append((words[2], _ply_type(words[1])))
        else:
            raise ValueError(f'PLY: unknown header line {line!r}')
    if fmt is None:
        raise ValueError('PLY: header has no format line')
    return fmt, elements


def _ply_ascii(tokens, pos, count, props):
    columns = {name: [] for name, _ in props}
    scalar_only = all(not isinstance(t, tuple) for _, t in props)
    if scalar_only:
        width = len(props)
        if pos + count * width > len(tokens):
            raise ValueError('PLY: ascii body ends early')
        block = np.array(tokens[pos:pos + count * width]).reshape(count, width) if count else np.empty((0, width))
        out = {name: block[:, k].astype(np.float64 if t.kind == 'f' else np.int64).astype(t) for k, (name, t) in enumerate(props)}
        return out, pos + count * width
    for _ in range(count):
        for name, t in props:
            if isinstance(t, tuple):
                n = int(tokens[pos])
                columns[name].append(np.array(tokens[pos + 1:pos + 1 + n], dtype=np.int64).astype(t[1]))
                pos += 1 + n
            else:
                columns[name].append(tokens[pos])
                pos += 1
    if pos > len(tokens):
        raise ValueError('PLY: ascii body ends early')
    out = {}
    for name, t in props:
        if isinstance(t, tuple):
            out[name] = _lists(columns[name], t[1])
        else:
            out[name] = np.array(columns[name]).astype(np.float64 if t.kind == 'f' else np.int64).astype(t)
    return out, pos


def _lists(rows, dtype):
    """Rows of one list property: an (n, L) array when every row has L
    entries, else an object array of 1-D arrays."""
    if rows and all(len(r) == len(rows[0]) for r in rows):
        return np.stack(rows).astype(dtype, copy=False)
    if not rows:
        return np.empty((0, 0), dtype=dtype)
    out = np.empty(len(rows), dtype=object)
    out[:] = rows
    return out


def _ply_binary(body, pos, count, props, order):
    scalar_only = all(not isinstance(t, tuple) for _, t in props)
    if scalar_only:
        dt = np.dtype([(name, t.newbyteorder(order)) for name, t in props])
        if pos + count * dt.itemsize > len(body):
            raise ValueError('PLY: binary body ends early')
        rec = np.frombuffer(body, dtype=dt, count=count, offset=pos)
        return {name: rec[name].astype(t) for name, t in props}, pos + count * dt.itemsize
    # fast path: every row's lists have the first row's lengths
    if count:
        lengths, p = [], pos
        for name, t in props:
            if isinstance(t, tuple):
                if p + t[0].itemsize > len(body):
                    raise ValueError('PLY: binary body ends early')
                n = int(np.frombuffer(body, dtype=t[0].newbyteorder(order), count=1, offset=p)[0])
                lengths.append(n)
                p += t[0].itemsize + n * t[1].itemsize
            else:
                p += t.itemsize
        fields, k = [], 0
        for name, t in props:
            if isinstance(t, tuple):
                fields.append(('__count_' + name, t[0].newbyteorder(order)))
                fields.append((name, t[1].newbyteorder(order), (lengths[k],)))
                k += 1
            else:
                fields.append((name, t.newbyteorder(order)))
        dt = np.dtype(fields)
        if pos + count * dt.itemsize <= len(body):
            rec = np.frombuffer(body, dtype=dt, count=count, offset=pos)
            k, uniform = 0, True
            for name, t in props:
                if isinstance(t, tuple):
                    uniform = uniform and bool(np.all(rec['__count_' + name] == lengths[k]))
                    k += 1
            if uniform:
                out = {}
                for name, t in props:
                    out[name] = rec[name].astype(t[1] if isinstance(t, tuple) else t)
                return out, pos + count * dt.itemsize
    columns = {name: [] for name, _ in props}
    for _ in range(count):
        for name, t in props:
            if isinstance(t, tuple):
                n = int(np.frombuffer(body, dtype=t[0].newbyteorder(order), count=1, offset=pos)[0])
                pos += t[0].itemsize
                columns[name].append(np.frombuffer(body, dtype=t[1].newbyteorder(order), count=n, offset=pos).astype(t[1]))
                pos += n * t[1].itemsize
            else:
                columns[name].append(np.frombuffer(body, dtype=t.newbyteorder(order), count=1, offset=pos)[0])
                pos += t.itemsize
    out = {}
    for name, t in props:
        out[name] = _lists(columns[name], t[1]) if isinstance(t, tuple) else np.array(columns[name], dtype=t)
    return out, pos


def read_ply(path):
    """Every element of a PLY file: {element: {property: values}}, in header
    order.  Scalar properties are 1-D arrays of their declared type; list
    properties are (n, L) arrays when all rows hold L entries, else object
    arrays of 1-D arrays."""
    with Path(path).open('rb') as stream:
        fmt, elements = _ply_header(stream)
        body = stream.read()
    out = {}
    if fmt == 'ascii':
        tokens = body.split()
        pos = 0
        for name, count, props in elements:
            out[name], pos = _ply_ascii(tokens, pos, count, props)
    else:
        order = '<' if fmt == 'binary_little_endian' else '>'
        pos = 0
        for name, count, props in elements:
            out[name], pos = _ply_binary(body, pos, count, props, order)
    return out


# ------------------------------------------------------------------ normals and loading
def mesh_vertex_normals(points, faces):
    """Area-weighted vertex normals that keep the source triangle winding
    (tools/datasets.py:97-124): each face's cross product (p1 - p0) x (p2 - p0)
    is added in float64 to its corner 0 vertices in face order, then corner 1,
    then corner 2, and every nonzero sum is normalized.  Zero sums stay zero;
    no vertex is dropped, welded or reoriented."""
    points = np.asarray(points, dtype=np.float64)
    faces = np.asarray(faces)
    if points.ndim != 2 or points.shape[1] != 3:
        raise ValueError('Mesh positions must be Nx3')
    if faces.ndim != 2 or faces.shape[1] != 3 or not np.issubdtype(faces.dtype, np.integer):
        raise ValueError('Normal derivation requires triangular integer faces')
    if np.any(faces < 0) or np.any(faces >= len(points)) or not np.isfinite(points).all():
        raise ValueError('Mesh has invalid indices or nonfinite positions')
    p0 = points[faces[:, 0]]
    area = np.cross(points[faces[:, 1]] - p0, points[faces[:, 2]] - p0)
    if not np.isfinite(area).all():
        raise ValueError('Triangle normal computation overflowed')
    normals = np.zeros_like(points)
    for corner in range(3):
        # unbuffered, in face order: repeated vertices accumulate one face at a time
        np.add.at(normals, faces[:, corner], area)
    length = np.linalg.norm(normals, axis=1)
    if not np.isfinite(length).all():
        raise ValueError('Vertex normal computation overflowed')
    np.divide(normals, length[:, None], out=normals, where=length[:, None] > 0)
    info = {'algorithm': 'area-weighted triangle-winding vertex normals v1; float64 corner-major face-order sums',
            'triangles': len(faces),
            'degenerate_triangles': int(np.count_nonzero(np.linalg.norm(area, axis=1) == 0)),
            'zero_vertex_normals': int(np.count_nonzero(length == 0)),
            'orientation': 'source face winding; no automatic reorientation',
            'vertex_policy': 'retain every source vertex in source order; no welding, filtering or resampling'}
    return normals, info


def _faces(face_element):
    for key in ('vertex_indices', 'vertex_index'):
        if key in face_element:
            lists = face_element[key]
            if lists.dtype == object:
                raise ValueError('Normal derivation requires triangular integer faces')
            return lists
    raise ValueError('PLY mesh has no vertex_indices face property')


def load_input(path, *, max_points=None, seed=2026, dataset=None):
    """(points, normals, metadata) from a local file (tools/datasets.py:127-166).

    PLY files supply x, y, z, nx, ny, nz vertex properties (other properties
    are reported as ignored), except the pinned mesh-derived datasets, whose
    normals come from `mesh_vertex_normals`; any other file is read as ASCII
    rows of six columns.  With `dataset` the file's SHA-256 and vertex count
    must match the pinned entry.  `max_points` keeps a seeded uniform subset
    in file order: numpy's default_rng(seed).choice(total, max_points,
    replace=False), sorted.  Arrays keep the file's value type (float32 for
    float PLY properties, float64 for ASCII rows and derived normals).
    """
    path = Path(path)
    ignored = []
    derived = None
    if path.suffix.lower() == '.ply':
        ply = read_ply(path)
        if 'vertex' not in ply:
            raise ValueError('PLY input must supply x, y, z, nx, ny, nz')
        vertex = ply['vertex']
        fields = ('x', 'y', 'z', 'nx', 'ny', 'nz')
        if dataset and DATASETS[dataset].get('normal_source') == 'triangle_winding':
            if not all(name in vertex for name in fields[:3]) or 'face' not in ply:
                raise ValueError(f'{dataset}: mesh input must supply x, y, z vertices and faces')
            points = np.column_stack([vertex[name] for name in fields[:3]])
            faces = _faces(ply['face'])
            if len(faces) != DATASETS[dataset]['faces']:
                raise ValueError(f'{dataset}: unexpected triangle count')
            normals, derived = mesh_vertex_normals(points, faces)
            data = np.column_stack([points, normals])
        else:
            if not all(name in vertex for name in fields):
                raise ValueError('PLY input must supply x, y, z, nx, ny, nz')
            data = np.column_stack([vertex[name] for name in fields])
        ignored = [name for name in vertex if name not in fields]
    else:
        data = np.loadtxt(path, ndmin=2)
        if data.shape[1] != 6:
            raise ValueError('ASCII oriented points must have exactly six columns')
    total = len(data)
    if not total:
        raise ValueError('Input point cloud is empty')
    metadata = {'source_file': path.name, 'source_sha256': checksum(path), 'original_count': total,
                'normal_policy': 'Use supplied normals; reconstruction applies its own reference normalization',
                'ignored_properties': ignored, 'selection': 'all rows in original file order'}
    if derived:
        metadata.update(input_kind='reconstructed_mesh_vertices', normal_policy=derived)
    if dataset:
        entry = DATASETS[dataset]
        if metadata['source_sha256'] != entry['sha256']:
            raise ValueError(f'{dataset}: point-cloud checksum mismatch')
        if total != entry['count']:
            raise ValueError(f'{dataset}: unexpected vertex count')
        metadata.update(dataset=dataset, url=entry.get('url', BASE + entry.get('archive', entry['file'])),
                        description=entry['description'], source_readme=entry.get('source_page', SOURCE))
        if 'archive' in entry:
            metadata.update(archive_sha256=entry['archive_sha256'], archive_member=entry.get('member', entry['file']))
    if max_points is not None:
        if max_points < 1:
            raise ValueError('max_points must be positive')
        if max_points < total:
            indices = np.sort(np.random.default_rng(seed).choice(total, max_points, replace=False))
            data = data[indices]
            metadata.update(selection='seeded uniform subset, retaining original file order', selection_seed=seed,
                            selection_indices_sha256=hashlib.sha256(indices.astype('<i8').tobytes()).hexdigest())
    metadata['selected_count'] = len(data)
    return data[:, :3], data[:, 3:], metadata
