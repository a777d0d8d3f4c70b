"""TEST INFRASTRUCTURE ONLY: CPU restatement of the marching-cubes extraction.

Used by tests/ to check the library's mesh kernels; nothing in the product
imports it.  It follows the reference's grid conventions (src/warp_psr/
meshing.py:29-59 corner order and cell classification, :88-94 corner points,
:126-146 float32 spacing) and restates the library's own case table from the
marching-cubes method: positive corners are those with value >= 0; on each
cube face the crossing edges are joined, an ambiguous face (two diagonal
positive corners) cutting each positive corner off by its own segment; the
closed loops are oriented toward their edges' positive endpoints (Newell
normal of the edge midpoints) and fanned from their first edge.  The
reference's Warp case table is third-party (warp-lang 1.19.0) and absent
here, so triangulations are pinned by invariants (closed, consistently
oriented surfaces; vertices on the zero crossings), not by Warp's output.
"""
import numpy as np

CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (0, 3, 7, 4), (1, 2, 6, 5)]


def _edge(a, b):
    for e, (p, q) in enumerate(EDGES):
        if {p, q} == {a, b}:
            return e
    raise KeyError((a, b))


def case_triangles(code):
    """Triangles (edge triples) of one corner-sign case."""
    pos = [(code >> c) & 1 == 1 for c in range(8)]
    adj = {}
    for q in FACES:
        e = [_edge(q[k], q[(k + 1) % 4]) for k in range(4)]
        cross = [e[k] for k in range(4) if pos[q[k]] != pos[q[(k + 1) % 4]]]
        pairs = []
        if len(cross) == 2:
            pairs = [tuple(cross)]
        elif len(cross) == 4:
            pairs = [(e[3], e[0]), (e[1], e[2])] if pos[q[0]] else [(e[0], e[1]), (e[2], e[3])]
        for a, b in pairs:
            adj.setdefault(a, []).append(b)
            adj.setdefault(b, []).append(a)
    seen, triangles = set(), []
    for start in range(12):
        if start in seen or start not in adj:
            continue
        loop, prev, cur = [], -1, start
        while True:
            loop.append(cur)
            seen.add(cur)
            nxt = adj[cur][0] if adj[cur][0] != prev else adj[cur][1]
            prev, cur = cur, nxt
            if cur == start:
                break
        mids = [(CORNERS[EDGES[e][0]] + CORNERS[EDGES[e][1]]) / 2.0 for e in loop]
        normal = sum(np.cross(mids[i], mids[(i + 1) % len(mids)]) for i in range(len(mids)))
        direction = np.zeros(3)
        for e in loop:
            lo, hi = EDGES[e]
            direction += (CORNERS[lo] - CORNERS[hi]) if pos[lo] else (CORNERS[hi] - CORNERS[lo])
        if float(np.dot(normal, direction)) < 0:
            loop.reverse()
        triangles += [(loop[0], loop[i], loop[i + 1]) for i in range(1, len(loop) - 1)]
    return triangles


def case_table():
    return [case_triangles(code) for code in range(256)]


def grid(shape, lower, upper):
    """float32 spacing (meshing.py:142-143)."""
    lower = np.asarray(lower, np.float32)
    upper = np.asarray(upper, np.float32)
    step = (upper - lower) / np.asarray([n - 1 for n in shape], np.float32)
    return lower, upper, step.astype(np.float32)


def node_points(shape, lower, upper):
    """Model-space grid nodes, x-major (meshing.py:37-41), shape (nx*ny*nz, 3)."""
    lower, upper, step = grid(shape, lower, upper)
    idx = np.indices(shape).reshape(3, -1).T.astype(np.float32)
    p = lower + step * idx
    return np.minimum(np.maximum(p, lower), upper).astype(np.float32)


def extract(values, shape, lower, upper, table=None):
    """Mesh of the zero set of node values (nx, ny, nz) float32: vertices
    sorted by grid-edge key (lower node index * 3 + axis), triangles in
    ascending cell order."""
    table = case_table() if table is None else table
    values = np.asarray(values, np.float32).reshape(shape)
    lower, upper, step = grid(shape, lower, upper)
    nx, ny, nz = shape
    edges = {}
    cells = []
    for x in range(nx - 1):
        for y in range(ny - 1):
            for z in range(nz - 1):
                v = [values[x + c[0], y + c[1], z + c[2]] for c in CORNERS]
                code = sum(1 << c for c in range(8) if v[c] >= 0)
                if code in (0, 255):
                    continue
                cells.append((x, y, z, code))
                for e, (lo, hi) in enumerate(EDGES):
                    if (v[lo] >= 0) == (v[hi] >= 0):
                        continue
                    node = CORNERS[lo] + (x, y, z)
                    axis = int(np.argmax(CORNERS[hi] - CORNERS[lo]))
                    key = (int(node[0]) * ny + int(node[1])) * nz + int(node[2])
                    key = key * 3 + axis
                    va, vb = np.float32(v[lo]), np.float32(v[hi])
                    t = np.float32(va / np.float32(va - vb))
                    g = node.astype(np.float32)
                    g[axis] = np.float32(g[axis] + t)
                    edges[key] = (lower + step * g).astype(np.float32)
    keys = sorted(edges)
    index = {k: i for i, k in enumerate(keys)}
    vertices = np.array([edges[k] for k in keys], np.float32).reshape(-1, 3)
    tris = []
    for x, y, z, code in cells:
        for tri in table[code]:
            out = []
            for e in tri:
                lo, hi = EDGES[e]
                node = CORNERS[lo] + (x, y, z)
                axis = int(np.argmax(CORNERS[hi] - CORNERS[lo]))
                key = ((int(node[0]) * ny + int(node[1])) * nz + int(node[2])) * 3 + axis
                out.append(index[key])
            tris.append(out)
    return vertices, np.array(tris, np.int32).reshape(-1)
