"""Fixture interchange: the text exports of the reference's PoissonRecon oracle
(reference/README.md:27-74), read, written and compared.

Every file is a list of rows ``tag v1 v2 ...`` with 17 significant digits
(reference/oracle.cpp:54-81, the export includes under reference/ and the
stage export of tools/build_reference.py:36-67).  ``read_rows`` reads any of
them the way tools/make_fixture.py:23-26 does: one float64 array per tag.

Writers produce the same files from a reconstruction of this library (or any
object with the same stage arrays), so a fixture directory made here is read
by the same tools as one made by PoissonRecon:

* ``oracle.txt`` / ``oracle_instrumented.txt``: ``iso``, ``transform`` (the
  cube-to-model matrix, column first), ``q`` (value and model-space gradient
  per query) and, instrumented, one ``node`` row per node of the complete tree
  (local depth and coordinates, flags, ghost mark, an index and the
  coefficient).  The header rows say the run was serial.
* ``topology.txt``: global and local identities, flags, ghost mark, parent and
  child count of every node.
* ``inputs.txt``: per input its status and accepted leaf.
* ``stages.txt``: resolved settings, density scale, aggregated samples,
  normal records, density coefficients and the unscreened FEM load.
* ``preprocess.txt``: aggregated samples in first-input order and density
  coefficients.
* ``splats.txt``: per normal splat the normalised position, raw depth,
  density weight, top level, blend fraction and weighted normal.
* ``operators.txt``: per level the right-hand side, the system matrix and the
  restriction, indexed by position in the concatenated active basis.  The
  ``point`` rows (interpolation points) are not written.

Node indices are incidental in both directions (reference/README.md:60-61):
the comparators key everything by ``(depth, x, y, z)``.  Comparators follow
the reference's own checks (tools/check_tree.py, tools/check_samples.py,
tools/check_tree_metadata.py, tests/test_preprocessing.py:43-65,
tools/check_reconstruction.py:31-41) and raise ``AssertionError`` on a
mismatch; ``exact=True`` asks for equal bits where the reference asks for a
tolerance.
"""
import json
from pathlib import Path

import numpy as np

FILES = ("oracle.txt", "oracle_instrumented.txt", "topology.txt", "inputs.txt", "stages.txt", "preprocess.txt",
         "splats.txt", "operators.txt")
DEPTH_OFFSET = 2


# ------------------------------------------------------------------ reading

def read_rows(path):
    """{tag: float64 array of the values of its rows} (tools/make_fixture.py:23-26)."""
    rows = [line.split() for line in Path(path).read_text().splitlines() if line.strip()]
    out = {}
    for tag in dict.fromkeys(r[0] for r in rows):
        out[tag] = np.array([[float(v) for v in r[1:]] for r in rows if r[0] == tag], dtype=np.float64)
    return out


def read_fixture(directory):
    """Every known file present in a fixture directory: ``manifest`` (dict),
    ``input`` ((n, 6) positions and normals), ``queries`` ((m, 3)) and one
    ``read_rows`` dict per text export, keyed by file name."""
    directory = Path(directory)
    out = {}
    if (directory / "manifest.json").exists():
        out["manifest"] = json.loads((directory / "manifest.json").read_text())
    for name, key in (("input.pts", "input"), ("queries.txt", "queries")):
        if (directory / name).exists():
            out[key] = np.loadtxt(directory / name, ndmin=2)
    for name in FILES:
        if (directory / name).exists():
            out[name] = read_rows(directory / name)
    return out


# ------------------------------------------------------------------ writing

def _fmt(v):
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return "%.17g" % float(v)


def _write(path, rows):
    with open(path, "w") as f:
        for tag, values in rows:
            f.write(" ".join([tag] + [_fmt(v) for v in values]) + "\n")


def _basis(k):
    k = np.asarray(k, dtype=np.int64)
    return np.stack([k >> 48, (k >> 32) & 0xFFFF, (k >> 16) & 0xFFFF, k & 0xFFFF], axis=-1)


def _tree(k):
    k = np.asarray(k, dtype=np.int64)
    return np.stack([((k >> s) & 1048575) - 524288 for s in (40, 20, 0)], axis=-1)


def _basis_key(d, x, y, z):
    return (int(d) << 48) | (int(x) << 32) | (int(y) << 16) | int(z)


def _tree_key(x, y, z):
    return ((int(x) + 524288) << 40) | ((int(y) + 524288) << 20) | (int(z) + 524288)


class Stages:
    """Stage arrays of one reconstruction: ``array(name, level=-1)`` and the
    resolved ``params``.  ``Stages.of(result)`` accepts a ``Reconstruction``
    (its diagnostics), a differentiable reconstruction (its forward result) or
    any object with ``array`` and ``params``."""

    def __init__(self, array, params):
        self.array = array
        self.params = dict(params)

    @classmethod
    def of(cls, result):
        if isinstance(result, cls):
            return result
        result = getattr(result, "_primal", result)  # a differentiable reconstruction's forward result
        diagnostics = getattr(result, "diagnostics", None)
        if diagnostics is not None:
            return cls(diagnostics.array, diagnostics.params)
        return cls(result.array, result.params)

    def basis_offsets(self):
        counts = [len(self.array("basis_keys", d)) for d in range(self.params["depth"] + 1)]
        return np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)


def _levels(metadata):
    """Complete-tree levels as (local_depth, keys, flags, ghost, parents,
    children) numpy tuples, from a TreeMetadata or such tuples."""
    levels = getattr(metadata, "levels", metadata)
    out = []
    for lv in levels:
        if hasattr(lv, "keys") and hasattr(lv, "local_depth"):
            arrays = [np.asarray(getattr(lv, a).cpu().numpy() if hasattr(getattr(lv, a), "cpu") else getattr(lv, a))
                      for a in ("keys", "flags", "ghost", "parents", "children")]
            out.append((int(lv.local_depth), *arrays))
        else:
            out.append((int(lv[0]), *[np.asarray(a) for a in lv[1:]]))
    return out


def node_rows(stages, metadata):
    """``node`` rows (reference/oracle.cpp:111-115) of the complete tree: the
    index is the node's position in the concatenated active basis (-1 for
    other nodes) and the coefficient is its solution value (0 otherwise)."""
    stages = Stages.of(stages)
    keys = stages.array("keys")
    coefficients = stages.array("coefficients")
    position = {int(k): i for i, k in enumerate(keys)}
    rows = []
    for d, tkeys, flags, ghost, _, _ in _levels(metadata):
        g = _tree(tkeys)
        local = g - (1 << (d + 1)) if d >= 0 else np.full_like(g, -1)
        for i in range(len(tkeys)):
            x, y, z = (int(v) for v in local[i])
            idx = position.get(_basis_key(d, x, y, z), -1) if d >= 0 and min(x, y, z) >= 0 else -1
            coef = coefficients[idx] if idx >= 0 else 0.0
            rows.append(("node", [d, x, y, z, int(flags[i]), int(ghost[i]), idx, coef]))
    return rows


def write_oracle(path, iso, cube_to_model, values=(), gradients=(), nodes=()):
    """``oracle.txt`` (reference/oracle.cpp:54-81); pass ``node_rows`` for the
    instrumented variant.  ``cube_to_model`` is the 4x4 matrix (model =
    M @ [cube, 1]); the file holds it column first."""
    m = np.asarray(cube_to_model, dtype=np.float64).reshape(4, 4)
    rows = [("parallel_async", [0]), ("threads", [1]), ("schedule", [0]), ("chunk_size", [0]),
            ("openmp_effective_threads", [0]), ("iso", [float(iso)]), ("transform", list(m.T.ravel()))]
    g = np.asarray(gradients, dtype=np.float64).reshape(-1, 3)
    for v, gr in zip(np.asarray(values, dtype=np.float64).ravel(), g):
        rows.append(("q", [v, *gr]))
    rows += list(nodes)
    _write(path, rows)


def write_topology(path, metadata):
    """``topology.txt`` (reference/oracle.cpp:116-126): global depth and
    coordinates, local depth and coordinates (-1 above the root), flags,
    ghost mark, parent global depth and coordinates (-1 for the top node) and
    child count."""
    levels = _levels(metadata)
    coords = [_tree(lv[1]) for lv in levels]
    rows = []
    for g, (d, _, flags, ghost, parents, children) in enumerate(levels):
        c = coords[g]
        local = c - (1 << (d + 1)) if d >= 0 else np.full_like(c, -1)
        for i in range(len(c)):
            if g:
                parent = [g - 1, *(int(v) for v in coords[g - 1][parents[i]])]
            else:
                parent = [-1, -1, -1, -1]
            count = 8 if (children[i] >= 0).any() else 0
            rows.append(("topology", [g, *(int(v) for v in c[i]), d, *(int(v) for v in local[i]), int(flags[i]),
                                      int(ghost[i]), *parent, count]))
    _write(path, rows)


def write_inputs(path, stages):
    """``inputs.txt`` (reference/input_export.inc): index, status, accepted leaf
    depth and coordinates; rejected inputs have depth -1 and zero
    coordinates."""
    stages = Stages.of(stages)
    status = stages.array("input_status")
    groups = stages.array("input_group_indices")
    leaf = _basis(stages.array("sample_keys"))
    rows = []
    for i in range(len(status)):
        if status[i] == 0:
            d, x, y, z = (int(v) for v in leaf[groups[i]])
        else:
            d, x, y, z = -1, 0, 0, 0
        rows.append(("input", [i, int(status[i]), d, x, y, z]))
    _write(path, rows)


def _first_input_order(stages):
    return np.argsort(stages.array("first_indices"), kind="stable")


def write_stages(path, stages):
    """``stages.txt`` (tools/build_reference.py:36-67)."""
    stages = Stages.of(stages)
    p = stages.params
    rows = [("effective", [int(p["depth"]), int(p["full_depth"]), int(p["base_depth"]), int(p["solve_depth"]),
                           int(p["kernel_depth"]), float(p["scale"]), float(p["samples_per_node"]),
                           float(p["point_weight"]), int(bool(p["exact_interpolation"])), int(p["iters"]),
                           float(p["cg_accuracy"])]),
            ("density_scale", list(stages.array("normal_average")[:2]))]
    order = _first_input_order(stages)
    leaf = _basis(stages.array("sample_keys"))
    sums = stages.array("position_sums").reshape(-1, 3)
    weights = stages.array("weights")
    for i in order:
        rows.append(("sample", [*(int(v) for v in leaf[i]), *sums[i], weights[i]]))
    nk = _basis(stages.array("normal_keys"))
    nv = stages.array("normal_values").reshape(-1, 3)
    for i in range(len(nk)):
        rows.append(("normal", [*(int(v) for v in nk[i]), *nv[i]]))
    for d in range(int(p["kernel_depth"]) + 1):
        dk = _basis(stages.array("density_keys", d))
        dv = stages.array("density_values", d)
        for i in range(len(dk)):
            rows.append(("density", [*(int(v) for v in dk[i]), dv[i]]))
    for d in range(int(p["depth"]) + 1):
        bk = _basis(stages.array("basis_keys", d))
        load = stages.array("normal_load", d)
        for i in range(len(load)):
            rows.append(("load", [*(int(v) for v in bk[i]), load[i]]))
    _write(path, rows)


def write_preprocess(path, stages):
    """``preprocess.txt`` (reference/preprocess_export.inc): aggregated samples
    in first-input order (index, leaf, position sums, weight, normal sums),
    then density coefficients."""
    stages = Stages.of(stages)
    order = _first_input_order(stages)
    leaf = _basis(stages.array("sample_keys"))
    sums = stages.array("position_sums").reshape(-1, 3)
    nsums = stages.array("normal_sums").reshape(-1, 3)
    weights = stages.array("weights")
    rows = []
    for row, i in enumerate(order):
        rows.append(("sample", [row, *(int(v) for v in leaf[i]), *sums[i], weights[i], *nsums[i]]))
    for d in range(int(stages.params["kernel_depth"]) + 1):
        dk = _basis(stages.array("density_keys", d))
        dv = stages.array("density_values", d)
        for i in range(len(dk)):
            rows.append(("density", [*(int(v) for v in dk[i]), dv[i]]))
    _write(path, rows)


def write_splats(path, stages):
    """``splats.txt`` (reference/splat_export.inc): per aggregated sample with
    a nonzero normal, in first-input order, the normalised position, raw
    depth, density weight, top level and blend fraction of
    src/warp_psr/normals.py:63-77, and the unit normal times the weight."""
    stages = Stages.of(stages)
    p = stages.params
    dtype = stages.array("weights").dtype.type
    order = _first_input_order(stages)
    sums = stages.array("position_sums").reshape(-1, 3)
    nsums = stages.array("normal_sums").reshape(-1, 3)
    weights = stages.array("weights")
    dw = stages.array("depths_and_weights").reshape(-1, 2)
    lo, hi = int(p["base_depth"]), int(p["depth"])
    rows = []
    for i in order:
        w = weights[i]
        normal = (nsums[i] / w).astype(dtype)
        length = np.sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2])
        if not length > 0:
            continue
        v = (normal / length * w).astype(dtype)
        position = (sums[i] / w).astype(dtype)
        depth = min(max(dw[i][0], dtype(lo)), dtype(hi))
        top = int(np.ceil(depth))
        dx = dtype(1.0) - (dtype(top) - depth)
        if top <= lo or top > hi:
            top, dx = min(max(top, lo), hi), dtype(1.0)
        rows.append(("splat", [*position, dw[i][0], dw[i][1], top, dx, *v]))
    _write(path, rows)


def write_operators(path, stages):
    """``operators.txt`` (reference/operator_export.inc) without the ``point``
    rows: per level up to the solve depth, ``level d begin rows``; from the
    base depth on (the levels this solve assembles; the coarser ones are
    Galerkin products inside the base multigrid) the right-hand side and the
    system matrix; above depth 0 the restriction to the coarser level (the
    transposed prolongation)."""
    stages = Stages.of(stages)
    begin = stages.basis_offsets()
    rows = []
    for d in range(int(stages.params["solve_depth"]) + 1):
        b = int(begin[d])
        n = int(begin[d + 1] - begin[d])
        rows.append(("level", [d, b, n]))
        rhs = stages.array("rhs", d)
        off, col, val = (stages.array("matrix_" + part, d) for part in ("offsets", "columns", "values"))
        for i in range(len(rhs)):
            rows.append(("rhs", [d, b + i, rhs[i]]))
            for e in range(off[i], off[i + 1]):
                rows.append(("A", [d, b + i, b + int(col[e]), val[e]]))
        if d:
            cb = int(begin[d - 1])
            off, col, val = (stages.array("prolong_" + part, d) for part in ("offsets", "columns", "values"))
            fine = np.repeat(np.arange(n), np.diff(off))
            for e in np.lexsort((fine, col)):
                rows.append(("R", [d, cb + int(col[e]), b + int(fine[e]), val[e]]))
    _write(path, rows)


def write_fixture(directory, reconstruction, queries=None, metadata=None):
    """Writes every export of a reconstruction of this library into
    ``directory``: oracle.txt (with ``queries``, model space), and, when
    ``metadata`` (``build_tree_metadata``) is given, oracle_instrumented.txt
    and topology.txt; inputs, stages, preprocess, splats and operators
    always."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    stages = Stages.of(reconstruction)
    values, gradients = (), ()
    if queries is not None:
        v, g, _ = reconstruction.evaluate(queries)
        values = v.detach().cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)
        gradients = g.detach().cpu().numpy() if hasattr(g, "cpu") else np.asarray(g)
    iso = float(stages.array("iso")[0])
    ctm = stages.array("cube_to_model")
    write_oracle(directory / "oracle.txt", iso, ctm, values, gradients)
    if metadata is not None:
        write_oracle(directory / "oracle_instrumented.txt", iso, ctm, values, gradients, node_rows(stages, metadata))
        write_topology(directory / "topology.txt", metadata)
    write_inputs(directory / "inputs.txt", stages)
    write_stages(directory / "stages.txt", stages)
    write_preprocess(directory / "preprocess.txt", stages)
    write_splats(directory / "splats.txt", stages)
    write_operators(directory / "operators.txt", stages)


# ------------------------------------------------------------------ comparing

def _close(actual, expected, exact, atol, rtol, message):
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    if exact:
        np.testing.assert_array_equal(actual, expected, err_msg=message)
    else:
        np.testing.assert_allclose(actual, expected, atol=atol, rtol=rtol, err_msg=message)


def _tolerances(stages):
    double = stages.array("weights").dtype == np.float64
    return (1e-10, 1e-9) if double else (1e-6, 1e-5)


def compare_tree(stages, nodes):
    """tools/check_tree.py:6-23: per depth, the non-ghost nodes are the
    support keys and the nodes with flag 2 the active basis keys."""
    stages = Stages.of(stages)
    nodes = np.asarray(nodes)
    for d in range(int(stages.params["depth"]) + 1):
        sel = nodes[(nodes[:, 0] == d) & (nodes[:, 5] == 0), 1:4].astype(np.int64)
        expected = np.sort(np.array([_tree_key(*v) for v in sel], dtype=np.int64))
        np.testing.assert_array_equal(stages.array("tree_keys", d), expected, err_msg=f"support keys differ at depth {d}")
        sel = nodes[(nodes[:, 0] == d) & ((nodes[:, 4].astype(np.int64) & 2) != 0), 1:4].astype(np.int64)
        expected = np.sort(np.array([_basis_key(d, *v) for v in sel], dtype=np.int64))
        np.testing.assert_array_equal(stages.array("basis_keys", d), expected, err_msg=f"active keys differ at depth {d}")


def compare_inputs(stages, inputs):
    """tools/check_samples.py:6-27: statuses, group membership, stable input
    order, group offsets, weights and first indices."""
    stages = Stages.of(stages)
    ref = np.asarray(inputs, dtype=np.int64)
    np.testing.assert_array_equal(ref[:, 0], np.arange(len(ref)))
    status = stages.array("input_status")
    np.testing.assert_array_equal(status, ref[:, 1])
    accepted = status == 0
    groups = stages.array("input_group_indices")
    np.testing.assert_array_equal(groups[~accepted], -1)
    expected = (ref[accepted, 2] << 48) | (ref[accepted, 3] << 32) | (ref[accepted, 4] << 16) | ref[accepted, 5]
    keys = stages.array("sample_keys")
    np.testing.assert_array_equal(keys[groups[accepted]], expected)
    per_input = np.full(len(ref), np.iinfo(np.int64).max, dtype=np.int64)
    per_input[accepted] = expected
    order = np.argsort(per_input, kind="stable")
    np.testing.assert_array_equal(stages.array("input_order"), order)
    counts = np.bincount(groups[accepted], minlength=len(keys))
    offsets = np.concatenate([[0], np.cumsum(counts)])
    np.testing.assert_array_equal(stages.array("group_offsets"), offsets)
    np.testing.assert_array_equal(stages.array("weights"), counts)
    np.testing.assert_array_equal(stages.array("first_indices"), order[offsets[:-1]])


def compare_topology(metadata, topology):
    """tools/check_tree_metadata.py:9-42: global identities, local depths,
    flags, ghost marks, parents and child counts of every level."""
    ref = np.asarray(topology, dtype=np.int64)
    assert len(np.unique(ref[:, :4], axis=0)) == len(ref), "duplicate global identities"
    np.testing.assert_array_equal(ref[:, 0] - ref[:, 4], DEPTH_OFFSET)
    levels = _levels(metadata)
    xyz = [_tree(lv[1]) for lv in levels]
    total = 0
    for g, (d, _, flags, ghost, parents, children) in enumerate(levels):
        rows = ref[ref[:, 0] == g]
        rows = rows[np.lexsort(rows[:, 1:4].T[::-1])]
        total += len(rows)
        np.testing.assert_array_equal(xyz[g], rows[:, 1:4])
        np.testing.assert_array_equal(rows[:, 4], d)
        np.testing.assert_array_equal(flags, rows[:, 8])
        np.testing.assert_array_equal(ghost, rows[:, 9])
        if g:
            np.testing.assert_array_equal(rows[:, 10], g - 1)
            np.testing.assert_array_equal(xyz[g - 1][parents], rows[:, 11:14])
        else:
            np.testing.assert_array_equal(parents, -1)
            np.testing.assert_array_equal(rows[:, 10:14], -1)
        np.testing.assert_array_equal((children >= 0).sum(axis=1), rows[:, 14])
    assert total == len(ref), "topology rows beyond the retained levels"


def _sorted_rows(rows, width=4):
    return rows[np.lexsort(rows[:, :width].T[::-1])]


def compare_preprocess(stages, preprocess, stage_rows, splats=None, exact=False):
    """tests/test_preprocessing.py:43-65: aggregated sums in first-input
    order, density coefficients per depth, normal records, the density scale
    and the splat depths and weights."""
    stages = Stages.of(stages)
    atol, rtol = _tolerances(stages)
    order = _first_input_order(stages)
    sample = preprocess["sample"]
    _close(stages.array("position_sums").reshape(-1, 3)[order], sample[:, 5:8], exact, atol, rtol, "position sums")
    _close(stages.array("normal_sums").reshape(-1, 3)[order], sample[:, 9:12], exact, atol, rtol, "normal sums")
    for d in range(int(stages.params["kernel_depth"]) + 1):
        rows = preprocess["density"]
        expected = _sorted_rows(rows[rows[:, 0] == d])
        keys = np.array([_basis_key(*v) for v in expected[:, :4].astype(np.int64)], dtype=np.int64)
        np.testing.assert_array_equal(stages.array("density_keys", d), keys, err_msg=f"density keys at depth {d}")
        _close(stages.array("density_values", d), expected[:, 4], exact, atol, rtol, f"density values at depth {d}")
    expected = _sorted_rows(stage_rows["normal"])
    keys = np.array([_basis_key(*v) for v in expected[:, :4].astype(np.int64)], dtype=np.int64)
    np.testing.assert_array_equal(stages.array("normal_keys"), keys, err_msg="normal record keys")
    _close(stages.array("normal_values").reshape(-1, 3), expected[:, 4:7], exact, atol, rtol, "normal record values")
    _close(stages.array("normal_average")[:2], stage_rows["density_scale"][0], exact, atol, rtol, "density scale")
    if splats is not None:
        dw = stages.array("depths_and_weights").reshape(-1, 2)[order]
        _close(dw, splats["splat"][:, 3:5], exact, atol, rtol, "splat depths and weights")


def compare_loads(stages, stage_rows, exact=False):
    """The unscreened FEM load of every active node from the base depth on
    against the ``load`` rows of stages.txt, keyed by (depth, x, y, z)."""
    stages = Stages.of(stages)
    atol, rtol = _tolerances(stages)
    expected = {_basis_key(*r[:4].astype(np.int64)): r[4] for r in stage_rows["load"]}
    for d in range(int(stages.params["base_depth"]), int(stages.params["depth"]) + 1):
        keys = stages.array("basis_keys", d)
        rows = {k for k in expected if k >> 48 == d}
        assert rows == {int(k) for k in keys}, f"load rows differ at depth {d}"
        want = np.array([expected[int(k)] for k in keys])
        _close(stages.array("normal_load", d), want, exact, atol, rtol, f"loads at depth {d}")


def compare_operators(stages, operators, nodes, exact=False):
    """Right-hand sides, system matrices and restrictions of operators.txt
    against this run's, with the file's node indices mapped to (depth, x, y,
    z) through the ``node`` rows of oracle_instrumented.txt (the indices
    themselves are incidental)."""
    stages = Stages.of(stages)
    atol, rtol = _tolerances(stages)
    identity = {int(r[6]): _basis_key(*r[:4].astype(np.int64)) for r in np.asarray(nodes) if r[6] >= 0 and r[0] >= 0}
    for d in range(int(stages.params["solve_depth"]) + 1):
        keys = stages.array("basis_keys", d)
        if d < int(stages.params["base_depth"]):
            if d and "R" in operators:
                _compare_restriction(stages, operators, identity, d, keys, exact, atol, rtol)
            continue
        rhs = dict((identity[int(r[1])], r[2]) for r in operators["rhs"] if r[0] == d)
        assert sorted(rhs) == sorted(int(k) for k in keys), f"right-hand side rows differ at depth {d}"
        _close(stages.array("rhs", d), [rhs[int(k)] for k in keys], exact, atol, rtol, f"right-hand side at depth {d}")
        off, col, val = (stages.array("matrix_" + part, d) for part in ("offsets", "columns", "values"))
        mine = {(int(keys[i]), int(keys[col[e]])): val[e] for i in range(len(keys)) for e in range(off[i], off[i + 1])}
        theirs = {(identity[int(r[1])], identity[int(r[2])]): r[3] for r in operators["A"] if r[0] == d}
        assert sorted(mine) == sorted(theirs), f"matrix pattern differs at depth {d}"
        order = sorted(mine)
        _close([mine[k] for k in order], [theirs[k] for k in order], exact, atol, rtol, f"matrix at depth {d}")
        if d and "R" in operators:
            _compare_restriction(stages, operators, identity, d, keys, exact, atol, rtol)


def _compare_restriction(stages, operators, identity, d, keys, exact, atol, rtol):
    coarse = stages.array("basis_keys", d - 1)
    off, col, val = (stages.array("prolong_" + part, d) for part in ("offsets", "columns", "values"))
    mine = {(int(coarse[col[e]]), int(keys[i])): val[e] for i in range(len(keys)) for e in range(off[i], off[i + 1])}
    theirs = {(identity[int(r[1])], identity[int(r[2])]): r[3] for r in operators["R"] if r[0] == d}
    assert sorted(mine) == sorted(theirs), f"restriction pattern differs at depth {d}"
    order = sorted(mine)
    _close([mine[k] for k in order], [theirs[k] for k in order], exact, atol, rtol, f"restriction at depth {d}")


def compare_field(values, gradients, iso, oracle_rows, double, exact=False):
    """tools/check_reconstruction.py:31-41 without the cell-width gradient
    scaling: query values, centered values, gradients and the isovalue."""
    atol, rtol = (1e-8, 1e-6) if double else (1e-5, 1e-4)
    q = oracle_rows["q"]
    reference_iso = oracle_rows["iso"][0, 0]
    values = np.asarray(values, dtype=np.float64)
    _close(values, q[:, 0], exact, atol, rtol, "field values")
    _close(values - iso, q[:, 0] - reference_iso, exact, atol, rtol, "centered values")
    _close(np.asarray(gradients, dtype=np.float64).reshape(-1, 3), q[:, 1:4], exact, atol, rtol, "gradients")
    _close([iso], [reference_iso], exact, atol, rtol, "isovalue")
