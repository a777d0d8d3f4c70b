"""Tree metadata (tree_metadata.py:95-153).

CPU: the restatement in oracle/tree_metadata.py, fed the oracle's trees, is
pinned by the reference's published complete-tree counts (nodes and ghost
nodes per level of the four topology fixtures of docs/results/tree-metadata.json,
where warp-psr matched PoissonRecon's instrumented topology exactly) and by what
the metadata must satisfy (tests/test_tree_metadata.py:56-73): every level is
made of whole broods and contains the active support nodes; the non-ghost nodes
of a depth are exactly its active nodes; parent and child links are mutually
consistent; the padding root holds the eight level -1 nodes; flags follow their
definitions.  GPU: the library's metadata equals the restatement fed the same
reconstruction's arrays, level by level.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import tree_metadata as tm
from warp_psr_amd import synthetic


def _inputs(r, depth):
    offsets = r.array("normal_offsets")
    nk = r.array("normal_keys")
    return ([r.array("tree_keys", d) for d in range(depth + 1)], r.array("sample_keys"),
            [r.array("density_keys", d) if d <= r.params["kernel_depth"] else np.zeros(0, np.int64) for d in range(depth + 1)],
            [nk[offsets[d]:offsets[d + 1]] for d in range(depth + 1)])


def _coords(keys):
    return np.stack([((keys >> s) & 1048575) - tm.OFF for s in (40, 20, 0)], axis=1)


def _check_sources(levels, depth, tree_keys, sample_keys, density_keys, normal_keys):
    """Every brood of a record of depth d (samples at the finest depth,
    density splats, normal records: tree_metadata.py:14-26) and every parent
    brood of level d + 1 is retained at level d, and every retained brood has
    one of those reasons or holds an active node."""
    for d in range(depth + 1):
        o = 1 << (d + 1)
        local = _coords(levels[d + 2][1]) - o
        have = {tuple(v) for v in local // 2}
        reasons = {tuple(v) for v in _coords(np.asarray(tree_keys[d], np.int64)) // 2}
        recs = [np.asarray(density_keys[d]), np.asarray(normal_keys[d])] + ([np.asarray(sample_keys)] if d == depth else [])
        for r in recs:
            r = r[(r >> 48) == d]
            b = {(int(x) // 2, int(y) // 2, int(z) // 2) for x, y, z in zip((r >> 32) & 0xFFFF, (r >> 16) & 0xFFFF, r & 0xFFFF)}
            assert b <= have
            reasons |= b
        if d < depth:
            finer = _coords(levels[d + 3][1]) - (1 << (d + 2))
            parent_broods = {tuple(v) for v in (finer // 2) // 2}
            assert parent_broods <= have
            reasons |= parent_broods
        assert have == reasons


def _check_invariants(levels, tree_keys):
    """Structural checks of the reference's own metadata test
    (tests/test_tree_metadata.py:56-73) plus the flag definitions, each
    checked from the other side of a link (ghost marks of the child level,
    coordinates decoded from the keys) rather than from the active sets."""
    assert [lv[0] for lv in levels] == list(range(-2, len(levels) - 2))
    assert len(levels[0][1]) == 1 and len(levels[1][1]) == 8
    # child k has offsets (k & 1, k >> 1 & 1, k >> 2 & 1); keys sort x-major
    np.testing.assert_array_equal(levels[0][5][0], [0, 4, 2, 6, 1, 5, 3, 7])
    np.testing.assert_array_equal(levels[0][3], [1])
    coords = [_coords(lv[1]) for lv in levels]
    for g, (d, keys, flags, ghost, parents, children) in enumerate(levels):
        c = coords[g]
        assert np.all(np.diff(keys) > 0)
        if g:
            assert len({tuple(v) for v in c // 2}) * 8 == len(keys)
            assert np.all(parents >= 0)
            np.testing.assert_array_equal(coords[g - 1][parents], c // 2)
        else:
            np.testing.assert_array_equal(parents, -1)
        assert np.all((children >= 0).sum(axis=1) % 8 == 0)
        child_ghost = np.ones(len(keys), bool)
        for k in range(8):
            present = children[:, k] >= 0
            if present.any():
                ch = children[present, k]
                np.testing.assert_array_equal(coords[g + 1][ch], 2 * c[present] + [k & 1, (k >> 1) & 1, (k >> 2) & 1])
                np.testing.assert_array_equal(levels[g + 1][4][ch], np.flatnonzero(present))
                child_ghost[present] &= levels[g + 1][3][ch] == 1
        has_child = (children >= 0).any(axis=1)
        bit64 = has_child & child_ghost if g else np.zeros(len(keys), bool)
        np.testing.assert_array_equal((flags & 64) != 0, bit64)
        if d >= 0:
            local = c - (1 << (d + 1))
            active = set(int(k) for k in tree_keys[d])
            assert active <= {tm.tree_key(*v) for v in local}
            real = ghost == 0
            assert int(real.sum()) == len(active)
            low = (local >= 0).all(axis=1) & real
            np.testing.assert_array_equal((flags & 1) != 0, low & (local.max(axis=1) < (1 << d)))
            np.testing.assert_array_equal((flags & 4) != 0, (flags & 1) != 0)
            np.testing.assert_array_equal((flags & 2) != 0, low & (local.max(axis=1) <= (1 << d)))
        else:
            np.testing.assert_array_equal(flags & 7, 0)


TOPOLOGY = json.loads((Path(__file__).parent / "golden" / "reference_reports.json").read_text())["complete_tree"]


def _topology_run(f):
    p = f["params"]
    pts, nrm = synthetic.fixture_input(p["shape"], p["count"], p["seed"], p["precision"], p["input_scale"],
                                       p["translation"])
    assert hashlib.sha256(synthetic.fixture_input_text(pts, nrm)).hexdigest() == f["input_sha256"]
    options = dict(depth=p["depth"], full_depth=p["full_depth"], base_depth=p["base_depth"],
                   solve_depth=p["solve_depth"], kernel_depth=p["kernel_depth"], samples_per_node=p["samples_per_node"],
                   scale=p["scale"], point_weight=p["weight"], exact_interpolation=p["exact"], iters=p["iters"],
                   base_v_cycles=p["base_v_cycles"], cg_accuracy=p["cg_accuracy"])
    return pts, nrm, options


def _counts(levels):
    return [{"local_depth": int(d), "nodes": len(keys), "ghosts": int(np.asarray(ghost).sum())}
            for d, keys, _, ghost, _, _ in levels]


@pytest.mark.parametrize("f", TOPOLOGY, ids=[f["name"] for f in TOPOLOGY])
def test_restatement_matches_reference_complete_tree(f):
    pts, nrm, options = _topology_run(f)
    r = oracle.reconstruct(pts, nrm, **options)
    depth = options["depth"]
    ins = _inputs(r, depth)
    levels = tm.build(depth, *ins)
    assert _counts(levels) == f["complete_tree"]
    # tools/check_samples.py's summary of the same run
    status = r.array("input_status")
    samples = f["samples"]
    assert len(status) == samples["input_count"] and int((status == 0).sum()) == samples["accepted"]
    assert len(r.array("sample_keys")) == samples["aggregates"]
    codes, counts = np.unique(status[status != 0], return_counts=True)
    assert {str(int(c)): int(k) for c, k in zip(codes, counts)} == samples["rejected_by_status"]
    for lv in f["tree"]:
        d = lv["depth"]
        assert hashlib.sha256(r.array("tree_keys", d).astype("<i8").tobytes()).hexdigest() == lv["support_sha256"]
    _check_invariants(levels, ins[0])
    _check_sources(levels, depth, *ins)


@pytest.mark.parametrize("shape,count,precision,params", [
    ("sphere", 512, "float", dict(depth=4, full_depth=2)),
    ("torus", 8192, "double", dict(depth=5, full_depth=3, point_weight=0.0)),
])
def test_restatement_invariants(shape, count, precision, params):
    p, n = synthetic.fixture_input(shape, count, synthetic.SEED, precision)
    r = oracle.reconstruct(p, n, **params)
    depth = params["depth"]
    ins = _inputs(r, depth)
    levels = tm.build(depth, *ins)
    _check_invariants(levels, ins[0])
    _check_sources(levels, depth, *ins)


@pytest.mark.gpu
def test_library_metadata_matches_restatement(cuda):
    import torch
    from warp_psr_amd import reconstruct
    from warp_psr_amd.tree_metadata import build_tree_metadata
    from parity import fixture_arrays
    for shape, count, precision, params in (("sphere", 512, "float", dict(depth=4, full_depth=2)),
                                            ("torus", 8192, "double", dict(depth=6, full_depth=3))):
        p, n = fixture_arrays(shape, count, precision)
        r = reconstruct(torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda), **params)
        depth = params["depth"]
        d = r.diagnostics
        offsets = d.array("normal_offsets")
        nk = d.array("normal_keys")
        kd = d.params["kernel_depth"]
        ins = ([d.array("tree_keys", k) for k in range(depth + 1)], d.array("sample_keys"),
               [d.array("density_keys", k) if k <= kd else np.zeros(0, np.int64) for k in range(depth + 1)],
               [nk[offsets[k]:offsets[k + 1]] for k in range(depth + 1)])
        expected = tm.build(depth, *ins)
        _check_invariants(expected, ins[0])
        print(f"tree metadata: {shape} {count} {precision} depth {depth}", flush=True)
        meta = build_tree_metadata(r)
        assert len(meta.levels) == depth + 3
        for level, (ld, keys, flags, ghost, parents, children) in zip(meta.levels, expected):
            assert level.local_depth == ld
            np.testing.assert_array_equal(level.keys.cpu().numpy(), keys)
            np.testing.assert_array_equal(level.flags.cpu().numpy(), flags)
            np.testing.assert_array_equal(level.ghost.cpu().numpy(), ghost)
            np.testing.assert_array_equal(level.parents.cpu().numpy(), parents)
            np.testing.assert_array_equal(level.children.cpu().numpy(), children)
    for f in TOPOLOGY:
        print(f"tree metadata: {f['name']}", flush=True)
        pts, nrm, options = _topology_run(f)
        r = reconstruct(torch.from_numpy(pts).to(cuda), torch.from_numpy(nrm).to(cuda), **options)
        meta = build_tree_metadata(r)
        counts = [{"local_depth": lv.local_depth, "nodes": int(lv.keys.numel()), "ghosts": int(lv.ghost.sum().item())}
                  for lv in meta.levels]
        assert counts == f["complete_tree"], f["name"]
