"""Fixture interchange (warp_psr_amd/interchange.py): the PoissonRecon oracle's
text exports (reference/README.md:27-74).

CPU: rows in the reference writer's layout parse as tools/make_fixture.py:23-26
reads them; a fixture written from an oracle reconstruction reads back to the
same bits and passes every comparator exactly; a mistake planted in one row
of a written file fails the comparator that owns it.  GPU: the library's
reconstruction writes the same files, byte for byte, as the oracle's run on
the same input, and passes the comparators against the oracle's fixture.
"""
import numpy as np
import pytest

import oracle
from oracle import tree_metadata as tm
from warp_psr_amd import interchange as ix
from warp_psr_amd import synthetic

CASE = ("sphere", 512, "double", dict(depth=4, full_depth=2))


def _metadata_inputs(array, params):
    depth = params["depth"]
    offsets = array("normal_offsets")
    nk = array("normal_keys")
    return ([array("tree_keys", d) for d in range(depth + 1)], array("sample_keys"),
            [array("density_keys", d) if d <= params["kernel_depth"] else np.zeros(0, np.int64) for d in range(depth + 1)],
            [nk[offsets[d]:offsets[d + 1]] for d in range(depth + 1)])


def _oracle_fixture(tmp_path, shape, count, precision, params):
    p, n = synthetic.fixture_input(shape, count, synthetic.SEED, precision)
    res = oracle.reconstruct(p, n, **params)
    q = synthetic.fixture_input(shape, 64, synthetic.SEED + 1, precision)[0] * 0.9
    meta = tm.build(params["depth"], *_metadata_inputs(res.array, res.params))
    ix.write_fixture(tmp_path, res, q, metadata=meta)
    return res, meta, q


def test_rows_in_the_reference_layout_parse():
    text = ("parallel_async 0\nthreads 1\niso 0.49815374612808228\n"
            "transform 2 0 0 0 0 2 0 0 0 0 2 0 -1 -1 -1 1\n"
            "q 0.25 1 -2 0.5\nq -0.125 0 0 3\n"
            "node -2 -1 -1 -1 0 1 -1 0\nnode 0 0 0 0 7 0 3 0.0625\n")
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "oracle.txt")
        open(path, "w").write(text)
        rows = ix.read_rows(path)
    assert rows["iso"].shape == (1, 1) and rows["iso"][0, 0] == 0.49815374612808228
    np.testing.assert_array_equal(rows["transform"].reshape(4, 4).T[:3, 3], [-1, -1, -1])
    np.testing.assert_array_equal(rows["q"], [[0.25, 1, -2, 0.5], [-0.125, 0, 0, 3]])
    np.testing.assert_array_equal(rows["node"][:, 4], [0, 7])


def test_oracle_fixture_round_trip(tmp_path):
    shape, count, precision, params = CASE
    res, meta, q = _oracle_fixture(tmp_path, shape, count, precision, params)
    f = ix.read_fixture(tmp_path)
    assert set(ix.FILES) <= set(f)
    o = f["oracle_instrumented.txt"]
    assert o["iso"][0, 0] == res.iso_value
    np.testing.assert_array_equal(o["transform"].reshape(4, 4).T, res.array("cube_to_model").reshape(4, 4))
    values, grads, _ = res.evaluate(q)
    ix.compare_field(values, grads, res.iso_value, o, double=True, exact=True)
    ix.compare_tree(res, o["node"])
    ix.compare_inputs(res, f["inputs.txt"]["input"])
    ix.compare_topology(meta, f["topology.txt"]["topology"])
    ix.compare_preprocess(res, f["preprocess.txt"], f["stages.txt"], f["splats.txt"], exact=True)
    ix.compare_loads(res, f["stages.txt"], exact=True)
    ix.compare_operators(res, f["operators.txt"], o["node"], exact=True)
    # node rows: one per retained node, coefficients of the active basis
    assert len(o["node"]) == sum(len(lv[1]) for lv in meta)
    coef = {int(r[6]): r[7] for r in o["node"] if r[6] >= 0}
    np.testing.assert_array_equal([coef[i] for i in sorted(coef)], res.array("coefficients")[sorted(coef)])
    # the settings row restates the resolved parameters
    eff = f["stages.txt"]["effective"][0]
    assert list(eff[:5]) == [params["depth"], params["full_depth"], 2, 4, 2]


def _plant(path, tag, index, column, value):
    lines = path.read_text().splitlines()
    hits = [i for i, l in enumerate(lines) if l.split()[0] == tag]
    parts = lines[hits[index]].split()
    parts[column + 1] = value(parts[column + 1])
    lines[hits[index]] = " ".join(parts)
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("file,tag,column,check", [
    ("oracle_instrumented.txt", "node", 4, "tree"),        # active flag of one node
    ("inputs.txt", "input", 1, "inputs"),                  # one status
    ("topology.txt", "topology", 14, "topology"),          # one child count
    ("preprocess.txt", "sample", 5, "preprocess"),         # one position sum
    ("stages.txt", "load", 4, "loads"),                    # one load
    ("operators.txt", "A", 3, "operators"),                # one matrix entry
])
def test_planted_row_mistakes_fail(tmp_path, file, tag, column, check):
    shape, count, precision, params = CASE
    res, meta, q = _oracle_fixture(tmp_path, shape, count, precision, params)
    if tag == "node":
        rows = ix.read_rows(tmp_path / file)["node"]
        index = int(np.flatnonzero((rows[:, 0] == 3) & (rows[:, 5] == 0))[0])
        change = lambda v: str(int(v) ^ 2)
    elif tag == "input":
        index, change = 5, lambda v: str(3 - int(v)) if int(v) in (0, 3) else "0"
    elif tag == "topology":
        rows = ix.read_rows(tmp_path / file)["topology"]
        index, change = int(np.flatnonzero(rows[:, 14] == 8)[0]), lambda v: "0"
    else:
        index, change = 7, lambda v: "%.17g" % np.nextafter(float(v), np.inf)
    _plant(tmp_path / file, tag, index, column, change)
    f = ix.read_fixture(tmp_path)
    with pytest.raises(AssertionError):
        if check == "tree":
            ix.compare_tree(res, f[file]["node"])
        elif check == "inputs":
            ix.compare_inputs(res, f[file]["input"])
        elif check == "topology":
            ix.compare_topology(meta, f[file]["topology"])
        elif check == "preprocess":
            ix.compare_preprocess(res, f[file], f["stages.txt"], exact=True)
        elif check == "loads":
            ix.compare_loads(res, f[file], exact=True)
        else:
            ix.compare_operators(res, f[file], f["oracle_instrumented.txt"]["node"], exact=True)


@pytest.mark.gpu
def test_library_fixture_equals_oracle_fixture(cuda, tmp_path):
    import torch
    from warp_psr_amd import reconstruct
    from warp_psr_amd.tree_metadata import build_tree_metadata
    for shape, count, precision, params in (CASE, ("torus", 8192, "float", dict(depth=6, full_depth=3))):
        print(f"interchange: {shape} {count} {precision} {params}", flush=True)
        ref_dir, gpu_dir = tmp_path / f"{shape}-oracle", tmp_path / f"{shape}-gpu"
        res, meta, q = _oracle_fixture(ref_dir, shape, count, precision, params)
        p, n = synthetic.fixture_input(shape, count, synthetic.SEED, precision)
        r = reconstruct(torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda), **params)
        ix.write_fixture(gpu_dir, r, torch.from_numpy(q).to(cuda), metadata=build_tree_metadata(r))
        for name in ix.FILES:
            assert (gpu_dir / name).read_text() == (ref_dir / name).read_text(), name
        f = ix.read_fixture(ref_dir)
        o = f["oracle_instrumented.txt"]
        v, g, _ = r.evaluate(torch.from_numpy(q).to(cuda))
        ix.compare_field(v.cpu().numpy(), g.cpu().numpy(), r.iso_value, o, double=precision == "double", exact=True)
        ix.compare_tree(r, o["node"])
        ix.compare_inputs(r, f["inputs.txt"]["input"])
        ix.compare_topology(build_tree_metadata(r), f["topology.txt"]["topology"])
        ix.compare_preprocess(r, f["preprocess.txt"], f["stages.txt"], f["splats.txt"], exact=True)
        ix.compare_loads(r, f["stages.txt"], exact=True)
        ix.compare_operators(r, f["operators.txt"], o["node"], exact=True)
