"""Tiled sliced Gauss-Seidel on the GPU (one launch per wave, one workgroup per
(x, y) tile of a slice, four colours behind workgroup barriers).

Every tile side, the default and the colour-per-launch form (relax_tile=-1)
must give the oracle's corrections and field bit for bit: odd and even sweep
counts (the odd count starts from the second buffer), float32, float64 and
mixed precision, and a level with more than 128 rows of one colour in a tile
(further passes inside a colour phase).
"""
import numpy as np
import pytest

import oracle
from parity import compare_stages, fixture_arrays, same

pytestmark = pytest.mark.gpu

STAGES = [("corrections", "solve"), ("rhs", "solve"), ("coefficients", False), ("cumulative", False), ("iso", False)]

CASES = [
    ("torus-float", "torus", 8192, "float", dict(depth=6, full_depth=3)),
    ("torus-mixed", "torus", 8192, "double", dict(depth=6, full_depth=3, mixed=True)),
    ("sphere-double-odd", "sphere", 4096, "double", dict(depth=6, full_depth=2, iters=3)),
    ("sheets-psr", "sheets", 4096, "float", dict(depth=6, full_depth=2, point_weight=0.0, iters=5)),
]


@pytest.mark.parametrize("name,shape,count,precision,params", CASES, ids=[c[0] for c in CASES])
def test_tiled_relaxation_bit_exact(cuda, name, shape, count, precision, params):
    import torch
    from warp_psr_amd import reconstruct
    p, n = fixture_arrays(shape, count, precision)
    ref = oracle.reconstruct(p, n, **params)
    tp, tn = torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda)
    for tile in (0, -1, 8, 24, 40):
        print(f"tiled relaxation {name} tile {tile}", flush=True)
        gpu = reconstruct(tp, tn, relax_tile=tile, **params)
        bad = compare_stages(gpu.diagnostics, ref, ref.params, STAGES)
        assert not bad, f"tile {tile}: " + "; ".join(f"{s}[{lv}] {m}" for s, lv, m in bad[:5])
        relax = gpu.metadata["relaxation"]
        n_levels = len(relax)
        assert n_levels >= 1
        for st in relax:
            waves = (1 << st["depth"]) + 2 * (params.get("iters", 8) - 1) + 1
            if tile == -1:
                assert st["tile"] == 0 and st["launches"] == 4 * waves
            else:
                assert (st["tile"] == tile if tile else 8 <= st["tile"] <= 40) and st["launches"] == waves
                assert st["tiles"] > 0 and st["tile_entries"] >= st["rows"]


def test_tiled_relaxation_large_level(cuda):
    """A depth-7 level of a 60,000-point ellipsoid union: many tiles per
    slice and tiles with more rows of a colour than one pass holds."""
    import torch
    from warp_psr_amd import reconstruct, synthetic
    p, n = synthetic.workload_ellipsoids(60000, dtype=np.float64)
    params = dict(depth=7, full_depth=3, mixed=True)
    ref = oracle.reconstruct(p, n, **params)
    tp, tn = torch.from_numpy(p).to(cuda), torch.from_numpy(n).to(cuda)
    for tile in (0, 40):
        print(f"tiled relaxation ellipsoids d7 tile {tile}", flush=True)
        gpu = reconstruct(tp, tn, relax_tile=tile, **params)
        bad = compare_stages(gpu.diagnostics, ref, ref.params, STAGES)
        assert not bad, f"tile {tile}: " + "; ".join(f"{s}[{lv}] {m}" for s, lv, m in bad[:5])
    assert same(gpu.diagnostics.array("corrections", 7), ref.array("corrections", 7))
