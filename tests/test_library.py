"""Host-side checks of the C ABI and the Python mirror (CPU only, no GPU compute).

* libwarp_psr_amd.so loads and exports every function include/*.h declares.
* Parameter validation happens before any device work and reproduces the
  reference's ValueError messages (src/warp_psr/reconstruction.py:160-212,
  tests/test_reconstruction.py:36-39).
"""
import ctypes
import re
from pathlib import Path

import pytest

from warp_psr_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        names += re.findall(r"\b(psr_[a-z_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.startswith("psr_evaluate_field")))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = declared_functions()
    assert "psr_reconstruct" in declared and "psr_relax_sliced" in declared
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert set(declared) == set(_lib.EXPORTED)


def _params(**kw):
    p = _lib.PsrParams()
    _lib.load().psr_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_params_default():
    p = _params()
    assert (p.depth, p.full_depth, p.base_depth, p.solve_depth, p.kernel_depth) == (6, 3, -1, -1, -1)
    assert (p.samples_per_node, p.scale, p.point_weight, p.iters, p.base_v_cycles, p.cg_accuracy) == (1.5, 1.1, 2.0, 8, 1, 1e-3)
    assert (p.keep_intermediates, p.float_operators, p.relax_tile) == (1, 0, 0)


@pytest.mark.parametrize("kw,message", [
    (dict(depth=16), "Require 0 <= base_depth <= full_depth"),
    (dict(depth=6, full_depth=7), "Require 0 <= base_depth <= full_depth"),
    (dict(depth=12, full_depth=3), "at most 8 integration levels"),
    (dict(iters=0), "iters and base_v_cycles must be positive integers"),
    (dict(point_weight=-1.0), "Require finite nonnegative point_weight"),
    (dict(cg_accuracy=0.0), "positive cg_accuracy"),
    (dict(scale=1.0), "scale > 1"),
    (dict(samples_per_node=float("nan")), "samples_per_node"),
    (dict(relax_tile=7), "relax_tile must be"),
    (dict(relax_tile=41), "relax_tile must be"),
    (dict(relax_tile=13), "relax_tile must be"),
    (dict(relax_tile=-2), "relax_tile must be"),
])
def test_c_abi_rejects_invalid_parameters_before_device_work(kw, message):
    lib = _lib.load()
    out = ctypes.c_void_p()
    rc = lib.psr_reconstruct(None, None, 1, 0, ctypes.byref(_params(**kw)), None, ctypes.byref(out))
    assert rc == 2
    assert message in lib.psr_last_error().decode()
    with pytest.raises(ValueError, match=re.escape(message)):
        _lib.check(rc)


def test_c_abi_rejects_empty_and_bad_dtype():
    lib = _lib.load()
    out = ctypes.c_void_p()
    assert lib.psr_reconstruct(None, None, 0, 0, ctypes.byref(_params()), None, ctypes.byref(out)) == 2
    assert "nonempty" in lib.psr_last_error().decode()
    assert lib.psr_reconstruct(None, None, 4, 7, ctypes.byref(_params()), None, ctypes.byref(out)) == 2
    assert lib.psr_relax_sliced(5, None, 1, None, None, None, None, None, 3, 8, None) == 3
    assert lib.psr_relax_scheduled(5, None, 1, None, None, None, None, None, 3, 8, 1, -1, None) == 3
    assert lib.psr_relax_scheduled(0, None, 0, None, None, None, None, None, 3, 8, 1, 9, None) == 2
    assert "slice z" in lib.psr_last_error().decode()


def test_c_abi_mixed_requires_float64():
    lib = _lib.load()
    out = ctypes.c_void_p()
    rc = lib.psr_reconstruct(None, None, 4, 0, ctypes.byref(_params(float_operators=1)), None, ctypes.byref(out))
    assert rc == 2
    assert "float64" in lib.psr_last_error().decode()


def test_python_validation_messages():
    import torch
    from warp_psr_amd import reconstruct
    p = torch.zeros((4, 3))
    with pytest.raises(ValueError, match="Depth parameters must be integers"):
        reconstruct(p, p, depth=6.0)
    with pytest.raises(ValueError, match="Require 0 <= base_depth"):
        reconstruct(p, p, depth=6, full_depth=7)
    with pytest.raises(ValueError, match="iters and base_v_cycles"):
        reconstruct(p, p, iters=0)
    # CPU tensors are rejected like non-CUDA Warp arrays (reconstruction.py:226-234)
    with pytest.raises(ValueError, match="matching float32/float64 3D vectors on the same CUDA device"):
        reconstruct(p, p)
    with pytest.raises(ValueError, match="matching float32/float64"):
        reconstruct(p, p.double())


def test_metadata_parser_shapes():
    """The C ABI's metadata JSON keys used by bench.py exist in the source (no device needed)."""
    text = (ROOT / "warp-psr_amd" / "csrc" / "psr_capi.hip").read_text()
    for key in ("basis_counts", "timings_seconds", "relaxation", "cube_to_model", "parameters", "iso_value"):
        assert f'\\"{key}\\"' in text


def test_differentiable_validates_tracked_inputs():
    """Something must be tracked (reconstruction_adjoint.py:70-85); options
    are constants."""
    import torch
    from warp_psr_amd import reconstruct_differentiable
    p = torch.zeros((4, 3), dtype=torch.float64, requires_grad=True)
    n = torch.ones((4, 3), dtype=torch.float64, requires_grad=True)
    with pytest.raises(TypeError, match="tensors"):
        reconstruct_differentiable(p.detach().numpy(), n)
    with pytest.raises(ValueError, match="requires_grad"):
        reconstruct_differentiable(p.detach(), n.detach())
    with pytest.raises(ValueError, match="constant scalar"):
        reconstruct_differentiable(p.detach(), n, point_weight=torch.tensor(1.0))


def test_c_abi_backward_inputs_validates_arguments():
    lib = _lib.load()
    assert lib.psr_backward_inputs(None, None, None, ctypes.c_double(0.0), None, None, None, 0, None, None, None) == 2
    assert lib.psr_field_vjp_inputs(None, None, 0, None, None, None, None, None, None, None) == 2


def test_c_abi_batch_validates_before_device_work():
    lib = _lib.load()
    count = 2
    vp = ctypes.c_void_p
    pts = (vp * count)(None, None)
    counts = (ctypes.c_int64 * count)(4, 0)
    params = (_lib.PsrParams * 1)(_params())
    out = (vp * count)()
    rc = lib.psr_reconstruct_batch(pts, pts, counts, count, 0, params, 1, 4, None, out)
    assert rc == 2 and "cloud 1" in lib.psr_last_error().decode()
    bad = (_lib.PsrParams * 1)(_params(depth=16))
    counts = (ctypes.c_int64 * count)(4, 4)
    assert lib.psr_reconstruct_batch(pts, pts, counts, count, 0, bad, 1, 4, None, out) == 2
    assert "Require 0 <= base_depth" in lib.psr_last_error().decode()
    assert lib.psr_reconstruct_batch(pts, pts, counts, count, 0, params, 3, 4, None, out) == 2
    assert lib.psr_reconstruct_batch(pts, pts, counts, 0, 0, params, 1, 4, None, out) == 0


def test_c_abi_tree_metadata_validates_arguments():
    lib = _lib.load()
    out = ctypes.c_void_p()
    count, depth, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
    assert lib.psr_tree_metadata(None, None, ctypes.byref(out)) == 2
    assert "null argument" in lib.psr_last_error().decode()
    assert lib.psr_tree_metadata_levels(None, ctypes.byref(count)) == 2
    assert lib.psr_tree_metadata_level(None, 0, ctypes.byref(depth), ctypes.byref(n)) == 2
    assert lib.psr_tree_metadata_copy(None, 0, None, None, None, None, None, None) == 2
    lib.psr_tree_metadata_free(None)


def _relax_operands(**changes):
    """A valid 3-row CSR operator on the CPU, with `changes` applied."""
    import torch
    ops = dict(
        keys=torch.tensor([1, 2, 3], dtype=torch.int64),
        offsets=torch.tensor([0, 1, 2, 3], dtype=torch.int32),
        columns=torch.tensor([0, 1, 2], dtype=torch.int32),
        values=torch.ones(3, dtype=torch.float64),
        b=torch.zeros(3, dtype=torch.float64),
        x=torch.zeros(3, dtype=torch.float64),
    )
    ops.update(changes)
    return ops


@pytest.mark.parametrize("fn", ["relax_sliced", "relax_scheduled"])
@pytest.mark.parametrize("change,message", [
    (lambda t: dict(columns=t.tensor([0, 1, 2], dtype=t.int64)), "columns must be int32"),
    (lambda t: dict(offsets=t.tensor([0, 1, 2, 3], dtype=t.int64)), "offsets must be int32"),
    (lambda t: dict(keys=t.tensor([1, 2, 3], dtype=t.int32)), "keys must be int64"),
    (lambda t: dict(b=t.zeros(3, dtype=t.float32)), "b must have the dtype of values"),
    (lambda t: dict(x=t.zeros(3, dtype=t.float32)), "x must have the dtype of values"),
    (lambda t: dict(values=t.ones(3, dtype=t.int32)), "values must be float32 or float64"),
    (lambda t: dict(offsets=t.tensor([0, 1, 3], dtype=t.int32)), "offsets must hold rows"),
    (lambda t: dict(x=t.zeros(4, dtype=t.float64)), "b and x must hold one value per row"),
    (lambda t: dict(columns=t.tensor([0, 1], dtype=t.int32)), "same length"),
    (lambda t: dict(x=t.zeros((3, 1), dtype=t.float64)), "one-dimensional"),
    (lambda t: dict(b=[0.0, 0.0, 0.0]), "b must be a torch tensor"),
    (lambda t: {}, "ROCm device"),
])
def test_relax_wrappers_reject_bad_operands(fn, change, message):
    """The relaxation wrappers check dtypes, lengths and the device before any
    pointer reaches the C ABI (which cannot see them)."""
    import torch
    from warp_psr_amd import reconstruction
    ops = _relax_operands(**change(torch))
    with pytest.raises(ValueError, match=message):
        getattr(reconstruction, fn)(ops["keys"], ops["offsets"], ops["columns"], ops["values"], ops["b"], ops["x"], 3, 2)
