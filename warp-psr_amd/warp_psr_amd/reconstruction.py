"""Host-side mirror of warp-psr's public API over the C ABI.

Reference: src/warp_psr/reconstruction.py:141-371 (reconstruct, Reconstruction),
src/warp_psr/field.py:122-184 (LevelField), src/warp_psr/device_field.py:18-142
(device handles).  Argument names, defaults, validation order and ValueError
messages follow the reference; inputs are PyTorch tensors instead of Warp arrays.
"""
import ctypes
import json
import math
from dataclasses import dataclass, field as dc_field

import numpy as np

from . import _lib

_ARRAY_TYPES = {
    "input_status": np.int32, "input_group_indices": np.int32, "input_order": np.int32, "first_indices": np.int32,
    "group_offsets": np.int32, "normal_offsets": np.int32, "matrix_offsets": np.int32, "matrix_columns": np.int32,
    "cross_offsets": np.int32, "cross_columns": np.int32, "prolong_offsets": np.int32, "prolong_columns": np.int32,
    "cg_runs": np.int32, "sample_keys": np.int64, "aggregate_keys": np.int64, "density_keys": np.int64,
    "normal_keys": np.int64, "tree_keys": np.int64, "basis_keys": np.int64, "keys": np.int64,
    "normal_load": np.float64, "point_load": np.float64,
}


def _torch():
    import torch
    return torch


def _stream_handle(device):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


@dataclass
class DeviceField:
    """Zero-copy device view of a field (reference device_field.py:18-29, 115-142).

    ``handle`` is the C ``PsrField`` struct; HIP kernels evaluate it with the
    header-only ``psr_evaluate_field`` of include/warp_psr_amd_field.h (the
    reference's ``evaluate_field(field, point)``, device_field.py:91-112).
    Keep the owning Reconstruction alive and unmodified while kernels use it.
    """
    handle: _lib.PsrField
    owner: object = dc_field(repr=False)
    dtype: object = None
    device: object = None

    @property
    def depth(self):
        return self.handle.depth

    @property
    def count(self):
        return self.handle.count

    def evaluate(self, queries):
        """psr_field_evaluate: the header-only device evaluation over queries."""
        torch = _torch()
        if (not isinstance(queries, torch.Tensor) or queries.dtype != self.dtype or queries.device != self.device
                or queries.ndim != 2 or queries.shape[1] != 3):
            raise ValueError("Query precision/device must match coefficients")
        q = queries.contiguous()
        m = q.shape[0]
        values = torch.empty(m, dtype=self.dtype, device=self.device)
        gradients = torch.empty((m, 3), dtype=self.dtype, device=self.device)
        valid = torch.empty(m, dtype=torch.int32, device=self.device)
        if m:
            _lib.check(_lib.load().psr_field_evaluate(ctypes.byref(self.handle), q.data_ptr(), m, values.data_ptr(),
                                                      gradients.data_ptr(), valid.data_ptr(),
                                                      _stream_handle(self.device)))
        return values, gradients, valid


class LevelField:
    """Per-level sparse degree-1 field (reference field.py:122-184).

    ``LevelField(keys, coefficients, linear, translation, depth, iso_value,
    cumulative=None)`` over an int64 key tensor and float32/float64
    coefficient tensors on one ROCm device, validated as the reference's
    dataclass (field.py:130-136).  With cumulative coefficients it evaluates
    the reference evaluation (field.py:364-421, leaf search, boundary nudge);
    without them the direct level summation (field.py:37-75, left derivatives
    at knots).  ``Reconstruction.field`` is a LevelField over the result's
    arrays.  Keep the tensors unmodified while the field or its device view is
    in use.
    """

    def __init__(self, keys, coefficients, linear, translation, depth, iso_value, cumulative=None):
        torch = _torch()
        if (not isinstance(keys, torch.Tensor) or not isinstance(coefficients, torch.Tensor)
                or coefficients.dtype not in (torch.float32, torch.float64) or keys.dtype != torch.int64):
            raise ValueError("Field requires float32/float64 coefficients and int64 keys")
        if not isinstance(depth, int) or not 0 <= depth <= 15:
            raise ValueError("Keys support depths 0..15, including upper boundary nodes")
        if keys.shape != coefficients.shape or keys.ndim != 1:
            raise ValueError("One key per coefficient is required")
        if cumulative is not None and (not isinstance(cumulative, torch.Tensor) or cumulative.dtype != coefficients.dtype
                                       or cumulative.shape != coefficients.shape):
            raise ValueError("Cumulative coefficients must match coefficient precision and shape")
        device = coefficients.device
        if device.type != "cuda" or keys.device != device or (cumulative is not None and cumulative.device != device):
            raise ValueError("Keys and coefficients must share one ROCm device")
        linear = np.asarray(linear, dtype=np.float64).reshape(-1)
        translation = np.asarray(translation, dtype=np.float64).reshape(-1)
        if linear.size != 9 or translation.size != 3:
            raise ValueError("linear must hold 9 values and translation 3")
        self._result = None
        self.dtype, self.device = coefficients.dtype, device
        self.keys, self.coefficients = keys.contiguous(), coefficients.contiguous()
        self.cumulative = None if cumulative is None else cumulative.contiguous()
        self.depth, self.iso_value = depth, float(iso_value)
        self.linear, self.translation = tuple(float(v) for v in linear), tuple(float(v) for v in translation)
        self._offsets = torch.empty(depth + 2, dtype=torch.int32, device=device)
        with torch.cuda.device(device):
            _lib.check(_lib.load().psr_field_level_offsets(self.keys.data_ptr(), self.keys.shape[0], depth,
                                                           self._offsets.data_ptr(), _stream_handle(device)))
        f = _lib.PsrField()
        f.keys, f.coefficients = self.keys.data_ptr(), self.coefficients.data_ptr()
        f.cumulative = None if self.cumulative is None else self.cumulative.data_ptr()
        f.level_offsets, f.count, f.depth = self._offsets.data_ptr(), self.keys.shape[0], depth
        f.dtype = 1 if self.dtype == torch.float64 else 0
        for i in range(9):
            f.linear[i] = self.linear[i]
        for i in range(3):
            f.translation[i] = self.translation[i]
        f.iso_value = self.iso_value
        self._field = f

    @classmethod
    def _from_result(cls, result, dtype, device):
        self = cls.__new__(cls)
        self._result = result
        self.dtype = dtype
        self.device = device
        field = _lib.PsrField()
        _lib.check(_lib.load().psr_result_field(result._handle, ctypes.byref(field)))
        self._field = field
        self.depth = field.depth
        self.iso_value = field.iso_value
        self.linear = tuple(field.linear)
        self.translation = tuple(field.translation)
        return self

    def evaluate(self, queries, *, requires_grad=False):
        """Values, model-space gradients and validity; NaN outside the cube."""
        torch = _torch()
        if requires_grad:
            if self._result is not None or self.cumulative is not None:
                raise ValueError("tracked cumulative fields require reconstruct_differentiable")
            raise ValueError("tracked direct-field evaluation is not supported; use reconstruct_differentiable")
        if (not isinstance(queries, torch.Tensor) or queries.dtype != self.dtype or queries.device != self.device
                or queries.ndim != 2 or queries.shape[1] != 3):
            raise ValueError("Query precision/device must match coefficients")
        if self._result is None:
            return self.device_field().evaluate(queries)
        q = queries.contiguous()
        m = q.shape[0]
        values = torch.empty(m, dtype=self.dtype, device=self.device)
        gradients = torch.empty((m, 3), dtype=self.dtype, device=self.device)
        valid = torch.empty(m, dtype=torch.int32, device=self.device)
        if m:
            _lib.check(_lib.load().psr_evaluate(self._result._handle, q.data_ptr(), m, values.data_ptr(),
                                                gradients.data_ptr(), valid.data_ptr(), _stream_handle(self.device)))
        return values, gradients, valid

    def device_field(self):
        return DeviceField(self._field, self._result if self._result is not None else self, self.dtype, self.device)


class _ResultHandle:
    def __init__(self, handle):
        self._handle = handle

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                _lib.load().psr_result_free(self._handle)
            except Exception:
                pass
            self._handle = None


class Reconstruction:
    """Reference Reconstruction (reconstruction.py:141-156)."""

    def __init__(self, handle, dtype, device, params):
        self._owner = _ResultHandle(handle)
        self.field = LevelField._from_result(self._owner, dtype, device)
        lib = _lib.load()
        need = ctypes.c_size_t()
        lib.psr_result_metadata_json(handle, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value)
        lib.psr_result_metadata_json(handle, buf, need.value, None)
        self.metadata = json.loads(buf.value.decode())
        self.metadata["device"] = str(device)
        self.diagnostics = _Diagnostics(self._owner, dtype, params)

    @property
    def iso_value(self):
        return self.field.iso_value

    def evaluate(self, queries):
        return self.field.evaluate(queries)

    def device_field(self):
        """Zero-copy handle for psr_evaluate_field in HIP kernels."""
        return self.field.device_field()


class _Diagnostics:
    """Host copies of stage arrays (reference ``Reconstruction.diagnostics``)."""

    def __init__(self, owner, dtype, params):
        self._owner = owner
        self._np_dtype = np.float32 if str(dtype).endswith("float32") else np.float64
        self.params = params

    def array(self, name, level=-1):
        data = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        _lib.check(_lib.load().psr_result_array(self._owner._handle, name.encode(), level, ctypes.byref(data),
                                                ctypes.byref(nbytes)))
        dt = _ARRAY_TYPES.get(name, self._np_dtype)
        count = nbytes.value // np.dtype(dt).itemsize
        if not count:
            return np.zeros(0, dtype=dt)
        buf = (ctypes.c_char * nbytes.value).from_address(data.value)
        return np.frombuffer(buf, dtype=dt, count=count).copy()


def _prepare(positions, normals, depth=6, full_depth=3, base_depth=None, solve_depth=None, kernel_depth=None,
             samples_per_node=1.5, scale=1.1, point_weight=2.0, exact_interpolation=False, iters=8, base_v_cycles=1,
             cg_accuracy=1e-3, mixed=False, relax_tile=0):
    """Reference validation (reconstruction.py:160-212), then the C parameters."""
    torch = _torch()
    if base_depth is None:
        base_depth = full_depth
    if solve_depth is None:
        solve_depth = depth
    if kernel_depth is None:
        kernel_depth = max(depth - 2, 0)
    depths = (depth, full_depth, base_depth, solve_depth, kernel_depth)
    if any(not isinstance(v, int) or isinstance(v, bool) for v in depths):
        raise ValueError("Depth parameters must be integers")
    if not 0 <= base_depth <= full_depth <= solve_depth <= depth <= 15 or depth < 1 or not 0 <= kernel_depth <= depth:
        raise ValueError("Require 0 <= base_depth <= full_depth <= solve_depth <= depth <= 15, depth >= 1, "
                         "and 0 <= kernel_depth <= depth")
    if depth - full_depth > 7:
        raise ValueError("Warp 1.19 AdaptiveNanogrid supports at most 8 integration levels; require depth - full_depth <= 7")
    if not isinstance(iters, int) or iters < 1 or not isinstance(base_v_cycles, int) or base_v_cycles < 1:
        raise ValueError("iters and base_v_cycles must be positive integers")
    if not math.isfinite(point_weight) or point_weight < 0 or not math.isfinite(cg_accuracy) or cg_accuracy <= 0:
        raise ValueError("Require finite nonnegative point_weight and positive cg_accuracy")
    if not math.isfinite(samples_per_node) or samples_per_node <= 0 or not math.isfinite(scale) or scale <= 1:
        raise ValueError("Require finite positive samples_per_node and scale > 1")
    if (not isinstance(positions, torch.Tensor) or not isinstance(normals, torch.Tensor)
            or positions.dtype not in (torch.float32, torch.float64) or normals.dtype != positions.dtype
            or positions.device.type != "cuda" or normals.device != positions.device
            or positions.ndim != 2 or positions.shape[1] != 3 or normals.shape != positions.shape):
        raise ValueError("Reconstruction requires matching float32/float64 3D vectors on the same CUDA device")
    if not positions.shape[0]:
        raise ValueError("points and normals must have equal nonempty 1D shapes")
    if mixed and positions.dtype != torch.float64:
        raise ValueError("mixed precision (float_operators) runs on float64 inputs")
    if not isinstance(relax_tile, int) or not (relax_tile in (0, -1) or (8 <= relax_tile <= 40 and relax_tile % 2 == 0)):
        raise ValueError("relax_tile must be 0 (default), -1 (one launch per colour) or an even tile side in [8, 40]")
    params = _lib.PsrParams()
    _lib.load().psr_params_default(ctypes.byref(params))
    params.depth, params.full_depth, params.base_depth = depth, full_depth, base_depth
    params.solve_depth, params.kernel_depth = solve_depth, kernel_depth
    params.samples_per_node, params.scale, params.point_weight = samples_per_node, scale, point_weight
    params.exact_interpolation = int(bool(exact_interpolation))
    params.iters, params.base_v_cycles, params.cg_accuracy = iters, base_v_cycles, cg_accuracy
    params.float_operators = int(bool(mixed))
    params.relax_tile = relax_tile
    resolved = dict(depth=depth, full_depth=full_depth, base_depth=base_depth, solve_depth=solve_depth,
                    kernel_depth=kernel_depth, samples_per_node=samples_per_node, scale=scale,
                    point_weight=point_weight, exact_interpolation=bool(exact_interpolation), iters=iters,
                    base_v_cycles=base_v_cycles, cg_accuracy=cg_accuracy, mixed=bool(mixed))
    return positions.contiguous(), normals.contiguous(), params, resolved


def reconstruct(positions, normals, *, depth=6, full_depth=3, base_depth=None, solve_depth=None, kernel_depth=None,
                samples_per_node=1.5, scale=1.1, point_weight=2.0, exact_interpolation=False, iters=8, base_v_cycles=1,
                cg_accuracy=1e-3, mixed=False, relax_tile=0):
    """Reconstruct oriented (n, 3) float32/float64 ROCm tensors; point_weight=0 selects PSR.

    Degree one, Neumann boundaries, confidence disabled, library alignDir=0.
    The surface is ``result.evaluate(q)[0] == result.iso_value``.
    ``mixed=True`` (float64 tensors only) stores the finer levels' Laplacian +
    screening operators as float32 while every vector, the octree stages and
    the base-level solve stay float64.
    ``relax_tile`` sets the side of the (x, y) tiles of the finer levels'
    sliced Gauss-Seidel (even, 8..40; 0 = library default; -1 = one launch per
    colour); every choice gives the same bits.
    """
    torch = _torch()
    p, n, params, resolved = _prepare(
        positions, normals, depth=depth, full_depth=full_depth, base_depth=base_depth, solve_depth=solve_depth,
        kernel_depth=kernel_depth, samples_per_node=samples_per_node, scale=scale, point_weight=point_weight,
        exact_interpolation=exact_interpolation, iters=iters, base_v_cycles=base_v_cycles, cg_accuracy=cg_accuracy,
        mixed=mixed, relax_tile=relax_tile)
    handle = ctypes.c_void_p()
    dtype_code = 1 if p.dtype == torch.float64 else 0
    with torch.cuda.device(p.device):
        _lib.check(_lib.load().psr_reconstruct(p.data_ptr(), n.data_ptr(), p.shape[0], dtype_code, ctypes.byref(params),
                                               _stream_handle(p.device), ctypes.byref(handle)))
    return Reconstruction(handle.value, p.dtype, p.device, resolved)


def reconstruct_batch(clouds, *, max_concurrency=8, **options):
    """Reconstruct many independent clouds concurrently (psr_reconstruct_batch).

    ``clouds`` is a sequence of (positions, normals) pairs of one dtype on one
    device; ``options`` are reconstruct()'s keywords, applied to every cloud,
    or a sequence of per-cloud option dicts under ``per_cloud=[...]``.  Each
    result equals reconstruct() of that cloud bit for bit.  No reference
    counterpart: the reference reconstructs one cloud per call.
    """
    torch = _torch()
    per_cloud = options.pop("per_cloud", None)
    clouds = list(clouds)
    if per_cloud is not None and len(per_cloud) != len(clouds):
        raise ValueError("per_cloud must hold one option dict per cloud")
    if not clouds:
        return []
    prepared = []
    for i, (positions, normals) in enumerate(clouds):
        opts = dict(options)
        if per_cloud is not None:
            opts.update(per_cloud[i])
        prepared.append(_prepare(positions, normals, **opts))
    dtype, device = prepared[0][0].dtype, prepared[0][0].device
    if any(p.dtype != dtype or p.device != device for p, _, _, _ in prepared):
        raise ValueError("batched clouds must share one dtype and device")
    count = len(prepared)
    pts = (ctypes.c_void_p * count)(*[p.data_ptr() for p, _, _, _ in prepared])
    nrm = (ctypes.c_void_p * count)(*[n.data_ptr() for _, n, _, _ in prepared])
    counts = (ctypes.c_int64 * count)(*[p.shape[0] for p, _, _, _ in prepared])
    params = (_lib.PsrParams * count)(*[prm for _, _, prm, _ in prepared])
    out = (ctypes.c_void_p * count)()
    with torch.cuda.device(device):
        _lib.check(_lib.load().psr_reconstruct_batch(pts, nrm, counts, count, 1 if dtype == torch.float64 else 0, params,
                                                     count, int(max_concurrency), _stream_handle(device), out))
    return [Reconstruction(out[i], dtype, device, prepared[i][3]) for i in range(count)]


def _relax_operands(name, keys, offsets, columns, values, b, x, depth, iters):
    """Checks a caller CSR operator before its raw pointers reach the C ABI,
    which cannot see dtypes, lengths or devices: int64 keys, int32 offsets and
    columns, float32/float64 values with b and x of the same dtype, rows+1
    offsets, offsets[-1] == len(columns) == len(values), columns in [0, rows),
    one ROCm device.  Returns the dtype code (0 float32, 1 float64).  The
    reference gets the same protection from Warp's typed arrays
    (relaxation.py:38-64)."""
    torch = _torch()
    tensors = dict(keys=keys, offsets=offsets, columns=columns, values=values, b=b, x=x)
    for label, t in tensors.items():
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name}: {label} must be a torch tensor")
        if t.dim() != 1 or not t.is_contiguous():
            raise ValueError(f"{name}: {label} must be a contiguous one-dimensional tensor")
    if keys.dtype != torch.int64:
        raise ValueError(f"{name}: keys must be int64 basis keys")
    for label in ("offsets", "columns"):
        if tensors[label].dtype != torch.int32:
            raise ValueError(f"{name}: {label} must be int32")
    if values.dtype not in (torch.float32, torch.float64):
        raise ValueError(f"{name}: values must be float32 or float64")
    for label in ("b", "x"):
        if tensors[label].dtype != values.dtype:
            raise ValueError(f"{name}: {label} must have the dtype of values ({values.dtype})")
    rows = keys.shape[0]
    if offsets.shape[0] != rows + 1:
        raise ValueError(f"{name}: offsets must hold rows + 1 = {rows + 1} entries")
    if b.shape[0] != rows or x.shape[0] != rows:
        raise ValueError(f"{name}: b and x must hold one value per row ({rows})")
    if columns.shape[0] != values.shape[0]:
        raise ValueError(f"{name}: columns and values must have the same length")
    if int(depth) < 0 or int(iters) < 0:
        raise ValueError(f"{name}: depth and iters must be non-negative")
    device = x.device
    if device.type != "cuda" or any(t.device != device for t in tensors.values()):
        raise ValueError(f"{name} requires every tensor on one ROCm device")
    if rows:
        if int(offsets[0]) != 0 or int(offsets[-1]) != columns.shape[0]:
            raise ValueError(f"{name}: offsets must start at 0 and end at len(columns)")
        if bool((offsets[1:] < offsets[:-1]).any()):
            raise ValueError(f"{name}: offsets must be non-decreasing")
        if columns.shape[0] and (int(columns.min()) < 0 or int(columns.max()) >= rows):
            raise ValueError(f"{name}: columns must lie in [0, rows)")
    return 1 if values.dtype == torch.float64 else 0


def relax_scheduled(keys, offsets, columns, values, b, x, depth, iters, forward=True, z=-1):
    """8-colour Gauss-Seidel on device CSR tensors (relaxation.py:261-271); updates x.

    ``z < 0`` sweeps the colour groups of ``row_schedule(keys, depth)``; ``z >= 0``
    relaxes slice ``z`` of the sliced schedule, as the reference's ``z`` argument.
    """
    dtype = _relax_operands("relax_scheduled", keys, offsets, columns, values, b, x, depth, iters)
    with _torch().cuda.device(x.device):
        _lib.check(_lib.load().psr_relax_scheduled(dtype, keys.data_ptr(), keys.shape[0], offsets.data_ptr(),
                                                   columns.data_ptr(), values.data_ptr(), b.data_ptr(), x.data_ptr(),
                                                   depth, iters, int(bool(forward)), int(z), _stream_handle(x.device)))
    return x


def relax_sliced(keys, offsets, columns, values, b, x, depth, iters):
    """Sliced Gauss-Seidel on device CSR tensors in the reference order (relaxation.py:189-210); updates x."""
    dtype = _relax_operands("relax_sliced", keys, offsets, columns, values, b, x, depth, iters)
    with _torch().cuda.device(x.device):
        _lib.check(_lib.load().psr_relax_sliced(dtype, keys.data_ptr(), keys.shape[0], offsets.data_ptr(),
                                                columns.data_ptr(), values.data_ptr(), b.data_ptr(), x.data_ptr(),
                                                depth, iters, _stream_handle(x.device)))
    return x
