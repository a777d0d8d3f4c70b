"""Marching-cubes mesh extraction (reference src/warp_psr/meshing.py).

``extract_isosurface(field, resolution=..., bounds=...)`` meshes the zero set of
a batched callback ``field(points) -> values`` ((m, 3) float32 model-space
points to (m,) float32 values, PyTorch tensors on one ROCm device);
``extract_mesh(reconstruction, ...)`` meshes a float32 reconstruction at its
isovalue.  Grid points, cell classification (a one-bit mask filled slab by
slab), compaction, vertices (one per crossing grid edge) and triangles run in
the library's HIP kernels (``psr_mesh_*``).  Triangles come from the library's
own case table (the reference uses Warp's): faces point toward increasing
values, values equal to the threshold count as nonnegative, vertices are not
welded.  When the callback's values are tracked by autograd, the vertices are
too: d vertex / d value through the edge interpolation, with cell selection,
cases and connectivity fixed (meshing.py:177-190).
"""
import ctypes
import math
import operator
from dataclasses import dataclass, field as dc_field

from . import _lib
from .reconstruction import _stream_handle, _torch


@dataclass
class Mesh:
    vertices: object
    indices: object
    metadata: dict = dc_field(default_factory=dict)


def _grid(resolution, bounds):
    """Validation and float32 spacing of meshing.py:126-146."""
    if isinstance(resolution, int) and not isinstance(resolution, bool):
        resolution = (resolution,) * 3
    try:
        shape = tuple(operator.index(v) for v in resolution)
    except (TypeError, ValueError):
        raise ValueError("resolution must be an integer or three integers >= 2") from None
    if len(shape) != 3 or any(isinstance(v, bool) for v in resolution) or min(shape) < 2:
        raise ValueError("resolution must be an integer or three integers >= 2")
    try:
        lo, hi = (tuple(float(v) for v in row) for row in bounds)
    except (TypeError, ValueError):
        raise ValueError("bounds must contain two finite 3D corners") from None
    if (len(lo) != 3 or len(hi) != 3 or not all(math.isfinite(v) for v in (*lo, *hi))
            or any(a >= b for a, b in zip(lo, hi))):
        raise ValueError("bounds must be finite with lower < upper on every axis")
    if math.prod(n - 1 for n in shape) > 2**31 - 1:
        raise ValueError("extraction cell count exceeds int32 indexing")
    if math.prod(shape) > 2**31 - 1:
        raise ValueError("grid node count exceeds int32 sampling indices")
    import numpy as np
    lower = np.asarray(lo, np.float32)
    upper = np.asarray(hi, np.float32)
    with np.errstate(all="ignore"):
        step = (upper - lower) / np.asarray([n - 1 for n in shape], np.float32)
    if not all(math.isfinite(float(s)) and s > 0 for s in step):
        raise ValueError("bounds/spacing must be representable in float32")
    return shape, lower, upper


class _MeshHandle:
    """Owns one PsrMesh; kept alive by the autograd graph for the backward."""

    def __init__(self, shape, lower, upper, stream):
        self._handle = ctypes.c_void_p()
        arr3i = (ctypes.c_int32 * 3)(*shape)
        lo = (ctypes.c_float * 3)(*[float(v) for v in lower])
        hi = (ctypes.c_float * 3)(*[float(v) for v in upper])
        _lib.check(_lib.load().psr_mesh_begin(arr3i, lo, hi, stream, ctypes.byref(self._handle)))

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                _lib.load().psr_mesh_free(self._handle)
            except Exception:
                pass
            self._handle = None


class _Vertices:
    fn = None


def _vertex_function():
    if _Vertices.fn is not None:
        return _Vertices.fn
    torch = _torch()

    class MeshVertices(torch.autograd.Function):
        @staticmethod
        def forward(ctx, corner_values, vertices, mesh, stream):
            ctx.mesh = mesh
            ctx.stream = stream
            ctx.save_for_backward(corner_values)
            return vertices.clone()

        @staticmethod
        def backward(ctx, grad_vertices):
            (corner_values,) = ctx.saved_tensors
            grad = torch.zeros_like(corner_values)
            g = grad_vertices.contiguous()
            if g.numel():
                _lib.check(_lib.load().psr_mesh_vertex_vjp(ctx.mesh._handle, corner_values.data_ptr(), g.data_ptr(),
                                                           grad.data_ptr()))
            return grad, None, None, None

    _Vertices.fn = MeshVertices
    return MeshVertices


def _check_values(values, points, device):
    torch = _torch()
    if (not isinstance(values, torch.Tensor) or values.dtype != torch.float32 or values.ndim != 1
            or values.shape[0] != points.shape[0] or values.device != device):
        raise ValueError("field must return contiguous float32 values, one per point, on the query device")
    return values.contiguous()


def extract_isosurface(field, *, resolution, bounds, backend="sparse", device=None, slab_size=16, lipschitz_bound=None,
                       classification_field=None):
    """Zero set of ``field(points) -> values`` (meshing.py:150-280).

    ``sparse`` classifies every cell slab by slab (``slab_size`` cell layers)
    and samples only the retained cells' corners again; ``dense`` classifies
    the whole grid in one slab; ``lipschitz`` requires a caller-proven bound
    and then classifies every cell as well (no octree pruning here), so all
    three give the same mesh.  ``classification_field`` (sparse only) supplies
    untracked values for the classification.
    """
    torch = _torch()
    if backend not in ("dense", "sparse", "lipschitz"):
        raise ValueError("backend must be dense, sparse, or lipschitz")
    if classification_field is not None and backend != "sparse":
        raise ValueError("classification_field is only used by the sparse backend")
    shape, lower, upper = _grid(resolution, bounds)
    if not isinstance(slab_size, int) or isinstance(slab_size, bool) or slab_size < 1:
        raise ValueError("slab_size must be a positive integer")
    if backend == "lipschitz":
        if lipschitz_bound is None or not math.isfinite(lipschitz_bound) or lipschitz_bound <= 0:
            raise ValueError("lipschitz backend requires a finite positive proven lipschitz_bound")
    elif lipschitz_bound is not None:
        raise ValueError("lipschitz_bound is only used by the lipschitz backend")
    device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if device.type != "cuda":
        raise ValueError("mesh extraction runs on a ROCm device")
    lib = _lib.load()
    samples = 0

    def sample(points, classification=False):
        nonlocal samples
        callback = classification_field if classification and classification_field is not None else field
        values = _check_values(callback(points), points, device)
        samples += points.shape[0]
        return values

    metadata = {"backend": backend, "resolution": shape,
                "bounds": (tuple(float(v) for v in lower), tuple(float(v) for v in upper)),
                "precision": "float32", "cell_selection": "fixed for backward"}
    with torch.cuda.device(device):
        stream = _stream_handle(device)
        mesh = _MeshHandle(shape, lower, upper, stream)
        nx = shape[0]
        slab = nx - 1 if backend != "sparse" else slab_size
        tracked = False
        peak = 0
        for start in range(0, nx - 1, slab):
            width = min(slab, nx - 1 - start)
            nodes = ctypes.c_int64()
            _lib.check(lib.psr_mesh_slab_nodes(mesh._handle, start, width, ctypes.byref(nodes)))
            points = torch.empty((nodes.value, 3), dtype=torch.float32, device=device)
            _lib.check(lib.psr_mesh_slab_points(mesh._handle, start, width, points.data_ptr()))
            values = sample(points, classification=True)
            tracked = tracked or values.requires_grad
            _lib.check(lib.psr_mesh_slab_values(mesh._handle, start, width, values.detach().data_ptr()))
            peak = max(peak, nodes.value)
            del points, values
        retained, invalid = ctypes.c_int64(), ctypes.c_int32()
        _lib.check(lib.psr_mesh_select(mesh._handle, ctypes.byref(retained), ctypes.byref(invalid)))
        if invalid.value:
            raise ValueError("field returned nonfinite values inside extraction bounds")
        count = retained.value
        metadata.update(retained_cells=count, slab_size=slab, selection_bytes=((math.prod(n - 1 for n in shape) + 31) // 32) * 12,
                        peak_slab_nodes=peak)
        if count:
            corners = torch.empty((8 * count, 3), dtype=torch.float32, device=device)
            _lib.check(lib.psr_mesh_corner_points(mesh._handle, corners.data_ptr()))
            corner_values = sample(corners)
            V, T = ctypes.c_int64(), ctypes.c_int64()
            _lib.check(lib.psr_mesh_triangulate(mesh._handle, corner_values.detach().data_ptr(), ctypes.byref(V),
                                                ctypes.byref(T)))
            vertices = torch.empty((V.value, 3), dtype=torch.float32, device=device)
            indices = torch.empty(3 * T.value, dtype=torch.int32, device=device)
            _lib.check(lib.psr_mesh_copy(mesh._handle, vertices.data_ptr(), indices.data_ptr()))
            if corner_values.requires_grad:
                vertices = _vertex_function().apply(corner_values, vertices, mesh, stream)
        else:
            vertices = torch.empty((0, 3), dtype=torch.float32, device=device, requires_grad=tracked)
            indices = torch.empty(0, dtype=torch.int32, device=device)
    metadata.update(field_evaluations=samples, vertices=int(vertices.shape[0]), triangles=int(indices.shape[0]) // 3)
    return Mesh(vertices, indices, metadata)


def extract_mesh(reconstruction, *, resolution, bounds, iso_value=None, differentiable=False, **options):
    """Mesh a float32 reconstruction at its isovalue (meshing.py:283-346).

    With ``differentiable=True`` and a result of ``reconstruct_differentiable``
    the vertices carry gradients to the input positions and normals (and to a
    tracked ``iso_value`` tensor).  Cell selection, cases and connectivity are
    fixed.
    """
    torch = _torch()
    from .differentiable import DifferentiableReconstruction
    from .reconstruction import LevelField, Reconstruction
    full_backward = isinstance(reconstruction, DifferentiableReconstruction)
    if full_backward:
        if not differentiable:
            raise ValueError("a differentiable reconstruction requires differentiable=True for meshing")
        primal = reconstruction._primal
    elif isinstance(reconstruction, Reconstruction):
        primal = reconstruction
    elif isinstance(reconstruction, LevelField):
        primal = None
    else:
        raise TypeError("expected a Reconstruction, LevelField or differentiable reconstruction")
    level_field = primal.field if primal is not None else reconstruction
    if level_field.dtype != torch.float32:
        raise ValueError("mesh extraction currently requires float32 fields")
    if differentiable and not full_backward:
        raise ValueError("differentiable meshing requires a differentiable reconstruction; "
                         "use reconstruct_differentiable for the reconstruction backward pass")
    device = level_field.device
    if "device" in options and torch.device(options["device"]) != device:
        raise ValueError("mesh device must match the field device")
    options["device"] = device
    explicit_iso = iso_value is not None
    if iso_value is None:
        iso_value = level_field.iso_value
    if isinstance(iso_value, torch.Tensor):
        if iso_value.dtype != torch.float32 or tuple(iso_value.shape) != (1,) or iso_value.device != device:
            raise ValueError("iso_value must be one float32 value on the field device")
        if iso_value.requires_grad and not differentiable:
            raise ValueError("tracked iso_value requires differentiable=True")
        iso = iso_value
    else:
        if not math.isfinite(float(iso_value)):
            raise ValueError("iso_value must be finite")
        iso = torch.full((1,), float(iso_value), dtype=torch.float32, device=device)

    def untracked(points):
        values, _, _ = level_field.evaluate(points)
        return values - iso.detach()

    if full_backward:
        def sample(points):
            if explicit_iso:
                values, _, _ = reconstruction.evaluate(points)
                return values - iso
            values, _, _ = reconstruction.evaluate(points, centered=True)
            return values
        if options.get("backend", "sparse") == "sparse":
            options.setdefault("classification_field", untracked)
    else:
        sample = untracked
    mesh = extract_isosurface(sample, resolution=resolution, bounds=bounds, **options)
    if differentiable and not len(mesh.vertices) and not mesh.vertices.requires_grad:
        mesh.vertices.requires_grad_(True)
    return mesh
