"""Differentiable reconstruction: field losses back to the input normals.

Mirrors ``warp_psr.reconstruct_differentiable`` (src/warp_psr/reconstruction_adjoint.py:61-185)
over PyTorch autograd instead of a Warp tape: the primal ``reconstruct`` runs
once, and every ``evaluate(queries, centered=...)`` is a ``torch.autograd.Function``
whose backward runs the library's hand-written reverse kernels
(``psr_field_vjp`` then ``psr_backward``): query values and model-space
gradients -> coefficients and cumulative coefficients -> isovalue -> cascadic
solve (reverse Gauss-Seidel sweeps, base multigrid) -> normal loads -> normal
splats -> unit normals -> input normals.  Discrete choices (octree, sample
groups, splat depths, solver branches) are held fixed, as in the reference.

Gradients with respect to positions go through the isovalue centroids, the
splats' basis, depth and weight, the density levels and aggregates, the
sample transform's extrema owners and, in screened runs, the constraint
points: their screening terms in every level's operator, cross operator and
point load, reached by replaying each relaxation once for the x it saw, the
Galerkin products of the base hierarchy and the normal average's weight.  Queries may require
grad (d value / d query is the field gradient; d gradient / d query uses the
trilinear second derivatives).
"""
import ctypes

from . import _lib
from .reconstruction import _stream_handle, _torch, reconstruct


class _FieldQuery:
    """Namespace holding the lazily created autograd Function."""

    fn = None


def _function():
    if _FieldQuery.fn is not None:
        return _FieldQuery.fn
    torch = _torch()

    class FieldQuery(torch.autograd.Function):
        @staticmethod
        def forward(ctx, positions, normals, queries, recon, centered):
            values, gradients, valid = recon._primal.evaluate(queries.detach())
            if centered:
                values = values - recon.iso_value
            ctx.recon = recon
            ctx.centered = centered
            ctx.save_for_backward(queries)
            ctx.mark_non_differentiable(valid)
            return values, gradients, valid

        @staticmethod
        def backward(ctx, gvalue, ggrad, gvalid):
            (queries,) = ctx.saved_tensors
            recon = ctx.recon
            dev = queries.device
            count = recon._primal.field._field.count
            gcoef = torch.zeros(count, dtype=torch.float64, device=dev)
            gcum = torch.zeros(count, dtype=torch.float64, device=dev)
            q = queries.detach().contiguous()
            gquery = torch.zeros_like(q) if ctx.needs_input_grad[2] else None
            gv = gvalue.contiguous() if gvalue is not None else None
            gg = ggrad.contiguous() if ggrad is not None else None
            lib = _lib.load()
            stream = _stream_handle(dev)
            handle = recon._primal._owner._handle
            want_p, want_n = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
            gtransform = torch.zeros(4, dtype=torch.float64, device=dev) if want_p else None
            _lib.check(lib.psr_field_vjp_inputs(
                handle, q.data_ptr(), q.shape[0], gv.data_ptr() if gv is not None else None,
                gg.data_ptr() if gg is not None else None, gcoef.data_ptr(), gcum.data_ptr(),
                gquery.data_ptr() if gquery is not None else None,
                gtransform.data_ptr() if gtransform is not None else None, stream))
            gpositions = gnormals = None
            if want_p or want_n:
                iso_seed = -float(gv.sum().item()) if (ctx.centered and gv is not None) else 0.0
                positions = recon._positions.detach().contiguous()
                normals = recon._normals.detach().contiguous()
                if want_p:
                    gpositions = torch.zeros_like(positions)
                if want_n:
                    gnormals = torch.zeros_like(normals)
                _lib.check(lib.psr_backward_inputs(
                    handle, gcoef.data_ptr(), gcum.data_ptr(), ctypes.c_double(iso_seed),
                    gtransform.data_ptr() if gtransform is not None else None, positions.data_ptr(),
                    normals.data_ptr(), normals.shape[0], gpositions.data_ptr() if want_p else None,
                    gnormals.data_ptr() if want_n else None, stream))
            return gpositions, gnormals, gquery, None, None

    _FieldQuery.fn = FieldQuery
    return FieldQuery


class DifferentiableReconstruction:
    """Reference DifferentiableReconstruction (reconstruction_adjoint.py:44-58)."""

    def __init__(self, primal, positions, normals):
        self._primal = primal
        self._positions = positions
        self._normals = normals
        self.metadata = dict(primal.metadata)
        self.metadata["differentiability"] = {
            "mode": "finite_algorithm_fixed_assignments", "inputs": ["positions", "normals"],
            "constant_options": ["point_weight", "scale", "samples_per_node", "cg_accuracy", "depths", "iterations"],
            "first_order_only": True}
        self.diagnostics = primal.diagnostics

    @property
    def iso_value(self):
        return self._primal.iso_value

    def evaluate(self, queries, *, centered=False):
        """(values, model-space gradients, valid) connected to the autograd graph."""
        torch = _torch()
        if (not isinstance(queries, torch.Tensor) or queries.dtype != self._normals.dtype
                or queries.device != self._normals.device or queries.ndim != 2 or queries.shape[1] != 3):
            raise ValueError("queries must match field precision and device")
        return _function().apply(self._positions, self._normals, queries, self, bool(centered))

    def device_field(self):
        return self._primal.device_field()


def reconstruct_differentiable(positions, normals, **options):
    """Reconstruct, then differentiate field losses with respect to the inputs
    (reconstruction_adjoint.py:61-185).

    ``positions`` and/or ``normals`` must require grad.  Options are
    reconstruct()'s keywords (constants).
    """
    torch = _torch()
    if not isinstance(positions, torch.Tensor) or not isinstance(normals, torch.Tensor):
        raise TypeError("positions and normals must be tensors")
    if not positions.requires_grad and not normals.requires_grad:
        raise ValueError("track positions and/or normals with requires_grad=True")
    for name, value in options.items():
        if isinstance(value, torch.Tensor):
            raise ValueError(f"{name} must be a constant scalar; only positions and normals support gradients")
    primal = reconstruct(positions.detach(), normals.detach(), **options)
    return DifferentiableReconstruction(primal, positions, normals)
