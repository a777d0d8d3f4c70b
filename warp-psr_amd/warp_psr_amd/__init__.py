"""MI355X screened Poisson surface reconstruction with warp-psr's API.

``reconstruct(points, normals, depth=..., ...)`` takes (n, 3) float32/float64
PyTorch tensors on one ROCm device and returns a :class:`Reconstruction` with
``iso_value``, ``evaluate(queries)`` and ``device_field()``, like
``warp_psr.reconstruct`` (src/warp_psr/reconstruction.py:137-349).  All compute
runs in hand-written HIP kernels of ``libwarp_psr_amd.so`` through its C ABI;
there is no CPU fallback.  ``reconstruct_differentiable`` differentiates field
losses with respect to the input positions and normals
(reconstruction_adjoint.py:61-185); ``meshing.extract_mesh`` extracts
marching-cubes meshes, differentiable through the vertices (meshing.py).
"""
from .differentiable import DifferentiableReconstruction, reconstruct_differentiable
from .reconstruction import reconstruct, reconstruct_batch, Reconstruction, LevelField, DeviceField

__all__ = ["reconstruct", "reconstruct_batch", "reconstruct_differentiable", "Reconstruction", "DifferentiableReconstruction", "LevelField",
           "DeviceField"]
