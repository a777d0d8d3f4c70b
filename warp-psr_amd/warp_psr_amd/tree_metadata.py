"""Tree metadata of a reconstruction (reference src/warp_psr/tree_metadata.py).

``build_tree_metadata(reconstruction)`` materialises on the GPU every retained
brood of every level (the active support nodes and the broods of the finer
level's parents, the samples, the density splat records and the normal
records), in global lattice coordinates after two padding levels, with each
node's flags, ghost mark, parent and children (tree_metadata.py:95-153).  The
reference takes the reconstruction's tree, samples, density and normals
diagnostics; here the reconstruction itself carries them.
"""
import ctypes
import time
from dataclasses import dataclass

from . import _lib
from .reconstruction import _stream_handle, _torch


@dataclass
class MetadataLevel:
    local_depth: int
    keys: object      # global lattice coordinates, tree-key encoded (int64)
    flags: object     # uint8
    ghost: object     # uint8
    parents: object   # int32 indices into the previous level; -1 for the root
    children: object  # (n, 8) int32 next-level indices, x bit first


@dataclass
class TreeMetadata:
    levels: list
    host_sync_seconds: float
    depth_offset: int = 2


def build_tree_metadata(reconstruction):
    torch = _torch()
    primal = getattr(reconstruction, "_primal", reconstruction)
    handle = primal._owner._handle
    device = primal.field.device
    lib = _lib.load()
    with torch.cuda.device(device):
        stream = _stream_handle(device)
        meta = ctypes.c_void_p()
        start = time.perf_counter()
        _lib.check(lib.psr_tree_metadata(handle, stream, ctypes.byref(meta)))
        sync = time.perf_counter() - start
        try:
            count = ctypes.c_int32()
            _lib.check(lib.psr_tree_metadata_levels(meta, ctypes.byref(count)))
            levels = []
            for g in range(count.value):
                depth, n = ctypes.c_int32(), ctypes.c_int64()
                _lib.check(lib.psr_tree_metadata_level(meta, g, ctypes.byref(depth), ctypes.byref(n)))
                n = n.value
                keys = torch.empty(n, dtype=torch.int64, device=device)
                flags = torch.empty(n, dtype=torch.uint8, device=device)
                ghost = torch.empty(n, dtype=torch.uint8, device=device)
                parents = torch.empty(n, dtype=torch.int32, device=device)
                children = torch.empty((n, 8), dtype=torch.int32, device=device)
                _lib.check(lib.psr_tree_metadata_copy(meta, g, keys.data_ptr(), flags.data_ptr(), ghost.data_ptr(),
                                                      parents.data_ptr(), children.data_ptr(), stream))
                levels.append(MetadataLevel(depth.value, keys, flags, ghost, parents, children))
            torch.cuda.current_stream(device).synchronize()
        finally:
            lib.psr_tree_metadata_free(meta)
    return TreeMetadata(levels, sync)
