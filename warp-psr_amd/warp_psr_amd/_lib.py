"""ctypes binding of libwarp_psr_amd.so (the C ABI in include/warp_psr_amd.h).

The library is required: importing the compute entry points without it raises,
there is no CPU fallback.
"""
import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libwarp_psr_amd.so"

EXPORTED = (
    "psr_params_default", "psr_reconstruct", "psr_evaluate", "psr_result_iso", "psr_result_field",
    "psr_result_metadata_json", "psr_result_array", "psr_result_free", "psr_relax_sliced", "psr_relax_scheduled", "psr_last_error",
    "psr_release_cache", "psr_field_evaluate", "psr_field_vjp", "psr_backward", "psr_field_vjp_inputs", "psr_backward_inputs",
    "psr_reconstruct_batch", "psr_mesh_begin", "psr_mesh_slab_nodes", "psr_mesh_slab_points", "psr_mesh_slab_values",
    "psr_mesh_select", "psr_mesh_corner_points", "psr_mesh_triangulate", "psr_mesh_copy", "psr_mesh_vertex_vjp",
    "psr_mesh_case_table", "psr_mesh_free", "psr_tree_metadata", "psr_tree_metadata_levels", "psr_tree_metadata_level",
    "psr_tree_metadata_copy", "psr_tree_metadata_free", "psr_field_level_offsets", "psr_isovalue",
    "psr_isovalue_noalloc",
)


class PsrParams(ctypes.Structure):
    _fields_ = [
        ("depth", ctypes.c_int), ("full_depth", ctypes.c_int), ("base_depth", ctypes.c_int),
        ("solve_depth", ctypes.c_int), ("kernel_depth", ctypes.c_int), ("samples_per_node", ctypes.c_double),
        ("scale", ctypes.c_double), ("point_weight", ctypes.c_double), ("exact_interpolation", ctypes.c_int),
        ("iters", ctypes.c_int), ("base_v_cycles", ctypes.c_int), ("cg_accuracy", ctypes.c_double),
        ("keep_intermediates", ctypes.c_int), ("float_operators", ctypes.c_int), ("relax_tile", ctypes.c_int),
    ]


class PsrField(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.c_void_p), ("coefficients", ctypes.c_void_p), ("cumulative", ctypes.c_void_p),
        ("level_offsets", ctypes.c_void_p), ("count", ctypes.c_int64), ("depth", ctypes.c_int),
        ("dtype", ctypes.c_int), ("linear", ctypes.c_double * 9), ("translation", ctypes.c_double * 3),
        ("iso_value", ctypes.c_double),
    ]


class LibraryMissing(RuntimeError):
    pass


_LIB = None


def load():
    """Load the HIP library; raise LibraryMissing when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise LibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, i64, c_int, c_double = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    lib.psr_params_default.argtypes = [ctypes.POINTER(PsrParams)]
    lib.psr_params_default.restype = None
    lib.psr_reconstruct.argtypes = [vp, vp, i64, c_int, ctypes.POINTER(PsrParams), vp, ctypes.POINTER(vp)]
    lib.psr_reconstruct.restype = c_int
    lib.psr_evaluate.argtypes = [vp, vp, i64, vp, vp, vp, vp]
    lib.psr_evaluate.restype = c_int
    lib.psr_result_iso.argtypes = [vp, ctypes.POINTER(c_double)]
    lib.psr_result_iso.restype = c_int
    lib.psr_result_field.argtypes = [vp, ctypes.POINTER(PsrField)]
    lib.psr_result_field.restype = c_int
    lib.psr_result_metadata_json.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.psr_result_metadata_json.restype = c_int
    lib.psr_result_array.argtypes = [vp, ctypes.c_char_p, c_int, ctypes.POINTER(vp), ctypes.POINTER(i64)]
    lib.psr_result_array.restype = c_int
    lib.psr_result_free.argtypes = [vp]
    lib.psr_result_free.restype = None
    lib.psr_relax_sliced.argtypes = [c_int, vp, i64, vp, vp, vp, vp, vp, c_int, c_int, vp]
    lib.psr_relax_scheduled.argtypes = [c_int, vp, i64, vp, vp, vp, vp, vp, c_int, c_int, c_int, c_int, vp]
    lib.psr_reconstruct_batch.restype = c_int
    lib.psr_reconstruct_batch.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i64), c_int, c_int,
                                          ctypes.POINTER(PsrParams), c_int, c_int, vp, ctypes.POINTER(vp)]
    lib.psr_field_vjp.restype = c_int
    lib.psr_field_vjp.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, vp]
    lib.psr_backward.restype = c_int
    lib.psr_backward.argtypes = [vp, vp, vp, c_double, vp, i64, vp, vp]
    lib.psr_backward_inputs.restype = c_int
    lib.psr_backward_inputs.argtypes = [vp, vp, vp, c_double, vp, vp, vp, i64, vp, vp, vp]
    lib.psr_field_vjp_inputs.restype = c_int
    lib.psr_field_vjp_inputs.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, vp, vp]
    lib.psr_field_evaluate.restype = c_int
    lib.psr_field_evaluate.argtypes = [ctypes.POINTER(PsrField), vp, i64, vp, vp, vp, vp]
    lib.psr_field_level_offsets.restype = c_int
    lib.psr_field_level_offsets.argtypes = [vp, i64, c_int, vp, vp]
    lib.psr_isovalue.restype = c_int
    lib.psr_isovalue.argtypes = [c_int, vp, vp, i64, vp, vp]
    lib.psr_isovalue_noalloc.restype = c_int
    lib.psr_isovalue_noalloc.argtypes = [c_int, vp, vp, i64, vp, vp]
    lib.psr_relax_sliced.restype = c_int
    lib.psr_relax_scheduled.restype = c_int
    i32p, i64p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(i64)
    lib.psr_mesh_begin.argtypes = [i32p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), vp,
                                   ctypes.POINTER(vp)]
    lib.psr_mesh_slab_nodes.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, i64p]
    lib.psr_mesh_slab_points.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
    lib.psr_mesh_slab_values.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
    lib.psr_mesh_select.argtypes = [vp, i64p, i32p]
    lib.psr_mesh_corner_points.argtypes = [vp, vp]
    lib.psr_mesh_triangulate.argtypes = [vp, vp, i64p, i64p]
    lib.psr_mesh_copy.argtypes = [vp, vp, vp]
    lib.psr_mesh_vertex_vjp.argtypes = [vp, vp, vp, vp]
    lib.psr_mesh_case_table.argtypes = [vp, vp]
    for name in ("psr_mesh_begin", "psr_mesh_slab_nodes", "psr_mesh_slab_points", "psr_mesh_slab_values", "psr_mesh_select",
                 "psr_mesh_corner_points", "psr_mesh_triangulate", "psr_mesh_copy", "psr_mesh_vertex_vjp",
                 "psr_mesh_case_table"):
        getattr(lib, name).restype = c_int
    lib.psr_tree_metadata.argtypes = [vp, vp, ctypes.POINTER(vp)]
    lib.psr_tree_metadata_levels.argtypes = [vp, i32p]
    lib.psr_tree_metadata_level.argtypes = [vp, ctypes.c_int32, i32p, i64p]
    lib.psr_tree_metadata_copy.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp, vp, vp]
    for name in ("psr_tree_metadata", "psr_tree_metadata_levels", "psr_tree_metadata_level", "psr_tree_metadata_copy"):
        getattr(lib, name).restype = c_int
    lib.psr_tree_metadata_free.argtypes = [vp]
    lib.psr_tree_metadata_free.restype = None
    lib.psr_mesh_free.argtypes = [vp]
    lib.psr_mesh_free.restype = None
    lib.psr_last_error.argtypes = []
    lib.psr_last_error.restype = ctypes.c_char_p
    lib.psr_release_cache.argtypes = []
    lib.psr_release_cache.restype = None
    _LIB = lib
    return lib


def check(rc):
    """Map C return codes to the reference's exception types."""
    if rc == 0:
        return
    message = load().psr_last_error().decode()
    if rc == 2:
        raise ValueError(message)
    if rc == 3:
        raise TypeError(message)
    raise RuntimeError(message)
