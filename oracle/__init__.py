"""Test infrastructure only: ctypes wrapper over the serial CPU parity oracle.

The oracle (``psr_oracle.cpp``) restates warp-psr's forward pipeline on the CPU
with the reference's arithmetic order.  Only ``tests/``, ``__graft_entry__.smoke``
(as the checker) and ``bench.py``'s ``cpu_baseline`` leg use it; the product
package never imports it.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIBS = {}

_INT_ARRAYS = {
    "input_status", "input_group_indices", "input_order", "first_indices", "group_offsets",
    "normal_offsets", "matrix_offsets", "matrix_columns", "cross_offsets", "cross_columns",
    "prolong_offsets", "prolong_columns", "cg_runs",
}
_I64_ARRAYS = {"sample_keys", "aggregate_keys", "density_keys", "normal_keys", "tree_keys", "basis_keys", "keys"}
_F64_ARRAYS = {"normal_load", "ordered_normal_load", "point_load"}


def build():
    """Compile the oracle libraries with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def _lib(openmp=False):
    name = "libpsr_oracle_omp.so" if openmp else "libpsr_oracle.so"
    if name not in _LIBS:
        path = _HERE / "_build" / name
        if not path.exists():
            build()
        lib = ctypes.CDLL(str(path))
        lib.oracle_reconstruct.restype = ctypes.c_int
        lib.oracle_reconstruct.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_free.argtypes = [ctypes.c_void_p]
        lib.oracle_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_evaluate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_relax_sliced.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_int]
        lib.oracle_relax_scheduled.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_hat_integrals.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p]
        lib.oracle_normal_integral.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p]
        _LIBS[name] = lib
    return _LIBS[name]


class OracleResult:
    """Stage arrays and field of one oracle reconstruction."""

    def __init__(self, handle, dtype, params, lib):
        self._h = handle
        self.dtype = dtype
        self.params = params
        self._lib = lib

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.oracle_free(self._h)
            self._h = None

    def array(self, name, level=-1):
        data = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        if self._lib.oracle_array(self._h, name.encode(), level, ctypes.byref(data), ctypes.byref(nbytes)):
            raise KeyError(self._lib.oracle_last_error().decode())
        if name in _INT_ARRAYS:
            dt = np.int32
        elif name in _I64_ARRAYS:
            dt = np.int64
        elif name in _F64_ARRAYS:
            dt = np.float64
        else:
            dt = self.dtype
        count = nbytes.value // np.dtype(dt).itemsize
        if not count:
            return np.zeros(0, dtype=dt)
        buf = (ctypes.c_char * nbytes.value).from_address(data.value)
        return np.frombuffer(buf, dtype=dt, count=count).copy()

    @property
    def iso_value(self):
        return float(self.array("iso")[0])

    @property
    def basis_counts(self):
        return [len(self.array("basis_keys", d)) for d in range(self.params["depth"] + 1)]

    def evaluate(self, queries):
        q = np.ascontiguousarray(queries, dtype=self.dtype).reshape(-1, 3)
        m = len(q)
        values = np.empty(m, dtype=self.dtype)
        grads = np.empty((m, 3), dtype=self.dtype)
        valid = np.empty(m, dtype=np.int32)
        self._lib.oracle_evaluate(self._h, q.ctypes.data, m, values.ctypes.data, grads.ctypes.data, valid.ctypes.data)
        return values, grads, valid


def resolve_params(depth=6, full_depth=3, base_depth=None, solve_depth=None, kernel_depth=None,
                   samples_per_node=1.5, scale=1.1, point_weight=2.0, exact_interpolation=False,
                   iters=8, base_v_cycles=1, cg_accuracy=1e-3):
    """Defaults as src/warp_psr/reconstruction.py:160-165."""
    if base_depth is None:
        base_depth = full_depth
    if solve_depth is None:
        solve_depth = depth
    if kernel_depth is None:
        kernel_depth = max(depth - 2, 0)
    return dict(depth=depth, full_depth=full_depth, base_depth=base_depth, solve_depth=solve_depth,
                kernel_depth=kernel_depth, samples_per_node=samples_per_node, scale=scale,
                point_weight=point_weight, exact_interpolation=bool(exact_interpolation), iters=iters,
                base_v_cycles=base_v_cycles, cg_accuracy=cg_accuracy)


def reconstruct(points, normals, *, openmp=False, mixed=False, **kwargs):
    """Run the oracle on (n,3) float32/float64 numpy arrays.

    mixed=True (float64 inputs only) rounds the finer levels' operator values
    to float32, as the device's mixed-precision mode stores them."""
    p = resolve_params(**kwargs)
    points = np.ascontiguousarray(points)
    dtype = np.float64 if points.dtype == np.float64 else np.float32
    if mixed and dtype != np.float64:
        raise ValueError("mixed precision runs on float64 inputs")
    points = np.ascontiguousarray(points, dtype=dtype).reshape(-1, 3)
    normals = np.ascontiguousarray(normals, dtype=dtype).reshape(-1, 3)
    lib = _lib(openmp)
    out = ctypes.c_void_p()
    rc = lib.oracle_reconstruct(points.ctypes.data, normals.ctypes.data, len(points), 1 if dtype == np.float64 else 0,
                                p["depth"], p["full_depth"], p["base_depth"], p["solve_depth"], p["kernel_depth"],
                                p["samples_per_node"], p["scale"], p["point_weight"], int(p["exact_interpolation"]),
                                p["iters"], p["base_v_cycles"], p["cg_accuracy"], int(bool(mixed)), ctypes.byref(out))
    if rc == 2:
        raise ValueError(lib.oracle_last_error().decode())
    if rc:
        raise RuntimeError(lib.oracle_last_error().decode())
    return OracleResult(out.value, dtype, p, lib)


def relax_sliced(keys, offsets, columns, values, b, x, depth, iters):
    """Serial sliced Gauss-Seidel on caller CSR arrays; returns the new x."""
    values = np.ascontiguousarray(values)
    dtype = values.dtype
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    columns = np.ascontiguousarray(columns, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.array(x, dtype=dtype, copy=True)
    _lib().oracle_relax_sliced(1 if dtype == np.float64 else 0, keys.ctypes.data, len(keys), offsets.ctypes.data,
                               columns.ctypes.data, values.ctypes.data, b.ctypes.data, x.ctypes.data, depth, iters)
    return x


def relax_scheduled(keys, offsets, columns, values, b, x, depth, iters, forward=True, z=-1):
    """Serial 8-colour Gauss-Seidel (or one slice of the sliced schedule) on caller CSR arrays; returns the new x."""
    values = np.ascontiguousarray(values)
    dtype = values.dtype
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    columns = np.ascontiguousarray(columns, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.array(x, dtype=dtype, copy=True)
    _lib().oracle_relax_scheduled(1 if dtype == np.float64 else 0, keys.ctypes.data, len(keys), offsets.ctypes.data,
                                  columns.ctypes.data, values.ctypes.data, b.ctypes.data, x.ctypes.data, depth, iters,
                                  int(bool(forward)), int(z))
    return x


def hat_integrals(depth, index, trial_depth, trial_index):
    out = np.zeros(2)
    _lib().oracle_hat_integrals(depth, index, trial_depth, trial_index, out.ctypes.data)
    return out


def normal_integral(test_depth, test_index, normal_depth, normal_index):
    out = np.zeros(2)
    _lib().oracle_normal_integral(test_depth, test_index, normal_depth, normal_index, out.ctypes.data)
    return out


def relax_sliced_tiled(keys, offsets, columns, values, b, x, depth, iters, tile, seed=0, plant=0):
    """The tile-parallel form of the sliced schedule (tiles of each wave in a
    seeded random order, two value buffers); returns the new x.  Equals
    relax_sliced bit for bit on degree-one operators.  ``plant`` (tests only)
    injects a known mistake: 1 = one halo ring short, 2 = one value buffer."""
    values = np.ascontiguousarray(values)
    dtype = values.dtype
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    columns = np.ascontiguousarray(columns, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.array(x, dtype=dtype, copy=True)
    lib = _lib()
    fn = lib.oracle_relax_sliced_tiled
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    if fn(1 if dtype == np.float64 else 0, keys.ctypes.data, len(keys), offsets.ctypes.data, columns.ctypes.data,
          values.ctypes.data, b.ctypes.data, x.ctypes.data, depth, iters, tile, seed, plant):
        raise RuntimeError(lib.oracle_last_error().decode())
    return x


# ---- known-answer hooks (stand-alone pipeline pieces on caller data)

def _code(dtype):
    return 1 if np.dtype(dtype) == np.float64 else 0


def field_evaluate(keys, coefficients, depth, linear, translation, queries, cumulative=None):
    """field.py:364-421 on a caller field; returns (values, gradients, valid)."""
    coefficients = np.ascontiguousarray(coefficients)
    dt = coefficients.dtype
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    q = np.ascontiguousarray(queries, dtype=dt)
    cum = None if cumulative is None else np.ascontiguousarray(cumulative, dtype=dt)
    m = len(q)
    values = np.empty(m, dt)
    grads = np.empty((m, 3), dt)
    valid = np.empty(m, np.int32)
    lin = np.ascontiguousarray(np.asarray(linear, np.float64).reshape(9))
    tr = np.ascontiguousarray(np.asarray(translation, np.float64).reshape(3))
    fn = _lib().oracle_field_evaluate
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p]
    fn(_code(dt), keys.ctypes.data, len(keys), coefficients.ctypes.data, None if cum is None else cum.ctypes.data, depth,
       lin.ctypes.data, tr.ctypes.data, q.ctypes.data, m, values.ctypes.data, grads.ctypes.data, valid.ctypes.data)
    return values, grads, valid


def cumulative(keys, coefficients, depth):
    """field.py:187-210 (cumulative_coefficients) on caller keys and coefficients."""
    coefficients = np.ascontiguousarray(coefficients)
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    out = np.empty_like(coefficients)
    fn = _lib().oracle_cumulative
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    fn(_code(coefficients.dtype), keys.ctypes.data, len(keys), coefficients.ctypes.data, depth, out.ctypes.data)
    return out


def isovalue(values, weights):
    """reconstruction.py:76-116 (_compute_isovalue) on caller values and weights."""
    values = np.ascontiguousarray(values)
    weights = np.ascontiguousarray(weights, dtype=values.dtype)
    out = np.zeros(1, values.dtype)
    fn = _lib().oracle_isovalue
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    fn(_code(values.dtype), values.ctypes.data, weights.ctypes.data, len(values), out.ctypes.data)
    return out[0]


def screening_dense(positions, weights, depth):
    """screening_stencils.py:80-101 on a uniform level (all nodes of `depth`,
    keys in sorted order): the dense screening matrix and the point load."""
    positions = np.ascontiguousarray(positions)
    weights = np.ascontiguousarray(weights, dtype=positions.dtype)
    k = ((1 << depth) + 1) ** 3
    dense = np.zeros((k, k))
    load = np.zeros(k)
    lib = _lib()
    fn = lib.oracle_screening_dense
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p]
    if fn(_code(positions.dtype), positions.ctypes.data, weights.ctypes.data, len(positions), depth, dense.ctypes.data,
          load.ctypes.data):
        raise RuntimeError(lib.oracle_last_error().decode())
    return dense, load


def _csr_args(offsets, columns, values):
    return (np.ascontiguousarray(offsets, dtype=np.int32), np.ascontiguousarray(columns, dtype=np.int32),
            np.ascontiguousarray(values))


def root_padded_apply(keys, offsets, columns, values, v):
    """solver.py:167-238: the 64-row padded root system applied with the
    constant-mode term to v (64 values)."""
    offsets, columns, values = _csr_args(offsets, columns, values)
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=values.dtype)
    out = np.empty(64, values.dtype)
    fn = _lib().oracle_root_padded_apply
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5
    fn(_code(values.dtype), keys.ctypes.data, len(keys), offsets.ctypes.data, columns.ctypes.data, values.ctypes.data,
       v.ctypes.data, out.ctypes.data)
    return out


def root_solve(keys, offsets, columns, values, b, iters=8, cycles=1, accuracy=1e-3, dc=True):
    """solver.py:241-317 with base depth 0 on a caller root operator."""
    offsets, columns, values = _csr_args(offsets, columns, values)
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=values.dtype)
    x = np.zeros(len(keys), values.dtype)
    fn = _lib().oracle_root_solve
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [
        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    fn(_code(values.dtype), keys.ctypes.data, len(keys), offsets.ctypes.data, columns.ctypes.data, values.ctypes.data,
       b.ctypes.data, x.ctypes.data, iters, cycles, accuracy, int(bool(dc)))
    return x
