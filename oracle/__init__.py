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
