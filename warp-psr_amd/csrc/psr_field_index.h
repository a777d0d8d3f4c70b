// Lookup structure of a concatenated multi-level field (all levels' sorted
// basis keys with per-level offsets and optional (x, y) column tables), shared
// by the field stage and its adjoint.
#pragma once
#include "psr_result.h"

namespace psr {

// keys of level d occupy [off[d], off[d+1]) of the concatenated sorted keys;
// start/end (optional) are the column runs of every level
struct FieldIndex {
  const int64_t* keys;
  const int* off;  // device, depth+2 entries
  const int* start;
  const int* end;
};

// first cell of level d's column table: sum over e < d of (2^e + 1)^2
PSR_HD int64_t column_base(int d) { return (((int64_t)1 << (2 * d)) - 1) / 3 + 2 * (((int64_t)1 << d) - 1) + d; }

// Absolute indices of keys (d, x, y, z0 + k), k < len (-1 when absent).
template <int MAXLEN>
__device__ __forceinline__ void level_run(const FieldIndex& f, int d, int x, int y, int z0, int len, int64_t* out) {
  if (f.start) {
    ColumnTable t;
    t.start = f.start + column_base(d);
    t.end = f.end + column_base(d);
    t.dim = (1 << d) + 1;
    find_run_table<MAXLEN>(f.keys, 0, t, d, x, y, z0, len, out);
    return;
  }
  const int64_t lo = f.off[d];
  find_run<MAXLEN>(f.keys + lo, (int64_t)f.off[d + 1] - lo, d, x, y, z0, len, out);
#pragma unroll
  for (int k = 0; k < MAXLEN; ++k)
    if (k < len && out[k] >= 0) out[k] += lo;
}

template <class T> inline FieldIndex field_index(const Result<T>& r) {
  FieldIndex f;
  f.keys = r.keys.data();
  f.off = r.offsets_dev.data();
  f.start = r.column_start.empty() ? nullptr : r.column_start.data();
  f.end = r.column_start.empty() ? nullptr : r.column_end.data();
  return f;
}

// model -> cube affine map of queries
template <class T> struct Affine { T L[9]; T t[3]; };

}  // namespace psr
