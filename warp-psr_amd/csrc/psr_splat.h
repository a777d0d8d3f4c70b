// Splat helpers shared by the forward splats and their adjoints: Warp's
// scalar min/max/clamp forms, the leaf cell index and one normal-splat record
// (src/warp_psr/normals.py:38-104).
#pragma once
#include "psr_result.h"

namespace psr {

// Warp's scalar min/max/clamp (a<b?a:b forms), kept for identical NaN/tie behaviour.
template <class T> PSR_HD T wmax(T a, T b) { return a > b ? a : b; }
template <class T> PSR_HD T wmin(T a, T b) { return a < b ? a : b; }
template <class T> PSR_HD T wclamp(T x, T a, T b) { return wmin(wmax(a, x), b); }

template <class T> PSR_HD int cell_index(T p, int n) {  // max(0, ceil(p*n)-1)
  return imax(0, (int)ceil(p * T(n)) - 1);
}

// normals.py:46-104 (_splats): one of the 54 records of aggregate `src`.
// `scale` = wgt/h^3 * fraction * basis, the factor each normal component is
// multiplied by (used by the adjoint; the values keep the forward's order).
template <class T> struct NormalRecord {
  bool valid;
  int top, x, y, z;
  T v0, v1, v2;
  double scale;
};

template <class T>
__device__ __forceinline__ NormalRecord<T> normal_record(int src, int k, const T* __restrict__ sums,
                                                         const T* __restrict__ nsums, const T* __restrict__ weights,
                                                         const T* __restrict__ dw, int minimum, int maximum, bool values) {
  NormalRecord<T> rec;
  rec.valid = false;
  const T w = weights[src];
  T n0 = nsums[3 * src] / w, n1 = nsums[3 * src + 1] / w, n2 = nsums[3 * src + 2] / w;
  const T length = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (!(length > T(0))) return rec;
  n0 = n0 / length * w; n1 = n1 / length * w; n2 = n2 / length * w;
  const T p0 = sums[3 * src] / w, p1 = sums[3 * src + 1] / w, p2 = sums[3 * src + 2] / w;
  const T depth = wclamp(dw[2 * src], T(minimum), T(maximum));
  int top = (int)ceil(depth);
  T dx = T(1) - (T(top) - depth);
  if (top <= minimum) { top = minimum; dx = T(1); }
  if (top > maximum) { top = maximum; dx = T(1); }
  const T fraction = k >= 27 ? T(1) - dx : dx;
  if (!(k < 27 || fabs(1.0 - double(dx)) > 1e-06)) return rec;
  const int n = 1 << top;
  const T h = T(1) / T(n);
  const int cx = cell_index(p0, n), cy = cell_index(p1, n), cz = cell_index(p2, n);
  const int qk = k % 27;
  const int x = qk / 9, y = qk / 3 % 3, z = qk % 3;
  const int tx = cx + x - 1, ty = cy + y - 1, tz = cz + z - 1;
  if (!(imin(imin(tx, ty), tz) >= 0 && imax(imax(tx, ty), tz) < n)) return rec;
  rec.valid = true;
  rec.top = top; rec.x = tx; rec.y = ty; rec.z = tz;
  if (values) {
    const Quad bx = quadratic(double((p0 - T(cx) * h) / h));
    const Quad by = quadratic(double((p1 - T(cy) * h) / h));
    const Quad bz = quadratic(double((p2 - T(cz) * h) / h));
    const T wgt = dw[2 * src + 1];
    const T hhh = h * h * h;
    const T basis = T(quad_at(bx, x) * quad_at(by, y) * quad_at(bz, z));
    rec.v0 = n0 * wgt / hhh * fraction * basis;
    rec.v1 = n1 * wgt / hhh * fraction * basis;
    rec.v2 = n2 * wgt / hhh * fraction * basis;
    rec.scale = double(wgt) / double(hhh) * double(fraction) * double(basis);
  }
  return rec;
}

}  // namespace psr
