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

// Derivatives of one splat record's scale s = wgt / h^3 * fraction * basis(p)
// (its value is n'' * s) with respect to the centroid p = sums / w, the raw
// splat depth dw[0] (through fraction, zero where the clamp or the integer
// cases of normals.py:67-77 hold it) and the weight dw[1].  The discrete
// choices (validity, top, cell) are normal_record's, from the same T
// arithmetic; the derivatives are float64.
struct RecordJacobian {
  bool valid;
  int top, x, y, z;
  double n[3];  // n'' = normalised normal sum times w
  double s;     // wgt / h^3 * fraction * basis
  double ds_dp[3];
  double ds_ddepth;
  double ds_dwgt;
};

template <class T>
PSR_HD RecordJacobian normal_record_jacobian(int src, int k, const T* __restrict__ sums, const T* __restrict__ nsums,
                                             const T* __restrict__ weights, const T* __restrict__ dw, int minimum,
                                             int maximum) {
  RecordJacobian J;
  J.valid = false;
  const T w = weights[src];
  T n0 = nsums[3 * src] / w, n1 = nsums[3 * src + 1] / w, n2 = nsums[3 * src + 2] / w;
  const T length = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (!(length > T(0))) return J;
  n0 = n0 / length * w; n1 = n1 / length * w; n2 = n2 / length * w;
  const T p[3] = {sums[3 * src] / w, sums[3 * src + 1] / w, sums[3 * src + 2] / w};
  const T raw = dw[2 * src];
  const T depth = wclamp(raw, T(minimum), T(maximum));
  int top = (int)ceil(depth);
  T dx = T(1) - (T(top) - depth);
  bool fixed = !(raw >= T(minimum) && raw <= T(maximum));  // clamp adjoint: zero outside
  if (top <= minimum) { top = minimum; dx = T(1); fixed = true; }
  if (top > maximum) { top = maximum; dx = T(1); fixed = true; }
  const T fraction = k >= 27 ? T(1) - dx : dx;
  if (!(k < 27 || fabs(1.0 - double(dx)) > 1e-06)) return J;
  const int nn = 1 << top;
  const T h = T(1) / T(nn);
  const int c[3] = {cell_index(p[0], nn), cell_index(p[1], nn), cell_index(p[2], nn)};
  const int qk = k % 27;
  const int o[3] = {qk / 9, qk / 3 % 3, qk % 3};
  const int t[3] = {c[0] + o[0] - 1, c[1] + o[1] - 1, c[2] + o[2] - 1};
  if (!(imin(imin(t[0], t[1]), t[2]) >= 0 && imax(imax(t[0], t[1]), t[2]) < nn)) return J;
  J.valid = true;
  J.top = top; J.x = t[0]; J.y = t[1]; J.z = t[2];
  J.n[0] = double(n0); J.n[1] = double(n1); J.n[2] = double(n2);
  double tv[3], bv[3], db[3];
  for (int a = 0; a < 3; ++a) {
    tv[a] = double((p[a] - T(c[a]) * h) / h);
    const Quad q = quadratic(tv[a]);
    bv[a] = quad_at(q, o[a]);
    db[a] = quad_deriv(tv[a], o[a]) / double(h);
  }
  const double basis = bv[0] * bv[1] * bv[2];
  const double wgt = double(dw[2 * src + 1]);
  const double hhh = double(h) * double(h) * double(h);
  const double f = double(fraction);
  J.s = wgt / hhh * f * basis;
  J.ds_dp[0] = wgt / hhh * f * db[0] * bv[1] * bv[2];
  J.ds_dp[1] = wgt / hhh * f * bv[0] * db[1] * bv[2];
  J.ds_dp[2] = wgt / hhh * f * bv[0] * bv[1] * db[2];
  J.ds_ddepth = fixed ? 0.0 : wgt / hhh * basis * (k >= 27 ? -1.0 : 1.0);
  J.ds_dwgt = f * basis / hhh;
  return J;
}

}  // namespace psr
