/* Header-only HIP device evaluation of a reconstructed field.
 *
 * psr_evaluate_field(field, point, &value, gradient) evaluates the implicit
 * function of a PsrField view (psr_result_field) at one model-space point from
 * inside any HIP kernel, with the same bits as psr_evaluate.  It mirrors the
 * reference's application-kernel entry point evaluate_field(field, point)
 * (src/warp_psr/device_field.py:91-112) with the reference evaluation
 * (src/warp_psr/field.py:289-359): leaf search from the finest level, the
 * 1e-6 cube-boundary nudge, and coefficient(leaf) + cumulative(leaf-1) +
 * coefficient(leaf+1), with the gradient mapped back to model space.  A view
 * whose cumulative pointer is NULL (a LevelField built without cumulative
 * coefficients) takes the reference's direct level summation
 * (src/warp_psr/field.py:37-75) instead, as the reference DeviceField does
 * (src/warp_psr/device_field.py:31-41).
 *
 * Returns 1 for a valid query; 0 outside the unit cube (value and gradient
 * NaN).  T must match field.dtype (float for 0, double for 1).  The field's
 * arrays are borrowed: keep the owning result alive and unmodified.
 * Build with -ffp-contract=off for bit-identical results.
 */
#ifndef WARP_PSR_AMD_FIELD_H
#define WARP_PSR_AMD_FIELD_H

#include <hip/hip_runtime.h>

#include "warp_psr_amd.h"

namespace psr_field_detail {

__device__ __forceinline__ int64_t pack(int d, int x, int y, int z) {
  return ((int64_t)d << 48) | ((int64_t)x << 32) | ((int64_t)y << 16) | (int64_t)z;
}

/* index of key in sorted keys[lo, hi), or -1 */
__device__ __forceinline__ int64_t find(const int64_t* keys, int64_t lo, int64_t hi, int64_t key) {
  const int64_t end = hi;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < end && keys[lo] == key) ? lo : -1;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* field.py:371-393: one level's trilinear value and cube-space derivatives */
template <class T>
__device__ __forceinline__ void add_level(const PsrField& f, const T* coef, const T p[3], int d, T res[4]) {
  const int n = 1 << d;
  int cell[3], base[3];
  T fr[3];
  for (int a = 0; a < 3; ++a) {
    const T t = p[a] * T(n);
    cell[a] = clampi((int)ceil(t) - 1, 0, n - 1);
    base[a] = (int)floor(t);
    fr[a] = t - T(base[a]);
  }
  const int64_t lo = f.level_offsets[d], hi = f.level_offsets[d + 1];
  for (int corner = 0; corner < 8; ++corner) {
    const int node[3] = {cell[0] + corner / 4, cell[1] + (corner / 2) % 2, cell[2] + corner % 2};
    const int m0 = node[0] - base[0], m1 = node[1] - base[1], m2 = node[2] - base[2];
    if ((m0 < m1 ? (m0 < m2 ? m0 : m2) : (m1 < m2 ? m1 : m2)) < 0) continue;
    const int64_t idx = find(f.keys, lo, hi, pack(d, node[0], node[1], node[2]));
    if (idx < 0) continue;
    T h[3], sg[3];
    for (int a = 0; a < 3; ++a) {
      h[a] = T(1) - fr[a];
      sg[a] = T(-1);
      if (node[a] > base[a]) { h[a] = fr[a]; sg[a] = T(1); }
    }
    const T c = coef[idx];
    res[0] = res[0] + c * h[0] * h[1] * h[2];
    res[1] = res[1] + c * T(n) * sg[0] * h[1] * h[2];
    res[2] = res[2] + c * T(n) * h[0] * sg[1] * h[2];
    res[3] = res[3] + c * T(n) * h[0] * h[1] * sg[2];
  }
}

/* field.py:37-75 (evaluation_function): direct summation over levels
 * 0..depth of a field without cumulative coefficients; left-owned cells, so
 * knots take left derivatives; no boundary nudge.  res = value and the
 * cube-space gradient. */
template <class T>
__device__ __forceinline__ void direct_levels(const PsrField& f, const T* coef, const T p[3], T res[4]) {
  for (int d = 0; d <= f.depth; ++d) {
    const int n = 1 << d;
    int cell[3];
    T fr[3];
    for (int a = 0; a < 3; ++a) {
      const T t = p[a] * T(n);
      cell[a] = clampi((int)ceil(t) - 1, 0, n - 1);
      fr[a] = t - T(cell[a]);
    }
    const int64_t lo = f.level_offsets[d], hi = f.level_offsets[d + 1];
    for (int corner = 0; corner < 8; ++corner) {
      const int o[3] = {corner / 4, (corner / 2) % 2, corner % 2};
      const int64_t idx = find(f.keys, lo, hi, pack(d, cell[0] + o[0], cell[1] + o[1], cell[2] + o[2]));
      if (idx < 0) continue;
      T h[3], sg[3];
      for (int a = 0; a < 3; ++a) {
        h[a] = o[a] ? fr[a] : T(1) - fr[a];
        sg[a] = o[a] ? T(1) : T(-1);
      }
      const T c = coef[idx];
      res[0] = res[0] + c * h[0] * h[1] * h[2];
      const T cn = c * T(n);
      res[1] = res[1] + cn * (sg[0] * h[1] * h[2]);
      res[2] = res[2] + cn * (h[0] * sg[1] * h[2]);
      res[3] = res[3] + cn * (h[0] * h[1] * sg[2]);
    }
  }
}

}  // namespace psr_field_detail

template <class T>
__device__ __forceinline__ int psr_evaluate_field(const PsrField& f, const T point[3], T* value, T gradient[3]) {
  using namespace psr_field_detail;
  T L[9], t[3];
  for (int k = 0; k < 9; ++k) L[k] = T(f.linear[k]);
  for (int k = 0; k < 3; ++k) t[k] = T(f.translation[k]);
  T p[3];
  for (int r = 0; r < 3; ++r) p[r] = L[3 * r] * point[0] + L[3 * r + 1] * point[1] + L[3 * r + 2] * point[2] + t[r];
  const T lo = p[1] < p[0] ? p[1] : p[0], lo3 = p[2] < lo ? p[2] : lo;
  const T hi = p[0] < p[1] ? p[1] : p[0], hi3 = hi < p[2] ? p[2] : hi;
  const bool ok = isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && lo3 >= T(0) && hi3 <= T(1);
  if (!ok) {
    *value = T(NAN);
    if (gradient) gradient[0] = gradient[1] = gradient[2] = T(NAN);
    return 0;
  }
  const T* coef = (const T*)f.coefficients;
  const T* cum = (const T*)f.cumulative;
  if (!cum) { /* a field without cumulative coefficients: direct summation */
    T res[4] = {T(0), T(0), T(0), T(0)};
    direct_levels(f, coef, p, res);
    *value = res[0];
    if (gradient)
      for (int c = 0; c < 3; ++c) gradient[c] = L[c] * res[1] + L[3 + c] * res[2] + L[6 + c] * res[3];
    return 1;
  }
  int leaf = 0;
  for (int d = f.depth; d >= 0; --d) {
    const int n = 1 << d;
    const int c0 = clampi((int)ceil(p[0] * T(n)) - 1, 0, n - 1);
    const int c1 = clampi((int)ceil(p[1] * T(n)) - 1, 0, n - 1);
    const int c2 = clampi((int)ceil(p[2] * T(n)) - 1, 0, n - 1);
    if (find(f.keys, f.level_offsets[d], f.level_offsets[d + 1], pack(d, c0, c1, c2)) >= 0) { leaf = d; break; }
  }
  for (int a = 0; a < 3; ++a) {  /* field.py:413-415 */
    if (p[a] == T(0)) p[a] = T(1e-6);
    else if (p[a] == T(1)) p[a] = T(1.0 - 1e-6);
  }
  T res[4] = {T(0), T(0), T(0), T(0)};
  add_level(f, coef, p, leaf, res);
  if (leaf > 0) {
    T r2[4] = {T(0), T(0), T(0), T(0)};
    add_level(f, cum, p, leaf - 1, r2);
    for (int a = 0; a < 4; ++a) res[a] = res[a] + r2[a];
  }
  if (leaf < f.depth) {
    T r2[4] = {T(0), T(0), T(0), T(0)};
    add_level(f, coef, p, leaf + 1, r2);
    for (int a = 0; a < 4; ++a) res[a] = res[a] + r2[a];
  }
  *value = res[0];
  if (gradient)
    for (int c = 0; c < 3; ++c) gradient[c] = L[c] * res[1] + L[3 + c] * res[2] + L[6 + c] * res[3];
  return 1;
}

#endif
