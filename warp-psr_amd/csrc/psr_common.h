// Shared device/host helpers for the MI355X screened Poisson reconstruction.
// Keys, coordinates and the deterministic arithmetic helpers follow the
// reference's definitions (cited inline); every floating-point expression
// keeps the reference's operation order and the whole library is built with
// -ffp-contract=off, IEEE division/sqrt and denormals preserved.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace psr {

constexpr int64_t KMAX = INT64_MAX;

#define PSR_HD __host__ __device__ __forceinline__

struct PsrError : std::runtime_error {
  int code;  // 1 internal/HIP failure, 2 invalid input (ValueError), 3 type (TypeError)
  PsrError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      throw ::psr::PsrError(1, std::string("HIP error ") + hipGetErrorString(_e) + " at " +     \
                                   __FILE__ + ":" + std::to_string(__LINE__) + " in " #expr);    \
  } while (0)

#define KERNEL_CHECK() HIP_CHECK(hipGetLastError())

// ------------------------------------------------------------------ keys
// src/warp_psr/field.py:13-16 -- d<<48 | x<<32 | y<<16 | z
PSR_HD int64_t basis_key(int d, int x, int y, int z) {
  return ((int64_t)d << 48) | ((int64_t)x << 32) | ((int64_t)y << 16) | (int64_t)z;
}
PSR_HD int key_x(int64_t k) { return (int)((k >> 32) & 65535); }
PSR_HD int key_y(int64_t k) { return (int)((k >> 16) & 65535); }
PSR_HD int key_z(int64_t k) { return (int)(k & 65535); }
PSR_HD int key_d(int64_t k) { return (int)(k >> 48); }
// src/warp_psr/tree.py:16-24 -- signed 20-bit coordinates offset by 2^19
PSR_HD int64_t tree_key(int x, int y, int z) {
  return ((int64_t)(x + 524288) << 40) | ((int64_t)(y + 524288) << 20) | (int64_t)(z + 524288);
}
PSR_HD int tkey_x(int64_t k) { return (int)((k >> 40) & 1048575) - 524288; }
PSR_HD int tkey_y(int64_t k) { return (int)((k >> 20) & 1048575) - 524288; }
PSR_HD int tkey_z(int64_t k) { return (int)(k & 1048575) - 524288; }

// src/warp_psr/field.py:19-32 -- lower-bound binary search, -1 when absent
PSR_HD int64_t find_key(const int64_t* __restrict__ keys, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < n && keys[lo] == key) ? lo : -1;
}
PSR_HD int64_t lower_bound(const int64_t* __restrict__ keys, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Indices of keys (d, x, y, z0+k), k < len, in one sorted key array: one
// lower-bound for the column (x, y), then a forward scan (keys are unique and
// z is the fastest coordinate).  out[k] = index or -1.  Same results as len
// find_key calls.
template <int MAXLEN>
PSR_HD void find_run(const int64_t* __restrict__ keys, int64_t n, int d, int x, int y, int z0, int len, int64_t* out) {
  int64_t lo = lower_bound(keys, n, basis_key(d, x, y, z0));
#pragma unroll
  for (int k = 0; k < MAXLEN; ++k) {
    if (k >= len) break;
    const int64_t key = basis_key(d, x, y, z0 + k);
    if (lo < n && keys[lo] == key) {
      out[k] = lo;
      ++lo;
    } else {
      out[k] = -1;
    }
  }
}

// (x, y) column index of one level's sorted keys: keys (d, x, y, *) occupy
// [start[x*dim + y], end[x*dim + y]).  Built for levels whose dim^2 table is
// small; a null table falls back to the full-array search.  Same results.
struct ColumnTable {
  const int* start = nullptr;
  const int* end = nullptr;
  int dim = 0;
};

template <int MAXLEN>
PSR_HD void find_run_table(const int64_t* __restrict__ keys, int64_t n, const ColumnTable& t, int d, int x, int y, int z0,
                           int len, int64_t* out) {
  if (!t.start) {
    find_run<MAXLEN>(keys, n, d, x, y, z0, len, out);
    return;
  }
#pragma unroll
  for (int k = 0; k < MAXLEN; ++k) out[k] = -1;
  if (x < 0 || y < 0 || x >= t.dim || y >= t.dim) return;
  const int c = x * t.dim + y;
  int64_t lo = t.start[c], hi = t.end[c];
  const int64_t end = hi;
  const int64_t key0 = basis_key(d, x, y, z0);
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key0) lo = mid + 1; else hi = mid;
  }
#pragma unroll
  for (int k = 0; k < MAXLEN; ++k) {
    if (k >= len) break;
    if (lo < end && keys[lo] == basis_key(d, x, y, z0 + k)) {
      out[k] = lo;
      ++lo;
    }
  }
}

PSR_HD int imin(int a, int b) { return a < b ? a : b; }
PSR_HD int imax(int a, int b) { return a > b ? a : b; }
template <class T> PSR_HD T tmin(T a, T b) { return b < a ? b : a; }
template <class T> PSR_HD T tmax(T a, T b) { return a < b ? b : a; }

// src/warp_psr/density.py:14-22 -- degree-2 B-spline weights (double)
struct Quad { double b0, b1, b2; };
PSR_HD Quad quadratic(double x) {
  Quad q;
  q.b0 = (1.0 - x) * (1.0 - x) * 0.5;
  q.b1 = ((1.0 - x) * (x + 1.0) + x * (2.0 - x)) * 0.5;
  q.b2 = x * x * 0.5;
  return q;
}
PSR_HD double quad_at(const Quad& q, int i) { return i == 0 ? q.b0 : (i == 1 ? q.b1 : q.b2); }
// d/dx of quadratic(x)'s three values: b0 = (1-x)^2/2, b1 = (1 + 2x - 2x^2)/2, b2 = x^2/2
PSR_HD double quad_deriv(double x, int i) { return i == 0 ? x - 1.0 : (i == 1 ? 1.0 - 2.0 * x : x); }

// Deterministic natural logarithm: double-double evaluation of
// k*ln2 + 2*atanh((m-1)/(m+1)) rounded once to double.  Identical on host and
// device (only IEEE +,-,*,/ and exactly-specified fma), correctly rounded
// except when the exact value lies within ~2^-100 of a rounding boundary.
// The reference calls CUDA's double log (density.py:31-35); this keeps host
// and device bit-identical instead of inheriting two libm's last-bit choices.
struct DD { double hi, lo; };
PSR_HD DD dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
PSR_HD DD dd_quick(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
PSR_HD DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  DD t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_quick(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick(s.hi, s.lo);
}
PSR_HD DD dd_mul(DD a, DD b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return dd_quick(p, e);
}
PSR_HD DD dd_div(DD a, DD b) {
  double q1 = a.hi / b.hi;
  DD r = dd_add(a, dd_mul(b, DD{-q1, 0.0}));
  double q2 = r.hi / b.hi;
  r = dd_add(r, dd_mul(b, DD{-q2, 0.0}));
  double q3 = r.hi / b.hi;
  DD q = dd_quick(q1, q2);
  return dd_add(q, DD{q3, 0.0});
}
PSR_HD double det_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  // split x = 2^k * m with m in [sqrt(1/2), sqrt(2))
  uint64_t bits = __builtin_bit_cast(uint64_t, x);
  int k = 0;
  if ((bits >> 52) == 0) {  // subnormal: scale into the normal range
    x *= 18014398509481984.0;  // 2^54
    bits = __builtin_bit_cast(uint64_t, x);
    k = -54;
  }
  k += (int)((bits >> 52) & 0x7ff) - 1023;
  uint64_t mbits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m = __builtin_bit_cast(double, mbits);
  if (m > 1.4142135623730951) { m *= 0.5; k += 1; }
  DD num = dd_two_sum(m, -1.0);
  DD den = dd_two_sum(m, 1.0);
  DD s = dd_div(num, den);
  DD s2 = dd_mul(s, s);
  // atanh(s)/s = sum_{j>=0} s^(2j)/(2j+1); |s| < 0.1716 -> 24 terms reach 2^-120
  DD acc = {1.0 / 49.0, 0.0};
  for (int j = 23; j >= 0; --j) {
    DD inv = dd_div(DD{1.0, 0.0}, DD{double(2 * j + 1), 0.0});
    acc = dd_add(dd_mul(acc, s2), inv);
    if (j == 0) break;
  }
  DD at = dd_mul(acc, s);
  at = dd_add(at, at);  // 2*atanh(s)
  const DD ln2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
  DD kl = dd_mul(ln2, DD{double(k), 0.0});
  DD r = dd_add(kl, at);
  return r.hi + r.lo;
}
// ln 2 rounded to double (the value log(2.0) returns when correctly rounded)
constexpr double LN2_D = 6.93147180559945286227e-01;

}  // namespace psr
