// TEST INFRASTRUCTURE ONLY -- the parity oracle.
//
// A serial CPU restatement of warp-psr's screened Poisson forward pipeline
// (degree 1, Neumann, confidence off, alignDir 0), written from the reference
// Python/Warp sources and cited function by function.  It is linked only by
// tests/, by __graft_entry__.smoke() (as the checker) and by bench.py's
// cpu_baseline leg.  The product (warp-psr_amd/) never links or calls it.
//
// Arithmetic rules: build with -ffp-contract=off and without fast-math
// (reference modules set fuse_fp=False, e.g. src/warp_psr/samples.py:16).
// Every floating-point expression keeps the reference's operation order.
//
// Parity status per part (see tests/test_oracle_golden.py for the checks):
//   samples, density, normal splats, tree      -> pinned (golden key SHA-256s,
//                                                 cube_to_model, known answers)
//   stiffness / prolongation / screening / normal-load integrals
//                                              -> pinned by closed forms and
//                                                 brute-force quadrature tests
//   base-level operators (reference: warp.fem quadrature), Galerkin product
//   order (reference: warp.sparse.bsr_mm), root CG (warp.optim.linear.cg),
//   isovalue partial-sum order (reference: wp.tile_sum)
//                                              -> tolerance-level parity only
//                                                 (isovalue goldens), bitwise
//                                                 parity unpinned.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
// Loops whose iterations write disjoint outputs and read nothing another
// iteration writes (per row, per query, per aggregate): the OpenMP build
// runs them on all threads with unchanged per-iteration arithmetic, so its
// arrays equal the serial build's bit for bit.
#define PSR_PARALLEL_FOR _Pragma("omp parallel for schedule(dynamic, 512)")
#else
#define PSR_PARALLEL_FOR
#endif

namespace {

constexpr int64_t KMAX = std::numeric_limits<int64_t>::max();

template <class T> inline T wmax_(T a, T b) { return a > b ? a : b; }
struct I3 {
  int v[3];
  int operator[](int a) const { return v[a]; }
  int& operator[](int a) { return v[a]; }
};

// src/warp_psr/field.py:13-16
inline int64_t basis_key(int d, int x, int y, int z) {
  return ((int64_t)d << 48) | ((int64_t)x << 32) | ((int64_t)y << 16) | (int64_t)z;
}
// src/warp_psr/keyed.py:14-20
inline I3 coords(int64_t k) {
  return {{(int)((k >> 32) & 65535), (int)((k >> 16) & 65535), (int)(k & 65535)}};
}
inline int key_depth(int64_t k) { return (int)(k >> 48); }
// src/warp_psr/tree.py:16-24 (signed 20-bit coordinates)
inline int64_t tree_key(I3 c) {
  return ((int64_t)(c[0] + 524288) << 40) | ((int64_t)(c[1] + 524288) << 20) |
         (int64_t)(c[2] + 524288);
}
inline I3 tree_coords(int64_t k) {
  return {{(int)((k >> 40) & 1048575) - 524288, (int)((k >> 20) & 1048575) - 524288,
           (int)(k & 1048575) - 524288}};
}
// src/warp_psr/tree.py:26-28: floor(c/2) == arithmetic shift
inline I3 tparent(I3 c) { return {{c[0] >> 1, c[1] >> 1, c[2] >> 1}}; }

// src/warp_psr/field.py:19-32
inline int64_t find_key(const int64_t* keys, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  if (lo < n && keys[lo] == key) return lo;
  return -1;
}
inline int64_t find_key(const std::vector<int64_t>& keys, int64_t key) {
  return find_key(keys.data(), (int64_t)keys.size(), key);
}
inline int64_t lower_bound_key(const std::vector<int64_t>& keys, int64_t key) {
  return std::lower_bound(keys.begin(), keys.end(), key) - keys.begin();
}

// src/warp_psr/keyed.py:23-29
inline int64_t morton(I3 c, int depth) {
  int64_t m = 0;
  for (int bit = 0; bit < depth; ++bit)
    for (int a = 0; a < 3; ++a) m |= (int64_t)((c[a] >> bit) & 1) << (3 * bit + a);
  return m;
}

// Stable sort of indices by key (Warp radix_sort_pairs is a stable sort).
std::vector<int> stable_order(const std::vector<int64_t>& keys) {
  std::vector<int> ids(keys.size());
  std::iota(ids.begin(), ids.end(), 0);
#ifdef _OPENMP
  // ties broken by position: the one stable order, from a parallel sort
  __gnu_parallel::sort(ids.begin(), ids.end(), [&](int a, int b) { return keys[a] < keys[b] || (keys[a] == keys[b] && a < b); });
#else
  std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return keys[a] < keys[b]; });
#endif
  return ids;
}

template <class T> struct V3 {
  T v[3];
  T operator[](int a) const { return v[a]; }
  T& operator[](int a) { return v[a]; }
};
template <class T> struct V4 {
  T v[4];
  T operator[](int a) const { return v[a]; }
  T& operator[](int a) { return v[a]; }
};
template <class T> inline V3<T> vzero3() { return {{T(0), T(0), T(0)}}; }
template <class T> inline V4<T> vzero4() { return {{T(0), T(0), T(0), T(0)}}; }
template <class T> inline void add3(V3<T>& a, const V3<T>& b) {
  a[0] = a[0] + b[0]; a[1] = a[1] + b[1]; a[2] = a[2] + b[2];
}
template <class T> inline void add4(V4<T>& a, const V4<T>& b) {
  for (int i = 0; i < 4; ++i) a[i] = a[i] + b[i];
}
// src/warp_psr/samples.py:19-27 (component-wise division)
template <class T> inline V3<T> point_div(const V3<T>& v, T s) {
  return {{v[0] / s, v[1] / s, v[2] / s}};
}
template <class T> inline T dot3(const V3<T>& a, const V3<T>& b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// src/warp_psr/density.py:14-22
struct Q3 { double b[3]; };
inline Q3 quadratic(double x) {
  const double one = 1.0, half = 0.5;
  return {{(one - x) * (one - x) * half, ((one - x) * (x + one) + x * (2.0 - x)) * half,
           x * x * half}};
}

template <class T> struct Records {
  std::vector<int64_t> keys;
  std::vector<V4<T>> values;
};

// src/warp_psr/keyed.py:111-142: stable sort by key, then a serial sum of each
// group's values in sorted order.  Sentinel KMAX records are dropped.
template <class T>
Records<T> reduce_records(const std::vector<int64_t>& keys, const std::vector<V4<T>>& values) {
  Records<T> out;
  std::vector<int> ids = stable_order(keys);
  size_t n = keys.size();
  size_t i = 0;
  while (i < n) {
    int64_t k = keys[ids[i]];
    if (k == KMAX) break;
    V4<T> total = vzero4<T>();
    size_t j = i;
    while (j < n && keys[ids[j]] == k) { add4(total, values[ids[j]]); ++j; }
    out.keys.push_back(k);
    out.values.push_back(total);
    i = j;
  }
  return out;
}

// src/warp_psr/keyed.py:102-108
std::vector<int> morton_order(const std::vector<int64_t>& keys, int depth) {
  std::vector<int64_t> codes(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) codes[i] = morton(coords(keys[i]), depth);
  return stable_order(codes);
}

// src/warp_psr/keyed.py:157-178
std::vector<int64_t> unique_keys(std::vector<int64_t> keys) {
  std::sort(keys.begin(), keys.end());
  std::vector<int64_t> out;
  for (size_t i = 0; i < keys.size(); ++i) {
    if (keys[i] == KMAX) break;
    if (i == 0 || keys[i] != keys[i - 1]) out.push_back(keys[i]);
  }
  return out;
}

template <class T> struct CSR {
  int rows = 0, cols = 0;
  std::vector<int> offsets, columns;
  std::vector<T> values;
};

struct Params {
  int depth, full_depth, base_depth, solve_depth, kernel_depth;
  int iters, base_v_cycles, exact;
  double samples_per_node, scale, point_weight, cg_accuracy;
  int threads;  // OpenMP threads for order-independent loops (1 = serial)
  int float_operators = 0;  // finer-level operators rounded to float (float64 runs)
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class T> struct State {
  Params prm;
  int64_t n = 0;
  // samples (src/warp_psr/samples.py)
  T transform[5];
  T bounds[6];
  std::vector<V3<T>> q, unit_normals;
  std::vector<int> input_status, input_group, input_order;
  std::vector<int64_t> sample_keys;
  std::vector<V3<T>> position_sums, normal_sums;
  std::vector<T> weights;
  std::vector<int> first_indices, group_offsets;
  // density (src/warp_psr/density.py)
  std::vector<Records<T>> aggregates, density_levels;
  std::vector<T> densities;  // G x (kernel_depth+1)
  std::vector<T> dw;         // G x 2
  T density_scale;
  // normals (src/warp_psr/normals.py)
  Records<T> normal_records;
  std::vector<int> normal_offsets;
  T normal_average[3];
  std::vector<int> normal_order;
  // tree (src/warp_psr/tree.py)
  std::vector<std::vector<int64_t>> tree_levels, basis;
  // solve (src/warp_psr/solver.py)
  std::vector<CSR<T>> matrices;       // per level (cast), base..solve
  std::vector<CSR<T>> cross;          // per level d>base
  std::vector<CSR<T>> prolong;        // index d-1 -> P_d
  std::vector<std::vector<double>> normal_loads, point_loads;
  std::vector<std::vector<T>> rhs, corrections;
  std::vector<int> cg_runs;
  // field (src/warp_psr/reconstruction.py:282-324)
  std::vector<int64_t> keys;
  std::vector<T> coefficients, cumulative, centroid_values;
  T iso;
  T cube_to_model[16], model_to_cube[16];
};

// ---------------------------------------------------------------- samples
// src/warp_psr/samples.py:54-88 (_transform); computed once on the host.
template <class T> int compute_transform(const T* b, double scale, T* t) {
  T width = wmax_(b[3] - b[0], wmax_(b[4] - b[1], b[5] - b[2])) * T(scale);
  if (!std::isfinite(width) || width <= T(0)) return 1;
  T inv = T(1.0 / double(width));
  t[0] = inv;
  T determinant = inv * (inv * inv);
  t[4] = inv * inv / determinant * T(std::pow(double(determinant), 1.0 / 3.0));
  for (int a = 0; a < 3; ++a) {
    int axis = (a + 1) % 3;
    T origin = (b[axis + 3] + b[axis]) / T(2) - width / T(2);
    t[a + 1] = -origin * inv;
  }
  return 0;
}

template <class T> void prepare_samples(State<T>& s, const T* pts, const T* nrm) {
  const int64_t n = s.n;
  const int depth = s.prm.depth;
  // samples.py:34-52 bounds over finite points
  T* b = s.bounds;
  for (int a = 0; a < 3; ++a) { b[a] = std::numeric_limits<T>::infinity(); b[a + 3] = -b[a]; }
  int status = 0;
  for (int64_t i = 0; i < n; ++i) {
    const T* p = pts + 3 * i;
    if (std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2])) {
      for (int a = 0; a < 3; ++a) {
        b[a] = std::min(b[a], p[a]);
        b[a + 3] = std::max(b[a + 3], p[a]);
      }
    } else {
      status++;
    }
  }
  if (status || compute_transform(b, s.prm.scale, s.transform))
    throw Error(2, "Nonfinite positions or degenerate/overflowing bounding cube");
  const T* t = s.transform;
  // samples.py:90-134 (_classify)
  s.q.resize(n); s.unit_normals.resize(n); s.input_status.resize(n);
  std::vector<int64_t> keys(n);
  const int res = 1 << depth;
  for (int64_t i = 0; i < n; ++i) {
    const T* P = pts + 3 * i;
    const T* N = nrm + 3 * i;
    T inv = t[0];
    V3<T> p = {{P[1] * inv + t[1], P[2] * inv + t[2], P[0] * inv + t[3]}};
    V3<T> nn = {{N[1] * t[4], N[2] * t[4], N[0] * t[4]}};
    T length2 = dot3(nn, nn);
    s.q[i] = p;
    keys[i] = KMAX;
    s.unit_normals[i] = vzero3<T>();
    s.input_status[i] = 1;
    bool ok = length2 > T(0) && std::isfinite(length2);
    if (ok) s.input_status[i] = 2;
    T pmin = std::min(std::min(p[0], p[1]), p[2]);
    T pmax = std::max(std::max(p[0], p[1]), p[2]);
    if (ok && pmin >= T(0) && pmax <= T(1)) {
      s.input_status[i] = 0;
      s.unit_normals[i] = point_div(nn, (T)std::sqrt(length2));
      int x = std::max(0, (int)std::ceil(p[0] * T(res)) - 1);
      int y = std::max(0, (int)std::ceil(p[1] * T(res)) - 1);
      int z = std::max(0, (int)std::ceil(p[2] * T(res)) - 1);
      keys[i] = basis_key(depth, x, y, z);
    }
  }
  // samples.py:269-306: stable sort, head marks, serial group sums
  std::vector<int> order = stable_order(keys);
  s.input_order = order;
  s.input_group.assign(n, -1);
  s.sample_keys.clear(); s.position_sums.clear(); s.normal_sums.clear();
  s.weights.clear(); s.first_indices.clear(); s.group_offsets.clear();
  int64_t i = 0;
  while (i < n) {
    int64_t k = keys[order[i]];
    if (k == KMAX) break;
    int g = (int)s.sample_keys.size();
    V3<T> p = vzero3<T>(), nsum = vzero3<T>();
    T weight = T(0);
    int64_t j = i;
    while (j < n && keys[order[j]] == k) {
      add3(p, s.q[order[j]]);
      add3(nsum, s.unit_normals[order[j]]);
      weight = weight + T(1);
      s.input_group[order[j]] = g;
      ++j;
    }
    s.sample_keys.push_back(k);
    s.position_sums.push_back(p);
    s.normal_sums.push_back(nsum);
    s.weights.push_back(weight);
    s.first_indices.push_back(order[i]);
    s.group_offsets.push_back((int)i);
    i = j;
  }
  if (s.sample_keys.empty()) throw Error(2, "No accepted samples with finite nonzero normals");
  s.group_offsets.push_back((int)i);
}

// ---------------------------------------------------------------- density
// src/warp_psr/density.py:37-50 (_scale)
template <class T> T density_scale() {
  T total = T(0);
  T dx = T(0.1);
  for (int x = 0; x < 10; ++x) {
    T p = dx / T(2) + dx * T(x);
    Q3 b = quadratic(double(p));
    T scale = T(1.0 / (b.b[0] * b.b[0] + b.b[1] * b.b[1] + b.b[2] * b.b[2]));
    for (int y = 0; y < 10; ++y)
      for (int z = 0; z < 10; ++z) total = total + scale;
  }
  return total / T(1000);
}

// Deterministic natural logarithm shared with the product: double-double
// k*ln2 + 2*atanh((m-1)/(m+1)), rounded once.  The reference calls CUDA's
// double log (density.py:31-35); both CUDA's and glibc's can differ from the
// correctly rounded value in the last bit, so host and device use this one.
struct DD { double hi, lo; };
inline DD dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
inline DD dd_quick(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
inline DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  DD t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_quick(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick(s.hi, s.lo);
}
inline DD dd_mul(DD a, DD b) {
  double p = a.hi * b.hi;
  double e = std::fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return dd_quick(p, e);
}
inline DD dd_div(DD a, DD b) {
  double q1 = a.hi / b.hi;
  DD r = dd_add(a, dd_mul(b, DD{-q1, 0.0}));
  double q2 = r.hi / b.hi;
  r = dd_add(r, dd_mul(b, DD{-q2, 0.0}));
  double q3 = r.hi / b.hi;
  DD q = dd_quick(q1, q2);
  return dd_add(q, DD{q3, 0.0});
}
double det_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
  if (std::isinf(x)) return x;
  uint64_t bits;
  std::memcpy(&bits, &x, 8);
  int k = 0;
  if ((bits >> 52) == 0) {
    x *= 18014398509481984.0;
    std::memcpy(&bits, &x, 8);
    k = -54;
  }
  k += (int)((bits >> 52) & 0x7ff) - 1023;
  uint64_t mbits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  std::memcpy(&m, &mbits, 8);
  if (m > 1.4142135623730951) { m *= 0.5; k += 1; }
  DD num = dd_two_sum(m, -1.0);
  DD den = dd_two_sum(m, 1.0);
  DD s = dd_div(num, den);
  DD s2 = dd_mul(s, s);
  DD acc = {1.0 / 49.0, 0.0};
  for (int j = 23; j >= 0; --j) {
    DD inv = dd_div(DD{1.0, 0.0}, DD{double(2 * j + 1), 0.0});
    acc = dd_add(dd_mul(acc, s2), inv);
  }
  DD at = dd_mul(acc, s);
  at = dd_add(at, at);
  const DD ln2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
  DD kl = dd_mul(ln2, DD{double(k), 0.0});
  DD r = dd_add(kl, at);
  return r.hi + r.lo;
}
constexpr double LN2_D = 6.93147180559945286227e-01;

// Warp's scalar min/max/clamp (a<b?a:b forms): same ties and NaN behaviour.
template <class T> inline T wmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T wmin(T a, T b) { return a < b ? a : b; }

template <class T> inline T real_log(T v) { return T(det_log(double(v))); }

// density.py:108-136 (density_value)
template <class T> T density_value(const Records<T>& lvl, V3<T> p, int depth) {
  int n = 1 << depth;
  I3 c = {{std::max(0, (int)std::ceil(p[0] * T(n)) - 1), std::max(0, (int)std::ceil(p[1] * T(n)) - 1),
           std::max(0, (int)std::ceil(p[2] * T(n)) - 1)}};
  T h = T(1) / T(n);
  Q3 bx = quadratic(double((p[0] - T(c[0]) * h) / h));
  Q3 by = quadratic(double((p[1] - T(c[1]) * h) / h));
  Q3 bz = quadratic(double((p[2] - T(c[2]) * h) / h));
  T value = T(0);
  for (int k = 0; k < 27; ++k) {
    int x = k / 9, y = k / 3 % 3, z = k % 3;
    I3 tg = {{c[0] + x - 1, c[1] + y - 1, c[2] + z - 1}};
    if (std::min(std::min(tg[0], tg[1]), tg[2]) >= 0 && std::max(std::max(tg[0], tg[1]), tg[2]) < n) {
      int64_t idx = find_key(lvl.keys, basis_key(depth, tg[0], tg[1], tg[2]));
      if (idx >= 0) value = value + T(bx.b[x] * by.b[y] * bz.b[z] * double(lvl.values[idx][0]));
    }
  }
  return value;
}

template <class T> void build_density(State<T>& s) {
  const int depth = s.prm.depth, kd = s.prm.kernel_depth;
  const size_t G = s.sample_keys.size();
  s.aggregates.assign(depth + 1, Records<T>());
  Records<T>& leaf = s.aggregates[depth];
  leaf.keys = s.sample_keys;
  leaf.values.resize(G);
  for (size_t i = 0; i < G; ++i)  // density.py:52-57
    leaf.values[i] = {{s.position_sums[i][0], s.position_sums[i][1], s.position_sums[i][2], s.weights[i]}};
  for (int d = depth; d > 0; --d) {  // density.py:228-240
    const Records<T>& src = s.aggregates[d];
    std::vector<int> order = morton_order(src.keys, d);
    std::vector<int64_t> keys(src.keys.size());
    std::vector<V4<T>> vals(src.keys.size());
    for (size_t i = 0; i < order.size(); ++i) {  // density.py:59-72
      int j = order[i];
      I3 c = coords(src.keys[j]);
      keys[i] = basis_key(d - 1, c[0] / 2, c[1] / 2, c[2] / 2);
      vals[i] = src.values[j];
    }
    s.aggregates[d - 1] = reduce_records(keys, vals);
  }
  s.density_scale = density_scale<T>();
  const T scale = s.density_scale;
  s.density_levels.clear();
  s.densities.assign(G * (kd + 1), T(0));
  for (int d = 0; d <= kd; ++d) {  // density.py:245-271
    const Records<T>& src = s.aggregates[d];
    std::vector<int> order = morton_order(src.keys, d);
    size_t m = src.keys.size();
    std::vector<int64_t> keys(27 * m, KMAX);
    std::vector<V4<T>> vals(27 * m, vzero4<T>());
    int n = 1 << d;
    PSR_PARALLEL_FOR
    for (size_t i = 0; i < m; ++i) {  // density.py:74-106 (_splat)
      int j = order[i];
      I3 c = coords(src.keys[j]);
      V4<T> v = src.values[j];
      V3<T> p = point_div(V3<T>{{v[0], v[1], v[2]}}, v[3]);
      T h = T(1) / T(n);
      Q3 bx = quadratic(double((p[0] - T(c[0]) * h) / h));
      Q3 by = quadratic(double((p[1] - T(c[1]) * h) / h));
      Q3 bz = quadratic(double((p[2] - T(c[2]) * h) / h));
      for (int k = 0; k < 27; ++k) {
        int x = k / 9, y = k / 3 % 3, z = k % 3;
        I3 tg = {{c[0] + x - 1, c[1] + y - 1, c[2] + z - 1}};
        if (std::min(std::min(tg[0], tg[1]), tg[2]) >= 0 &&
            std::max(std::max(tg[0], tg[1]), tg[2]) < n) {
          T weight = v[3] * scale;
          T val = T(double(weight) * bx.b[x] * by.b[y] * bz.b[z]);
          keys[i * 27 + k] = basis_key(d, tg[0], tg[1], tg[2]);
          vals[i * 27 + k] = {{val, T(0), T(0), T(0)}};
        }
      }
    }
    s.density_levels.push_back(reduce_records(keys, vals));
    PSR_PARALLEL_FOR
    for (size_t i = 0; i < G; ++i)  // density.py:138-150 (_evaluate)
      s.densities[i * (kd + 1) + d] =
          density_value(s.density_levels.back(), point_div(s.position_sums[i], s.weights[i]), d);
  }
  // density.py:152-182 (_depths)
  const T target = T(s.prm.samples_per_node);
  s.dw.assign(2 * G, T(0));
  const double two_ln2 = LN2_D * 2.0;
  PSR_PARALLEL_FOR
  for (size_t i = 0; i < G; ++i) {
    int d = kd;
    T density = s.densities[i * (kd + 1) + d];
    T raw;
    if (density >= target) {
      raw = T(double(d) + double(real_log(density / target)) / two_ln2);
    } else {
      T fine = density;
      while (density < target && d > 0) {
        fine = density;
        d -= 1;
        density = s.densities[i * (kd + 1) + d];
      }
      raw = T(T(d) + real_log(density / target) / real_log(density / fine));
    }
    T h = T(1) / T(1 << d);
    s.dw[2 * i] = raw;
    s.dw[2 * i + 1] = h * h / density;
  }
}

// ---------------------------------------------------------------- normals
template <class T> inline T clampT(T x, T a, T b) { return wmin(wmax(a, x), b); }

template <class T> void build_normal_field(State<T>& s) {
  const int minimum = s.prm.base_depth, maximum = s.prm.depth;
  const size_t G = s.sample_keys.size();
  std::vector<int64_t> first(G);
  for (size_t i = 0; i < G; ++i) first[i] = s.first_indices[i];
  std::vector<int> order = stable_order(first);  // normals.py:190-198
  s.normal_order = order;
  std::vector<int64_t> keys(54 * G, KMAX);
  std::vector<V4<T>> vals(54 * G, vzero4<T>());
  PSR_PARALLEL_FOR
  for (size_t i = 0; i < G; ++i) {  // normals.py:46-104 (_splats)
    int src = order[i];
    T w = s.weights[src];
    V3<T> normal = point_div(s.normal_sums[src], w);
    T length = std::sqrt(dot3(normal, normal));
    if (!(length > T(0))) continue;
    normal = point_div(normal, length);
    normal = {{normal[0] * w, normal[1] * w, normal[2] * w}};
    V3<T> p = point_div(s.position_sums[src], w);
    T depth = clampT(s.dw[2 * src], T(minimum), T(maximum));
    int top = (int)std::ceil(depth);
    T dx = T(1) - (T(top) - depth);
    if (top <= minimum) { top = minimum; dx = T(1); }
    if (top > maximum) { top = maximum; dx = T(1); }
    const int n = 1 << top;
    const T h = T(1) / T(n);
    I3 c = {{std::max(0, (int)std::ceil(p[0] * T(n)) - 1), std::max(0, (int)std::ceil(p[1] * T(n)) - 1),
             std::max(0, (int)std::ceil(p[2] * T(n)) - 1)}};
    Q3 bx = quadratic(double((p[0] - T(c[0]) * h) / h));
    Q3 by = quadratic(double((p[1] - T(c[1]) * h) / h));
    Q3 bz = quadratic(double((p[2] - T(c[2]) * h) / h));
    const bool second = std::fabs(1.0 - double(dx)) > 1e-06;
    for (int k = 0; k < 54; ++k) {
      T fraction = k >= 27 ? T(1) - dx : dx;
      if (!(k < 27 || second)) continue;
      int qk = k % 27;
      int x = qk / 9, y = qk / 3 % 3, z = qk % 3;
      I3 tg = {{c[0] + x - 1, c[1] + y - 1, c[2] + z - 1}};
      if (std::min(std::min(tg[0], tg[1]), tg[2]) >= 0 && std::max(std::max(tg[0], tg[1]), tg[2]) < n) {
        T hhh = h * h * h;
        T wgt = s.dw[2 * src + 1];
        V3<T> v;
        for (int a = 0; a < 3; ++a) v[a] = normal[a] * wgt / hhh * fraction;
        T basis = T(bx.b[x] * by.b[y] * bz.b[z]);
        for (int a = 0; a < 3; ++a) v[a] = v[a] * basis;
        keys[i * 54 + k] = basis_key(top, tg[0], tg[1], tg[2]);
        vals[i * 54 + k] = {{v[0], v[1], v[2], T(0)}};
      }
    }
  }
  s.normal_records = reduce_records(keys, vals);
  for (auto& v : s.normal_records.values)  // normals.py:106-109 (_negative)
    for (int a = 0; a < 4; ++a) v[a] = -v[a];
  // normals.py:111-163 (_average_contributions + serial _average_tiled)
  V3<T> total = vzero3<T>();
  for (size_t i = 0; i < G; ++i) {
    int src = order[i];
    V3<T> nn = point_div(s.normal_sums[src], s.weights[src]);
    V3<T> value = vzero3<T>();
    if (std::sqrt(dot3(nn, nn)) > T(0)) {
      value = {{s.dw[2 * src] * s.weights[src], s.dw[2 * src + 1] * s.weights[src], s.weights[src]}};
    }
    if (value[2] > T(0)) add3(total, value);
  }
  s.normal_average[0] = s.normal_average[1] = s.normal_average[2] = T(0);
  if (total[2] > T(0)) {
    s.normal_average[0] = total[0] / total[2];
    s.normal_average[1] = total[1] / total[2];
    s.normal_average[2] = total[2];
  }
  // normals.py:17-29 (_depth_offsets)
  s.normal_offsets.resize(maximum + 2);
  for (int d = 0; d <= maximum + 1; ++d)
    s.normal_offsets[d] = (int)lower_bound_key(s.normal_records.keys, (int64_t)d << 48);
}

// ---------------------------------------------------------------- tree
// src/warp_psr/tree.py:152-201
template <class T> void build_tree(State<T>& s) {
  const int depth = s.prm.depth, full = s.prm.full_depth, base = s.prm.base_depth;
  s.tree_levels.assign(depth + 1, {});
  s.basis.assign(depth + 1, {});
  bool have_requests = false;
  std::vector<int64_t> requests, broods;
  for (int d = depth; d >= 0; --d) {
    const int64_t r0 = s.normal_offsets[d], r1 = s.normal_offsets[d + 1];
    std::vector<int64_t> parents = broods;
    if (d > base) {
      requests = parents;
    } else if (have_requests) {
      std::vector<int64_t> rk(requests.size());
      for (size_t i = 0; i < requests.size(); ++i) rk[i] = tree_key(tparent(tree_coords(requests[i])));
      requests = unique_keys(rk);
    } else {
      requests = parents;
    }
    have_requests = true;
    const int cc = d <= full ? (d == 0 ? 2 : (1 << (d - 1)) + 2) : 0;
    std::vector<int64_t> cand;
    cand.reserve((r1 - r0) + (size_t)cc * cc * cc + requests.size() * 27 + parents.size());
    for (int64_t i = r0; i < r1; ++i) {  // tree.py:41-50 (_seed_broods)
      int64_t key = s.normal_records.keys[i];
      const V4<T>& v = s.normal_records.values[i];
      int nd = key_depth(key);
      T len2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
      if (nd >= d && double(len2) > 0.0) {
        I3 c = coords(key);
        c = {{c[0] >> (nd - d), c[1] >> (nd - d), c[2] >> (nd - d)}};
        cand.push_back(tree_key(tparent(c)));
      } else {
        cand.push_back(KMAX);
      }
    }
    for (int i = 0; i < cc * cc * cc; ++i)  // tree.py:30-34 (_coarse, lo=-1)
      cand.push_back(tree_key(I3{{i / (cc * cc) - 1, (i / cc) % cc - 1, i % cc - 1}}));
    const int lim = 1 << d;
    auto support = [&](const std::vector<int64_t>& src, int radius) {  // tree.py:52-64
      const int width = 2 * radius + 1;
      for (int64_t key : src) {
        I3 p = tparent(tree_coords(key));
        for (int j = 0; j < width * width * width; ++j) {
          I3 c = {{p[0] + j / (width * width) - radius, p[1] + (j / width) % width - radius,
                   p[2] + j % width - radius}};
          if (std::min(std::min(c[0], c[1]), c[2]) >= -lim && std::max(std::max(c[0], c[1]), c[2]) < lim)
            cand.push_back(tree_key(c));
          else
            cand.push_back(KMAX);
        }
      }
    };
    support(requests, 1);
    support(parents, 0);
    broods = unique_keys(cand);
    // tree.py:78-103: the sorted set of all eight children of every brood
    std::vector<int64_t> children;
    children.reserve(8 * broods.size());
    for (int64_t b : broods) {
      I3 c = tree_coords(b);
      for (int k = 0; k < 8; ++k)
        children.push_back(tree_key(I3{{2 * c[0] + ((k >> 2) & 1), 2 * c[1] + ((k >> 1) & 1), 2 * c[2] + (k & 1)}}));
    }
    std::sort(children.begin(), children.end());
    s.tree_levels[d] = children;
    // tree.py:105-143 (compact_basis): nodes with 0 <= c <= 2^d
    for (int64_t k : children) {
      I3 c = tree_coords(k);
      if (std::min(std::min(c[0], c[1]), c[2]) >= 0 && std::max(std::max(c[0], c[1]), c[2]) <= lim)
        s.basis[d].push_back(basis_key(d, c[0], c[1], c[2]));
    }
  }
}

// ---------------------------------------------------------------- operators
struct D2 { double v[2]; };

// src/warp_psr/stiffness.py:14-35
D2 hat_integrals(int depth, int index, int trial_depth, int trial_index) {
  const int resolution = 1 << depth;
  const int ratio = 1 << (depth - trial_depth);
  const int first = std::max(0, std::max(index - 1, (trial_index - 1) * ratio));
  const int end = std::min(resolution, std::min(index + 1, (trial_index + 1) * ratio));
  D2 r = {{0.0, 0.0}};
  for (int cell = first; cell < end; ++cell) {
    double t0 = double(cell - index);
    double c0 = double(cell) / double(ratio) - double(trial_index);
    double ts = cell >= index ? -1.0 : 1.0;
    double cs = 1.0 / double(ratio);
    if (cell >= trial_index * ratio) cs = -cs;
    double tv = 1.0 - std::fabs(t0);
    double cv = 1.0 - std::fabs(c0);
    r.v[0] += (tv * cv + (tv * cs + ts * cv) * 0.5 + ts * cs / 3.0) / double(resolution);
    r.v[1] += ts * cs * double(resolution);
  }
  return r;
}

// src/warp_psr/stiffness.py:86-109 (float64 values, sorted rows)
CSR<double> assemble_stiffness(const std::vector<int64_t>& row_keys, const std::vector<int64_t>& col_keys,
                               int depth, int trial_depth) {
  CSR<double> m;
  m.rows = (int)row_keys.size();
  m.cols = (int)col_keys.size();
  const int shift = depth - trial_depth;
  const int res = 1 << depth;
  std::vector<D2> table((res + 1) * 3);
  for (int i = 0; i <= res; ++i)
    for (int k = 0; k < 3; ++k) table[i * 3 + k] = hat_integrals(depth, i, trial_depth, (i >> shift) + k - 1);
  m.offsets.assign(m.rows + 1, 0);
  // present neighbours of row i in slot order; a count pass, then a fill pass
  auto row = [&](int i, bool fill) {
    I3 c = coords(row_keys[i]);
    I3 center = {{c[0] >> shift, c[1] >> shift, c[2] >> shift}};
    int e = fill ? m.offsets[i] : 0;
    for (int k = 0; k < 27; ++k) {
      I3 q = {{center[0] + k / 9 - 1, center[1] + (k / 3) % 3 - 1, center[2] + k % 3 - 1}};
      int64_t index = -1;
      if (std::min(std::min(q[0], q[1]), q[2]) >= 0 && std::max(std::max(q[0], q[1]), q[2]) <= (1 << trial_depth))
        index = find_key(col_keys, basis_key(trial_depth, q[0], q[1], q[2]));
      if (index >= 0) {
        if (fill) {
          const D2& x = table[c[0] * 3 + k / 9];
          const D2& y = table[c[1] * 3 + (k / 3) % 3];
          const D2& z = table[c[2] * 3 + k % 3];
          m.values[e] = x.v[1] * y.v[0] * z.v[0] + x.v[0] * y.v[1] * z.v[0] + x.v[0] * y.v[0] * z.v[1];
          m.columns[e] = (int)index;
        }
        ++e;
      }
    }
    return e;
  };
  PSR_PARALLEL_FOR
  for (int i = 0; i < m.rows; ++i) m.offsets[i + 1] = row(i, false);
  for (int i = 0; i < m.rows; ++i) m.offsets[i + 1] += m.offsets[i];
  m.columns.resize(m.offsets[m.rows]);
  m.values.resize(m.offsets[m.rows]);
  PSR_PARALLEL_FOR
  for (int i = 0; i < m.rows; ++i) row(i, true);
  return m;
}

// src/warp_psr/transfer_stencils.py:15-66
template <class T> CSR<T> assemble_prolongation(const std::vector<int64_t>& coarse, const std::vector<int64_t>& fine, int depth) {
  CSR<T> m;
  m.rows = (int)fine.size();
  m.cols = (int)coarse.size();
  m.offsets.assign(m.rows + 1, 0);
  for (int i = 0; i < m.rows; ++i) {
    I3 f = coords(fine[i]);
    for (int k = 0; k < 8; ++k) {
      I3 c = {{f[0] / 2 + k / 4, f[1] / 2 + (k / 2) % 2, f[2] / 2 + k % 2}};
      T h[3];
      for (int a = 0; a < 3; ++a) h[a] = T(1) - std::fabs(T(f[a]) * T(0.5) - T(c[a]));
      T weight = h[0] * h[1] * h[2];
      if (weight > T(0)) {
        int64_t column = find_key(coarse, basis_key(depth - 1, c[0], c[1], c[2]));
        if (column >= 0) {
          m.columns.push_back((int)column);
          m.values.push_back(weight);
        }
      }
    }
    m.offsets[i + 1] = (int)m.columns.size();
  }
  return m;
}

template <class T> struct Points {
  std::vector<V3<T>> positions;
  std::vector<T> weights;
  std::vector<int64_t> bin_keys;
  std::vector<int> bin_ids;
};

// src/warp_psr/constraints.py:20-84 (level-clustered constraints)
template <class T> Points<T> point_constraints(const State<T>& s, int depth, double point_weight) {
  Points<T> pts;
  const Records<T>& src = s.aggregates[depth];
  const std::vector<int64_t>& active = s.tree_levels[depth];
  size_t m = src.keys.size();
  pts.positions.resize(m);
  pts.weights.resize(m);
  const T pw = T(point_weight);
  for (size_t i = 0; i < m; ++i) {
    const V4<T>& v = src.values[i];
    pts.positions[i] = point_div(V3<T>{{v[0], v[1], v[2]}}, v[3]);
    I3 c = coords(src.keys[i]);
    T w = v[3] * T(1 << depth);
    w = w * (pw * s.normal_average[1]);
    if (depth == 0 || find_key(active, tree_key(c)) < 0) w = T(0);
    pts.weights[i] = w;
  }
  // screening_stencils.py:16-26,80-88 (bin_points + stable sort)
  const int n = 1 << depth;
  std::vector<int64_t> keys(m);
  for (size_t i = 0; i < m; ++i) {
    I3 c;
    for (int a = 0; a < 3; ++a) c[a] = std::min(std::max((int)std::floor(pts.positions[i][a] * T(n)), 0), n - 1);
    keys[i] = basis_key(depth, c[0], c[1], c[2]);
  }
  std::vector<int> ids = stable_order(keys);
  pts.bin_keys.resize(m);
  for (size_t i = 0; i < m; ++i) pts.bin_keys[i] = keys[ids[i]];
  pts.bin_ids = ids;
  return pts;
}

// screening_stencils.py:28-32
template <class T> inline T hat(const V3<T>& p, I3 c, int depth) {
  T h[3];
  for (int a = 0; a < 3; ++a) {
    T t = p[a] * T(1 << depth) - T(c[a]);
    h[a] = wmax(T(1) - std::fabs(t), T(0));
  }
  return h[0] * h[1] * h[2];
}

// screening_stencils.py:34-70 (add)
template <class T>
std::vector<double> add_screening(CSR<double>& A, const std::vector<int64_t>& row_keys, const std::vector<int64_t>& col_keys,
                                  const Points<T>& pts, int depth, int trial_depth) {
  std::vector<double> load(row_keys.size(), 0.0);
  const int n = 1 << depth;
  PSR_PARALLEL_FOR
  for (size_t i = 0; i < row_keys.size(); ++i) {
    I3 c = coords(row_keys[i]);
    const int o0 = A.offsets[i], o1 = A.offsets[i + 1];
    double acc[27];
    for (int e = 0; e < 27; ++e) acc[e] = 0.0;
    double rhs = 0.0;
    for (int k = 0; k < 8; ++k) {
      I3 cell = {{c[0] - k / 4, c[1] - (k / 2) % 2, c[2] - k % 2}};
      if (std::min(std::min(cell[0], cell[1]), cell[2]) >= 0 && std::max(std::max(cell[0], cell[1]), cell[2]) < n) {
        int64_t key = basis_key(depth, cell[0], cell[1], cell[2]);
        int64_t first = find_key(pts.bin_keys, key);
        if (first < 0) continue;
        for (int64_t j = first; j < (int64_t)pts.bin_keys.size() && pts.bin_keys[j] == key; ++j) {
          int index = pts.bin_ids[j];
          const V3<T>& p = pts.positions[index];
          T weight = pts.weights[index];
          if (weight != T(0)) {
            T phi = hat(p, c, depth);
            rhs += double(phi * weight);
            for (int e = o0; e < o1; ++e) {
              I3 qq = coords(col_keys[A.columns[e]]);
              T trial = hat(p, qq, trial_depth);
              acc[e - o0] += double(phi * trial * weight);
            }
          }
        }
      }
    }
    load[i] = rhs;
    for (int e = o0; e < o1; ++e) A.values[e] += acc[e - o0];
  }
  return load;
}

// src/warp_psr/ordered_normal_loads.py:15-84
inline void quadratic_piece(double t0, double step, int offset, double* b) {
  b[0] = b[1] = b[2] = 0.0;
  if (offset == -1) {
    b[0] = 0.5 * (1.0 - t0) * (1.0 - t0);
    b[1] = -(1.0 - t0) * step;
    b[2] = 0.5 * step * step;
  } else if (offset == 0) {
    b[0] = 0.5 + t0 - t0 * t0;
    b[1] = (1.0 - 2.0 * t0) * step;
    b[2] = -step * step;
  } else if (offset == 1) {
    b[0] = 0.5 * t0 * t0;
    b[1] = t0 * step;
    b[2] = 0.5 * step * step;
  }
}

D2 normal_integral(int test_depth, int test_index, int normal_depth, int normal_index) {
  const int nt = 1 << test_depth, nn = 1 << normal_depth;
  const int resolution = std::max(nt, nn);
  const double test_step = double(nt) / double(resolution);
  const double normal_step = double(nn) / double(resolution);
  const int first = std::max(0, std::max((test_index - 1) * (resolution / nt), (normal_index - 1) * (resolution / nn)));
  const int end = std::min(resolution, std::min((test_index + 1) * (resolution / nt), (normal_index + 2) * (resolution / nn)));
  D2 out = {{0.0, 0.0}};
  for (int cell = first; cell < end; ++cell) {
    double tx = double(cell) * test_step - double(test_index);
    double sign = -1.0;
    if (tx + 0.5 * test_step < 0.0) sign = 1.0;
    double h0 = 1.0 + sign * tx;
    double h1 = sign * test_step;
    int normal_cell = (int)std::floor((double(cell) + 0.5) * normal_step);
    double t0 = double(cell) * normal_step - double(normal_cell);
    double b[3], r[3];
    quadratic_piece(t0, normal_step, normal_index - normal_cell, b);
    quadratic_piece(t0, normal_step, -1 - normal_index - normal_cell, r);
    for (int a = 0; a < 3; ++a) b[a] = b[a] - r[a];
    quadratic_piece(t0, normal_step, 2 * nn - 1 - normal_index - normal_cell, r);
    for (int a = 0; a < 3; ++a) b[a] = b[a] - r[a];
    out.v[0] += (h0 * b[0] + (h0 * b[1] + h1 * b[0]) * 0.5 + (h0 * b[2] + h1 * b[1]) / 3.0 + h1 * b[2] * 0.25) /
                double(resolution);
    out.v[1] += (b[0] + b[1] * 0.5 + b[2] / 3.0) * sign * test_step;
  }
  return out;
}

// src/warp_psr/normal_stencils.py:13-70
inline int normal_center(int index, int test_depth, int normal_depth) {
  if (test_depth > normal_depth) return index >> 1;
  if (test_depth < normal_depth) return index << 1;
  return index;
}

template <class T>
std::vector<double> normal_stencil_load(const std::vector<int64_t>& keys, const int64_t* nkeys, const V4<T>* nvals,
                                        int64_t ncount, int test_depth, int normal_depth) {
  std::vector<double> out(keys.size(), 0.0);
  if (!ncount) return out;
  const int width = test_depth < normal_depth ? 6 : 4;
  const int first = test_depth < normal_depth ? -3 : -2;
  const int tn = (1 << test_depth) + 1;
  std::vector<D2> table((size_t)tn * width);
  for (int i = 0; i < tn; ++i)
    for (int k = 0; k < width; ++k)
      table[i * width + k] = normal_integral(test_depth, i, normal_depth, normal_center(i, test_depth, normal_depth) + first + k);
  const int lim = 1 << normal_depth;
  PSR_PARALLEL_FOR
  for (size_t i = 0; i < keys.size(); ++i) {
    I3 c = coords(keys[i]);
    I3 center = {{normal_center(c[0], test_depth, normal_depth), normal_center(c[1], test_depth, normal_depth),
                  normal_center(c[2], test_depth, normal_depth)}};
    double value = 0.0;
    for (int k = 0; k < width * width * width; ++k) {
      int x = k / (width * width), y = (k / width) % width, z = k % width;
      I3 nn = {{center[0] + x + first, center[1] + y + first, center[2] + z + first}};
      if (std::min(std::min(nn[0], nn[1]), nn[2]) >= 0 && std::max(std::max(nn[0], nn[1]), nn[2]) < lim) {
        int64_t j = find_key(nkeys, ncount, basis_key(normal_depth, nn[0], nn[1], nn[2]));
        if (j >= 0) {
          const D2& ix = table[c[0] * width + x];
          const D2& iy = table[c[1] * width + y];
          const D2& iz = table[c[2] * width + z];
          const V4<T>& nv = nvals[j];
          value += double(nv[0]) * ix.v[1] * iy.v[0] * iz.v[0];
          value += double(nv[1]) * ix.v[0] * iy.v[1] * iz.v[0];
          value += double(nv[2]) * ix.v[0] * iy.v[0] * iz.v[1];
        }
      }
    }
    out[i] = value;
  }
  return out;
}

// src/warp_psr/ordered_normal_loads.py:86-94 (normal_stencil) and :137-152
// (contribution): the record's three values against the float64 stencil
// rounded to T, summed in T
template <class T>
T ordered_contribution(int test_depth, const I3& test, int normal_depth, const I3& n, const int64_t* keys, int64_t count,
                       const T* values, int stride) {
  T dot = T(0);
  if (std::min(std::min(n[0], n[1]), n[2]) >= 0 && std::max(std::max(n[0], n[1]), n[2]) < (1 << normal_depth)) {
    const int64_t j = find_key(keys, count, basis_key(normal_depth, n[0], n[1], n[2]));
    if (j >= 0) {
      const D2 x = normal_integral(test_depth, test[0], normal_depth, n[0]);
      const D2 y = normal_integral(test_depth, test[1], normal_depth, n[1]);
      const D2 z = normal_integral(test_depth, test[2], normal_depth, n[2]);
      const double st[3] = {x.v[1] * y.v[0] * z.v[0], x.v[0] * y.v[1] * z.v[0], x.v[0] * y.v[0] * z.v[1]};
      for (int a = 0; a < 3; ++a) dot = dot + values[stride * j + a] * T(st[a]);
    }
  }
  return dot;
}

// src/warp_psr/ordered_normal_loads.py:97-103 (morton_xy)
inline int64_t morton_xy(const I3& c) {
  int64_t key = 0;
  for (int bit = 0; bit < 15; ++bit) {
    key |= (int64_t)((c[0] >> bit) & 1) << (2 * bit);
    key |= (int64_t)((c[1] >> bit) & 1) << (2 * bit + 1);
  }
  return key;
}

// src/warp_psr/ordered_normal_loads.py:263-337 (ordered_normal_loads): the
// loads of a float32 run with base depth 0 (normal_loads.py:84-92), summed in
// T in the pinned serial order, then widened to float64.
template <class T>
std::vector<std::vector<double>> ordered_normal_loads(const State<T>& s, int max_depth, int base_depth) {
  const auto& rk = s.normal_records.keys;
  const auto& rv = s.normal_records.values;
  std::vector<T> rvals(4 * rv.size());
  for (size_t i = 0; i < rv.size(); ++i)
    for (int a = 0; a < 4; ++a) rvals[4 * i + a] = rv[i][a];
  auto own_keys = [&](int d) { return rk.data() + s.normal_offsets[d]; };
  auto own_count = [&](int d) { return (int64_t)(s.normal_offsets[d + 1] - s.normal_offsets[d]); };
  auto own_vals = [&](int d) { return rvals.data() + 4 * s.normal_offsets[d]; };
  std::vector<std::vector<T>> loads(max_depth + 1);
  std::vector<T> finer(s.basis[max_depth].size(), T(0));
  for (int d = max_depth; d >= base_depth; --d) {
    const std::vector<int64_t>& keys = s.basis[d];
    std::vector<T> load(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) {  // :156-171 (same)
      const I3 c = coords(keys[i]);
      T value = finer[i];
      for (int k = 0; k < 64; ++k) {
        const I3 n = {{c[0] + k / 16 - 2, c[1] + (k / 4) % 4 - 2, c[2] + k % 4 - 2}};
        value = value + ordered_contribution<T>(d, c, d, n, own_keys(d), own_count(d), own_vals(d), 4);
      }
      load[i] = value;
    }
    loads[d] = load;
    if (d > base_depth) {
      const std::vector<int64_t>& coarse = s.basis[d - 1];
      std::vector<T> next(coarse.size(), T(0));
      if (own_count(d)) {
        for (size_t i = 0; i < coarse.size(); ++i) {  // :105-122 (order) and :173-188 (parent)
          const I3 c = coords(coarse[i]);
          std::vector<std::pair<int64_t, int>> order(216);
          for (int k = 0; k < 216; ++k) {
            const I3 cc = {{2 * c[0] + k % 6 - 3, 2 * c[1] + (k / 6) % 6 - 3, 2 * c[2] + k / 36 - 3}};
            order[k] = {std::min(std::min(cc[0], cc[1]), cc[2]) >= 0 ? morton_xy(cc) : (int64_t)2147483647, k};
          }
          for (int z = 0; z < 6; ++z) std::stable_sort(order.begin() + 36 * z, order.begin() + 36 * (z + 1),
                                                       [](const auto& a, const auto& b) { return a.first < b.first; });
          T value = T(0);
          for (int j = 0; j < 216; ++j) {
            const int k = order[j].second;
            const I3 n = {{2 * c[0] + k % 6 - 3, 2 * c[1] + (k / 6) % 6 - 3, 2 * c[2] + k / 36 - 3}};
            value = value + ordered_contribution<T>(d - 1, c, d, n, own_keys(d), own_count(d), own_vals(d), 4);
          }
          next[i] = value;
        }
      }
      if (d < max_depth) {
        for (size_t i = 0; i < coarse.size(); ++i) {  // :190-212 (restrict)
          const I3 c = coords(coarse[i]);
          T value = next[i];
          for (int k = 0; k < 27; ++k) {
            const I3 off = {{k / 9 - 1, (k / 3) % 3 - 1, k % 3 - 1}};
            const I3 n = {{2 * c[0] + off[0], 2 * c[1] + off[1], 2 * c[2] + off[2]}};
            if (std::min(std::min(n[0], n[1]), n[2]) >= 0 && std::max(std::max(n[0], n[1]), n[2]) <= (1 << d)) {
              const int64_t j = find_key(keys, basis_key(d, n[0], n[1], n[2]));
              if (j >= 0) {
                T weight = T(1);
                for (int a = 0; a < 3; ++a)
                  if (off[a] != 0) weight = weight * T(0.5);
                value = value + finer[j] * weight;
              }
            }
          }
          next[i] = value;
        }
      }
      finer = std::move(next);
    }
  }
  std::vector<int64_t> previous_keys;
  std::vector<T> previous;
  for (int d = 0; d < max_depth; ++d) {
    const std::vector<int64_t>& keys = s.basis[d];
    std::vector<T> cumulative(4 * keys.size());
    for (size_t i = 0; i < keys.size(); ++i) {  // :214-238 (prolong)
      const I3 c = coords(keys[i]);
      T value[4] = {T(0), T(0), T(0), T(0)};
      const int64_t own = find_key(own_keys(d), own_count(d), keys[i]);
      if (own >= 0)
        for (int a = 0; a < 4; ++a) value[a] = own_vals(d)[4 * own + a];
      if (d > 0) {
        for (int k = 0; k < 8; ++k) {
          const I3 p = {{c[0] / 2 + (k >> 2), c[1] / 2 + ((k >> 1) & 1), c[2] / 2 + (k & 1)}};
          T weight = T(1);
          for (int a = 0; a < 3; ++a) weight = weight * (T(1) - std::fabs(T(c[a]) * T(0.5) - T(p[a])));
          if (weight > T(0)) {
            const int64_t j = find_key(previous_keys, basis_key(d - 1, p[0], p[1], p[2]));
            if (j >= 0)
              for (int a = 0; a < 4; ++a) value[a] = value[a] + previous[4 * j + a] * weight;
          }
        }
      }
      for (int a = 0; a < 4; ++a) cumulative[4 * i + a] = value[a];
    }
    if (d + 1 >= base_depth) {  // :240-258 (from_coarse)
      const std::vector<int64_t>& fine = s.basis[d + 1];
      for (size_t i = 0; i < fine.size(); ++i) {
        const I3 c = coords(fine[i]);
        T value = T(0);
        for (int k = 0; k < 64; ++k) {
          const I3 n = {{c[0] / 2 + k / 16 - 2, c[1] / 2 + (k / 4) % 4 - 2, c[2] / 2 + k % 4 - 2}};
          value = value + ordered_contribution<T>(d + 1, c, d, n, keys.data(), (int64_t)keys.size(), cumulative.data(), 4);
        }
        loads[d + 1][i] = loads[d + 1][i] + value;
      }
    }
    previous_keys = keys;
    previous = std::move(cumulative);
  }
  std::vector<std::vector<double>> out(max_depth + 1);
  for (int d = base_depth; d <= max_depth; ++d) out[d].assign(loads[d].begin(), loads[d].end());
  return out;
}

// src/warp_psr/normal_loads.py:66-190 (stencil_loads=True path); float32 runs
// with base depth 0 take the ordered path (normal_loads.py:84-92)
template <class T> std::vector<std::vector<double>> normal_loads(const State<T>& s, int max_depth, int base_depth) {
  if (std::is_same<T, float>::value && base_depth == 0) return ordered_normal_loads(s, max_depth, base_depth);
  std::vector<std::vector<double>> same(max_depth + 1), from_child(max_depth + 1), from_parent(max_depth + 1);
  std::vector<bool> has_parent(max_depth + 1, false);
  std::vector<int64_t> previous_keys;
  std::vector<V4<T>> previous;
  bool has_cumulative = false;
  const auto& rk = s.normal_records.keys;
  const auto& rv = s.normal_records.values;
  for (int d = 0; d <= max_depth; ++d) {
    const int64_t r0 = s.normal_offsets[d], r1 = s.normal_offsets[d + 1];
    has_cumulative = has_cumulative || (r1 > r0);
    if (d >= base_depth) {
      same[d] = normal_stencil_load(s.basis[d], rk.data() + r0, rv.data() + r0, r1 - r0, d, d);
      if (d > base_depth)
        from_child[d - 1] = normal_stencil_load(s.basis[d - 1], rk.data() + r0, rv.data() + r0, r1 - r0, d - 1, d);
    }
    if (d < max_depth) {
      const std::vector<int64_t>& keys = s.basis[d];
      std::vector<V4<T>> cumulative(keys.size());
      for (size_t i = 0; i < keys.size(); ++i) {  // normal_loads.py:27-63 (prolong)
        I3 c = coords(keys[i]);
        V4<T> value = vzero4<T>();
        int64_t own = find_key(rk.data() + r0, r1 - r0, keys[i]);
        if (own >= 0) value = rv[r0 + own];
        if (d > 0) {
          for (int corner = 0; corner < 8; ++corner) {
            I3 p = {{c[0] / 2, c[1] / 2, c[2] / 2}};
            T weight = T(1);
            for (int axis = 0; axis < 3; ++axis) {
              if ((corner >> axis) & 1) p[axis] += 1;
              weight = weight * (T(1) - std::fabs(T(c[axis]) * T(0.5) - T(p[axis])));
            }
            if (weight > T(0)) {
              int64_t j = find_key(previous_keys, basis_key(d - 1, p[0], p[1], p[2]));
              if (j >= 0)
                for (int a = 0; a < 4; ++a) value[a] = value[a] + previous[j][a] * weight;
            }
          }
        }
        cumulative[i] = value;
      }
      if (d >= base_depth && has_cumulative) {
        from_parent[d + 1] = normal_stencil_load(s.basis[d + 1], keys.data(), cumulative.data(), (int64_t)keys.size(), d + 1, d);
        has_parent[d + 1] = true;
      }
      previous_keys = keys;
      previous = std::move(cumulative);
    }
  }
  std::vector<std::vector<double>> loads(max_depth + 1);
  std::vector<double> finer(s.basis[max_depth].size(), 0.0);
  for (int d = max_depth; d >= base_depth; --d) {
    std::vector<double> load = same[d];
    for (size_t i = 0; i < load.size(); ++i) load[i] = finer[i] + load[i];
    if (has_parent[d])
      for (size_t i = 0; i < load.size(); ++i) load[i] = from_parent[d][i] + load[i];
    loads[d] = load;
    if (d > base_depth) {
      // next_finer += P_d^T finer (warp.sparse.bsr_mv over the transposed
      // prolongation; per coarse row, fine rows in ascending order).
      std::vector<double> next = from_child[d - 1];
      const CSR<T>& P = s.prolong[d - 1];
      std::vector<double> acc(next.size(), 0.0);
      for (int i = 0; i < P.rows; ++i)
        for (int e = P.offsets[i]; e < P.offsets[i + 1]; ++e) acc[P.columns[e]] += double(P.values[e]) * finer[i];
      for (size_t r = 0; r < next.size(); ++r) next[r] = acc[r] + next[r];
      finer = std::move(next);
    }
  }
  return loads;
}

template <class T> CSR<T> cast_csr(const CSR<double>& a) {
  CSR<T> m;
  m.rows = a.rows; m.cols = a.cols; m.offsets = a.offsets; m.columns = a.columns;
  m.values.resize(a.values.size());
  PSR_PARALLEL_FOR
  for (size_t i = 0; i < a.values.size(); ++i) m.values[i] = T(a.values[i]);
  return m;
}

// warp.sparse.bsr_mv semantics: v = sum_k A_k x[col_k] (CSR order); v *= alpha; v += beta*y
template <class T> void spmv(const CSR<T>& A, const std::vector<T>& x, std::vector<T>& y, T alpha, T beta) {
  PSR_PARALLEL_FOR
  for (int i = 0; i < A.rows; ++i) {
    T v = T(0);
    for (int e = A.offsets[i]; e < A.offsets[i + 1]; ++e) v = v + A.values[e] * x[A.columns[e]];
    v = v * alpha;
    if (beta != T(0)) v = v + beta * y[i];
    y[i] = v;
  }
}

template <class T> CSR<T> transpose(const CSR<T>& a) {
  CSR<T> t;
  t.rows = a.cols; t.cols = a.rows;
  t.offsets.assign(t.rows + 1, 0);
  for (int c : a.columns) t.offsets[c + 1]++;
  for (int i = 0; i < t.rows; ++i) t.offsets[i + 1] += t.offsets[i];
  t.columns.resize(a.columns.size());
  t.values.resize(a.values.size());
  std::vector<int> fill(t.offsets.begin(), t.offsets.end() - 1);
  for (int i = 0; i < a.rows; ++i)
    for (int e = a.offsets[i]; e < a.offsets[i + 1]; ++e) {
      int p = fill[a.columns[e]]++;
      t.columns[p] = i;
      t.values[p] = a.values[e];
    }
  return t;
}

// Galerkin product order (row-by-row, k ascending): reference uses warp.sparse.bsr_mm,
// whose internal summation order is not visible here (bitwise parity unpinned).
template <class T> CSR<T> spgemm(const CSR<T>& a, const CSR<T>& b) {
  CSR<T> c;
  c.rows = a.rows; c.cols = b.cols;
  c.offsets.assign(c.rows + 1, 0);
  std::map<int, T> acc;
  for (int i = 0; i < a.rows; ++i) {
    acc.clear();
    for (int e = a.offsets[i]; e < a.offsets[i + 1]; ++e) {
      int k = a.columns[e];
      for (int f = b.offsets[k]; f < b.offsets[k + 1]; ++f) {
        auto it = acc.find(b.columns[f]);
        T prod = a.values[e] * b.values[f];
        if (it == acc.end()) acc.emplace(b.columns[f], T(0) + prod);
        else it->second = it->second + prod;
      }
    }
    for (auto& kv : acc) { c.columns.push_back(kv.first); c.values.push_back(kv.second); }
    c.offsets[i + 1] = (int)c.columns.size();
  }
  return c;
}

// relaxation.py:38-64 (_relax): one Gauss-Seidel row update
template <class T> inline void gs_row(const CSR<T>& A, const std::vector<T>& b, std::vector<T>& x, int i) {
  T r = b[i];
  T diagonal = T(0);
  for (int e = A.offsets[i]; e < A.offsets[i + 1]; ++e) {
    int col = A.columns[e];
    r = r - x[col] * A.values[e];
    if (col == i) diagonal = A.values[e];
  }
  if (diagonal != T(0)) x[i] = x[i] + r * (T(1) / diagonal);
}

struct Schedule {
  std::vector<int> rows, offsets;
};

// relaxation.py:221-238 (row_schedule)
Schedule row_schedule(const std::vector<int64_t>& keys, int depth, bool sliced) {
  const int groups = sliced ? 4 * ((1 << depth) + 1) : 8;
  std::vector<int64_t> g(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) {
    I3 c = coords(keys[i]);
    g[i] = sliced ? (c[0] & 1) + 2 * (c[1] & 1) + 4 * c[2] : (c[0] & 1) + 2 * (c[1] & 1) + 4 * (c[2] & 1);
  }
  Schedule s;
  s.rows = stable_order(g);
  s.offsets.assign(groups + 1, 0);
  for (auto v : g) s.offsets[v + 1]++;
  for (int i = 0; i < groups; ++i) s.offsets[i + 1] += s.offsets[i];
  return s;
}

template <class T> void relax_group(const CSR<T>& A, const std::vector<T>& b, std::vector<T>& x, const Schedule& s, int group) {
  // rows of one colour group are not coupled: their updates commute
  const int m = s.offsets[group + 1] - s.offsets[group];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) if (m > 2048)
#endif
  for (int j = s.offsets[group]; j < s.offsets[group + 1]; ++j) gs_row(A, b, x, s.rows[j]);
  (void)m;
}

// relaxation.py:261-271 (8-colour)
template <class T> void relax_scheduled(const CSR<T>& A, const std::vector<T>& b, std::vector<T>& x, const Schedule& s,
                                        int iters, bool forward) {
  for (int it = 0; it < iters; ++it)
    for (int k = 0; k < 8; ++k) relax_group(A, b, x, s, forward ? k : 7 - k);
}

// relaxation.py:121-125 (sliced_groups order) and :189-210
template <class T> void relax_sliced(const CSR<T>& A, const std::vector<T>& b, std::vector<T>& x, const Schedule& s,
                                     int depth, int iters) {
  const int n = 1 << depth;
#ifdef _OPENMP
  if (omp_get_max_threads() > 1) {
    // relaxation.py:128-143 (wavefront_groups): the slices of one wave are two
    // apart and uncoupled, so a colour of all of them runs as one parallel
    // loop; every row keeps its inputs (the reference's own wavefront proof)
    std::vector<int> rows;
    for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
      const int dq = q - n + 1;
      const int first = std::max(0, dq >= 0 ? dq / 2 : -((-dq + 1) / 2)), last = std::min(iters - 1, q / 2);
      if (last < first) continue;
      const int z0 = n - q + 2 * first;
      for (int color = 0; color < 4; ++color) {
        rows.clear();
        for (int t = 0; t <= last - first; ++t) {
          const int g = 4 * (z0 + 2 * t) + color;
          for (int j = s.offsets[g]; j < s.offsets[g + 1]; ++j) rows.push_back(s.rows[j]);
        }
        const int m = (int)rows.size();
#pragma omp parallel for schedule(dynamic, 256) if (m > 2048)
        for (int k = 0; k < m; ++k) gs_row(A, b, x, rows[k]);
      }
    }
    return;
  }
#endif
  for (int front = n; front > -iters; --front)
    for (int z = std::max(front, 0); z < std::min(front + iters, n + 1); ++z)
      for (int color = 0; color < 4; ++color) relax_group(A, b, x, s, z * 4 + color);
}

// Tile-parallel execution of the sliced schedule, the form the HIP library
// runs (not a reference function).  It is checked here against relax_sliced
// above, bit for bit, to show that the decomposition keeps every update's
// inputs and arithmetic order.
//
// Wave q of wavefront_groups (relaxation.py:128-143) visits slices z two
// apart (visit t = (q + z - n) / 2 of slice z), so during a wave the slices
// z - 1 and z + 1 are not written.  Each slice's (x, y) plane is cut into
// ts x ts tiles; a tile updates colours 0..3 of its rows in order.  Colour c
// at Chebyshev ring r outside the tile needs colours < c at ring r + 1, so a
// tile recomputes colour c on rings up to 3 - c from old values it loaded on
// rings up to 4, and writes only its interior rows.  Two buffers keep loads
// and stores of one wave apart: after t visits a slice's values sit in
// buffer (iters - t) % 2, so visit t of slice z reads slices z - 1 and z from
// buffer (iters - t) % 2, slice z + 1 (t + 1 visits done) from the other,
// and writes slice z there.  Tiles of a wave are run in a seeded random
// order; the result must not depend on it.  Needs same-level couplings
// within one cell in x, y and z (degree-one operators); throws otherwise.
// `plant` injects a known mistake for the tests that show the check has
// teeth: 1 computes each colour on one ring fewer, 2 uses one buffer.
template <class T> void relax_sliced_tiled(const std::vector<int64_t>& keys, const CSR<T>& A, const std::vector<T>& b,
                                           std::vector<T>& x, int depth, int iters, int ts, uint64_t seed, int plant) {
  if (ts < 8) throw std::runtime_error("tile size must be at least 8");
  const int n = 1 << depth, R = (int)keys.size(), nt = (n + 1 + ts - 1) / ts, g = ts + 8;
  std::vector<I3> c(R);
  std::vector<std::vector<int>> slice(n + 1);
  for (int i = 0; i < R; ++i) {
    c[i] = coords(keys[i]);
    slice[c[i][2]].push_back(i);
  }
  std::vector<T> buf[2] = {x, x};
  auto cur = [&](int t) -> std::vector<T>& { return buf[plant == 2 ? 0 : ((iters - t) % 2 + 2) % 2]; };
  uint64_t rng = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return rng; };
  auto ring = [&](int v, int t0) { return std::max(0, std::max(t0 - v, v - (t0 + ts - 1))); };
  for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
    const int d = q - n + 1;  // Python's (q - n + 1) // 2
    const int first = std::max(0, d >= 0 ? d / 2 : -((-d + 1) / 2)), last = std::min(iters - 1, q / 2);
    if (last < first) continue;
    const int z0 = n - q + 2 * first;
    // every tile of every slice of the wave, in a random order
    std::vector<std::pair<int, int>> tiles;  // (slice index, tile)
    std::vector<std::vector<char>> has(last - first + 1, std::vector<char>((size_t)nt * nt, 0));
    for (int si = 0; si <= last - first; ++si)
      for (int i : slice[z0 + 2 * si]) has[si][(size_t)(c[i][1] / ts) * nt + c[i][0] / ts] = 1;
    for (int si = 0; si <= last - first; ++si)
      for (int k = 0; k < nt * nt; ++k)
        if (has[si][k]) tiles.push_back({si, k});
    for (size_t k = tiles.size(); k > 1; --k) std::swap(tiles[k - 1], tiles[next() % k]);
    for (auto [si, tile] : tiles) {
      const int z = z0 + 2 * si, t = first + si;
      const std::vector<T>& P = cur(t);
      std::vector<T>& Q = cur(t + 1);
      const int tx0 = (tile % nt) * ts, ty0 = (tile / nt) * ts;
      std::vector<T> grid((size_t)g * g, T(0));
      std::vector<char> loaded((size_t)g * g, 0);
      auto cell = [&](int i) { return (size_t)(c[i][1] - ty0 + 4) * g + (c[i][0] - tx0 + 4); };
      std::vector<int> rows;
      for (int i : slice[z]) {
        const int r = std::max(ring(c[i][0], tx0), ring(c[i][1], ty0));
        if (r > 4) continue;
        grid[cell(i)] = P[i];
        loaded[cell(i)] = 1;
        rows.push_back(i);
      }
      for (int color = 0; color < 4; ++color) {
        std::vector<std::pair<int, T>> upd;
        for (int i : rows) {
          if ((c[i][0] & 1) + 2 * (c[i][1] & 1) != color) continue;
          const int reach = plant == 1 ? std::max(0, 2 - color) : 3 - color;
          if (std::max(ring(c[i][0], tx0), ring(c[i][1], ty0)) > reach) continue;
          T r = b[i];
          T diagonal = T(0);
          for (int e = A.offsets[i]; e < A.offsets[i + 1]; ++e) {
            const int col = A.columns[e];
            const I3 cc = c[col];
            if (std::abs(cc[0] - c[i][0]) > 1 || std::abs(cc[1] - c[i][1]) > 1 || std::abs(cc[2] - z) > 1)
              throw std::runtime_error("tiled relaxation needs couplings within one cell");
            T xv;
            if (cc[2] == z) {
              if (!loaded[cell(col)]) throw std::runtime_error("tiled relaxation: neighbour not loaded");
              xv = grid[cell(col)];
            } else {
              xv = cc[2] < z ? P[col] : Q[col];
            }
            r = r - xv * A.values[e];
            if (col == i) diagonal = A.values[e];
          }
          T v = grid[cell(i)];
          if (diagonal != T(0)) v = v + r * (T(1) / diagonal);
          upd.push_back({i, v});
        }
        for (auto [i, v] : upd) {
          grid[cell(i)] = v;
          if (c[i][0] / ts == tx0 / ts && c[i][1] / ts == ty0 / ts) Q[i] = v;
        }
      }
    }
  }
  x = cur(iters);
}

// A caller CSR operator (host arrays) as the oracle's CSR.
template <class T> CSR<T> csr_from(int64_t rows, const int* offsets, const int* columns, const void* values) {
  CSR<T> A;
  A.rows = A.cols = (int)rows;
  A.offsets.assign(offsets, offsets + rows + 1);
  A.columns.assign(columns, columns + offsets[rows]);
  A.values.assign((const T*)values, (const T*)values + offsets[rows]);
  return A;
}

// solver.py:30-33
inline int root_row(int64_t key) {
  I3 c = coords(key);
  c = {{c[0] + 2, c[1] + 2, c[2] + 2}};
  return c[2] * 16 + (c[0] & 1) + 2 * (c[1] & 1) + 2 * (c[0] & 2) + 4 * (c[1] & 2);
}

// warp.optim.linear.cg restated from its published algorithm (unpreconditioned CG,
// relative tolerance tol*|b| or absolute atol, whichever is larger, residual checked
// every iteration).  Bitwise parity unpinned.
template <class T, class Op>
void conjugate_gradient(Op apply, const std::vector<T>& b, std::vector<T>& x, double tol, double atol, int maxiter) {
  const size_t n = b.size();
  std::vector<T> r(n), p(n), ap(n);
  apply(x, ap);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - ap[i];
  double bnorm = 0.0;
  for (size_t i = 0; i < n; ++i) bnorm += double(b[i]) * double(b[i]);
  const double threshold = std::max(tol * std::sqrt(bnorm), atol);
  double rz = 0.0;
  for (size_t i = 0; i < n; ++i) rz += double(r[i]) * double(r[i]);
  p = r;
  for (int it = 0; it < maxiter; ++it) {
    if (std::sqrt(rz) <= threshold) break;
    apply(p, ap);
    double pap = 0.0;
    for (size_t i = 0; i < n; ++i) pap += double(p[i]) * double(ap[i]);
    if (pap == 0.0) break;
    T alpha = T(rz / pap);
    for (size_t i = 0; i < n; ++i) { x[i] = x[i] + alpha * p[i]; r[i] = r[i] - alpha * ap[i]; }
    double rz_new = 0.0;
    for (size_t i = 0; i < n; ++i) rz_new += double(r[i]) * double(r[i]);
    T beta = T(rz_new / rz);
    rz = rz_new;
    for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
  }
}

// solver.py:167-202 (padded_root_system): the root operator's rows and
// columns moved to root_row(key) in a 64-row system, row entries in column
// order (as a triplet build sorts them).
template <class T> CSR<T> padded_root_system(const CSR<T>& K, const std::vector<int64_t>& keys) {
  CSR<T> pad;
  pad.rows = pad.cols = 64;
  std::vector<std::vector<std::pair<int, T>>> rowsv(64);
  for (size_t i = 0; i < keys.size(); ++i)
    for (int e = K.offsets[i]; e < K.offsets[i + 1]; ++e)
      rowsv[root_row(keys[i])].push_back({root_row(keys[K.columns[e]]), K.values[e]});
  pad.offsets.assign(65, 0);
  for (int r = 0; r < 64; ++r) {
    std::sort(rowsv[r].begin(), rowsv[r].end(), [](const std::pair<int, T>& a, const std::pair<int, T>& b) { return a.first < b.first; });
    for (auto& e : rowsv[r]) { pad.columns.push_back(e.first); pad.values.push_back(e.second); }
    pad.offsets[r + 1] = (int)pad.columns.size();
  }
  return pad;
}

// solver.py:205-238 (constant_mode_operator): out = M v, plus mean(v) in
// every row when the constant-mode term is on.
template <class T> void constant_mode_apply(const CSR<T>& M, bool dc, const std::vector<T>& v, std::vector<T>& out) {
  for (int i = 0; i < M.rows; ++i) {
    T a = T(0);
    for (int e = M.offsets[i]; e < M.offsets[i + 1]; ++e) a = a + M.values[e] * v[M.columns[e]];
    out[i] = a;
  }
  if (dc) {
    T mean = T(0);
    for (int i = 0; i < M.rows; ++i) mean = mean + v[i];
    mean = mean / T(M.rows);
    for (int i = 0; i < M.rows; ++i) out[i] = out[i] + T(1) * mean;
  }
}

// solver.py:241-317 (regular_multigrid)
template <class T> std::vector<T> regular_multigrid(State<T>& s, const CSR<T>& A, const std::vector<T>& b, int base) {
  std::vector<CSR<T>> mats(base + 1);
  mats[base] = A;
  std::vector<CSR<T>> restr(base);
  for (int d = 0; d < base; ++d) restr[d] = transpose(s.prolong[d]);
  for (int d = base; d > 0; --d) mats[d - 1] = spgemm(spgemm(restr[d - 1], mats[d]), s.prolong[d - 1]);
  std::vector<std::vector<T>> rhs(base + 1), x(base + 1);
  for (int d = 0; d <= base; ++d) {
    rhs[d].assign(s.basis[d].size(), T(0));
    x[d].assign(s.basis[d].size(), T(0));
  }
  rhs[base] = b;
  std::vector<Schedule> sched(base + 1);
  for (int d = 0; d <= base; ++d) sched[d] = row_schedule(s.basis[d], d, false);
  const bool dc = s.prm.base_depth == 0 && s.prm.point_weight == 0.0;
  const int iters = s.prm.iters;
  const double acc = s.prm.cg_accuracy;
  // Root operator; with the constant-mode term the 64-row padded system of
  // solver.py:167-202 (rows/columns mapped by root_row, row entries in column
  // order as a triplet build sorts them, plus mean(x)).
  const CSR<T>& K = mats[0];
  const size_t m0 = x[0].size();
  CSR<T> pad;
  std::vector<T> pad_rhs, pad_x;
  if (dc) {
    pad = padded_root_system(K, s.basis[0]);
    pad_rhs.assign(64, T(0));
    pad_x.assign(64, T(0));
    for (size_t i = 0; i < m0; ++i) pad_rhs[root_row(s.basis[0][i])] = rhs[0][i];
  }
  auto apply = [&](const std::vector<T>& v, std::vector<T>& out) { constant_mode_apply(dc ? pad : K, dc, v, out); };
  for (int cycle = 0; cycle < s.prm.base_v_cycles; ++cycle) {
    for (int d = base; d > 0; --d) {
      relax_scheduled(mats[d], rhs[d], x[d], sched[d], iters, true);
      std::vector<T> residual = rhs[d];
      spmv(mats[d], x[d], residual, T(-1), T(1));
      spmv(restr[d - 1], residual, rhs[d - 1], T(1), T(0));
      std::fill(x[d - 1].begin(), x[d - 1].end(), T(0));
    }
    // Root: skip CG when |r|^2 <= eps^2 (solver.py:276-297)
    std::vector<T>& root_rhs = dc ? pad_rhs : rhs[0];
    std::vector<T>& root_x = dc ? pad_x : x[0];
    std::vector<T> residual = root_rhs;
    if (base == 0 && cycle > 0) {
      std::vector<T> ax(root_rhs.size());
      apply(root_x, ax);
      for (size_t i = 0; i < residual.size(); ++i) residual[i] = root_rhs[i] - ax[i];
    }
    double norm2 = 0.0;
    for (auto v : residual) norm2 += double(v) * double(v);
    if (norm2 > acc * acc) {
      s.cg_runs.push_back(cycle);
      conjugate_gradient<T>(apply, root_rhs, root_x, base == 0 ? 0.0 : acc, base == 0 ? acc * std::sqrt(norm2) : 0.0,
                            (int)s.basis[0].size());
    }
    for (int d = 1; d <= base; ++d) {
      spmv(s.prolong[d - 1], x[d - 1], x[d], T(1), T(1));
      std::fill(x[d - 1].begin(), x[d - 1].end(), T(0));
      relax_scheduled(mats[d], rhs[d], x[d], sched[d], iters, false);
    }
  }
  if (dc)
    for (size_t i = 0; i < m0; ++i) x[0][i] = pad_x[root_row(s.basis[0][i])];
  return x[base];
}

// Mixed precision: finer-level operator values stored in float and widened
// back exactly (the device keeps them as float; vectors stay double).
template <class T> void round_to_float(CSR<T>& a) {
  for (auto& v : a.values) v = T(float(v));
}

template <class T> void solve_levels(State<T>& s) {
  const int finest = s.prm.depth, base = s.prm.base_depth;
  s.prolong.clear();
  for (int d = 1; d <= finest; ++d) s.prolong.push_back(assemble_prolongation<T>(s.basis[d - 1], s.basis[d], d));
  s.corrections.assign(finest + 1, {});
  for (int d = 0; d <= finest; ++d) s.corrections[d].assign(s.basis[d].size(), T(0));
  s.normal_loads = normal_loads(s, finest, base);
  s.matrices.assign(finest + 1, CSR<T>());
  s.cross.assign(finest + 1, CSR<T>());
  s.rhs.assign(finest + 1, {});
  s.point_loads.assign(finest + 1, {});
  const bool screened = s.prm.point_weight > 0 && !s.prm.exact;
  std::vector<T> cumulative;
  for (int d = base; d <= s.prm.solve_depth; ++d) {
    if (s.basis[d].empty()) break;
    const std::vector<double>& load = s.normal_loads[d];
    CSR<double> A = assemble_stiffness(s.basis[d], s.basis[d], d, d);
    Points<T> pts;
    std::vector<double> point_load(s.basis[d].size(), 0.0);
    if (screened) {
      pts = point_constraints(s, d, s.prm.point_weight);
      point_load = add_screening(A, s.basis[d], s.basis[d], pts, d, d);
    }
    s.point_loads[d] = point_load;
    std::vector<T> b(s.basis[d].size());
    for (size_t i = 0; i < b.size(); ++i) b[i] = T(load[i]) + T(0.5) * T(point_load[i]);  // solver.py:119-126
    CSR<T> At = cast_csr<T>(A);
    if (s.prm.float_operators && d != base) round_to_float(At);
    if (d == base) {
      s.corrections[d] = regular_multigrid(s, At, b, base);
      cumulative = s.corrections[d];
    } else {
      CSR<double> cross = assemble_stiffness(s.basis[d], s.basis[d - 1], d, d - 1);
      if (screened) add_screening(cross, s.basis[d], s.basis[d - 1], pts, d, d - 1);
      CSR<T> ct = cast_csr<T>(cross);
      if (s.prm.float_operators) round_to_float(ct);
      spmv(ct, cumulative, b, T(-1), T(1));
      Schedule sched = row_schedule(s.basis[d], d, true);
      relax_sliced(At, b, s.corrections[d], sched, d, s.prm.iters);
      std::vector<T> next = s.corrections[d];
      spmv(s.prolong[d - 1], cumulative, next, T(1), T(1));
      cumulative = std::move(next);
      s.cross[d] = std::move(ct);
    }
    s.rhs[d] = std::move(b);
    s.matrices[d] = std::move(At);
  }
}

// ---------------------------------------------------------------- field
// src/warp_psr/field.py:187-210
template <class T> void cumulative_coefficients(State<T>& s) {
  s.cumulative = s.coefficients;
  const int64_t K = (int64_t)s.keys.size();
  for (int d = 1; d <= s.prm.depth; ++d) {
    PSR_PARALLEL_FOR
    for (int64_t i = 0; i < K; ++i) {
      int64_t key = s.keys[i];
      if (key_depth(key) != d) continue;
      I3 c = coords(key);
      T value = T(0);
      for (int corner = 0; corner < 8; ++corner) {
        int a = corner / 4, b = (corner / 2) % 2, cz = corner % 2;
        T h = (T(1) - std::fabs(T(c[0]) * T(.5) - T(c[0] / 2 + a))) *
              (T(1) - std::fabs(T(c[1]) * T(.5) - T(c[1] / 2 + b))) *
              (T(1) - std::fabs(T(c[2]) * T(.5) - T(c[2] / 2 + cz)));
        if (h > T(0)) {
          int64_t idx = find_key(s.keys, basis_key(d - 1, c[0] / 2 + a, c[1] / 2 + b, c[2] / 2 + cz));
          if (idx >= 0) value = value + s.cumulative[idx] * h;
        }
      }
      s.cumulative[i] = value + s.coefficients[i];
    }
  }
}

// field.py:371-393 (add_level)
template <class T> void add_level(const std::vector<int64_t>& keys, const std::vector<T>& coef, const V3<T>& p, int d, T* result) {
  const int n = 1 << d;
  V3<T> t = {{p[0] * T(n), p[1] * T(n), p[2] * T(n)}};
  I3 cell, base;
  V3<T> f;
  for (int a = 0; a < 3; ++a) {
    cell[a] = std::min(std::max((int)std::ceil(t[a]) - 1, 0), n - 1);
    base[a] = (int)std::floor(t[a]);
    f[a] = t[a] - T(base[a]);
  }
  for (int corner = 0; corner < 8; ++corner) {
    I3 node = {{cell[0] + corner / 4, cell[1] + (corner / 2) % 2, cell[2] + corner % 2}};
    if (std::min(std::min(node[0] - base[0], node[1] - base[1]), node[2] - base[2]) >= 0) {
      int64_t idx = find_key(keys, basis_key(d, node[0], node[1], node[2]));
      if (idx >= 0) {
        T h[3], sg[3];
        for (int a = 0; a < 3; ++a) {
          h[a] = T(1) - f[a];
          sg[a] = T(-1);
          if (node[a] > base[a]) { h[a] = f[a]; sg[a] = T(1); }
        }
        T c = coef[idx];
        result[0] = result[0] + c * h[0] * h[1] * h[2];
        result[1] = result[1] + c * T(n) * sg[0] * h[1] * h[2];
        result[2] = result[2] + c * T(n) * h[0] * sg[1] * h[2];
        result[3] = result[3] + c * T(n) * h[0] * h[1] * sg[2];
      }
    }
  }
}

// field.py:37-75 (evaluation_function): direct level summation of a field
// without cumulative coefficients -- every level 0..depth, the left-owned
// cell (ceil - 1) and its local coordinates, so knots take left derivatives;
// no boundary nudge.
template <class T>
void evaluate_direct(const State<T>& s, const V3<T>& q, const T* L, const T* tr, T* value, T* grad, int* valid) {
  V3<T> p;
  for (int i = 0; i < 3; ++i) p[i] = L[3 * i] * q[0] + L[3 * i + 1] * q[1] + L[3 * i + 2] * q[2] + tr[i];
  bool ok = std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]) &&
            std::min(std::min(p[0], p[1]), p[2]) >= T(0) && std::max(std::max(p[0], p[1]), p[2]) <= T(1);
  *valid = ok ? 1 : 0;
  if (!ok) {
    *value = std::numeric_limits<T>::quiet_NaN();
    grad[0] = grad[1] = grad[2] = *value;
    return;
  }
  T v = T(0), g[3] = {T(0), T(0), T(0)};
  for (int d = 0; d <= s.prm.depth; ++d) {
    const int n = 1 << d;
    int c[3];
    T f[3];
    for (int a = 0; a < 3; ++a) {
      const T t = p[a] * T(n);
      c[a] = std::min(std::max((int)std::ceil(t) - 1, 0), n - 1);
      f[a] = t - T(c[a]);
    }
    for (int corner = 0; corner < 8; ++corner) {
      const int o[3] = {corner / 4, (corner / 2) % 2, corner % 2};
      const int64_t idx = find_key(s.keys, basis_key(d, c[0] + o[0], c[1] + o[1], c[2] + o[2]));
      if (idx < 0) continue;
      T h[3], sg[3];
      for (int a = 0; a < 3; ++a) {
        h[a] = o[a] ? f[a] : T(1) - f[a];
        sg[a] = o[a] ? T(1) : T(-1);
      }
      const T co = s.coefficients[idx];
      v = v + co * h[0] * h[1] * h[2];
      const T cn = co * T(n);
      g[0] = g[0] + cn * (sg[0] * h[1] * h[2]);
      g[1] = g[1] + cn * (h[0] * sg[1] * h[2]);
      g[2] = g[2] + cn * (h[0] * h[1] * sg[2]);
    }
  }
  *value = v;
  for (int i = 0; i < 3; ++i) grad[i] = L[i] * g[0] + L[3 + i] * g[1] + L[6 + i] * g[2];
}

// field.py:395-420 (reference evaluation kernel)
template <class T>
void evaluate_one(const State<T>& s, const V3<T>& q, const T* L, const T* tr, T* value, T* grad, int* valid) {
  V3<T> p;
  for (int i = 0; i < 3; ++i) p[i] = L[3 * i] * q[0] + L[3 * i + 1] * q[1] + L[3 * i + 2] * q[2] + tr[i];
  bool ok = std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]) &&
            std::min(std::min(p[0], p[1]), p[2]) >= T(0) && std::max(std::max(p[0], p[1]), p[2]) <= T(1);
  *valid = ok ? 1 : 0;
  if (!ok) {
    *value = std::numeric_limits<T>::quiet_NaN();
    grad[0] = grad[1] = grad[2] = *value;
    return;
  }
  const int depth = s.prm.depth;
  int leaf = 0;
  for (int d = depth; d >= 0; --d) {
    const int n = 1 << d;
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = std::min(std::max((int)std::ceil(p[a] * T(n)) - 1, 0), n - 1);
    if (find_key(s.keys, basis_key(d, c[0], c[1], c[2])) >= 0) { leaf = d; break; }
  }
  for (int a = 0; a < 3; ++a) {
    if (p[a] == T(0)) p[a] = T(1e-6);
    else if (p[a] == T(1)) p[a] = T(1.0 - 1e-6);
  }
  T r[4] = {T(0), T(0), T(0), T(0)};
  add_level(s.keys, s.coefficients, p, leaf, r);
  if (leaf > 0) {
    T r2[4] = {T(0), T(0), T(0), T(0)};
    add_level(s.keys, s.cumulative, p, leaf - 1, r2);
    for (int a = 0; a < 4; ++a) r[a] = r[a] + r2[a];
  }
  if (leaf < depth) {
    T r2[4] = {T(0), T(0), T(0), T(0)};
    add_level(s.keys, s.coefficients, p, leaf + 1, r2);
    for (int a = 0; a < 4; ++a) r[a] = r[a] + r2[a];
  }
  *value = r[0];
  for (int i = 0; i < 3; ++i) grad[i] = L[i] * r[1] + L[3 + i] * r[2] + L[6 + i] * r[3];
}

// transforms.py:13-79 (cofactor inverse in the input precision)
template <class T> T det3(const T m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
template <class T> T minor4(const T* m, int row, int col) {
  T out[3][3];
  int rr = 0;
  for (int r = 0; r < 4; ++r) {
    if (r == row) continue;
    int cc = 0;
    for (int c = 0; c < 4; ++c) {
      if (c == col) continue;
      out[rr][cc++] = m[r * 4 + c];
    }
    rr++;
  }
  return det3(out);
}
template <class T> void reference_inverse(const T* m, T* out) {
  T det = T(0);
  for (int c = 0; c < 4; ++c) {
    if (c % 2 == 0) det = det + m[c] * minor4(m, 0, c);
    else det = det - m[c] * minor4(m, 0, c);
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      T value = minor4(m, r, c);
      if ((r + c) % 2 != 0) value = -value;
      out[c * 4 + r] = value / det;
    }
}

// Isovalue (reconstruction.py:101-116).  <=1024 aggregates: serial double sum,
// as the reference.  Otherwise per-1024 block partial sums in this fixed order:
// 256 lanes each add elements lane, lane+256, lane+512, lane+768, then a halving
// tree; the partials are combined the same way (lane strides over blocks).
// The reference's tile_sum order is not visible; bitwise parity unpinned.
inline double tree256(double* v) {
  for (int stride = 128; stride > 0; stride >>= 1)
    for (int t = 0; t < stride; ++t) v[t] = v[t] + v[t + stride];
  return v[0];
}
template <class T> T isovalue(const std::vector<T>& values, const std::vector<T>& weights) {
  const size_t G = values.size();
  if (G <= 1024) {
    double total = 0.0, weight = 0.0;
    for (size_t i = 0; i < G; ++i) {
      total += double(values[i] * weights[i]);
      weight += double(weights[i]);
    }
    return T(total / weight);
  }
  const size_t blocks = (G + 1023) / 1024;
  std::vector<double> tp(blocks), mp(blocks);
  double lt[256], lm[256];
  for (size_t blk = 0; blk < blocks; ++blk) {
    for (int t = 0; t < 256; ++t) {
      double a = 0.0, m = 0.0;
      for (int k = 0; k < 4; ++k) {
        size_t i = blk * 1024 + t + 256 * k;
        if (i < G) { a += double(values[i] * weights[i]); m += double(weights[i]); }
      }
      lt[t] = a; lm[t] = m;
    }
    tp[blk] = tree256(lt);
    mp[blk] = tree256(lm);
  }
  for (int t = 0; t < 256; ++t) {
    double a = 0.0, m = 0.0;
    for (size_t i = t; i < blocks; i += 256) { a += tp[i]; m += mp[i]; }
    lt[t] = a; lm[t] = m;
  }
  double total = tree256(lt), mass = tree256(lm);
  return T(total / mass);
}

template <class T> void finish_field(State<T>& s) {
  const int depth = s.prm.depth;
  s.keys.clear(); s.coefficients.clear();
  for (int d = 0; d <= depth; ++d) {
    s.keys.insert(s.keys.end(), s.basis[d].begin(), s.basis[d].end());
    s.coefficients.insert(s.coefficients.end(), s.corrections[d].begin(), s.corrections[d].end());
  }
  cumulative_coefficients(s);
  const size_t G = s.sample_keys.size();
  s.centroid_values.resize(G);
  const T I[9] = {T(1), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(1)};
  const T Z[3] = {T(0), T(0), T(0)};
  for (size_t i = 0; i < G; ++i) {
    V3<T> c = point_div(s.position_sums[i], s.weights[i]);
    T g[3];
    int ok;
    evaluate_one(s, c, I, Z, &s.centroid_values[i], g, &ok);
  }
  s.iso = isovalue(s.centroid_values, s.weights);
  const T sc = s.transform[0];
  const T m[16] = {T(0), sc, T(0), s.transform[1], T(0), T(0), sc, s.transform[2],
                   sc, T(0), T(0), s.transform[3], T(0), T(0), T(0), T(1)};
  reference_inverse(m, s.cube_to_model);
  reference_inverse(s.cube_to_model, s.model_to_cube);
}

template <class T> State<T>* run(const T* pts, const T* nrm, int64_t n, const Params& prm) {
  State<T>* s = new State<T>();
  s->prm = prm;
  s->n = n;
  try {
    prepare_samples(*s, pts, nrm);
    build_density(*s);
    build_normal_field(*s);
    build_tree(*s);
    solve_levels(*s);
    finish_field(*s);
  } catch (...) {
    delete s;
    throw;
  }
  return s;
}

struct Handle {
  int dtype;  // 0 float32, 1 float64
  void* state;
  std::map<std::string, std::vector<char>> cache;
};

thread_local std::string g_error;

template <class X> void put(std::vector<char>& buf, const X* p, size_t n) {
  buf.resize(n * sizeof(X));
  if (n) std::memcpy(buf.data(), p, n * sizeof(X));
}

// Flattened export of one named stage array (level -1 = not per level).
template <class T> bool export_array(State<T>& s, const std::string& name, int level, std::vector<char>& out) {
  auto lvl_ok = [&](size_t n) { return level >= 0 && (size_t)level < n; };
  if (name == "transform") { put(out, s.transform, 5); return true; }
  if (name == "bounds") { put(out, s.bounds, 6); return true; }
  if (name == "positions") { put(out, &s.q[0][0], 3 * s.q.size()); return true; }
  if (name == "unit_normals") { put(out, &s.unit_normals[0][0], 3 * s.unit_normals.size()); return true; }
  if (name == "input_status") { put(out, s.input_status.data(), s.input_status.size()); return true; }
  if (name == "input_group_indices") { put(out, s.input_group.data(), s.input_group.size()); return true; }
  if (name == "input_order") { put(out, s.input_order.data(), s.input_order.size()); return true; }
  if (name == "sample_keys") { put(out, s.sample_keys.data(), s.sample_keys.size()); return true; }
  if (name == "position_sums") { put(out, &s.position_sums[0][0], 3 * s.position_sums.size()); return true; }
  if (name == "normal_sums") { put(out, &s.normal_sums[0][0], 3 * s.normal_sums.size()); return true; }
  if (name == "weights") { put(out, s.weights.data(), s.weights.size()); return true; }
  if (name == "first_indices") { put(out, s.first_indices.data(), s.first_indices.size()); return true; }
  if (name == "group_offsets") { put(out, s.group_offsets.data(), s.group_offsets.size()); return true; }
  if (name == "aggregate_keys" && lvl_ok(s.aggregates.size())) { auto& r = s.aggregates[level]; put(out, r.keys.data(), r.keys.size()); return true; }
  if (name == "aggregate_values" && lvl_ok(s.aggregates.size())) { auto& r = s.aggregates[level]; put(out, r.values.empty() ? (T*)nullptr : &r.values[0][0], 4 * r.values.size()); return true; }
  if (name == "density_keys" && lvl_ok(s.density_levels.size())) { auto& r = s.density_levels[level]; put(out, r.keys.data(), r.keys.size()); return true; }
  if (name == "density_values" && lvl_ok(s.density_levels.size())) {
    auto& r = s.density_levels[level];
    std::vector<T> v(r.values.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = r.values[i][0];
    put(out, v.data(), v.size());
    return true;
  }
  if (name == "densities") { put(out, s.densities.data(), s.densities.size()); return true; }
  if (name == "depths_and_weights") { put(out, s.dw.data(), s.dw.size()); return true; }
  if (name == "density_scale") { put(out, &s.density_scale, 1); return true; }
  if (name == "normal_keys") { put(out, s.normal_records.keys.data(), s.normal_records.keys.size()); return true; }
  if (name == "normal_values") {
    std::vector<T> v(3 * s.normal_records.values.size());
    for (size_t i = 0; i < s.normal_records.values.size(); ++i)
      for (int a = 0; a < 3; ++a) v[3 * i + a] = s.normal_records.values[i][a];
    put(out, v.data(), v.size());
    return true;
  }
  if (name == "normal_average") { put(out, s.normal_average, 3); return true; }
  if (name == "normal_offsets") { put(out, s.normal_offsets.data(), s.normal_offsets.size()); return true; }
  if (name == "tree_keys" && lvl_ok(s.tree_levels.size())) { put(out, s.tree_levels[level].data(), s.tree_levels[level].size()); return true; }
  if (name == "basis_keys" && lvl_ok(s.basis.size())) { put(out, s.basis[level].data(), s.basis[level].size()); return true; }
  if (name == "normal_load" && lvl_ok(s.normal_loads.size())) { put(out, s.normal_loads[level].data(), s.normal_loads[level].size()); return true; }
  if (name == "ordered_normal_load" && lvl_ok(s.basis.size())) {  // the ordered path at any precision, base depth 0
    std::vector<std::vector<double>> loads = ordered_normal_loads(s, (int)s.basis.size() - 1, 0);
    put(out, loads[level].data(), loads[level].size());
    return true;
  }
  if (name == "point_load" && lvl_ok(s.point_loads.size())) { put(out, s.point_loads[level].data(), s.point_loads[level].size()); return true; }
  if (name == "rhs" && lvl_ok(s.rhs.size())) { put(out, s.rhs[level].data(), s.rhs[level].size()); return true; }
  if (name == "corrections" && lvl_ok(s.corrections.size())) { put(out, s.corrections[level].data(), s.corrections[level].size()); return true; }
  if (name == "matrix_offsets" && lvl_ok(s.matrices.size())) { put(out, s.matrices[level].offsets.data(), s.matrices[level].offsets.size()); return true; }
  if (name == "matrix_columns" && lvl_ok(s.matrices.size())) { put(out, s.matrices[level].columns.data(), s.matrices[level].columns.size()); return true; }
  if (name == "matrix_values" && lvl_ok(s.matrices.size())) { put(out, s.matrices[level].values.data(), s.matrices[level].values.size()); return true; }
  if (name == "cross_offsets" && lvl_ok(s.cross.size())) { put(out, s.cross[level].offsets.data(), s.cross[level].offsets.size()); return true; }
  if (name == "cross_columns" && lvl_ok(s.cross.size())) { put(out, s.cross[level].columns.data(), s.cross[level].columns.size()); return true; }
  if (name == "cross_values" && lvl_ok(s.cross.size())) { put(out, s.cross[level].values.data(), s.cross[level].values.size()); return true; }
  if (name == "prolong_offsets" && lvl_ok(s.prolong.size() + 1) && level >= 1) { put(out, s.prolong[level - 1].offsets.data(), s.prolong[level - 1].offsets.size()); return true; }
  if (name == "prolong_columns" && lvl_ok(s.prolong.size() + 1) && level >= 1) { put(out, s.prolong[level - 1].columns.data(), s.prolong[level - 1].columns.size()); return true; }
  if (name == "prolong_values" && lvl_ok(s.prolong.size() + 1) && level >= 1) { put(out, s.prolong[level - 1].values.data(), s.prolong[level - 1].values.size()); return true; }
  if (name == "keys") { put(out, s.keys.data(), s.keys.size()); return true; }
  if (name == "coefficients") { put(out, s.coefficients.data(), s.coefficients.size()); return true; }
  if (name == "cumulative") { put(out, s.cumulative.data(), s.cumulative.size()); return true; }
  if (name == "centroid_values") { put(out, s.centroid_values.data(), s.centroid_values.size()); return true; }
  if (name == "iso") { put(out, &s.iso, 1); return true; }
  if (name == "cube_to_model") { put(out, s.cube_to_model, 16); return true; }
  if (name == "model_to_cube") { put(out, s.model_to_cube, 16); return true; }
  if (name == "cg_runs") { put(out, s.cg_runs.data(), s.cg_runs.size()); return true; }
  return false;
}

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

// Returns 0 on success, 2 for an input rejected as the reference rejects it
// (ValueError), 1 for any other failure.
int oracle_reconstruct(const void* points, const void* normals, int64_t n, int dtype, int depth, int full_depth,
                       int base_depth, int solve_depth, int kernel_depth, double samples_per_node, double scale,
                       double point_weight, int exact, int iters, int base_v_cycles, double cg_accuracy,
                       int float_operators, void** out) {
  Params p{depth, full_depth, base_depth, solve_depth, kernel_depth, iters, base_v_cycles, exact, samples_per_node,
           scale, point_weight, cg_accuracy, 1, float_operators};
  try {
    Handle* h = new Handle();
    h->dtype = dtype;
    if (dtype == 0) h->state = run<float>((const float*)points, (const float*)normals, n, p);
    else h->state = run<double>((const double*)points, (const double*)normals, n, p);
    *out = h;
    return 0;
  } catch (const Error& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

void oracle_free(void* handle) {
  Handle* h = (Handle*)handle;
  if (!h) return;
  if (h->dtype == 0) delete (State<float>*)h->state; else delete (State<double>*)h->state;
  delete h;
}

// Borrow a pointer to a named stage array; *nbytes receives its size in bytes.
int oracle_array(void* handle, const char* name, int level, const void** data, int64_t* nbytes) {
  Handle* h = (Handle*)handle;
  std::string key = std::string(name) + "#" + std::to_string(level);
  auto it = h->cache.find(key);
  if (it == h->cache.end()) {
    std::vector<char> buf;
    bool ok = h->dtype == 0 ? export_array(*(State<float>*)h->state, name, level, buf)
                            : export_array(*(State<double>*)h->state, name, level, buf);
    if (!ok) { g_error = std::string("unknown array ") + name; return 1; }
    it = h->cache.emplace(key, std::move(buf)).first;
  }
  *data = it->second.data();
  *nbytes = (int64_t)it->second.size();
  return 0;
}

// Evaluate the reconstructed field at model-space queries (field.py:395-420
// with the model_to_cube transform, reconstruction.py:316-324).
int oracle_evaluate(void* handle, const void* queries, int64_t m, void* values, void* grads, int32_t* valid) {
  Handle* h = (Handle*)handle;
  auto go = [&](auto* s, auto tag) {
    using T = decltype(tag);
    T L[9], tr[3];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) L[3 * r + c] = s->model_to_cube[r * 4 + c];
      tr[r] = s->model_to_cube[r * 4 + 3];
    }
    const T* q = (const T*)queries;
    PSR_PARALLEL_FOR
    for (int64_t i = 0; i < m; ++i) {
      V3<T> qq = {{q[3 * i], q[3 * i + 1], q[3 * i + 2]}};
      int ok;
      evaluate_one(*s, qq, L, tr, (T*)values + i, (T*)grads + 3 * i, &ok);
      valid[i] = ok;
    }
  };
  if (h->dtype == 0) go((State<float>*)h->state, float());
  else go((State<double>*)h->state, double());
  return 0;
}

// Stand-alone sliced Gauss-Seidel on caller CSR arrays (relaxation.py:121-210),
// used to produce golden vectors for the relaxation kernel.
int oracle_relax_sliced(int dtype, const int64_t* keys, int64_t rows, const int* offsets, const int* columns,
                        const void* values, const void* b, void* x, int depth, int iters) {
  std::vector<int64_t> k(keys, keys + rows);
  Schedule s = row_schedule(k, depth, true);
  auto go = [&](auto tag) {
    using T = decltype(tag);
    CSR<T> A;
    A.rows = (int)rows;
    A.offsets.assign(offsets, offsets + rows + 1);
    A.columns.assign(columns, columns + offsets[rows]);
    A.values.assign((const T*)values, (const T*)values + offsets[rows]);
    std::vector<T> bb((const T*)b, (const T*)b + rows), xx((T*)x, (T*)x + rows);
    relax_sliced(A, bb, xx, s, depth, iters);
    std::copy(xx.begin(), xx.end(), (T*)x);
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// Stand-alone 8-colour Gauss-Seidel (relaxation.py:261-271, relax_scheduled):
// z < 0 sweeps the 8 colour groups of row_schedule(keys, depth) (0..7 or
// 7..0) iters times; z >= 0 relaxes slice z of the sliced schedule, taking the
// colours whose bit 2 equals z & 1 in the same order (group z*4 + (colour & 3)).
int oracle_relax_scheduled(int dtype, const int64_t* keys, int64_t rows, const int* offsets, const int* columns,
                           const void* values, const void* b, void* x, int depth, int iters, int forward, int z) {
  std::vector<int64_t> k(keys, keys + rows);
  Schedule s = row_schedule(k, depth, z >= 0);
  auto go = [&](auto tag) {
    using T = decltype(tag);
    CSR<T> A;
    A.rows = (int)rows;
    A.offsets.assign(offsets, offsets + rows + 1);
    A.columns.assign(columns, columns + offsets[rows]);
    A.values.assign((const T*)values, (const T*)values + offsets[rows]);
    std::vector<T> bb((const T*)b, (const T*)b + rows), xx((T*)x, (T*)x + rows);
    if (z < 0) {
      relax_scheduled(A, bb, xx, s, iters, forward != 0);
    } else {
      for (int it = 0; it < iters; ++it)
        for (int k8 = 0; k8 < 8; ++k8) {
          const int color = forward ? k8 : 7 - k8;
          if ((color >> 2) != (z & 1)) continue;
          relax_group(A, bb, xx, s, z * 4 + (color & 3));
        }
    }
    std::copy(xx.begin(), xx.end(), (T*)x);
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// The tile-parallel sliced schedule (relax_sliced_tiled) on caller CSR arrays.
int oracle_relax_sliced_tiled(int dtype, const int64_t* keys, int64_t rows, const int* offsets, const int* columns,
                              const void* values, const void* b, void* x, int depth, int iters, int tile, uint64_t seed,
                              int plant) {
  try {
    std::vector<int64_t> k(keys, keys + rows);
    auto go = [&](auto tag) {
      using T = decltype(tag);
      CSR<T> A;
      A.rows = (int)rows;
      A.offsets.assign(offsets, offsets + rows + 1);
      A.columns.assign(columns, columns + offsets[rows]);
      A.values.assign((const T*)values, (const T*)values + offsets[rows]);
      std::vector<T> bb((const T*)b, (const T*)b + rows), xx((T*)x, (T*)x + rows);
      relax_sliced_tiled(k, A, bb, xx, depth, iters, tile, seed, plant);
      std::copy(xx.begin(), xx.end(), (T*)x);
    };
    if (dtype == 0) go(float()); else go(double());
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

// ---- known-answer hooks: stand-alone pieces of the pipeline on caller data

// LevelField.evaluate (field.py:150-184) on a caller field (sorted keys of
// levels 0..depth, level corrections, model -> cube linear map and
// translation): the reference evaluation (field.py:364-421) with cumulative
// coefficients, the direct level summation (field.py:37-75) without.
int oracle_field_evaluate(int dtype, const int64_t* keys, int64_t count, const void* coefficients, const void* cumulative,
                          int depth, const double* linear, const double* translation, const void* queries, int64_t m,
                          void* values, void* grads, int32_t* valid) {
  auto go = [&](auto tag) {
    using T = decltype(tag);
    State<T> st;
    st.prm.depth = depth;
    st.keys.assign(keys, keys + count);
    st.coefficients.assign((const T*)coefficients, (const T*)coefficients + count);
    if (cumulative) st.cumulative.assign((const T*)cumulative, (const T*)cumulative + count);
    T L[9], tr[3];
    for (int i = 0; i < 9; ++i) L[i] = T(linear[i]);
    for (int i = 0; i < 3; ++i) tr[i] = T(translation[i]);
    const T* q = (const T*)queries;
    for (int64_t i = 0; i < m; ++i) {
      V3<T> qq = {{q[3 * i], q[3 * i + 1], q[3 * i + 2]}};
      int ok;
      if (cumulative) evaluate_one(st, qq, L, tr, (T*)values + i, (T*)grads + 3 * i, &ok);
      else evaluate_direct(st, qq, L, tr, (T*)values + i, (T*)grads + 3 * i, &ok);
      valid[i] = ok;
    }
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// field.py:187-210 on caller keys (sorted, levels 0..depth) and coefficients.
int oracle_cumulative(int dtype, const int64_t* keys, int64_t count, const void* coefficients, int depth, void* out) {
  auto go = [&](auto tag) {
    using T = decltype(tag);
    State<T> st;
    st.prm.depth = depth;
    st.keys.assign(keys, keys + count);
    st.coefficients.assign((const T*)coefficients, (const T*)coefficients + count);
    cumulative_coefficients(st);
    std::copy(st.cumulative.begin(), st.cumulative.end(), (T*)out);
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// reconstruction.py:76-116 (_compute_isovalue) on caller values and weights.
int oracle_isovalue(int dtype, const void* values, const void* weights, int64_t n, void* out) {
  auto go = [&](auto tag) {
    using T = decltype(tag);
    std::vector<T> v((const T*)values, (const T*)values + n), w((const T*)weights, (const T*)weights + n);
    *(T*)out = isovalue(v, w);
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// screening_stencils.py:80-101 on one uniform level: every node of depth
// `depth` as a row, the 27-neighbour CSR pattern, points binned by floor cell
// (positions in the unit cube); writes the dense count x count screening
// matrix (row-major, keys in sorted order) and the point load.
int oracle_screening_dense(int dtype, const void* positions, const void* weights, int64_t m, int depth, double* dense,
                           double* load) {
  try {
    const int n = 1 << depth;
    std::vector<int64_t> keys;
    for (int x = 0; x <= n; ++x)
      for (int y = 0; y <= n; ++y)
        for (int z = 0; z <= n; ++z) keys.push_back(basis_key(depth, x, y, z));
    const size_t K = keys.size();
    CSR<double> A;
    A.rows = A.cols = (int)K;
    A.offsets.assign(K + 1, 0);
    for (size_t i = 0; i < K; ++i) {
      I3 c = coords(keys[i]);
      for (int k = 0; k < 27; ++k) {
        const int64_t nb = find_key(keys, basis_key(depth, c[0] + k / 9 - 1, c[1] + (k / 3) % 3 - 1, c[2] + k % 3 - 1));
        if (nb >= 0) { A.columns.push_back((int)nb); A.values.push_back(0.0); }
      }
      A.offsets[i + 1] = (int)A.columns.size();
    }
    auto go = [&](auto tag) {
      using T = decltype(tag);
      Points<T> pts;
      const T* p = (const T*)positions;
      for (int64_t i = 0; i < m; ++i) {
        pts.positions.push_back(V3<T>{{p[3 * i], p[3 * i + 1], p[3 * i + 2]}});
        pts.weights.push_back(((const T*)weights)[i]);
      }
      std::vector<int64_t> bins(m);
      for (int64_t i = 0; i < m; ++i) {
        I3 c;
        for (int a = 0; a < 3; ++a) c[a] = std::min(std::max((int)std::floor(pts.positions[i][a] * T(n)), 0), n - 1);
        bins[i] = basis_key(depth, c[0], c[1], c[2]);
      }
      std::vector<int> ids = stable_order(bins);
      for (int64_t i = 0; i < m; ++i) pts.bin_keys.push_back(bins[ids[i]]);
      pts.bin_ids = ids;
      std::vector<double> l = add_screening(A, keys, keys, pts, depth, depth);
      std::copy(l.begin(), l.end(), load);
    };
    if (dtype == 0) go(float()); else go(double());
    std::fill(dense, dense + K * K, 0.0);
    for (size_t i = 0; i < K; ++i)
      for (int e = A.offsets[i]; e < A.offsets[i + 1]; ++e) dense[i * K + A.columns[e]] = A.values[e];
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

// solver.py:167-238 on a caller root operator: the 64-row padded system of
// `keys` (depth-0 nodes) applied with the constant-mode term to v (64
// values): out = P v + mean(v).
int oracle_root_padded_apply(int dtype, const int64_t* keys, int64_t rows, const int* offsets, const int* columns,
                             const void* values, const void* v, void* out) {
  auto go = [&](auto tag) {
    using T = decltype(tag);
    CSR<T> K = csr_from<T>(rows, offsets, columns, values);
    CSR<T> pad = padded_root_system(K, std::vector<int64_t>(keys, keys + rows));
    std::vector<T> vv((const T*)v, (const T*)v + 64), o(64);
    constant_mode_apply(pad, true, vv, o);
    std::copy(o.begin(), o.end(), (T*)out);
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// solver.py:241-317 with base depth 0: `cycles` root solves (skip test, then
// CG on the padded system with the constant-mode term when dc) of a caller
// root operator; x receives the solution.
int oracle_root_solve(int dtype, const int64_t* keys, int64_t rows, const int* offsets, const int* columns,
                      const void* values, const void* b, void* x, int iters, int cycles, double accuracy, int dc) {
  auto go = [&](auto tag) {
    using T = decltype(tag);
    State<T> st;
    st.prm.depth = 0;
    st.prm.base_depth = 0;
    st.prm.point_weight = dc ? 0.0 : 2.0;
    st.prm.iters = iters;
    st.prm.base_v_cycles = cycles;
    st.prm.cg_accuracy = accuracy;
    st.basis.assign(1, std::vector<int64_t>(keys, keys + rows));
    CSR<T> K = csr_from<T>(rows, offsets, columns, values);
    std::vector<T> bb((const T*)b, (const T*)b + rows);
    std::vector<T> sol = regular_multigrid(st, K, bb, 0);
    std::copy(sol.begin(), sol.end(), (T*)x);
  };
  if (dtype == 0) go(float()); else go(double());
  return 0;
}

// 1D integral tables, exported for closed-form checks.
void oracle_hat_integrals(int depth, int index, int trial_depth, int trial_index, double* out) {
  D2 r = hat_integrals(depth, index, trial_depth, trial_index);
  out[0] = r.v[0]; out[1] = r.v[1];
}
void oracle_normal_integral(int test_depth, int test_index, int normal_depth, int normal_index, double* out) {
  D2 r = normal_integral(test_depth, test_index, normal_depth, normal_index);
  out[0] = r.v[0]; out[1] = r.v[1];
}

}  // extern "C"

extern "C" double oracle_det_log(double x) { return det_log(x); }
