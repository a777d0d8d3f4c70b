// Address- and undefined-behaviour-sanitizer run of the CPU oracle (test
// infrastructure; built by `make -C oracle asan` with g++, never linked by the
// product).  Reconstructs seeded clouds in float32, float64 and mixed
// precision, screened and unscreened, with base depth 0 and with V-cycles,
// evaluates the fields, exports every stage array, and runs the stand-alone
// relaxation entry points; any invalid access aborts with the sanitizer's
// report.  Prints "asan ok" on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
int oracle_reconstruct(const void*, const void*, int64_t, int, int, int, int, int, int, double, double, double, int, int,
                       int, double, int, void**);
void oracle_free(void*);
int oracle_array(void*, const char*, int, const void**, int64_t*);
int oracle_evaluate(void*, const void*, int64_t, void*, void*, int32_t*);
int oracle_relax_sliced(int, const int64_t*, int64_t, const int*, const int*, const void*, const void*, void*, int, int);
int oracle_relax_sliced_tiled(int, const int64_t*, int64_t, const int*, const int*, const void*, const void*, void*, int,
                              int, int, uint64_t, int);
}

namespace {
const char* kArrays[] = {"transform", "bounds", "positions", "unit_normals", "input_status", "input_order",
                         "input_group_indices", "sample_keys", "position_sums", "normal_sums", "weights",
                         "first_indices", "group_offsets", "aggregate_keys", "aggregate_values", "density_keys",
                         "density_values", "densities", "depths_and_weights", "density_scale", "normal_keys",
                         "normal_values", "normal_average", "normal_offsets", "tree_keys", "basis_keys",
                         "normal_load", "point_load", "rhs", "matrix_offsets", "matrix_columns", "matrix_values",
                         "cross_offsets", "cross_columns", "cross_values", "prolong_offsets", "prolong_columns",
                         "prolong_values", "corrections", "keys", "coefficients", "cumulative", "centroid_values",
                         "iso", "cube_to_model", "model_to_cube", "cg_runs"};

template <class T> int run(int count, int depth, int full_depth, int base_depth, double weight, int cycles, int mixed) {
  std::mt19937_64 rng(count + depth);
  std::normal_distribution<double> g(0.0, 1.0);
  std::vector<T> p(3 * count), n(3 * count);
  for (int i = 0; i < count; ++i) {  // a torus with a noisy sphere inside
    const double u = 6.283185307179586 * (i % 997) / 997.0, v = 6.283185307179586 * (i % 991) / 991.0;
    const bool torus = i % 3;
    double x[3] = {(0.7 + 0.26 * std::cos(v)) * std::cos(u), (0.7 + 0.26 * std::cos(v)) * std::sin(u), 0.26 * std::sin(v)};
    double nn[3] = {std::cos(v) * std::cos(u), std::cos(v) * std::sin(u), std::sin(v)};
    if (!torus) {
      double d[3] = {g(rng), g(rng), g(rng)};
      const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int a = 0; a < 3; ++a) { nn[a] = d[a] / r; x[a] = 0.3 * nn[a]; }
    }
    for (int a = 0; a < 3; ++a) { p[3 * i + a] = T(x[a] + 1e-3 * g(rng)); n[3 * i + a] = T(nn[a]); }
  }
  const int dtype = sizeof(T) == 8 ? 1 : 0;
  void* h = nullptr;
  int rc = oracle_reconstruct(p.data(), n.data(), count, dtype, depth, full_depth, base_depth, depth,
                              depth > 2 ? depth - 2 : 0, 1.5, 1.1, weight, 0, 8, cycles, 1e-3, mixed, &h);
  if (rc) return rc;
  for (const char* name : kArrays)
    for (int level = -1; level <= depth; ++level) {
      const void* data = nullptr;
      int64_t bytes = 0;
      (void)oracle_array(h, name, level, &data, &bytes);
    }
  std::vector<T> q(3 * 1000), values(1000), grads(3000);
  std::vector<int32_t> valid(1000);
  std::uniform_real_distribution<double> U(-1.2, 1.2);
  for (auto& c : q) c = T(U(rng));
  rc = oracle_evaluate(h, q.data(), 1000, values.data(), grads.data(), valid.data());
  // the finest level's operator through the stand-alone relaxations
  const void *keys = nullptr, *off = nullptr, *cols = nullptr, *vals = nullptr;
  int64_t kb = 0, ob = 0, cb = 0, vb = 0;
  if (!rc && !oracle_array(h, "basis_keys", depth, &keys, &kb) && !oracle_array(h, "matrix_offsets", depth, &off, &ob) &&
      !oracle_array(h, "matrix_columns", depth, &cols, &cb) && !oracle_array(h, "matrix_values", depth, &vals, &vb)) {
    const int64_t rows = kb / 8;
    std::vector<T> b(rows, T(1)), x0(rows, T(0)), x1(rows, T(0));
    if (vb / int64_t(sizeof(T)) == cb / 4) {
      rc |= oracle_relax_sliced(dtype, (const int64_t*)keys, rows, (const int*)off, (const int*)cols, vals, b.data(),
                                x0.data(), depth, 3);
      rc |= oracle_relax_sliced_tiled(dtype, (const int64_t*)keys, rows, (const int*)off, (const int*)cols, vals, b.data(),
                                      x1.data(), depth, 3, 8, 1, 0);
      for (int64_t i = 0; i < rows && !rc; ++i)
        if (x0[i] != x1[i] && !(std::isnan(double(x0[i])) && std::isnan(double(x1[i])))) rc = 3;
    }
  }
  oracle_free(h);
  return rc;
}
}  // namespace

int main() {
  int rc = 0;
  rc |= run<float>(6000, 6, 3, 3, 2.0, 1, 0);
  rc |= run<double>(6000, 6, 3, 3, 0.0, 1, 0);
  rc |= run<double>(6000, 6, 3, 3, 2.0, 1, 1);
  rc |= run<float>(3000, 5, 2, 0, 2.0, 2, 0);
  rc |= run<double>(2000, 4, 0, 0, 0.0, 1, 0);
  std::printf(rc ? "asan failed (%d)\n" : "asan ok\n", rc);
  return rc;
}
