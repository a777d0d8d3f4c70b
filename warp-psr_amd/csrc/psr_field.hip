// Stage 6: cumulative coefficients (src/warp_psr/field.py:187-210), the
// reference evaluation (field.py:364-421), the isovalue
// (reconstruction.py:76-138) and the cube/model transforms (transforms.py).
#include "psr_field_index.h"

#include <cmath>

namespace psr {

namespace {

__global__ void k_field_columns(const int64_t* __restrict__ keys, int64_t n, int* __restrict__ start, int* __restrict__ end) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto cell = [&](int64_t k) {
    const int d = key_d(k);
    return column_base(d) + (int64_t)key_x(k) * ((1 << d) + 1) + key_y(k);
  };
  const int64_t c = cell(keys[i]);
  if (i == 0 || cell(keys[i - 1]) != c) start[c] = (int)i;
  if (i == n - 1 || cell(keys[i + 1]) != c) end[c] = (int)(i + 1);
}

template <class T>
__global__ void k_cumulative(FieldIndex f, int d, const T* __restrict__ coef, T* __restrict__ cum) {
  const int64_t i = f.off[d] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.off[d + 1]) return;
  const int64_t key = f.keys[i];
  const int x = key_x(key), y = key_y(key), z = key_z(key);
  T value = T(0);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {  // corners a*4 + b*2 + c in order; one lookup per (x, y) column
      int64_t found[2];
      level_run<2>(f, d - 1, x / 2 + a, y / 2 + b, z / 2, 2, found);
      for (int c = 0; c < 2; ++c) {
        const T h = (T(1) - fabs(T(x) * T(.5) - T(x / 2 + a))) * (T(1) - fabs(T(y) * T(.5) - T(y / 2 + b))) *
                    (T(1) - fabs(T(z) * T(.5) - T(z / 2 + c)));
        if (h > T(0) && found[c] >= 0) value = value + cum[found[c]] * h;
      }
    }
  cum[i] = value + coef[i];
}

template <class T> PSR_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// field.py:371-393 (add_level)
template <class T>
__device__ __forceinline__ void add_level(const FieldIndex& f, const T* coef, T p0, T p1, T p2, int d, T* res) {
  const int n = 1 << d;
  const T t[3] = {p0 * T(n), p1 * T(n), p2 * T(n)};
  int cell[3], base[3];
  T fr[3];
  for (int a = 0; a < 3; ++a) {
    cell[a] = clampi<T>((int)ceil(t[a]) - 1, 0, n - 1);
    base[a] = (int)floor(t[a]);
    fr[a] = t[a] - T(base[a]);
  }
  int64_t found[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) level_run<2>(f, d, cell[0] + a, cell[1] + b, cell[2], 2, found[a][b]);
  for (int corner = 0; corner < 8; ++corner) {
    const int node[3] = {cell[0] + corner / 4, cell[1] + (corner / 2) % 2, cell[2] + corner % 2};
    if (imin(imin(node[0] - base[0], node[1] - base[1]), node[2] - base[2]) < 0) continue;
    const int64_t idx = found[corner / 4][(corner / 2) % 2][corner % 2];
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

// field.py:395-420 (reference evaluation kernel)
template <class T>
__global__ void k_evaluate(FieldIndex f, const T* __restrict__ coef, const T* __restrict__ cum, int depth, const T* __restrict__ queries, int64_t m, Affine<T> A,
                           int point_div_mode, const T* __restrict__ divisors, T* __restrict__ values, T* __restrict__ grads,
                           int* __restrict__ valid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  T q0 = queries[3 * i], q1 = queries[3 * i + 1], q2 = queries[3 * i + 2];
  if (point_div_mode) {  // reconstruction.py:23-29 (_centroids)
    const T w = divisors[i];
    q0 = q0 / w; q1 = q1 / w; q2 = q2 / w;
  }
  T p[3];
  for (int r = 0; r < 3; ++r) p[r] = A.L[3 * r] * q0 + A.L[3 * r + 1] * q1 + A.L[3 * r + 2] * q2 + A.t[r];
  const bool ok = isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && tmin(tmin(p[0], p[1]), p[2]) >= T(0) &&
                  tmax(tmax(p[0], p[1]), p[2]) <= T(1);
  if (valid) valid[i] = ok ? 1 : 0;
  if (!ok) {
    values[i] = T(NAN);
    if (grads) { grads[3 * i] = T(NAN); grads[3 * i + 1] = T(NAN); grads[3 * i + 2] = T(NAN); }
    return;
  }
  int leaf = 0;
  for (int d = depth; d >= 0; --d) {
    const int n = 1 << d;
    const int c0 = clampi<T>((int)ceil(p[0] * T(n)) - 1, 0, n - 1);
    const int c1 = clampi<T>((int)ceil(p[1] * T(n)) - 1, 0, n - 1);
    const int c2 = clampi<T>((int)ceil(p[2] * T(n)) - 1, 0, n - 1);
    int64_t idx;
    level_run<1>(f, d, c0, c1, c2, 1, &idx);
    if (idx >= 0) { leaf = d; break; }
  }
  for (int a = 0; a < 3; ++a) {
    if (p[a] == T(0)) p[a] = T(1e-6);
    else if (p[a] == T(1)) p[a] = T(1.0 - 1e-6);
  }
  T res[4] = {T(0), T(0), T(0), T(0)};
  add_level(f, coef, p[0], p[1], p[2], leaf, res);
  if (leaf > 0) {
    T r2[4] = {T(0), T(0), T(0), T(0)};
    add_level(f, cum, p[0], p[1], p[2], leaf - 1, r2);
    for (int a = 0; a < 4; ++a) res[a] = res[a] + r2[a];
  }
  if (leaf < depth) {
    T r2[4] = {T(0), T(0), T(0), T(0)};
    add_level(f, coef, p[0], p[1], p[2], leaf + 1, r2);
    for (int a = 0; a < 4; ++a) res[a] = res[a] + r2[a];
  }
  values[i] = res[0];
  if (grads)
    for (int c = 0; c < 3; ++c) grads[3 * i + c] = A.L[c] * res[1] + A.L[3 + c] * res[2] + A.L[6 + c] * res[3];
}

// Isovalue: <=1024 aggregates one serial double chain (reconstruction.py:66-83);
// otherwise per-1024 partials: 256 lanes add elements lane+256k (k=0..3) in
// order, then a halving tree; partials combined the same way.
template <class T>
__global__ void __launch_bounds__(256) k_iso_partials(const T* __restrict__ v, const T* __restrict__ w, int64_t G,
                                                      double* __restrict__ tp, double* __restrict__ mp) {
  __shared__ double st[256], sm[256];
  double a = 0.0, m = 0.0;
  for (int k = 0; k < 4; ++k) {
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x + 256 * k;
    if (i < G) { a += double(v[i] * w[i]); m += double(w[i]); }
  }
  st[threadIdx.x] = a;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      st[threadIdx.x] = st[threadIdx.x] + st[threadIdx.x + stride];
      sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { tp[blockIdx.x] = st[0]; mp[blockIdx.x] = sm[0]; }
}

template <class T>
__global__ void __launch_bounds__(256) k_iso_finish(const double* __restrict__ tp, const double* __restrict__ mp, int64_t blocks,
                                                    T* __restrict__ out) {
  __shared__ double st[256], sm[256];
  double a = 0.0, m = 0.0;
  for (int64_t i = threadIdx.x; i < blocks; i += 256) { a += tp[i]; m += mp[i]; }
  st[threadIdx.x] = a;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      st[threadIdx.x] = st[threadIdx.x] + st[threadIdx.x + stride];
      sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = T(st[0] / sm[0]);
}

template <class T>
__global__ void k_iso_serial(const T* __restrict__ v, const T* __restrict__ w, int64_t G, T* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double total = 0.0, weight = 0.0;
  for (int64_t i = 0; i < G; ++i) {
    total += double(v[i] * w[i]);
    weight += double(w[i]);
  }
  *out = T(total / weight);
}

}  // namespace

template <class T>
void isovalue_launch(const T* v, const T* w, int64_t G, double* tp, double* mp, T* out, hipStream_t s) {
  if (G <= 1024) {
    hipLaunchKernelGGL(k_iso_serial<T>, dim3(1), dim3(64), 0, s, v, w, G, out);
  } else {
    const int64_t blocks = isovalue_partials(G);
    hipLaunchKernelGGL(k_iso_partials<T>, dim3((unsigned)blocks), dim3(256), 0, s, v, w, G, tp, mp);
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_iso_finish<T>, dim3(1), dim3(256), 0, s, (const double*)tp, (const double*)mp, blocks, out);
  }
  KERNEL_CHECK();
}

template <class T> void compute_isovalue(const T* v, const T* w, int64_t G, T* out, hipStream_t s) {
  const int64_t blocks = G > 1024 ? isovalue_partials(G) : 0;
  DArray<double> tp(blocks, s), mp(blocks, s);
  isovalue_launch(v, w, G, tp.data(), mp.data(), out, s);
}

// The isovalue reduction without allocation, for callers that cannot call
// hipMalloc (HIP graph capture, tight serving loops): the double partials
// live in a file-scope device array, shared by every call on a GPU (calls
// that overlap on one GPU race on it).  Every partial a call reads is written
// by the same call first, so nothing needs clearing.  `scratch` is a
// host-side stand-in for the array in host code; its device address is
// looked up once per device (with that device current) and only that
// address is passed to kernels.
constexpr int ISO_SCRATCH_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[2 * ISOVALUE_NOALLOC_PARTIALS];

static double* scratch_address(int dev) {
  static std::mutex mu;
  static double* addr[ISO_SCRATCH_MAX_DEVICES] = {};
  if (dev < 0 || dev >= ISO_SCRATCH_MAX_DEVICES) throw PsrError(1, "device index beyond the supported range");
  std::lock_guard<std::mutex> g(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      throw PsrError(1, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    addr[dev] = (double*)p;
  }
  return addr[dev];
}

template <class T> void isovalue_noalloc(const T* v, const T* w, int64_t G, T* out, hipStream_t s) {
  if (G > ISOVALUE_NOALLOC_MAX)
    throw PsrError(2, "psr_isovalue_noalloc handles at most " + std::to_string(ISOVALUE_NOALLOC_MAX) +
                          " values (its partial sums live in a fixed device array)");
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  double* tp = scratch_address(dev);
  isovalue_launch(v, w, G, tp, tp + ISOVALUE_NOALLOC_PARTIALS, out, s);
}

template void isovalue_noalloc<float>(const float*, const float*, int64_t, float*, hipStream_t);
template void isovalue_noalloc<double>(const double*, const double*, int64_t, double*, hipStream_t);
template void isovalue_launch<float>(const float*, const float*, int64_t, double*, double*, float*, hipStream_t);
template void isovalue_launch<double>(const double*, const double*, int64_t, double*, double*, double*, hipStream_t);
template void compute_isovalue<float>(const float*, const float*, int64_t, float*, hipStream_t);
template void compute_isovalue<double>(const double*, const double*, int64_t, double*, hipStream_t);

// transforms.py:13-52 (cofactor inverse in the input precision)
template <class T> static T det3(const T m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
template <class T> static T minor4(const T* m, int row, int col) {
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


template <class T>
void evaluate_field(const Result<T>& r, const T* queries, int64_t m, const T* linear, const T* translation, T* values, T* grads,
                    int* valid, hipStream_t s) {
  if (m <= 0) return;
  Affine<T> A;
  for (int k = 0; k < 9; ++k) A.L[k] = linear[k];
  for (int k = 0; k < 3; ++k) A.t[k] = translation[k];
  hipLaunchKernelGGL(k_evaluate<T>, dim3(blocks_for(m, 256)), dim3(256), 0, s, field_index(r),
                     (const T*)r.coefficients.data(), (const T*)r.cumulative.data(), r.prm.depth,
                     queries, m, A, 0, (const T*)nullptr, values, grads, valid);
  KERNEL_CHECK();
}

template <class T> void stage_field(Result<T>& r) {
  hipStream_t s = r.s;
  const int D = r.prm.depth;
  r.join_average();
  read_bytes(r.normal_average_host, r.normal_average.data(), sizeof(T) * 3, s);
  r.average_contrib.free();
  r.level_offsets.assign(D + 2, 0);
  for (int d = 0; d <= D; ++d) r.level_offsets[d + 1] = r.level_offsets[d] + r.levels[d].basis_count;
  const int64_t K = r.level_offsets[D + 1];
  r.keys.reset(K, s);
  r.coefficients.reset(K, s);
  for (int d = 0; d <= D; ++d) {
    const LevelState<T>& L = r.levels[d];
    const int64_t o = r.level_offsets[d];
    if (!L.basis_count) continue;
    HIP_CHECK(hipMemcpyAsync(r.keys.data() + o, L.basis_keys.data(), sizeof(int64_t) * L.basis_count, hipMemcpyDeviceToDevice, s));
    if (L.correction.size())
      HIP_CHECK(hipMemcpyAsync(r.coefficients.data() + o, L.correction.data(), sizeof(T) * L.basis_count, hipMemcpyDeviceToDevice, s));
    else
      HIP_CHECK(hipMemsetAsync(r.coefficients.data() + o, 0, sizeof(T) * L.basis_count, s));
  }
  r.offsets_dev.reset(D + 2, s);
  HIP_CHECK(hipMemcpyAsync(r.offsets_dev.data(), r.level_offsets.data(), sizeof(int) * (D + 2), hipMemcpyHostToDevice, s));
  if (D <= field_table_max_depth() && K) {
    const size_t cells = (size_t)column_base(D + 1);
    r.column_start.reset(cells, s);
    r.column_end.reset(cells, s);
    r.column_start.zero();
    r.column_end.zero();
    hipLaunchKernelGGL(k_field_columns, dim3(blocks_for(K, 256)), dim3(256), 0, s, (const int64_t*)r.keys.data(), K,
                       r.column_start.data(), r.column_end.data());
    KERNEL_CHECK();
  } else {
    r.column_start.free();
    r.column_end.free();
  }
  const FieldIndex f = field_index(r);
  r.cumulative.reset(K, s);
  if (K) HIP_CHECK(hipMemcpyAsync(r.cumulative.data(), r.coefficients.data(), sizeof(T) * K, hipMemcpyDeviceToDevice, s));
  for (int d = 1; d <= D; ++d) {
    const int cnt = r.levels[d].basis_count;
    if (!cnt) continue;
    hipLaunchKernelGGL(k_cumulative<T>, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, f, d, (const T*)r.coefficients.data(),
                       r.cumulative.data());
    KERNEL_CHECK();
  }
  // centroid values with the cube-space field (reconstruction.py:295-303)
  const int G = r.G;
  r.centroid_values.reset(G, s);
  Affine<T> I;
  for (int k = 0; k < 9; ++k) I.L[k] = (k % 4 == 0) ? T(1) : T(0);
  for (int k = 0; k < 3; ++k) I.t[k] = T(0);
  hipLaunchKernelGGL(k_evaluate<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, f, (const T*)r.coefficients.data(),
                     (const T*)r.cumulative.data(), D,
                     (const T*)r.position_sums.data(), (int64_t)G, I, 1, (const T*)r.weights.data(), r.centroid_values.data(),
                     (T*)nullptr, (int*)nullptr);
  KERNEL_CHECK();
  DArray<T> iso(1, s);
  compute_isovalue((const T*)r.centroid_values.data(), (const T*)r.weights.data(), (int64_t)G, iso.data(), s);
  read_bytes(&r.iso, iso.data(), sizeof(T), s);
  // transforms.py:54-79 (host scalar work)
  const T sc = r.transform[0];
  const T M[16] = {T(0), sc, T(0), r.transform[1], T(0), T(0), sc, r.transform[2],
                   sc, T(0), T(0), r.transform[3], T(0), T(0), T(0), T(1)};
  reference_inverse(M, r.cube_to_model);
  reference_inverse(r.cube_to_model, r.model_to_cube);
}

template void stage_field<float>(Result<float>&);
template void stage_field<double>(Result<double>&);
template void evaluate_field<float>(const Result<float>&, const float*, int64_t, const float*, const float*, float*, float*, int*,
                                    hipStream_t);
template void evaluate_field<double>(const Result<double>&, const double*, int64_t, const double*, const double*, double*,
                                     double*, int*, hipStream_t);
template void reference_inverse<float>(const float*, float*);
template void reference_inverse<double>(const double*, double*);

}  // namespace psr
