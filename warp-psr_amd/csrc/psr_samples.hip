// Stages 1-3: sample normalisation and grouping (src/warp_psr/samples.py),
// density and splat depth (src/warp_psr/density.py) and the normal-field
// splat (src/warp_psr/normals.py).
//
// Every keyed reduction is "stable sort by key, then one lane sums its run
// serially" so sums keep the reference order.  Splat records are sorted as
// (packed key, record id) pairs only; the reducing lane recomputes each
// record's value from its id instead of storing and re-reading it.
#include "psr_splat.h"

#include <cmath>

namespace psr {

// ------------------------------------------------------------------ samples
template <class T>
__global__ void k_bounds_partial(const T* __restrict__ pts, int64_t n, T* __restrict__ partial, int* __restrict__ bad) {
  __shared__ T sm[6][256];
  __shared__ int sb[256];
  T lo[3], hi[3];
  for (int a = 0; a < 3; ++a) { lo[a] = T(INFINITY); hi[a] = -T(INFINITY); }
  int invalid = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T p0 = pts[3 * i], p1 = pts[3 * i + 1], p2 = pts[3 * i + 2];
    if (isfinite(p0) && isfinite(p1) && isfinite(p2)) {
      lo[0] = tmin(lo[0], p0); hi[0] = tmax(hi[0], p0);
      lo[1] = tmin(lo[1], p1); hi[1] = tmax(hi[1], p1);
      lo[2] = tmin(lo[2], p2); hi[2] = tmax(hi[2], p2);
    } else {
      invalid++;
    }
  }
  for (int a = 0; a < 3; ++a) { sm[a][threadIdx.x] = lo[a]; sm[a + 3][threadIdx.x] = hi[a]; }
  sb[threadIdx.x] = invalid;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      for (int a = 0; a < 3; ++a) {
        sm[a][threadIdx.x] = tmin(sm[a][threadIdx.x], sm[a][threadIdx.x + st]);
        sm[a + 3][threadIdx.x] = tmax(sm[a + 3][threadIdx.x], sm[a + 3][threadIdx.x + st]);
      }
      sb[threadIdx.x] += sb[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int a = 0; a < 6; ++a) partial[blockIdx.x * 6 + a] = sm[a][0];
    bad[blockIdx.x] = sb[0];
  }
}

// One workgroup folds the per-block partials (min/max of finite values are
// order-independent).
template <class T>
__global__ void __launch_bounds__(256) k_bounds_final(const T* __restrict__ partial, const int* __restrict__ bad, int blocks,
                                                      T* __restrict__ out, int* __restrict__ out_bad) {
  __shared__ T sm[6][256];
  __shared__ int sb[256];
  T b[6];
  for (int a = 0; a < 3; ++a) { b[a] = T(INFINITY); b[a + 3] = -T(INFINITY); }
  int invalid = 0;
  for (int k = threadIdx.x; k < blocks; k += 256) {
    for (int a = 0; a < 3; ++a) {
      b[a] = tmin(b[a], partial[k * 6 + a]);
      b[a + 3] = tmax(b[a + 3], partial[k * 6 + a + 3]);
    }
    invalid += bad[k];
  }
  for (int a = 0; a < 6; ++a) sm[a][threadIdx.x] = b[a];
  sb[threadIdx.x] = invalid;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      for (int a = 0; a < 3; ++a) {
        sm[a][threadIdx.x] = tmin(sm[a][threadIdx.x], sm[a][threadIdx.x + st]);
        sm[a + 3][threadIdx.x] = tmax(sm[a + 3][threadIdx.x], sm[a + 3][threadIdx.x + st]);
      }
      sb[threadIdx.x] += sb[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int a = 0; a < 6; ++a) out[a] = sm[a][0];
    *out_bad = sb[0];
  }
}

template <class T> struct Xform { T t[5]; };

// samples.py:90-134 (_classify); packed leaf key x<<2D | y<<D | z, sentinel 1<<3D
template <class T>
__global__ void k_classify(const T* __restrict__ pts, const T* __restrict__ nrm, int64_t n, Xform<T> tf, int depth,
                           T* __restrict__ q, T* __restrict__ unit, uint64_t* __restrict__ keys, int* __restrict__ order,
                           int* __restrict__ status) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* t = tf.t;
  const T inv = t[0];
  T P0 = pts[3 * i], P1 = pts[3 * i + 1], P2 = pts[3 * i + 2];
  T N0 = nrm[3 * i], N1 = nrm[3 * i + 1], N2 = nrm[3 * i + 2];
  T p0 = P1 * inv + t[1], p1 = P2 * inv + t[2], p2 = P0 * inv + t[3];
  T n0 = N1 * t[4], n1 = N2 * t[4], n2 = N0 * t[4];
  T length2 = n0 * n0 + n1 * n1 + n2 * n2;
  q[3 * i] = p0; q[3 * i + 1] = p1; q[3 * i + 2] = p2;
  order[i] = (int)i;
  uint64_t key = (uint64_t)1 << (3 * depth);
  T u0 = T(0), u1 = T(0), u2 = T(0);
  int st = 1;
  const bool ok = length2 > T(0) && isfinite(length2);
  if (ok) st = 2;
  T pmin = wmin(wmin(p0, p1), p2), pmax = wmax(wmax(p0, p1), p2);
  if (ok && pmin >= T(0) && pmax <= T(1)) {
    st = 0;
    T l = sqrt(length2);
    u0 = n0 / l; u1 = n1 / l; u2 = n2 / l;
    const int res = 1 << depth;
    int x = cell_index(p0, res), y = cell_index(p1, res), z = cell_index(p2, res);
    key = ((uint64_t)x << (2 * depth)) | ((uint64_t)y << depth) | (uint64_t)z;
  }
  unit[3 * i] = u0; unit[3 * i + 1] = u1; unit[3 * i + 2] = u2;
  keys[i] = key;
  status[i] = st;
}

// samples.py:149-188 (_groups): serial sums over each run of equal leaf keys.
// A run is [head_pos[g], end): the next head, or for the last group the end
// of its equal keys.
__device__ __forceinline__ int64_t group_end(const uint64_t* __restrict__ keys, const int* __restrict__ head_pos, int g, int G,
                                             int64_t n) {
  if (g + 1 < G) return head_pos[g + 1];
  const uint64_t key = keys[head_pos[g]];
  int64_t lo = head_pos[g], hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] <= key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <class T>
__device__ __forceinline__ void group_fields(const uint64_t key, int g, int G, int64_t i, int64_t end, int depth, const int* order,
                                             int64_t* unique, int* first, int* goff) {
  const uint64_t mask = ((uint64_t)1 << depth) - 1;
  unique[g] = basis_key(depth, (int)(key >> (2 * depth)), (int)((key >> depth) & mask), (int)(key & mask));
  first[g] = order[i];
  goff[g] = (int)i;
  if (g == G - 1) goff[G] = (int)end;
}

constexpr int GROUP_SERIAL_MAX = 16;  // longer runs are summed by a wave (k_groups_wide)

template <class T>
__global__ void k_groups(const uint64_t* __restrict__ keys, const int* __restrict__ order, const int* __restrict__ head_pos,
                         int G, int64_t n, int depth, const T* __restrict__ q, const T* __restrict__ unit,
                         int64_t* __restrict__ unique, T* __restrict__ sums, T* __restrict__ nsums, T* __restrict__ weights,
                         int* __restrict__ first, int* __restrict__ goff, int* __restrict__ group_of, int* __restrict__ wide,
                         int* __restrict__ wide_count) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int64_t i = head_pos[g];
  const int64_t end = group_end(keys, head_pos, g, G, n);
  if (end - i > GROUP_SERIAL_MAX) {
    wide[atomicAdd(wide_count, 1)] = g;
    return;
  }
  T p0 = T(0), p1 = T(0), p2 = T(0), s0 = T(0), s1 = T(0), s2 = T(0), w = T(0);
  for (int64_t j = i; j < end; ++j) {
    const int o = order[j];
    p0 = p0 + q[3 * o]; p1 = p1 + q[3 * o + 1]; p2 = p2 + q[3 * o + 2];
    s0 = s0 + unit[3 * o]; s1 = s1 + unit[3 * o + 1]; s2 = s2 + unit[3 * o + 2];
    w = w + T(1);
    group_of[o] = g;
  }
  group_fields<T>(keys[i], g, G, i, end, depth, order, unique, first, goff);
  sums[3 * g] = p0; sums[3 * g + 1] = p1; sums[3 * g + 2] = p2;
  nsums[3 * g] = s0; nsums[3 * g + 1] = s1; nsums[3 * g + 2] = s2;
  weights[g] = w;
}

// Long runs, one wave each: the lanes load 64 members at a time and every
// lane adds them in run order through lane broadcasts, so the sums equal the
// serial ones bit for bit.
template <class T>
__global__ void __launch_bounds__(64) k_groups_wide(const uint64_t* __restrict__ keys, const int* __restrict__ order,
                                                    const int* __restrict__ head_pos, int G, int64_t n, int depth,
                                                    const T* __restrict__ q, const T* __restrict__ unit,
                                                    int64_t* __restrict__ unique, T* __restrict__ sums, T* __restrict__ nsums,
                                                    T* __restrict__ weights, int* __restrict__ first, int* __restrict__ goff,
                                                    int* __restrict__ group_of, const int* __restrict__ wide,
                                                    const int* __restrict__ wide_count) {
  const int lane = threadIdx.x;
  const int count = *wide_count;
  for (int k = blockIdx.x; k < count; k += gridDim.x) {
    const int g = wide[k];
    const int64_t i = head_pos[g];
    const int64_t end = group_end(keys, head_pos, g, G, n);
    T p0 = T(0), p1 = T(0), p2 = T(0), s0 = T(0), s1 = T(0), s2 = T(0), w = T(0);
    for (int64_t base = i; base < end; base += 64) {
      const int64_t j = base + lane;
      T v[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
      if (j < end) {
        const int o = order[j];
        v[0] = q[3 * o]; v[1] = q[3 * o + 1]; v[2] = q[3 * o + 2];
        v[3] = unit[3 * o]; v[4] = unit[3 * o + 1]; v[5] = unit[3 * o + 2];
        group_of[o] = g;
      }
      const int m = (int)(end - base < 64 ? end - base : 64);
      for (int t = 0; t < m; ++t) {
        p0 = p0 + __shfl(v[0], t, 64); p1 = p1 + __shfl(v[1], t, 64); p2 = p2 + __shfl(v[2], t, 64);
        s0 = s0 + __shfl(v[3], t, 64); s1 = s1 + __shfl(v[4], t, 64); s2 = s2 + __shfl(v[5], t, 64);
        w = w + T(1);
      }
    }
    if (lane == 0) {
      group_fields<T>(keys[i], g, G, i, end, depth, order, unique, first, goff);
      sums[3 * g] = p0; sums[3 * g + 1] = p1; sums[3 * g + 2] = p2;
      nsums[3 * g] = s0; nsums[3 * g + 1] = s1; nsums[3 * g + 2] = s2;
      weights[g] = w;
    }
  }
}

__global__ void k_fill_i32(int* p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class T> void stage_samples(Result<T>& r, const T* pts, const T* nrm) {
  hipStream_t s = r.s;
  const int64_t n = r.n;
  const int depth = r.prm.depth;
  // bounds over finite points (samples.py:34-52)
  const int blocks = (int)std::min<int64_t>(512, std::max<int64_t>(1, (n + 255) / 256));
  DArray<T> partial(6 * blocks, s), bounds(6, s);
  DArray<int> bad(blocks, s), bad_total(1, s);
  hipLaunchKernelGGL(k_bounds_partial<T>, dim3(blocks), dim3(256), 0, s, pts, n, partial.data(), bad.data());
  KERNEL_CHECK();
  hipLaunchKernelGGL(k_bounds_final<T>, dim3(1), dim3(256), 0, s, (const T*)partial.data(), (const int*)bad.data(), blocks,
                     bounds.data(), bad_total.data());
  KERNEL_CHECK();
  struct { T b[6]; int bad; } host;
  read_bytes(host.b, bounds.data(), sizeof(host.b), s);
  read_bytes(&host.bad, bad_total.data(), sizeof(int), s);
  for (int a = 0; a < 6; ++a) r.bounds[a] = host.b[a];
  if (host.bad || compute_transform(r.bounds, r.prm.scale, r.transform))
    throw PsrError(2, "Nonfinite positions or degenerate/overflowing bounding cube");
  Xform<T> tf;
  for (int a = 0; a < 5; ++a) tf.t[a] = r.transform[a];
  // classify + stable sort by leaf key (samples.py:90-134, 269)
  r.q.reset(3 * n, s);
  r.unit_normals.reset(3 * n, s);
  r.input_status.reset(n, s);
  DArray<uint64_t> keys(n, s);
  DArray<int> order(n, s);
  hipLaunchKernelGGL(k_classify<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, pts, nrm, n, tf, depth, r.q.data(),
                     r.unit_normals.data(), keys.data(), order.data(), r.input_status.data());
  KERNEL_CHECK();
  const uint64_t sentinel = (uint64_t)1 << (3 * depth);
  sort_pairs_u64(keys.data(), order.data(), n, 3 * depth + 1, s);
  // head marks + scan (samples.py:136-147, 270-279)
  DArray<int> heads(n, s), prefix(n, s), total(1, s);
  mark_heads_u64(keys.data(), n, sentinel, heads.data(), s);
  scan_exclusive_i32(heads.data(), prefix.data(), n, total.data(), s);
  const int G = read_int(total.data(), s);
  if (!G) throw PsrError(2, "No accepted samples with finite nonzero normals");
  r.G = G;
  DArray<int> head_pos(G, s);
  compact_heads(heads.data(), prefix.data(), n, head_pos.data(), s);
  r.sample_keys.reset(G, s);
  r.position_sums.reset(3 * G, s);
  r.normal_sums.reset(3 * G, s);
  r.weights.reset(G, s);
  r.first_indices.reset(G, s);
  r.group_offsets.reset(G + 1, s);
  r.input_group.reset(n, s);
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, r.input_group.data(), n, -1);
  KERNEL_CHECK();
  DArray<int> wide(G, s), wide_count(1, s);
  wide_count.zero();
  hipLaunchKernelGGL(k_groups<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const uint64_t*)keys.data(),
                     (const int*)order.data(), (const int*)head_pos.data(), G, n, depth, (const T*)r.q.data(),
                     (const T*)r.unit_normals.data(), r.sample_keys.data(), r.position_sums.data(), r.normal_sums.data(),
                     r.weights.data(), r.first_indices.data(), r.group_offsets.data(), r.input_group.data(), wide.data(),
                     wide_count.data());
  KERNEL_CHECK();
  hipLaunchKernelGGL(k_groups_wide<T>, dim3(256), dim3(64), 0, s, (const uint64_t*)keys.data(), (const int*)order.data(),
                     (const int*)head_pos.data(), G, n, depth, (const T*)r.q.data(), (const T*)r.unit_normals.data(),
                     r.sample_keys.data(), r.position_sums.data(), r.normal_sums.data(), r.weights.data(),
                     r.first_indices.data(), r.group_offsets.data(), r.input_group.data(), (const int*)wide.data(),
                     (const int*)wide_count.data());
  KERNEL_CHECK();
  r.input_order = std::move(order);
}

// ------------------------------------------------------------------ density
// density.py:59-72 (_parents) sorted as (parent lexicographic, child octant):
// the reference's Morton-ordered stable sort puts siblings in octant order.
__global__ void k_parent_records(const int64_t* __restrict__ keys, int64_t m, int d, uint64_t* __restrict__ out,
                                 int* __restrict__ ids) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t k = keys[i];
  const int x = key_x(k), y = key_y(k), z = key_z(k);
  const int pb = d - 1;
  const uint64_t parent = ((uint64_t)(x >> 1) << (2 * pb)) | ((uint64_t)(y >> 1) << pb) | (uint64_t)(z >> 1);
  out[i] = (parent << 3) | (uint64_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
  ids[i] = (int)i;
}

__global__ void k_parent_heads(const uint64_t* __restrict__ keys, int64_t m, int* __restrict__ heads) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  heads[i] = (i == 0 || (keys[i] >> 3) != (keys[i - 1] >> 3)) ? 1 : 0;
}

template <class T>
__global__ void k_parent_reduce(const uint64_t* __restrict__ keys, const int* __restrict__ ids, const int* __restrict__ head_pos,
                                int count, int64_t m, int pd, const T* __restrict__ vals, int64_t* __restrict__ out_keys,
                                T* __restrict__ out_vals) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  int64_t j = head_pos[g];
  const uint64_t parent = keys[j] >> 3;
  T a = T(0), b = T(0), c = T(0), w = T(0);
  while (j < m && (keys[j] >> 3) == parent) {
    const T* v = vals + 4 * (int64_t)ids[j];
    a = a + v[0]; b = b + v[1]; c = c + v[2]; w = w + v[3];
    ++j;
  }
  const uint64_t mask = ((uint64_t)1 << pd) - 1;
  out_keys[g] = basis_key(pd, (int)(parent >> (2 * pd)), (int)((parent >> pd) & mask), (int)(parent & mask));
  out_vals[4 * g] = a; out_vals[4 * g + 1] = b; out_vals[4 * g + 2] = c; out_vals[4 * g + 3] = w;
}

template <class T>
__global__ void k_leaf_values(const T* __restrict__ sums, const T* __restrict__ w, int G, T* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  out[4 * i] = sums[3 * i]; out[4 * i + 1] = sums[3 * i + 1]; out[4 * i + 2] = sums[3 * i + 2]; out[4 * i + 3] = w[i];
}

// keyed.py:23-38 Morton codes (x is the lowest bit of each triple)
__global__ void k_morton(const int64_t* __restrict__ keys, int64_t m, int depth, uint64_t* __restrict__ codes,
                         int* __restrict__ ids) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t k = keys[i];
  const int c[3] = {key_x(k), key_y(k), key_z(k)};
  uint64_t code = 0;
  for (int bit = 0; bit < depth; ++bit)
    for (int a = 0; a < 3; ++a) code |= (uint64_t)((c[a] >> bit) & 1) << (3 * bit + a);
  codes[i] = code;
  ids[i] = (int)i;
}

// density.py:74-106 (_splat), one record: target key and value.
template <class T> struct SplatSource {
  int cx, cy, cz;
  T w;
  Quad bx, by, bz;
};
template <class T> __device__ __forceinline__ SplatSource<T> density_source(const int64_t key, const T* v, int d) {
  SplatSource<T> s;
  s.cx = key_x(key); s.cy = key_y(key); s.cz = key_z(key);
  const int n = 1 << d;
  T p0 = v[0] / v[3], p1 = v[1] / v[3], p2 = v[2] / v[3];
  T h = T(1) / T(n);
  s.bx = quadratic(double((p0 - T(s.cx) * h) / h));
  s.by = quadratic(double((p1 - T(s.cy) * h) / h));
  s.bz = quadratic(double((p2 - T(s.cz) * h) / h));
  s.w = v[3];
  return s;
}

__global__ void k_density_splat_keys(const int64_t* __restrict__ keys, const int* __restrict__ order, int64_t m, int d,
                                     uint64_t* __restrict__ out, int* __restrict__ ids) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 27 * m) return;
  const int64_t i = idx / 27;
  const int k = (int)(idx % 27);
  const int64_t key = keys[order[i]];
  const int n = 1 << d;
  const int tx = key_x(key) + k / 9 - 1, ty = key_y(key) + k / 3 % 3 - 1, tz = key_z(key) + k % 3 - 1;
  uint64_t packed = (uint64_t)1 << (3 * d);
  if (imin(imin(tx, ty), tz) >= 0 && imax(imax(tx, ty), tz) < n)
    packed = ((uint64_t)tx << (2 * d)) | ((uint64_t)ty << d) | (uint64_t)tz;
  out[idx] = packed;
  ids[idx] = (int)idx;
}

template <class T>
__global__ void k_density_splat_reduce(const uint64_t* __restrict__ keys, const int* __restrict__ ids,
                                       const int* __restrict__ head_pos, int count, int64_t total, int d,
                                       const int64_t* __restrict__ src_keys, const T* __restrict__ src_vals,
                                       const int* __restrict__ order, T scale, int64_t* __restrict__ out_keys,
                                       T* __restrict__ out_vals) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  int64_t j = head_pos[g];
  const uint64_t key = keys[j];
  T acc = T(0);
  while (j < total && keys[j] == key) {
    const int id = ids[j];
    const int i = id / 27, k = id % 27;
    const int src = order[i];
    SplatSource<T> sp = density_source<T>(src_keys[src], src_vals + 4 * (int64_t)src, d);
    const T weight = sp.w * scale;
    const T val = T(double(weight) * quad_at(sp.bx, k / 9) * quad_at(sp.by, k / 3 % 3) * quad_at(sp.bz, k % 3));
    acc = acc + val;
    ++j;
  }
  const uint64_t mask = ((uint64_t)1 << d) - 1;
  out_keys[g] = basis_key(d, (int)(key >> (2 * d)), (int)((key >> d) & mask), (int)(key & mask));
  out_vals[g] = acc;
}

// density.py:108-150 (density_value at each aggregate centroid)
template <class T>
__global__ void k_density_eval(const int64_t* __restrict__ lkeys, const T* __restrict__ lvals, int64_t lcount, ColumnTable ct,
                               const T* __restrict__ sums, const T* __restrict__ weights, int G, int d, int stride,
                               T* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const T w = weights[i];
  const T p0 = sums[3 * i] / w, p1 = sums[3 * i + 1] / w, p2 = sums[3 * i + 2] / w;
  const int n = 1 << d;
  const int cx = cell_index(p0, n), cy = cell_index(p1, n), cz = cell_index(p2, n);
  const T h = T(1) / T(n);
  const Quad bx = quadratic(double((p0 - T(cx) * h) / h));
  const Quad by = quadratic(double((p1 - T(cy) * h) / h));
  const Quad bz = quadratic(double((p2 - T(cz) * h) / h));
  T value = T(0);
  const int zlo = imax(cz - 1, 0), zhi = imin(cz + 1, n - 1);
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) {  // k = 9x + 3y + z, one lookup per (x, y) column
      const int tx = cx + x - 1, ty = cy + y - 1;
      if (tx < 0 || ty < 0 || tx >= n || ty >= n || zlo > zhi) continue;
      int64_t found[3];
      find_run_table<3>(lkeys, lcount, ct, d, tx, ty, zlo, zhi - zlo + 1, found);
      for (int z = 0; z < 3; ++z) {
        const int tz = cz + z - 1;
        if (tz < zlo || tz > zhi) continue;
        const int64_t idx = found[tz - zlo];
        if (idx >= 0) value = value + T(quad_at(bx, x) * quad_at(by, y) * quad_at(bz, z) * double(lvals[idx]));
      }
    }
  out[(int64_t)i * stride + d] = value;
}

// density.py:152-182 (_depths)
template <class T>
__global__ void k_depths(const T* __restrict__ dens, int G, int kd, T target, T* __restrict__ dw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const int stride = kd + 1;
  int d = kd;
  T density = dens[(int64_t)i * stride + d];
  T raw;
  if (density >= target) {
    raw = T(double(d) + double(T(det_log(double(density / target)))) / (LN2_D * 2.0));
  } else {
    T fine = density;
    while (density < target && d > 0) {
      fine = density;
      d -= 1;
      density = dens[(int64_t)i * stride + d];
    }
    raw = T(T(d) + T(det_log(double(density / target))) / T(det_log(double(density / fine))));
  }
  const T h = T(1) / T(1 << d);
  dw[2 * i] = raw;
  dw[2 * i + 1] = h * h / density;
}

// density.py:37-50 (_scale), evaluated once on the host.
template <class T> T density_scale_host() {
  T total = T(0);
  const T dx = T(0.1);
  for (int x = 0; x < 10; ++x) {
    T p = dx / T(2) + dx * T(x);
    Quad b = quadratic(double(p));
    T scale = T(1.0 / (b.b0 * b.b0 + b.b1 * b.b1 + b.b2 * b.b2));
    for (int y = 0; y < 10; ++y)
      for (int z = 0; z < 10; ++z) total = total + scale;
  }
  return total / T(1000);
}

template <class T> void stage_density(Result<T>& r) {
  hipStream_t s = r.s;
  const int depth = r.prm.depth, kd = r.prm.kernel_depth;
  const int G = r.G;
  r.levels.resize(depth + 1);
  for (auto& L : r.levels) (void)L;
  LevelState<T>& leaf = r.levels[depth];
  leaf.agg_keys.reset(G, s);
  HIP_CHECK(hipMemcpyAsync(leaf.agg_keys.data(), r.sample_keys.data(), G * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
  leaf.agg_vals.reset(4 * (size_t)G, s);
  hipLaunchKernelGGL(k_leaf_values<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const T*)r.position_sums.data(),
                     (const T*)r.weights.data(), G, leaf.agg_vals.data());
  KERNEL_CHECK();
  // aggregates d -> d-1 (density.py:228-240)
  for (int d = depth; d > 0; --d) {
    LevelState<T>& src = r.levels[d];
    const int64_t m = (int64_t)src.agg_keys.size();
    DArray<uint64_t> keys(m, s);
    DArray<int> ids(m, s), heads(m, s), prefix(m, s), total(1, s);
    hipLaunchKernelGGL(k_parent_records, dim3(blocks_for(m, 256)), dim3(256), 0, s, (const int64_t*)src.agg_keys.data(), m,
                       d, keys.data(), ids.data());
    KERNEL_CHECK();
    sort_pairs_u64(keys.data(), ids.data(), m, 3 * d, s);
    hipLaunchKernelGGL(k_parent_heads, dim3(blocks_for(m, 256)), dim3(256), 0, s, (const uint64_t*)keys.data(), m,
                       heads.data());
    KERNEL_CHECK();
    scan_exclusive_i32(heads.data(), prefix.data(), m, total.data(), s);
    const int count = read_int(total.data(), s);
    DArray<int> head_pos(count, s);
    compact_heads(heads.data(), prefix.data(), m, head_pos.data(), s);
    LevelState<T>& dst = r.levels[d - 1];
    dst.agg_keys.reset(count, s);
    dst.agg_vals.reset(4 * (size_t)count, s);
    hipLaunchKernelGGL(k_parent_reduce<T>, dim3(blocks_for(count, 256)), dim3(256), 0, s, (const uint64_t*)keys.data(),
                       (const int*)ids.data(), (const int*)head_pos.data(), count, m, d - 1, (const T*)src.agg_vals.data(),
                       dst.agg_keys.data(), dst.agg_vals.data());
    KERNEL_CHECK();
  }
  r.density_scale = density_scale_host<T>();
  r.densities.reset((size_t)G * (kd + 1), s);
  // density.py:245-271.  The levels are independent: every level's records
  // are sorted and their heads counted first, the counts come to the host in
  // one read, then each level is reduced and evaluated.
  struct Splat {
    int64_t m = 0, total_records = 0;
    DArray<int> order, rids, heads, prefix;
    DArray<uint64_t> rkeys;
  };
  std::vector<Splat> lv(kd + 1);
  DArray<int> totals(kd + 1, s);
  for (int d = 0; d <= kd; ++d) {
    LevelState<T>& L = r.levels[d];
    Splat& v = lv[d];
    v.m = (int64_t)L.agg_keys.size();
    DArray<uint64_t> codes(v.m, s);
    v.order.reset(v.m, s);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(v.m, 256)), dim3(256), 0, s, (const int64_t*)L.agg_keys.data(), v.m, d,
                       codes.data(), v.order.data());
    KERNEL_CHECK();
    sort_pairs_u64(codes.data(), v.order.data(), v.m, 3 * d, s);
    v.total_records = 27 * v.m;
    v.rkeys.reset(v.total_records, s);
    v.rids.reset(v.total_records, s);
    v.heads.reset(v.total_records, s);
    v.prefix.reset(v.total_records, s);
    hipLaunchKernelGGL(k_density_splat_keys, dim3(blocks_for(v.total_records, 256)), dim3(256), 0, s,
                       (const int64_t*)L.agg_keys.data(), (const int*)v.order.data(), v.m, d, v.rkeys.data(), v.rids.data());
    KERNEL_CHECK();
    const uint64_t sentinel = (uint64_t)1 << (3 * d);
    sort_pairs_u64(v.rkeys.data(), v.rids.data(), v.total_records, 3 * d + 1, s);
    mark_heads_u64(v.rkeys.data(), v.total_records, sentinel, v.heads.data(), s);
    scan_exclusive_i32(v.heads.data(), v.prefix.data(), v.total_records, totals.data() + d, s);
  }
  std::vector<int> counts(kd + 1);
  read_bytes(counts.data(), totals.data(), sizeof(int) * (kd + 1), s);
  for (int d = 0; d <= kd; ++d) {
    LevelState<T>& L = r.levels[d];
    Splat& v = lv[d];
    const int count = counts[d];
    DArray<int> head_pos(count, s);
    compact_heads(v.heads.data(), v.prefix.data(), v.total_records, head_pos.data(), s);
    L.dens_keys.reset(count, s);
    L.dens_vals.reset(count, s);
    hipLaunchKernelGGL(k_density_splat_reduce<T>, dim3(blocks_for(count, 256)), dim3(256), 0, s,
                       (const uint64_t*)v.rkeys.data(), (const int*)v.rids.data(), (const int*)head_pos.data(), count,
                       v.total_records, d, (const int64_t*)L.agg_keys.data(), (const T*)L.agg_vals.data(),
                       (const int*)v.order.data(), r.density_scale, L.dens_keys.data(), L.dens_vals.data());
    KERNEL_CHECK();
    ColumnIndex columns;
    build_column_index((const int64_t*)L.dens_keys.data(), count, (1 << d) + 1, columns, s);
    hipLaunchKernelGGL(k_density_eval<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const int64_t*)L.dens_keys.data(),
                       (const T*)L.dens_vals.data(), (int64_t)count, columns.view(), (const T*)r.position_sums.data(),
                       (const T*)r.weights.data(), G, d, kd + 1, r.densities.data());
    KERNEL_CHECK();
  }
  r.dw.reset(2 * (size_t)G, s);
  hipLaunchKernelGGL(k_depths<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const T*)r.densities.data(), G, kd,
                     T(r.prm.samples_per_node), r.dw.data());
  KERNEL_CHECK();
}

// ------------------------------------------------------------------ normals
__global__ void k_first_keys(const int* __restrict__ first, int G, uint64_t* __restrict__ keys, int* __restrict__ ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  keys[i] = (uint64_t)first[i];
  ids[i] = i;
}

template <class T>
__device__ __forceinline__ bool normal_record_valid(int64_t idx, int G, const int* __restrict__ order, const T* __restrict__ sums,
                                                    const T* __restrict__ nsums, const T* __restrict__ weights,
                                                    const T* __restrict__ dw, int minimum, int maximum,
                                                    NormalRecord<T>* rec) {
  if (idx >= 54 * (int64_t)G) return false;
  *rec = normal_record<T>(order[idx / 54], (int)(idx % 54), sums, nsums, weights, dw, minimum, maximum, false);
  return rec->valid;
}

// Valid splat records per 256-record block (normals.py:46-104 skips records
// off the grid and the second depth when the fraction is 1).
template <class T>
__global__ void __launch_bounds__(256) k_normal_splat_count(const int* __restrict__ order, int G, const T* __restrict__ sums,
                                                            const T* __restrict__ nsums, const T* __restrict__ weights,
                                                            const T* __restrict__ dw, int minimum, int maximum,
                                                            int* __restrict__ block_counts) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  NormalRecord<T> rec;
  const bool valid = normal_record_valid(idx, G, order, sums, nsums, weights, dw, minimum, maximum, &rec);
  const int count = __syncthreads_count(valid);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = count;
}

// The valid records, compacted in record order: key top<<3D | x<<2D | y<<D | z,
// the compacted position (the sort payload) and the record's value.
template <class T, class K>
__global__ void __launch_bounds__(256) k_normal_splat_keys(const int* __restrict__ order, int G, const T* __restrict__ sums,
                                                           const T* __restrict__ nsums, const T* __restrict__ weights,
                                                           const T* __restrict__ dw, int minimum, int maximum,
                                                           const int* __restrict__ block_offsets, int64_t capacity,
                                                           K* __restrict__ out, int* __restrict__ ids, T* __restrict__ vals) {
  __shared__ int wave_count[4];
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  NormalRecord<T> rec;
  rec.valid = false;
  if (idx < 54 * (int64_t)G)
    rec = normal_record<T>(order[idx / 54], (int)(idx % 54), sums, nsums, weights, dw, minimum, maximum, true);
  const bool valid = rec.valid;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t ballot = __ballot(valid);
  if (lane == 0) wave_count[wave] = __popcll(ballot);
  __syncthreads();
  if (!valid) return;
  int pos = block_offsets[blockIdx.x] + __popcll(ballot & ((lane == 0) ? 0ull : ((~0ull) >> (64 - lane))));
  for (int w = 0; w < wave; ++w) pos += wave_count[w];
  if (pos >= capacity) return;  // the counting pass saw the same records
  const int D = maximum;
  out[pos] = ((K)rec.top << (3 * D)) | ((K)rec.x << (2 * D)) | ((K)rec.y << D) | (K)rec.z;
  ids[pos] = pos;
  vals[3 * (int64_t)pos] = rec.v0;
  vals[3 * (int64_t)pos + 1] = rec.v1;
  vals[3 * (int64_t)pos + 2] = rec.v2;
}

template <class T, class K>
__global__ void k_normal_splat_reduce(const K* __restrict__ keys, const int* __restrict__ ids, const T* __restrict__ vals,
                                      const int* __restrict__ head_pos, int count, int64_t total, int maximum,
                                      int64_t* __restrict__ out_keys, T* __restrict__ out_vals) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  int64_t j = head_pos[g];
  const K key = keys[j];
  T a = T(0), b = T(0), c = T(0);
  while (j < total && keys[j] == key) {
    const T* v = vals + 3 * (int64_t)ids[j];
    a = a + v[0]; b = b + v[1]; c = c + v[2];
    ++j;
  }
  const int D = maximum;
  const K mask = ((K)1 << D) - 1;
  out_keys[g] = basis_key((int)(key >> (3 * D)), (int)((key >> (2 * D)) & mask), (int)((key >> D) & mask), (int)(key & mask));
  // normals.py:106-109 (_negative)
  out_vals[3 * g] = -a; out_vals[3 * g + 1] = -b; out_vals[3 * g + 2] = -c;
}

// normals.py:111-163: contributions in first-index order, then one serial
// chain; stored component-major (component c of aggregate i at c * G + i).
template <class T>
__global__ void k_average_contributions(const int* __restrict__ order, int G, const T* __restrict__ nsums,
                                        const T* __restrict__ weights, const T* __restrict__ dw, T* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const int s = order[i];
  const T w = weights[s];
  const T n0 = nsums[3 * s] / w, n1 = nsums[3 * s + 1] / w, n2 = nsums[3 * s + 2] / w;
  T v0 = T(0), v1 = T(0), v2 = T(0);
  if (sqrt(n0 * n0 + n1 * n1 + n2 * n2) > T(0)) {
    v0 = dw[2 * s] * w;
    v1 = dw[2 * s + 1] * w;
    v2 = w;
  }
  out[i] = v0; out[G + i] = v1; out[2 * (int64_t)G + i] = v2;
}

// One component chain per wave (three waves on three SIMDs), in element
// order.  A skipped contribution is exactly (0,0,0); the running sums start at
// +0 and never become -0, so adding it (or the +0 padding of the last chunk)
// is the identity and the reference's `> 0` test (normals.py:160-163) needs no
// branch.  Each wave stages AVG_CHUNK values of its component in LDS and reads
// them back with one broadcast LDS load per value, so a step of the chain is
// that load and one add; the next chunk's global loads are in flight
// meanwhile.
constexpr int AVG_PER_LANE = 16;
constexpr int AVG_CHUNK = 64 * AVG_PER_LANE;

template <class T>
__global__ void __launch_bounds__(192) k_average_serial(const T* __restrict__ contrib, int G, T* __restrict__ out) {
  __shared__ T stage[2][3][AVG_CHUNK];
  __shared__ T totals[3];
  const int comp = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (G <= 0) {
    if (threadIdx.x == 0) out[0] = out[1] = out[2] = T(0);
    return;
  }
  const T* __restrict__ src = contrib + (int64_t)comp * G;
  T pre[AVG_PER_LANE];
  auto load = [&](int first) {  // clamped addresses (G >= 1)
#pragma unroll
    for (int k = 0; k < AVG_PER_LANE; ++k) {
      const int e = first + k * 64 + lane;
      pre[k] = src[e < G ? e : G - 1];
    }
  };
  auto keep = [&](int b, int first) {  // +0 past the end
#pragma unroll
    for (int k = 0; k < AVG_PER_LANE; ++k) stage[b][comp][k * 64 + lane] = first + k * 64 + lane < G ? pre[k] : T(0);
  };
  T acc = T(0);
  load(0);
  keep(0, 0);
  __syncthreads();
  int b = 0;
  for (int first = 0; first < G; first += AVG_CHUNK) {
    const bool more = first + AVG_CHUNK < G;
    if (more) load(first + AVG_CHUNK);
    const T* chunk = stage[b][comp];
    // the next 16 values are read while the current 16 are added
    T cur[16], nxt[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) cur[u] = chunk[u];
    for (int k = 0; k < AVG_CHUNK; k += 16) {
      if (k + 16 < AVG_CHUNK) {
#pragma unroll
        for (int u = 0; u < 16; ++u) nxt[u] = chunk[k + 16 + u];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) acc = acc + cur[u];
#pragma unroll
      for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
    }
    // the staging stores wait for the next chunk's loads: keep them after the chain
    __asm__ __volatile__("" ::: "memory");
    if (more) keep(b ^ 1, first + AVG_CHUNK);
    __syncthreads();
    b ^= 1;
  }
  if (lane == 0) totals[comp] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const T t0 = totals[0], t1 = totals[1], t2 = totals[2];
    out[0] = out[1] = out[2] = T(0);
    if (t2 > T(0)) {
      out[0] = t0 / t2;
      out[1] = t1 / t2;
      out[2] = t2;
    }
  }
}

__global__ void k_depth_offsets(const int64_t* __restrict__ keys, int64_t n, int count, int* __restrict__ out) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= count) return;
  out[d] = (int)lower_bound(keys, n, (int64_t)d << 48);
}

// normals.py:46-109: valid splat records compacted, sorted by key (ties in
// record order), reduced per key in that order.
template <class T, class K> void normal_splats(Result<T>& r, int minimum, int maximum, K sentinel) {
  hipStream_t s = r.s;
  const int G = r.G;
  const int64_t records = 54 * (int64_t)G;
  const int blocks = (int)blocks_for(records, 256);
  DArray<int> block_counts(blocks, s), block_offsets(blocks, s), total(1, s);
  hipLaunchKernelGGL(k_normal_splat_count<T>, dim3(blocks), dim3(256), 0, s, (const int*)r.normal_order.data(), G,
                     (const T*)r.position_sums.data(), (const T*)r.normal_sums.data(), (const T*)r.weights.data(),
                     (const T*)r.dw.data(), minimum, maximum, block_counts.data());
  KERNEL_CHECK();
  scan_exclusive_i32(block_counts.data(), block_offsets.data(), blocks, total.data(), s);
  const int64_t valid = read_int(total.data(), s);
  DArray<K> rkeys(valid, s);
  DArray<int> rids(valid, s), heads(valid, s), prefix(valid, s);
  DArray<T> rvals(3 * (size_t)valid, s);
  hipLaunchKernelGGL((k_normal_splat_keys<T, K>), dim3(blocks), dim3(256), 0, s, (const int*)r.normal_order.data(), G,
                     (const T*)r.position_sums.data(), (const T*)r.normal_sums.data(), (const T*)r.weights.data(),
                     (const T*)r.dw.data(), minimum, maximum, (const int*)block_offsets.data(), valid, rkeys.data(),
                     rids.data(), rvals.data());
  KERNEL_CHECK();
  sort_pairs(rkeys.data(), rids.data(), valid, bits_for((uint64_t)sentinel), s);
  mark_heads(rkeys.data(), valid, sentinel, heads.data(), s);
  scan_exclusive_i32(heads.data(), prefix.data(), valid, total.data(), s);
  const int count = valid ? read_int(total.data(), s) : 0;
  DArray<int> head_pos(count, s);
  compact_heads(heads.data(), prefix.data(), valid, head_pos.data(), s);
  r.normal_keys.reset(count, s);
  r.normal_vals.reset(3 * (size_t)count, s);
  if (count) {
    hipLaunchKernelGGL((k_normal_splat_reduce<T, K>), dim3(blocks_for(count, 256)), dim3(256), 0, s,
                       (const K*)rkeys.data(), (const int*)rids.data(), (const T*)rvals.data(), (const int*)head_pos.data(),
                       count, valid, maximum, r.normal_keys.data(), r.normal_vals.data());
    KERNEL_CHECK();
  }
}

template <class T> void stage_normals(Result<T>& r) {
  hipStream_t s = r.s;
  const int G = r.G;
  const int minimum = r.prm.base_depth, maximum = r.prm.depth;
  // order aggregates by first input index (normals.py:190-198)
  {
    DArray<uint64_t> keys(G, s);
    r.normal_order.reset(G, s);
    hipLaunchKernelGGL(k_first_keys, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const int*)r.first_indices.data(), G,
                       keys.data(), r.normal_order.data());
    KERNEL_CHECK();
    sort_pairs_u64(keys.data(), r.normal_order.data(), G, bits_for((uint64_t)r.n), s);
  }
  const uint64_t sentinel = (uint64_t)(maximum + 1) << (3 * maximum);
  if (bits_for(sentinel) <= 32)
    normal_splats<T, uint32_t>(r, minimum, maximum, (uint32_t)sentinel);
  else
    normal_splats<T, uint64_t>(r, minimum, maximum, sentinel);
  // average (normals.py:221-227): contributions here, the serial chain on a
  // side stream (joined before the first use of the average)
  r.average_contrib.reset(3 * (size_t)G, s);
  r.normal_average.reset(3, s);
  hipLaunchKernelGGL(k_average_contributions<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s,
                     (const int*)r.normal_order.data(), G, (const T*)r.normal_sums.data(), (const T*)r.weights.data(),
                     (const T*)r.dw.data(), r.average_contrib.data());
  KERNEL_CHECK();
  if (!r.side) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    r.side = companion_stream(dev, s, 0);
  }
  if (!r.average_done) HIP_CHECK(hipEventCreateWithFlags(&r.average_done, hipEventDisableTiming));
  {
    hipEvent_t ready;
    HIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(ready, s));
    HIP_CHECK(hipStreamWaitEvent(r.side, ready, 0));
    HIP_CHECK(hipEventDestroy(ready));
  }
  hipLaunchKernelGGL(k_average_serial<T>, dim3(1), dim3(192), 0, r.side, (const T*)r.average_contrib.data(), G,
                     r.normal_average.data());
  KERNEL_CHECK();
  HIP_CHECK(hipEventRecord(r.average_done, r.side));
  r.average_pending = true;
  // per-depth record offsets (normals.py:17-29, 228-234)
  DArray<int> offsets(maximum + 2, s);
  hipLaunchKernelGGL(k_depth_offsets, dim3(1), dim3(64), 0, s, (const int64_t*)r.normal_keys.data(),
                     (int64_t)r.normal_keys.size(), maximum + 2, offsets.data());
  KERNEL_CHECK();
  r.normal_offsets.resize(maximum + 2);
  read_bytes(r.normal_offsets.data(), offsets.data(), sizeof(int) * (maximum + 2), s);
}

// samples.py:54-88 (_transform), host scalar code shared by the stages
template <class T> int compute_transform(const T* b, double scale, T* t) {
  T width = wmax(b[3] - b[0], wmax(b[4] - b[1], b[5] - b[2])) * T(scale);
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

#define INSTANTIATE(T)                                                   \
  template void stage_samples<T>(Result<T>&, const T*, const T*);        \
  template void stage_density<T>(Result<T>&);                            \
  template void stage_normals<T>(Result<T>&);                            \
  template int compute_transform<T>(const T*, double, T*);
INSTANTIATE(float)
INSTANTIATE(double)

}  // namespace psr
