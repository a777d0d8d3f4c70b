// Stage 5: operators and the cascadic solve (src/warp_psr/solver.py:320-466).
//
// Matrices are 27-slot stencils: slot k of row i holds the neighbour at
// offset (k/9-1, k/3%3-1, k%3-1) around the row node (same level) or around its
// parent (coarser trial level); absent neighbours have column -1.  Iterating
// present slots in k order is exactly the reference's ascending-column CSR
// order, so every row sum keeps the reference order.
//
// Gauss-Seidel runs on a copy of the operator gathered into schedule order
// (rows of one colour/slice group contiguous, slot-major), so the loads of one
// phase are coalesced; x stays in canonical order for the neighbour gathers.
#include "psr_result.h"

#include <algorithm>
#include <condition_variable>
#include <exception>
#include <functional>
#include <thread>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace psr {

namespace {

struct D2 { double v0, v1; };

// stiffness.py:14-35
PSR_HD D2 hat_integrals(int depth, int index, int trial_depth, int trial_index) {
  const int resolution = 1 << depth;
  const int ratio = 1 << (depth - trial_depth);
  const int first = imax(0, imax(index - 1, (trial_index - 1) * ratio));
  const int end = imin(resolution, imin(index + 1, (trial_index + 1) * ratio));
  D2 r = {0.0, 0.0};
  for (int cell = first; cell < end; ++cell) {
    const double t0 = double(cell - index);
    const double c0 = double(cell) / double(ratio) - double(trial_index);
    const double ts = cell >= index ? -1.0 : 1.0;
    double cs = 1.0 / double(ratio);
    if (cell >= trial_index * ratio) cs = -cs;
    const double tv = 1.0 - fabs(t0);
    const double cv = 1.0 - fabs(c0);
    r.v0 += (tv * cv + (tv * cs + ts * cv) * 0.5 + ts * cs / 3.0) / double(resolution);
    r.v1 += ts * cs * double(resolution);
  }
  return r;
}

// ordered_normal_loads.py:15-84
PSR_HD void quadratic_piece(double t0, double step, int offset, double* b) {
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

PSR_HD D2 normal_integral(int test_depth, int test_index, int normal_depth, int normal_index) {
  const int nt = 1 << test_depth, nn = 1 << normal_depth;
  const int resolution = imax(nt, nn);
  const double test_step = double(nt) / double(resolution);
  const double normal_step = double(nn) / double(resolution);
  const int first = imax(0, imax((test_index - 1) * (resolution / nt), (normal_index - 1) * (resolution / nn)));
  const int end = imin(resolution, imin((test_index + 1) * (resolution / nt), (normal_index + 2) * (resolution / nn)));
  D2 out = {0.0, 0.0};
  for (int cell = first; cell < end; ++cell) {
    const double tx = double(cell) * test_step - double(test_index);
    double sign = -1.0;
    if (tx + 0.5 * test_step < 0.0) sign = 1.0;
    const double h0 = 1.0 + sign * tx;
    const double h1 = sign * test_step;
    const int normal_cell = (int)floor((double(cell) + 0.5) * normal_step);
    const double t0 = double(cell) * normal_step - double(normal_cell);
    double b[3], rr[3];
    quadratic_piece(t0, normal_step, normal_index - normal_cell, b);
    quadratic_piece(t0, normal_step, -1 - normal_index - normal_cell, rr);
    for (int a = 0; a < 3; ++a) b[a] = b[a] - rr[a];
    quadratic_piece(t0, normal_step, 2 * nn - 1 - normal_index - normal_cell, rr);
    for (int a = 0; a < 3; ++a) b[a] = b[a] - rr[a];
    out.v0 += (h0 * b[0] + (h0 * b[1] + h1 * b[0]) * 0.5 + (h0 * b[2] + h1 * b[1]) / 3.0 + h1 * b[2] * 0.25) /
              double(resolution);
    out.v1 += (b[0] + b[1] * 0.5 + b[2] / 3.0) * sign * test_step;
  }
  return out;
}

PSR_HD int normal_center(int index, int test_depth, int normal_depth) {
  if (test_depth > normal_depth) return index >> 1;
  if (test_depth < normal_depth) return index << 1;
  return index;
}

// ---------------------------------------------------------------- tables
__global__ void k_hat_table(int depth, int trial_depth, D2* __restrict__ table) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int res = 1 << depth;
  if (idx >= (res + 1) * 3) return;
  const int i = idx / 3, k = idx % 3;
  table[idx] = hat_integrals(depth, i, trial_depth, (i >> (depth - trial_depth)) + k - 1);
}

__global__ void k_normal_table(int test_depth, int normal_depth, int first, int width, D2* __restrict__ table) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int rows = (1 << test_depth) + 1;
  if (idx >= rows * width) return;
  const int i = idx / width, k = idx % width;
  table[idx] = normal_integral(test_depth, i, normal_depth, normal_center(i, test_depth, normal_depth) + first + k);
}

// ---------------------------------------------------------------- stiffness (stiffness.py:38-109)
// z-major dense index of one level's basis keys: dense[(z*dim + x)*dim + y] =
// key index or -1.  Rows taken in sliced schedule order share a z slice, so
// their neighbour reads fall in a few contiguous (x, y) lines.
__global__ void k_dense_fill(const int64_t* __restrict__ keys, int n, int dim, int* __restrict__ dense) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = keys[i];
  dense[((int64_t)key_z(k) * dim + key_x(k)) * dim + key_y(k)] = i;
}

// One row's 27 stiffness slots: slot k = kx*9 + ky*3 + kz holds the trial
// function (p + (kx-1, ky-1, kz-1)), p the row's coordinates at trial_depth;
// emit(k, index, value) receives the trial function's index (-1 if absent)
// and the entry (0 if absent), slots in order.
template <class Emit>
__device__ __forceinline__ void stiffness_row(int64_t key, const int64_t* __restrict__ cols, int64_t C, ColumnTable ct,
                                              const int* __restrict__ dense, int dense_dim, int depth, int trial_depth,
                                              const D2* __restrict__ table, Emit&& emit) {
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  const int shift = depth - trial_depth;
  const int px = cx >> shift, py = cy >> shift, pz = cz >> shift;
  const int lim = 1 << trial_depth;
  const int zlo = imax(pz - 1, 0), zhi = imin(pz + 1, lim);
#pragma unroll
  for (int kx = 0; kx < 3; ++kx) {
    const int qx = px + kx - 1;
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
      const int qy = py + ky - 1;
      int64_t found[3] = {-1, -1, -1};
      if (qx >= 0 && qx <= lim && qy >= 0 && qy <= lim && zlo <= zhi) {
        if (dense) {
          for (int z = zlo; z <= zhi; ++z) found[z - zlo] = dense[((int64_t)z * dense_dim + qx) * dense_dim + qy];
        } else {
          find_run_table<3>(cols, C, ct, trial_depth, qx, qy, zlo, zhi - zlo + 1, found);
        }
      }
#pragma unroll
      for (int kz = 0; kz < 3; ++kz) {
        const int qz = pz + kz - 1;
        const int k = kx * 9 + ky * 3 + kz;
        const int64_t index = (qz >= zlo && qz <= zhi) ? found[qz - zlo] : -1;
        double v = 0.0;
        if (index >= 0) {
          const D2 x = table[cx * 3 + kx], y = table[cy * 3 + ky], z = table[cz * 3 + kz];
          v = x.v1 * y.v0 * z.v0 + x.v0 * y.v1 * z.v0 + x.v0 * y.v0 * z.v1;
        }
        emit(k, index, v);
      }
    }
  }
}

__global__ void k_stiffness(const int64_t* __restrict__ rows, int R, const int* __restrict__ row_order,
                            const int64_t* __restrict__ cols, int64_t C, ColumnTable ct, const int* __restrict__ dense,
                            int dense_dim, int depth, int trial_depth, const D2* __restrict__ table, int* __restrict__ out_cols,
                            double* __restrict__ out_vals) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // output position
  if (j >= R) return;
  const int i = row_order ? row_order[j] : j;
  stiffness_row(rows[i], cols, C, ct, dense, dense_dim, depth, trial_depth, table, [&](int k, int64_t index, double v) {
    out_cols[(int64_t)k * R + j] = (int)index;
    out_vals[(int64_t)k * R + j] = v;
  });
}

// ---------------------------------------------------------------- constraints (constraints.py:20-84)
template <class T>
__global__ void k_points(const int64_t* __restrict__ akeys, const T* __restrict__ avals, int P, const int64_t* __restrict__ active,
                         int64_t nactive, const T* __restrict__ average, T point_weight, int depth, T* __restrict__ pos,
                         T* __restrict__ w, uint64_t* __restrict__ bins, int* __restrict__ ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const T* v = avals + 4 * (int64_t)i;
  const T p0 = v[0] / v[3], p1 = v[1] / v[3], p2 = v[2] / v[3];
  pos[3 * i] = p0; pos[3 * i + 1] = p1; pos[3 * i + 2] = p2;
  const int64_t key = akeys[i];
  T weight = v[3] * T(1 << depth);
  weight = weight * (point_weight * average[1]);
  if (depth == 0 || find_key(active, nactive, tree_key(key_x(key), key_y(key), key_z(key))) < 0) weight = T(0);
  w[i] = weight;
  // screening_stencils.py:16-26 (bin_points: floor cell, clamped)
  const int n = 1 << depth;
  const int bx = imin(imax((int)floor(p0 * T(n)), 0), n - 1);
  const int by = imin(imax((int)floor(p1 * T(n)), 0), n - 1);
  const int bz = imin(imax((int)floor(p2 * T(n)), 0), n - 1);
  bins[i] = ((uint64_t)bx << (2 * depth)) | ((uint64_t)by << depth) | (uint64_t)bz;
  ids[i] = i;
}

// screening_stencils.py:28-32 (Warp max(a, 0) = a > 0 ? a : 0)
template <class T> PSR_HD T hat3(T p0, T p1, T p2, int cx, int cy, int cz, int depth) {
  const T s = T(1 << depth);
  T t0 = p0 * s - T(cx), t1 = p1 * s - T(cy), t2 = p2 * s - T(cz);
  T h0 = T(1) - fabs(t0), h1 = T(1) - fabs(t1), h2 = T(1) - fabs(t2);
  h0 = h0 > T(0) ? h0 : T(0);
  h1 = h1 > T(0) ? h1 : T(0);
  h2 = h2 > T(0) ? h2 : T(0);
  return h0 * h1 * h2;
}

// Column runs of the sorted point bins (bx << 2d | by << d | bz): points of
// column (bx, by) occupy [start[bx*n + by], end[bx*n + by]).
__global__ void k_bin_columns(const uint64_t* __restrict__ bins, int P, int depth, int* __restrict__ start,
                              int* __restrict__ end) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint64_t c = bins[i] >> depth;
  if (i == 0 || (bins[i - 1] >> depth) != c) start[c] = i;
  if (i == P - 1 || (bins[i + 1] >> depth) != c) end[c] = i + 1;
}

// Dense z-major ranges of the sorted point bins: points of cell (bx, by, bz)
// occupy [start[c], end[c]), c = (bz*n + bx)*n + by (empty cells: 0, 0).
__global__ void k_bin_cells(const uint64_t* __restrict__ bins, int P, int depth, int* __restrict__ start,
                            int* __restrict__ end) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint64_t b = bins[i];
  const int64_t n = (int64_t)1 << depth;
  const int64_t bx = (int64_t)(b >> (2 * depth)), by = (int64_t)((b >> depth) & (uint64_t)(n - 1)),
                bz = (int64_t)(b & (uint64_t)(n - 1));
  const int64_t c = (bz * n + bx) * n + by;
  if (i == 0 || bins[i - 1] != b) start[c] = i;
  if (i == P - 1 || bins[i + 1] != b) end[c] = i + 1;
}

// screening_stencils.py:34-70 (add): per row, points binned in the 8 cells at
// its vertex.  One pass over the points with the 27 slot sums (and the point
// load) held in registers; every sum receives its terms in the reference's
// (cell, point) order.  Present slot k holds the trial function
// (p + (kx-1, ky-1, kz-1)) with p the row's coordinates at trial_depth (the
// stiffness slot layout), so its hat factors are shared per axis.
// Sorted point bins of one level: the points of cell (bx, by, bz) are found
// through the dense z-major cell ranges (depth <= 9), else through the (bx, by)
// column runs and a search of the sorted bin keys.
template <class T> struct PointBins {
  const uint64_t* bin_keys;
  const int* bin_ids;
  int P;
  ColumnTable bin_cols;
  const int* cell_start;
  const int* cell_end;
  const T* pos;
  const T* w;
};

// One row's screening sums (27 slots in the stiffness slot layout) and point
// load, from the points binned in the 8 cells at the row's vertex.  Every
// sum receives its terms in the reference's (cell, point) order.
template <class T>
__device__ __forceinline__ void screening_row(int64_t key, int depth, int trial_depth, const PointBins<T>& pb,
                                              double (&acc)[27], double& rhs) {
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  const int n = 1 << depth;
  // point ranges of the 8 cells at the vertex, in the reference's cell order
  // k = 0..7 -> cell (cx - k/4, cy - k/2%2, cz - k%2); bins are sorted, so the
  // bins (cz-1) and (cz) of one column are adjacent and found with one search
  int rb[8], re[8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      const int k0 = a * 4 + bb * 2;
      rb[k0] = re[k0] = rb[k0 + 1] = re[k0 + 1] = 0;
      const int bx = cx - a, by = cy - bb;
      if (bx < 0 || by < 0 || bx >= n || by >= n) continue;
      if (pb.cell_start) {  // dense z-major cell ranges
        if (cz >= 1) {
          const int64_t c = ((int64_t)(cz - 1) * n + bx) * n + by;
          rb[k0 + 1] = pb.cell_start[c];
          re[k0 + 1] = pb.cell_end[c];
        }
        if (cz < n) {
          const int64_t c = ((int64_t)cz * n + bx) * n + by;
          rb[k0] = pb.cell_start[c];
          re[k0] = pb.cell_end[c];
        }
        continue;
      }
      const uint64_t col = ((uint64_t)bx << (2 * depth)) | ((uint64_t)by << depth);
      const int zfirst = cz >= 1 ? cz - 1 : cz;
      int lo = 0, hi = pb.P;
      if (pb.bin_cols.start) {
        lo = pb.bin_cols.start[bx * n + by];
        hi = pb.bin_cols.end[bx * n + by];
      }
      const int stop = hi;
      const uint64_t first_key = col | (uint64_t)zfirst;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pb.bin_keys[mid] < first_key) lo = mid + 1; else hi = mid;
      }
      if (cz >= 1) {
        rb[k0 + 1] = lo;
        while (lo < stop && pb.bin_keys[lo] == (col | (uint64_t)(cz - 1))) ++lo;
        re[k0 + 1] = lo;
      }
      if (cz < n) {
        rb[k0] = lo;
        while (lo < stop && pb.bin_keys[lo] == (col | (uint64_t)cz)) ++lo;
        re[k0] = lo;
      }
    }
  const int shift = depth - trial_depth;
  const int px = (cx >> shift) - 1, py = (cy >> shift) - 1, pz = (cz >> shift) - 1;
  const T st = T(1 << trial_depth);
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.0;
  rhs = 0.0;
  for (int kk = 0; kk < 8; ++kk)
    for (int jj = rb[kk]; jj < re[kk]; ++jj) {
      const int index = pb.bin_ids[jj];
      const T weight = pb.w[index];
      if (weight != T(0)) {
        const T p0 = pb.pos[3 * index], p1 = pb.pos[3 * index + 1], p2 = pb.pos[3 * index + 2];
        const T phi = hat3(p0, p1, p2, cx, cy, cz, depth);
        rhs += double(phi * weight);
        T hx[3], hy[3], hz[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const T h0 = T(1) - fabs(p0 * st - T(px + a));
          const T h1 = T(1) - fabs(p1 * st - T(py + a));
          const T h2 = T(1) - fabs(p2 * st - T(pz + a));
          hx[a] = h0 > T(0) ? h0 : T(0);
          hy[a] = h1 > T(0) ? h1 : T(0);
          hz[a] = h2 > T(0) ? h2 : T(0);
        }
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          const T trial = hx[k / 9] * hy[(k / 3) % 3] * hz[k % 3];
          acc[k] += double(phi * trial * weight);
        }
      }
    }
}

// screening_stencils.py:34-70 (add): per row, the screening sums added to
// the present slots of an assembled stencil, and the point load.
template <class T>
__global__ void __launch_bounds__(128) k_screening(const int64_t* __restrict__ rows, int R, const int* __restrict__ row_order,
                                                   int depth, int trial_depth, PointBins<T> pb,
                                                   const int* __restrict__ scols, double* __restrict__ svals,
                                                   double* __restrict__ load) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // output position
  if (j >= R) return;
  const int i = row_order ? row_order[j] : j;
  double acc[27], rhs;
  screening_row(rows[i], depth, trial_depth, pb, acc, rhs);
  if (load) load[j] = rhs;
#pragma unroll
  for (int k = 0; k < 27; ++k)
    if (scols[(int64_t)k * R + j] >= 0) svals[(int64_t)k * R + j] += acc[k];
}

// One row of a finer level's operator for the tiled relaxation: its 27 slot
// values, diagonal flag and 1/A_ii, at schedule position j, so a row's
// operator is a few consecutive 16-byte words.
template <class T, class V> struct alignas(16) RowRecord {
  V v[27];
  int has;
  T inv;
};

typedef int record_word __attribute__((ext_vector_type(4)));

template <class T, class V> union RowRecordWords {
  record_word q[sizeof(RowRecord<T, V>) / sizeof(record_word)];
  RowRecord<T, V> r;
};

// stiffness.py:38-109 with screening_stencils.py:34-70 for the sliced finer
// levels, in one pass per schedule position j: the 27 stiffness entries,
// plus (SCREEN) the row's screening sums, rounded once to the stored type V
// -- the same value the separate stiffness, screening and cast passes give.
// For the level's own operator (inv_diag != nullptr) the diagonal factors
// 1/A_ii and the count of present slots are taken from the same registers;
// with sched_cols (the colour-per-launch relaxation's input) the columns are
// also written as schedule positions (inv: the schedule's inverse), and with
// records (the tiled relaxation's input) each row's record.
template <class T, class V, bool SCREEN>
__global__ void __launch_bounds__(256) k_level_operator(const int64_t* __restrict__ rows, int R, const int* __restrict__ row_order,
                                                        const int64_t* __restrict__ cols, int64_t C, ColumnTable ct,
                                                        const int* __restrict__ dense, int dense_dim, int depth, int trial_depth,
                                                        const D2* __restrict__ table, PointBins<T> pb, double* __restrict__ load,
                                                        int* __restrict__ out_cols, V* __restrict__ out_vals,
                                                        const int* __restrict__ inv, int* __restrict__ sched_cols,
                                                        T* __restrict__ inv_diag, unsigned char* __restrict__ has_diag,
                                                        RowRecord<T, V>* __restrict__ records,
                                                        unsigned long long* __restrict__ present) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // output position
  int count = 0;
  if (j < R) {
    const int i = row_order ? row_order[j] : j;
    const int64_t key = rows[i];
    double acc[27], rhs = 0.0;
    if (SCREEN) {
      screening_row(key, depth, trial_depth, pb, acc, rhs);
      if (load) load[j] = rhs;
    }
    T diagonal = T(0);
    RowRecordWords<T, V> rec;
    stiffness_row(key, cols, C, ct, dense, dense_dim, depth, trial_depth, table, [&](int k, int64_t index, double v) {
      const bool found = index >= 0;
      const V stored = V(found && SCREEN ? v + acc[k] : v);
      out_cols[(int64_t)k * R + j] = (int)index;
      out_vals[(int64_t)k * R + j] = stored;
      rec.r.v[k] = stored;
      if (inv_diag) {
        if (sched_cols) sched_cols[(int64_t)k * R + j] = found ? inv[index] : -1;
        if (index == i) diagonal = T(stored);
        count += found ? 1 : 0;
      }
    });
    if (inv_diag) {
      has_diag[j] = diagonal != T(0);
      inv_diag[j] = diagonal != T(0) ? T(1) / diagonal : T(0);
      if (records) {
        rec.r.has = diagonal != T(0);
        rec.r.inv = diagonal != T(0) ? T(1) / diagonal : T(0);
        record_word* dst = (record_word*)(records + j);
#pragma unroll
        for (int q = 0; q < (int)(sizeof(rec.q) / sizeof(record_word)); ++q) dst[q] = rec.q[q];
      }
    }
  }
  if (present) {
    __shared__ unsigned long long block_count;
    if (threadIdx.x == 0) block_count = 0;
    __syncthreads();
    if (count) atomicAdd(&block_count, (unsigned long long)count);
    __syncthreads();
    if (threadIdx.x == 0 && block_count) atomicAdd(present, block_count);
  }
}

template <class T>
__global__ void k_cast(const double* __restrict__ in, T* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = T(in[i]);
}

// solver.py:119-126 (_rhs)
template <class T>
__global__ void k_rhs(const double* __restrict__ load, const double* __restrict__ point_load, int R, T* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < R) b[i] = T(load[i]) + T(0.5) * T(point_load[i]);
}

// ---------------------------------------------------------------- prolongation (transfer_stencils.py:15-66)
template <class T>
__global__ void k_prolongation(const int64_t* __restrict__ fine, int R, const int64_t* __restrict__ coarse, int64_t C, ColumnTable ct, int depth,
                               int* __restrict__ cols, T* __restrict__ vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = fine[i];
  const int f[3] = {key_x(key), key_y(key), key_z(key)};
  T h[3][2];
  for (int a = 0; a < 3; ++a)
    for (int o = 0; o < 2; ++o) h[a][o] = T(1) - fabs(T(f[a]) * T(0.5) - T(f[a] / 2 + o));
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const T wz0 = h[0][a] * h[1][b] * h[2][0], wz1 = h[0][a] * h[1][b] * h[2][1];
      int64_t found[2] = {-1, -1};
      if (wz0 > T(0) || wz1 > T(0)) find_run_table<2>(coarse, C, ct, depth - 1, f[0] / 2 + a, f[1] / 2 + b, f[2] / 2, 2, found);
      for (int c = 0; c < 2; ++c) {
        const int k = a * 4 + b * 2 + c;
        const T weight = c ? wz1 : wz0;
        const int64_t column = weight > T(0) ? found[c] : -1;
        cols[(int64_t)k * R + i] = (int)column;
        vals[(int64_t)k * R + i] = column >= 0 ? weight : T(0);
      }
    }
}

// Restriction P^T as a gather: coarse row r takes the fine nodes 2r+o,
// o in {-1,0,1}^3 (ascending fine index = lexicographic o), weight 1/2 per
// nonzero axis offset; exactly the rows of the transposed prolongation.
__global__ void k_restriction(const int64_t* __restrict__ coarse, int C, const int64_t* __restrict__ fine, int64_t F,
                              ColumnTable ct, int fine_depth, int* __restrict__ cols) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  const int64_t key = coarse[r];
  const int x = 2 * key_x(key), y = 2 * key_y(key), z = 2 * key_z(key);
  const int lim = 1 << fine_depth;
  const int zlo = imax(z - 1, 0), zhi = imin(z + 1, lim);
  for (int kx = 0; kx < 3; ++kx)
    for (int ky = 0; ky < 3; ++ky) {
      const int fx = x + kx - 1, fy = y + ky - 1;
      int64_t found[3] = {-1, -1, -1};
      if (fx >= 0 && fx <= lim && fy >= 0 && fy <= lim && zlo <= zhi)
        find_run_table<3>(fine, F, ct, fine_depth, fx, fy, zlo, zhi - zlo + 1, found);
      for (int kz = 0; kz < 3; ++kz) {
        const int fz = z + kz - 1;
        cols[(int64_t)(kx * 9 + ky * 3 + kz) * C + r] = (fz >= zlo && fz <= zhi) ? (int)found[fz - zlo] : -1;
      }
    }
}

PSR_HD double restriction_weight(int k) {
  double w = 1.0;
  if (k / 9 != 1) w *= 0.5;
  if ((k / 3) % 3 != 1) w *= 0.5;
  if (k % 3 != 1) w *= 0.5;
  return w;
}

// y[r] = (sum_k R_rk x[k]) (*alpha) (+ beta*y[r]) in warp.sparse.bsr_mv order
template <class V>
__global__ void k_restrict_apply(const int* __restrict__ cols, int C, const V* __restrict__ x, V* __restrict__ y, int accumulate) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  int c[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) c[k] = cols[(int64_t)k * C + r];
  V xv[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    // x may be empty (a finer level without nodes): absent slots read nothing
    xv[k] = V(0);
    if (c[k] >= 0) xv[k] = x[c[k]];
  }
  V v = V(0);
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const V t = v + V(restriction_weight(k)) * xv[k];
    v = c[k] >= 0 ? t : v;
  }
  if (accumulate) v = v + y[r];
  y[r] = v;
}

// y = alpha*(A x) + beta*y with A a stencil (slots in column order)
// (values of type V, widened exactly to T)
// The slot sum with every load issued before the additions (S known at
// compile time); absent slots are skipped by a select, so the sum is the
// sequential one.
template <class T, class V, int S>
__device__ __forceinline__ T stencil_row_sum(const int* __restrict__ cols, const V* __restrict__ vals, int R, int i,
                                             const T* __restrict__ x) {
  int c[S];
#pragma unroll
  for (int k = 0; k < S; ++k) c[k] = cols[(int64_t)k * R + i];
  V w[S];
  T xv[S];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    w[k] = vals[(int64_t)k * R + i];
    // x may be empty (a level without nodes): absent slots read nothing
    xv[k] = T(0);
    if (c[k] >= 0) xv[k] = x[c[k]];
  }
  T v = T(0);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const T t = v + T(w[k]) * xv[k];
    v = c[k] >= 0 ? t : v;
  }
  return v;
}

template <class T, class V = T>
__global__ void k_spmv(const int* __restrict__ cols, const V* __restrict__ vals, int R, int slots, const T* __restrict__ x,
                       T* __restrict__ y, T alpha, T beta, const T* __restrict__ y_in) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  T v = T(0);
  if (slots == 27) {
    v = stencil_row_sum<T, V, 27>(cols, vals, R, i, x);
  } else if (slots == 8) {
    v = stencil_row_sum<T, V, 8>(cols, vals, R, i, x);
  } else {
    for (int k = 0; k < slots; ++k) {
      const int c = cols[(int64_t)k * R + i];
      if (c >= 0) v = v + T(vals[(int64_t)k * R + i]) * x[c];
    }
  }
  v = v * alpha;
  if (beta != T(0)) v = v + beta * (y_in ? y_in[i] : y[i]);  // y_in: the beta term's input when not y
  y[i] = v;
}

// ---------------------------------------------------------------- normal loads (normal_stencils.py, normal_loads.py)
// Dense index of a key set: zmajor ? (z*dim + x)*dim + y : (x*dim + y)*dim + z
struct DenseIndex {
  const int* idx = nullptr;
  int dim = 0;
  int zmajor = 0;
};

__global__ void k_dense_fill_xyz(const int64_t* __restrict__ keys, int64_t n, int dim, int* __restrict__ dense) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = keys[i];
  dense[((int64_t)key_x(k) * dim + key_y(k)) * dim + key_z(k)] = (int)i;
}

template <class V>
__global__ void k_normal_load(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ nkeys, int64_t ncount, ColumnTable ct,
                              DenseIndex dn, const V* __restrict__ nvals, int nstride, int test_depth, int normal_depth, int first,
                              int width, const D2* __restrict__ table, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int c0 = key_x(key), c1 = key_y(key), c2 = key_z(key);
  const int e0 = normal_center(c0, test_depth, normal_depth), e1 = normal_center(c1, test_depth, normal_depth),
            e2 = normal_center(c2, test_depth, normal_depth);
  const int lim = 1 << normal_depth;
  const int zlo = imax(e2 + first, 0), zhi = imin(e2 + first + width - 1, lim - 1);
  double value = 0.0;
  for (int x = 0; x < width; ++x) {
    const int nx = e0 + x + first;
    if (nx < 0 || nx >= lim) continue;
    for (int y = 0; y < width; ++y) {
      const int ny = e1 + y + first;
      if (ny < 0 || ny >= lim || zlo > zhi) continue;
      int64_t found[6];
      if (dn.idx) {
        for (int z = zlo; z <= zhi; ++z)
          found[z - zlo] = dn.zmajor ? dn.idx[((int64_t)z * dn.dim + nx) * dn.dim + ny] : dn.idx[((int64_t)nx * dn.dim + ny) * dn.dim + z];
      } else {
        find_run_table<6>(nkeys, ncount, ct, normal_depth, nx, ny, zlo, zhi - zlo + 1, found);
      }
      for (int z = 0; z < width; ++z) {
        const int nz = e2 + z + first;
        if (nz < zlo || nz > zhi) continue;
        const int64_t jn = found[nz - zlo];
        if (jn < 0) continue;
        const D2 ix = table[c0 * width + x], iy = table[c1 * width + y], iz = table[c2 * width + z];
        const V* nv = nvals + (int64_t)nstride * jn;
        value += double(nv[0]) * ix.v1 * iy.v0 * iz.v0;
        value += double(nv[1]) * ix.v0 * iy.v1 * iz.v0;
        value += double(nv[2]) * ix.v0 * iy.v0 * iz.v1;
      }
    }
  }
  out[i] = value;
}

// k_normal_load over a dense record index with each (x, y) column's loads
// batched: the column's W dense entries, then all W records' values, then the
// terms in the same (x, y, z, component) order as k_normal_load, so the sum is
// identical.  Unconditional, clamped addresses keep the loads independent of
// the presence tests (k_normal_load issues them one record at a time).
template <class V, int W>
__global__ void __launch_bounds__(128) k_normal_load_dense(const int64_t* __restrict__ keys, int R, DenseIndex dn,
                                                           const V* __restrict__ nvals, int nstride, int test_depth,
                                                           int normal_depth, int first, const D2* __restrict__ table,
                                                           double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int c0 = key_x(key), c1 = key_y(key), c2 = key_z(key);
  const int e0 = normal_center(c0, test_depth, normal_depth), e1 = normal_center(c1, test_depth, normal_depth),
            e2 = normal_center(c2, test_depth, normal_depth);
  const int lim = 1 << normal_depth;
  const int zlo = imax(e2 + first, 0), zhi = imin(e2 + first + W - 1, lim - 1);
  D2 iz[W];
#pragma unroll
  for (int k = 0; k < W; ++k) iz[k] = table[c2 * W + k];
  const int64_t dim = dn.dim;
  double value = 0.0;
  for (int x = 0; x < W; ++x) {
    const int nx = e0 + x + first;
    if (nx < 0 || nx >= lim) continue;
    const D2 ix = table[c0 * W + x];
    for (int y = 0; y < W; ++y) {
      const int ny = e1 + y + first;
      if (ny < 0 || ny >= lim) continue;
      const D2 iy = table[c1 * W + y];
      int jn[W];
#pragma unroll
      for (int z = 0; z < W; ++z) {
        const int nz = e2 + z + first;
        const bool in = nz >= zlo && nz <= zhi;
        int64_t cell = dn.zmajor ? ((int64_t)nz * dim + nx) * dim + ny : ((int64_t)nx * dim + ny) * dim + nz;
        cell = in ? cell : 0;
        jn[z] = in ? dn.idx[cell] : -1;
      }
      V v[W][3];
#pragma unroll
      for (int z = 0; z < W; ++z) {
        const int64_t j = jn[z] >= 0 ? jn[z] : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) v[z][c] = jn[z] >= 0 ? nvals[(int64_t)nstride * j + c] : V(0);
      }
#pragma unroll
      for (int z = 0; z < W; ++z) {
        if (jn[z] < 0) continue;
        value += double(v[z][0]) * ix.v1 * iy.v0 * iz[z].v0;
        value += double(v[z][1]) * ix.v0 * iy.v1 * iz[z].v0;
        value += double(v[z][2]) * ix.v0 * iy.v0 * iz[z].v1;
      }
    }
  }
  out[i] = value;
}

// normal_loads.py:27-63 (prolong degree-2 coefficients with degree-1 weights, issue #35)
template <class T>
__global__ void k_normal_prolong(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ own_keys, int64_t own_count,
                                 ColumnTable own_ct, const T* __restrict__ own_vals, const int64_t* __restrict__ prev_keys, int64_t prev_count, ColumnTable ct,
                                 const T* __restrict__ prev_vals, int depth, T* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int c[3] = {key_x(key), key_y(key), key_z(key)};
  T v0 = T(0), v1 = T(0), v2 = T(0);
  int64_t own;
  find_run_table<1>(own_keys, own_count, own_ct, depth, c[0], c[1], c[2], 1, &own);
  if (own >= 0) { v0 = own_vals[3 * own]; v1 = own_vals[3 * own + 1]; v2 = own_vals[3 * own + 2]; }
  if (depth > 0) {
    int64_t idx[2][2][2];  // [x bit][y bit][z bit]
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) find_run_table<2>(prev_keys, prev_count, ct, depth - 1, c[0] / 2 + a, c[1] / 2 + b, c[2] / 2, 2, idx[a][b]);
    for (int corner = 0; corner < 8; ++corner) {  // x is bit 0 here (normal_loads.py:285-297)
      int p[3] = {c[0] / 2, c[1] / 2, c[2] / 2};
      T weight = T(1);
      for (int axis = 0; axis < 3; ++axis) {
        if ((corner >> axis) & 1) p[axis] += 1;
        weight = weight * (T(1) - fabs(T(c[axis]) * T(0.5) - T(p[axis])));
      }
      if (weight > T(0)) {
        const int64_t j = idx[corner & 1][(corner >> 1) & 1][(corner >> 2) & 1];
        if (j >= 0) {
          v0 = v0 + prev_vals[3 * j] * weight;
          v1 = v1 + prev_vals[3 * j + 1] * weight;
          v2 = v2 + prev_vals[3 * j + 2] * weight;
        }
      }
    }
  }
  out[3 * i] = v0; out[3 * i + 1] = v1; out[3 * i + 2] = v2;
}

__global__ void k_add_f64(double* __restrict__ y, const double* __restrict__ x, int n) {  // y = x + y
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + y[i];
}

// ---------------------------------------------------------------- Galerkin R*A*P (base levels)
// Row r of R*A*P with the reference spgemm order: contributions to an entry
// arrive in ascending middle index (R row order, then (RA) row order).
template <class T>
__global__ void __launch_bounds__(128) k_galerkin(const int* __restrict__ rcols, int C, const int* __restrict__ acols,
                                                  const T* __restrict__ avals, int F, const int* __restrict__ pcols,
                                                  const T* __restrict__ pvals, int* __restrict__ out_cols, T* __restrict__ out_vals,
                                                  const int64_t* __restrict__ ckeys) {
  // one workgroup per coarse row r.  (R A)_rt for the 125 fine offsets t:
  // thread t sums its contributions in ascending R-row order (one A slot per
  // R entry reaches t), then thread s sums row s of (R A) P over t ascending
  // and P slot ascending -- the sequential order of both products.
  __shared__ T ra[125];
  __shared__ int rak[125];
  __shared__ int cs[125][8];
  __shared__ int cn[125][8];  // the coarse column a contribution lands in
  __shared__ T cv[125][8];
  __shared__ int landed[128];  // contributions per fine offset
  const int r = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t rkey = ckeys[r];
  const int rx = key_x(rkey), ry = key_y(rkey), rz = key_z(rkey);
  int n = 0;
  if (t < 125) {
    // Each dependent step (R columns, A columns, A values; P columns, coarse
    // keys, P values) is issued for every slot before any is used, then the
    // sums run in the sequential order.
    const int tx = t / 25, ty = (t / 5) % 5, tz = t % 5;
    int ac[27];
    int64_t at[27];
#pragma unroll
    for (int kr = 0; kr < 27; ++kr) {  // fine k = 2r + o_kr, ascending
      const int k = rcols[(int64_t)kr * C + r];
      const int jx = tx - kr / 9, jy = ty - (kr / 3) % 3, jz = tz - kr % 3;
      const bool in = k >= 0 && (unsigned)jx <= 2u && (unsigned)jy <= 2u && (unsigned)jz <= 2u;
      at[kr] = in ? (int64_t)(jx * 9 + jy * 3 + jz) * F + k : -1;
    }
#pragma unroll
    for (int kr = 0; kr < 27; ++kr) ac[kr] = at[kr] >= 0 ? acols[at[kr]] : -1;
    T av[27];
#pragma unroll
    for (int kr = 0; kr < 27; ++kr) av[kr] = ac[kr] >= 0 ? avals[at[kr]] : T(0);
    T acc = T(0);
    int col = -1;
#pragma unroll
    for (int kr = 0; kr < 27; ++kr) {
      if (ac[kr] < 0) continue;
      acc = acc + T(restriction_weight(kr)) * av[kr];
      col = ac[kr];
    }
    ra[t] = acc;
    rak[t] = col;
    int pc[8];
#pragma unroll
    for (int kp = 0; kp < 8; ++kp) pc[kp] = col >= 0 ? pcols[(int64_t)kp * F + col] : -1;
    int64_t pkey[8];
#pragma unroll
    for (int kp = 0; kp < 8; ++kp) pkey[kp] = pc[kp] >= 0 ? ckeys[pc[kp]] : int64_t(0);
#pragma unroll
    for (int kp = 0; kp < 8; ++kp) {
      int sidx = -1;
      T pv = T(0);
      if (pc[kp] >= 0) {
        const int ox = key_x(pkey[kp]) - rx + 1, oy = key_y(pkey[kp]) - ry + 1, oz = key_z(pkey[kp]) - rz + 1;
        if ((unsigned)ox <= 2u && (unsigned)oy <= 2u && (unsigned)oz <= 2u) {  // always for trilinear P
          sidx = ox * 9 + oy * 3 + oz;
          pv = pvals[(int64_t)kp * F + col];
        }
      }
      cs[t][kp] = sidx;
      cn[t][kp] = pc[kp];
      cv[t][kp] = pv;
      n += sidx >= 0 ? 1 : 0;
    }
  }
  if (t < 128) landed[t] = n;
  __syncthreads();
  if (t < 64) {
    // Slot t (coarse offset o per axis) is reached only from fine offsets
    // 2o-1..2o+1 per axis (trilinear P: fine f feeds coarse f/2 when f is
    // even, (f-1)/2 and (f+1)/2 when odd), so each slot scans those 27 in
    // ascending order -- the order of the full scan.  Every contribution is
    // counted once: if the candidates miss any, the full scan runs instead.
    // The column of slot t is the coarse index of any contribution landing
    // there (keys are unique).
    int total = landed[t] + landed[t + 64];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) total += __shfl_xor(total, m);
    T out = T(0);
    int col = -1, found = 0;
    if (t < 27) {
      const int ox = t / 9, oy = (t / 3) % 3, oz = t % 3;
      for (int a = 0; a < 3; ++a) {
        const int ux = 2 * ox - 1 + a;
        if (ux < 0 || ux > 4) continue;
        for (int b = 0; b < 3; ++b) {
          const int uy = 2 * oy - 1 + b;
          if (uy < 0 || uy > 4) continue;
          for (int c = 0; c < 3; ++c) {
            const int uz = 2 * oz - 1 + c;
            if (uz < 0 || uz > 4) continue;
            const int u = ux * 25 + uy * 5 + uz;
            if (rak[u] < 0) continue;
#pragma unroll
            for (int kp = 0; kp < 8; ++kp)
              if (cs[u][kp] == t) {
                out = out + ra[u] * cv[u][kp];
                col = cn[u][kp];
                ++found;
              }
          }
        }
      }
    }
    int all = found;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) all += __shfl_xor(all, m);
    if (all != total && t < 27) {
      out = T(0);
      col = -1;
      for (int u = 0; u < 125; ++u) {
        if (rak[u] < 0) continue;
        for (int kp = 0; kp < 8; ++kp)
          if (cs[u][kp] == t) {
            out = out + ra[u] * cv[u][kp];
            col = cn[u][kp];
          }
      }
    }
    if (t < 27) {
      out_cols[(int64_t)t * C + r] = col;
      out_vals[(int64_t)t * C + r] = out;
    }
  }
}

// ---------------------------------------------------------------- relaxation
// relaxation.py:38-64 (_relax) on the schedule-ordered operator.  Operator
// values have type V (T, or float operators under T = double), widened
// exactly to T before use.
template <class T, class V>
__device__ __forceinline__ void gs_update(int j, int R, const int* __restrict__ order, const int* __restrict__ cols,
                                          const V* __restrict__ vals, const T* __restrict__ b, const T* __restrict__ inv_diag,
                                          const unsigned char* __restrict__ has_diag, T* __restrict__ x) {
  // All column loads, then all value/x loads in flight, then the subtraction
  // chain in slot order; absent slots are skipped by a select (no rounding).
  // An absent slot loads x[0], which exists: row j belongs to this level, so
  // the level (and x) holds at least one row.
  int c[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) c[k] = cols[(int64_t)k * R + j];
  V v[27];
  T xv[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    v[k] = vals[(int64_t)k * R + j];
    xv[k] = x[c[k] >= 0 ? c[k] : 0];
  }
  T r = b[j];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const T t = r - xv[k] * T(v[k]);
    r = c[k] >= 0 ? t : r;
  }
  if (has_diag[j]) {
    const int i = order ? order[j] : j;  // no order: x is kept in schedule order
    x[i] = x[i] + r * inv_diag[j];
  }
}

// One launch of relaxation.py:100-116 (wavefront_update): `slices` slices two
// apart, same colour.
// waves_per_eu(1,4): the row update keeps all 81 loads of a row in flight
// (needs ~100 VGPRs); occupancy is not what bounds these small launches.
template <class T, class V>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_relax_wave(const int* __restrict__ order, const int* __restrict__ goff, const int* __restrict__ cols,
             const V* __restrict__ vals, const T* __restrict__ b, const T* __restrict__ inv_diag,
             const unsigned char* __restrict__ has_diag, T* __restrict__ x, int R, int first_z, int color) {
  const int group = 4 * (first_z + 2 * (int)blockIdx.y) + color;
  const int j = goff[group] + blockIdx.x * blockDim.x + threadIdx.x;
  if (j < goff[group + 1]) gs_update(j, R, order, cols, vals, b, inv_diag, has_diag, x);
}

// Buffers of one sliced relaxation, passed by value in every launch.
template <class T, class V> struct RelaxArgs {
  const int* order;
  const int* cols;
  const V* vals;
  const T* b;
  const T* inv_diag;
  const unsigned char* has_diag;
  T* x;
  int R;
};

// Row ranges of one launch's slices (a wavefront holds at most `iters`
// slices, two apart), in schedule positions: slice t is [start[t], end[t]).
constexpr int RELAX_MAX_SLICES = 16;
struct SliceRanges {
  int start[RELAX_MAX_SLICES], end[RELAX_MAX_SLICES];
};

// The same launch with everything it needs in its arguments: the first loads
// of a row depend on nothing but the kernel argument segment.
template <class T, class V, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_relax_wave_args(RelaxArgs<T, V> d, SliceRanges rg) {
  const int t = blockIdx.y;
  const int j = rg.start[t] + (int)blockIdx.x * BS + (int)threadIdx.x;
  if (j < rg.end[t]) gs_update(j, d.R, d.order, d.cols, d.vals, d.b, d.inv_diag, d.has_diag, d.x);
}

// relaxation.py:261-271 (8-colour, all sweeps) inside one workgroup: the
// barrier between colour groups replaces the kernel boundary.
template <class T>
__global__ void __launch_bounds__(1024) k_relax_block(const int* __restrict__ order, const int* __restrict__ goff,
                                                      const int* __restrict__ cols, const T* __restrict__ vals,
                                                      const T* __restrict__ b, const T* __restrict__ inv_diag,
                                                      const unsigned char* __restrict__ has_diag, T* __restrict__ x, int R,
                                                      int iters, int forward) {
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < 8; ++c) {
      const int g = forward ? c : 7 - c;
      for (int j = goff[g] + threadIdx.x; j < goff[g + 1]; j += blockDim.x)
        gs_update(j, R, order, cols, vals, b, inv_diag, has_diag, x);
      __syncthreads();
    }
  }
}

// The same with x resident in LDS for the whole call (levels whose x fits):
// the matrix rows are streamed from L2 while every x gather and update stays
// on-chip; the updates and their order are unchanged.
template <class T>
__global__ void __launch_bounds__(1024) k_relax_block_lds(const int* __restrict__ order, const int* __restrict__ goff,
                                                          const int* __restrict__ cols, const T* __restrict__ vals,
                                                          const T* __restrict__ b, const T* __restrict__ inv_diag,
                                                          const unsigned char* __restrict__ has_diag, T* __restrict__ x,
                                                          int R, int iters, int forward) {
  extern __shared__ unsigned char lds_raw[];
  T* xs = (T*)lds_raw;
  for (int i = threadIdx.x; i < R; i += blockDim.x) xs[i] = x[i];
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < 8; ++c) {
      const int g = forward ? c : 7 - c;
      for (int j = goff[g] + threadIdx.x; j < goff[g + 1]; j += blockDim.x) {
        int cc[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) cc[k] = cols[(int64_t)k * R + j];
        T v[27], xv[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          v[k] = vals[(int64_t)k * R + j];
          xv[k] = xs[cc[k] >= 0 ? cc[k] : 0];
        }
        T r = b[j];
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          const T t = r - xv[k] * v[k];
          r = cc[k] >= 0 ? t : r;
        }
        if (has_diag[j]) {
          const int i = order[j];
          xs[i] = xs[i] + r * inv_diag[j];
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < R; i += blockDim.x) x[i] = xs[i];
}

// schedule: group per row
__global__ void k_groups_of(const int64_t* __restrict__ keys, int R, int sliced, uint64_t* __restrict__ g, int* __restrict__ ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t k = keys[i];
  const int x = key_x(k), y = key_y(k), z = key_z(k);
  g[i] = sliced ? (uint64_t)((x & 1) + 2 * (y & 1) + 4 * z) : (uint64_t)((x & 1) + 2 * (y & 1) + 4 * (z & 1));
  ids[i] = i;
}

__global__ void k_group_offsets(const uint64_t* __restrict__ g, int R, int groups, int* __restrict__ off) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > groups) return;
  int lo = 0, hi = R;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (g[mid] < (uint64_t)t) lo = mid + 1; else hi = mid;
  }
  off[t] = lo;
}

template <class T>
__global__ void k_gather_operator(const int* __restrict__ order, int R, const int* __restrict__ cols, const T* __restrict__ vals,
                                  const T* __restrict__ b, int* __restrict__ scols, T* __restrict__ svals, T* __restrict__ sb,
                                  T* __restrict__ inv_diag, unsigned char* __restrict__ has_diag) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const int i = order[j];
  T diagonal = T(0);
  for (int k = 0; k < 27; ++k) {
    const int c = cols[(int64_t)k * R + i];
    const T v = vals[(int64_t)k * R + i];
    scols[(int64_t)k * R + j] = c;
    svals[(int64_t)k * R + j] = v;
    if (c == i) diagonal = v;
  }
  sb[j] = b[i];
  has_diag[j] = diagonal != T(0);
  inv_diag[j] = diagonal != T(0) ? T(1) / diagonal : T(0);
}

// Root level (<= 64 rows): padded row of node i when the mean term is on
// (solver.py:30-33), and the operator applied by the root CG.
__device__ __forceinline__ int root_row(const int64_t* __restrict__ keys, int i) {
  const int64_t key = keys[i];
  const int cx = key_x(key) + 2, cy = key_y(key) + 2, cz = key_z(key) + 2;
  return cz * 16 + (cx & 1) + 2 * (cy & 1) + 2 * (cx & 2) + 4 * (cy & 2);
}

template <class T>
__device__ void root_apply(const int* __restrict__ cols, const T* __restrict__ vals, int R, const int64_t* __restrict__ keys, int dc,
                           const T* v, T* out) {
  if (!dc) {
    for (int i = 0; i < R; ++i) {
      T a = T(0);
      for (int k = 0; k < 27; ++k) {
        const int c = cols[k * R + i];
        if (c >= 0) a = a + vals[k * R + i] * v[c];
      }
      out[i] = a;
    }
    return;
  }
  // padded rows: entries sorted by padded column
  for (int i = 0; i < 64; ++i) out[i] = T(0);
  for (int i = 0; i < R; ++i) {
    int pc[27];
    T pv[27];
    int cnt = 0;
    for (int k = 0; k < 27; ++k) {
      const int c = cols[k * R + i];
      if (c < 0) continue;
      int q = root_row(keys, c);
      int pos = cnt++;
      while (pos > 0 && pc[pos - 1] > q) { pc[pos] = pc[pos - 1]; pv[pos] = pv[pos - 1]; --pos; }
      pc[pos] = q;
      pv[pos] = vals[k * R + i];
    }
    T a = T(0);
    for (int e = 0; e < cnt; ++e) a = a + pv[e] * v[pc[e]];
    out[root_row(keys, i)] = a;
  }
  T mean = T(0);
  for (int i = 0; i < 64; ++i) mean = mean + v[i];
  mean = mean / T(64);
  for (int i = 0; i < 64; ++i) out[i] = out[i] + T(1) * mean;
}

// Root solve on one lane: skip when |r|^2 <= eps^2, else unpreconditioned CG
// (solver.py:276-297; Warp's cg restated, bitwise parity unpinned).
template <class T>
__global__ void k_root_solve(const int* __restrict__ cols, const T* __restrict__ vals, int R, const int64_t* __restrict__ keys,
                             const T* __restrict__ rhs, T* __restrict__ x, T* __restrict__ pad_x, T* __restrict__ work,
                             int dc, int base, int cycle, double acc, int maxiter, int* __restrict__ ran) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int m = dc ? 64 : R;
  T* b = work;
  T* r = work + 64;
  T* p = work + 128;
  T* ap = work + 192;
  T* xx = dc ? pad_x : x;
  for (int i = 0; i < m; ++i) b[i] = T(0);
  for (int i = 0; i < R; ++i) b[dc ? root_row(keys, i) : i] = rhs[i];
  auto apply = [&](const T* v, T* out) { root_apply<T>(cols, vals, R, keys, dc, v, out); };
  for (int i = 0; i < m; ++i) r[i] = b[i];
  if (base == 0 && cycle > 0) {
    apply(xx, ap);
    for (int i = 0; i < m; ++i) r[i] = b[i] - ap[i];
  }
  double norm2 = 0.0;
  for (int i = 0; i < m; ++i) norm2 += double(r[i]) * double(r[i]);
  *ran = 0;
  if (!(norm2 > acc * acc)) return;
  *ran = 1;
  const double tol = base == 0 ? 0.0 : acc;
  const double atol = base == 0 ? acc * sqrt(norm2) : 0.0;
  apply(xx, ap);
  for (int i = 0; i < m; ++i) r[i] = b[i] - ap[i];
  double bnorm = 0.0;
  for (int i = 0; i < m; ++i) bnorm += double(b[i]) * double(b[i]);
  const double threshold = fmax(tol * sqrt(bnorm), atol);
  double rz = 0.0;
  for (int i = 0; i < m; ++i) rz += double(r[i]) * double(r[i]);
  for (int i = 0; i < m; ++i) p[i] = r[i];
  for (int it = 0; it < maxiter; ++it) {
    if (sqrt(rz) <= threshold) break;
    apply(p, ap);
    double pap = 0.0;
    for (int i = 0; i < m; ++i) pap += double(p[i]) * double(ap[i]);
    if (pap == 0.0) break;
    const T alpha = T(rz / pap);
    for (int i = 0; i < m; ++i) { xx[i] = xx[i] + alpha * p[i]; r[i] = r[i] - alpha * ap[i]; }
    double rz_new = 0.0;
    for (int i = 0; i < m; ++i) rz_new += double(r[i]) * double(r[i]);
    const T beta = T(rz_new / rz);
    rz = rz_new;
    for (int i = 0; i < m; ++i) p[i] = r[i] + beta * p[i];
  }
}

// Reverse of k_root_solve when its conjugate gradient ran (solver_adjoint.py
// finite_root_cg, with the iteration count and the skip branch fixed): the
// forward iterates are recomputed with the forward's arithmetic and stored,
// then the iteration map is reversed in float64.  gx (layout of the solve:
// padded 64 rows with the mean term, else R) enters as the gradient w.r.t.
// the solution and leaves as the gradient w.r.t. the initial x; grhs
// (canonical) accumulates the right-hand-side gradient.
// With gvals, the root operator's value gradients (its canonical 27-slot
// layout) are accumulated from every A-product of the iteration.
template <class T>
__global__ void k_root_cg_vjp(const int* __restrict__ cols, const T* __restrict__ vals, int R, const int64_t* __restrict__ keys,
                              const T* __restrict__ rhs, const T* __restrict__ x_init, int dc, int base, int cycle, double acc,
                              int maxiter, double* __restrict__ store, double* __restrict__ gx, double* __restrict__ grhs,
                              double* __restrict__ gvals) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int m = dc ? 64 : R;
  T b[64], xx[64], r[64], p[64], ap[64];
  for (int i = 0; i < m; ++i) b[i] = T(0);
  for (int i = 0; i < R; ++i) b[dc ? root_row(keys, i) : i] = rhs[i];
  for (int i = 0; i < m; ++i) xx[i] = x_init[i];
  auto apply = [&](const T* v, T* out) { root_apply<T>(cols, vals, R, keys, dc, v, out); };
  for (int i = 0; i < m; ++i) r[i] = b[i];
  if (base == 0 && cycle > 0) {
    apply(xx, ap);
    for (int i = 0; i < m; ++i) r[i] = b[i] - ap[i];
  }
  double norm2 = 0.0;
  for (int i = 0; i < m; ++i) norm2 += double(r[i]) * double(r[i]);
  if (!(norm2 > acc * acc)) return;  // skipped: x unchanged
  const double tol = base == 0 ? 0.0 : acc;
  const double atol = base == 0 ? acc * sqrt(norm2) : 0.0;
  apply(xx, ap);
  for (int i = 0; i < m; ++i) r[i] = b[i] - ap[i];
  double bnorm = 0.0;
  for (int i = 0; i < m; ++i) bnorm += double(b[i]) * double(b[i]);
  const double threshold = fmax(tol * sqrt(bnorm), atol);
  double rz = 0.0;
  for (int i = 0; i < m; ++i) rz += double(r[i]) * double(r[i]);
  for (int i = 0; i < m; ++i) p[i] = r[i];
  // store: per iteration k, r_k | p_k | ap_k (64 each) and rz_k, pap_k, alpha_k, beta_k
  auto vec = [&](int k, int which) { return store + ((size_t)k * 3 + which) * 64; };
  double* scal = store + (size_t)(maxiter + 1) * 3 * 64;
  for (int i = 0; i < m; ++i) { vec(0, 0)[i] = double(r[i]); vec(0, 1)[i] = double(p[i]); }
  scal[0] = rz;
  int K = 0;
  for (int it = 0; it < maxiter; ++it) {
    if (sqrt(rz) <= threshold) break;
    apply(p, ap);
    double pap = 0.0;
    for (int i = 0; i < m; ++i) pap += double(p[i]) * double(ap[i]);
    if (pap == 0.0) break;
    const T alpha = T(rz / pap);
    for (int i = 0; i < m; ++i) { xx[i] = xx[i] + alpha * p[i]; r[i] = r[i] - alpha * ap[i]; }
    double rz_new = 0.0;
    for (int i = 0; i < m; ++i) rz_new += double(r[i]) * double(r[i]);
    const T beta = T(rz_new / rz);
    for (int i = 0; i < m; ++i) vec(it, 2)[i] = double(ap[i]);
    scal[4 * it + 1] = pap;
    scal[4 * it + 2] = double(alpha);
    scal[4 * it + 3] = double(beta);
    rz = rz_new;
    for (int i = 0; i < m; ++i) p[i] = r[i] + beta * p[i];
    for (int i = 0; i < m; ++i) { vec(it + 1, 0)[i] = double(r[i]); vec(it + 1, 1)[i] = double(p[i]); }
    scal[4 * (it + 1)] = rz;
    K = it + 1;
  }
  // A^T v in float64 (the padded operator's mean term is symmetric)
  auto apply_t = [&](const double* v, double* out) {
    for (int i = 0; i < m; ++i) out[i] = 0.0;
    for (int i = 0; i < R; ++i)
      for (int k = 0; k < 27; ++k) {
        const int c = cols[k * R + i];
        if (c < 0) continue;
        if (dc) out[root_row(keys, c)] += double(vals[k * R + i]) * v[root_row(keys, i)];
        else out[c] += double(vals[k * R + i]) * v[i];
      }
    if (dc) {
      double mean = 0.0;
      for (int i = 0; i < 64; ++i) mean += v[i];
      mean /= 64.0;
      for (int i = 0; i < 64; ++i) out[i] += mean;
    }
  };
  // gvals[k][i] += a * u[row(i)] * v[row(col)] for out = A v contributions
  auto outer = [&](const double* u, const double* v, double a) {
    for (int i = 0; i < R; ++i)
      for (int k = 0; k < 27; ++k) {
        const int c = cols[k * R + i];
        if (c < 0) continue;
        gvals[k * R + i] += a * u[dc ? root_row(keys, i) : i] * v[dc ? root_row(keys, c) : c];
      }
  };
  double xb[64], rb[64], pb[64], pk[64], apb[64], tmp[64];
  for (int i = 0; i < m; ++i) { xb[i] = gx[i]; rb[i] = 0.0; pb[i] = 0.0; }
  double rzb = 0.0;
  for (int k = K - 1; k >= 0; --k) {
    const double* rk1 = vec(k + 1, 0);
    const double* pk_ = vec(k, 1);
    const double* apk = vec(k, 2);
    const double rzk = scal[4 * k], rzk1 = scal[4 * (k + 1)], papk = scal[4 * k + 1], alpha = scal[4 * k + 2],
                 beta = scal[4 * k + 3];
    // p_{k+1} = r_{k+1} + beta_k p_k
    double betab = 0.0;
    for (int i = 0; i < m; ++i) { rb[i] += pb[i]; betab += pb[i] * pk_[i]; pk[i] = beta * pb[i]; }
    // beta_k = rz_{k+1} / rz_k
    rzb += betab / rzk;
    double rzbk = -betab * rzk1 / (rzk * rzk);
    // rz_{k+1} = r_{k+1} . r_{k+1}
    for (int i = 0; i < m; ++i) rb[i] += 2.0 * rzb * rk1[i];
    // r_{k+1} = r_k - alpha ap_k ; x_{k+1} = x_k + alpha p_k
    double alphab = 0.0;
    for (int i = 0; i < m; ++i) {
      alphab += -rb[i] * apk[i] + xb[i] * pk_[i];
      apb[i] = -alpha * rb[i];
      pk[i] += alpha * xb[i];
    }
    // alpha_k = rz_k / pap_k ; pap_k = p_k . ap_k
    rzbk += alphab / papk;
    const double papb = -alphab * rzk / (papk * papk);
    for (int i = 0; i < m; ++i) { pk[i] += papb * apk[i]; apb[i] += papb * pk_[i]; }
    // ap_k = A p_k
    apply_t(apb, tmp);
    for (int i = 0; i < m; ++i) pb[i] = pk[i] + tmp[i];
    if (gvals) outer(apb, pk_, 1.0);
    rzb = rzbk;
  }
  // p_0 = r_0, rz_0 = r_0 . r_0, r_0 = b - A x_init
  const double* r0 = vec(0, 0);
  for (int i = 0; i < m; ++i) rb[i] += pb[i] + 2.0 * rzb * r0[i];
  apply_t(rb, tmp);
  if (gvals) {
    double xi[64];
    for (int i = 0; i < m; ++i) xi[i] = double(x_init[i]);
    outer(rb, xi, -1.0);
  }
  for (int i = 0; i < m; ++i) gx[i] = xb[i] - tmp[i];
  for (int i = 0; i < R; ++i) grhs[i] += rb[dc ? root_row(keys, i) : i];
}

// final solution of a padded root solve: gradient w.r.t. pad_x
__global__ void k_root_pad_grad(const int64_t* __restrict__ keys, int R, const double* __restrict__ gx, double* __restrict__ gpad) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < 64; ++i) gpad[i] = 0.0;
  for (int i = 0; i < R; ++i) gpad[root_row(keys, i)] += gx[i];
}

template <class T>
__global__ void k_root_extract(const int64_t* __restrict__ keys, int R, const T* __restrict__ pad_x, T* __restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int cx = key_x(key) + 2, cy = key_y(key) + 2, cz = key_z(key) + 2;
  x[i] = pad_x[cz * 16 + (cx & 1) + 2 * (cy & 1) + 2 * (cx & 2) + 4 * (cy & 2)];
}

// A finer level's right-hand side at schedule position j (solver.py:119-126
// and :343-350): b_j = load_i + 0.5 * point_load_j rounded to T (i the row
// at j), then -1 * (cross * cumulative)_i + 1 * b_j as k_spmv combines with
// alpha = -1, beta = 1; and x_j = 0, the level's start value.
template <class T>
__global__ void k_level_rhs(const double* __restrict__ load, const double* __restrict__ point_load_sched,
                            const int* __restrict__ order, int R, const T* __restrict__ sum, T* __restrict__ b,
                            T* __restrict__ x) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const int i = order[j];
  const T bj = T(load[i]) + T(0.5) * T(point_load_sched[j]);
  T v = sum[i] * T(-1);
  v = v + T(1) * bj;
  b[j] = v;
  x[j] = T(0);
}



__global__ void k_inverse_permutation(const int* __restrict__ order, int R, int* __restrict__ inv) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < R) inv[order[j]] = j;
}

__global__ void k_remap_columns(const int* __restrict__ cols, const int* __restrict__ inv, int64_t n, int* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int c = cols[i];
    out[i] = c >= 0 ? inv[c] : -1;
  }
}

template <class T>
__global__ void k_scatter_rows(const int* __restrict__ order, int R, const T* __restrict__ xs, T* __restrict__ x) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < R) x[order[j]] = xs[j];
}

// 1/A_ii of schedule-ordered rows (the entry whose column is the row itself)
template <class T, class V = T>
__global__ void k_diag_sched(const int* __restrict__ order, int R, const int* __restrict__ cols, const V* __restrict__ vals,
                             T* __restrict__ inv_diag, unsigned char* __restrict__ has_diag) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const int i = order[j];
  T diagonal = T(0);
  for (int k = 0; k < 27; ++k)
    if (cols[(int64_t)k * R + j] == i) diagonal = T(vals[(int64_t)k * R + j]);
  has_diag[j] = diagonal != T(0);
  inv_diag[j] = diagonal != T(0) ? T(1) / diagonal : T(0);
}

// ---------------------------------------------------------------- tiled sliced relaxation
// Wave q of relaxation.py:128-143 updates slices two apart (visit t of slice
// z at q = 2t + 2^d - z), so slices z - 1 and z + 1 are constant during the
// wave.  One launch per wave: each workgroup takes one ts x ts (x, y) tile of
// one slice and runs its four colours in order with workgroup barriers
// between them.  Colour c at Chebyshev ring r outside the tile reads colours
// < c at ring r + 1, so the workgroup recomputes colour c on rings up to
// 3 - c and stores only its interior rows.  Old values of slices z - 1, z and
// z + 1 on rings up to 4 are staged in three LDS grids ((ts + 8)^2 cells
// each), so a row update reads its 27 neighbours from LDS by offset and needs
// no column indices: absent neighbours have no grid entry (the grid is
// cleared to +0) and their operator slots hold +0, so their term
// r - (+0)(+0) leaves r unchanged, bit for bit, as skipping it does.  Every
// update keeps its inputs and slot order, so the result equals the
// colour-per-launch form.  Two value buffers keep the loads and stores of one
// wave apart: after t visits a slice lives in buffer (iters - t) % 2, so
// visit t of slice z reads slices z - 1 and z from buffer (iters - t) % 2 and
// slice z + 1 from the other one, which it writes; the last visit lands in
// buffer 0.
//
// Tile tables, built per level from the schedule, for every tile whose
// ring-4 region holds rows of its slice: entries (schedule position, lx |
// ly << 8 | interior << 16, lx and ly the row's grid coordinates, the tile
// interior starting at 4) grouped per tile by category -- rows computed for colour
// 0..3 (categories 0..3; tiles without interior rows have none), then the
// staged values of the tile's slice (4) -- and per tile 12 ints: the five
// category starts, the end, the staged ranges of the tiles at the same
// (x, y) in slices z + 1 and z - 1 (empty when absent), which are exactly the
// values the tile needs from those slices.
constexpr int TILE_DESC_INTS = 12;

__device__ __forceinline__ int ring_of(int v, int t0, int ts) {
  const int lo = t0 - v, hi = v - (t0 + ts - 1);
  return lo > 0 ? lo : (hi > 0 ? hi : 0);
}

// Tiles of row j's slice whose ring-4 region holds it: flags (FLAG), counts,
// or the sort keys (compact tile << 3 | category) and values (j) from
// count[j] on (EMIT).
template <int MODE>  // 0 flag, 1 count, 2 emit
__global__ void k_tile_members(const int* __restrict__ order, const int64_t* __restrict__ keys, int R, int ts, int nt,
                               int* __restrict__ flags, const int* __restrict__ compact, int* __restrict__ count,
                               uint32_t* __restrict__ out_keys, int* __restrict__ out_vals) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const int64_t k = keys[order[j]];
  const int x = key_x(k), y = key_y(k), z = key_z(k);
  const int colour = (x & 1) + 2 * (y & 1);
  const int tx0 = max(0, (x - 4) / ts), tx1 = min(nt - 1, (x + 4) / ts);
  const int ty0 = max(0, (y - 4) / ts), ty1 = min(nt - 1, (y + 4) / ts);
  int c = 0, pos = MODE == 2 ? count[j] : 0;
  for (int ty = ty0; ty <= ty1; ++ty)
    for (int tx = tx0; tx <= tx1; ++tx) {
      const int r = max(ring_of(x, tx * ts, ts), ring_of(y, ty * ts, ts));
      if (r > 4) continue;
      const int64_t dense = ((int64_t)z * nt + ty) * nt + tx;
      if (MODE == 0) {
        flags[dense] = 1;
        continue;
      }
      const int n_here = r <= 3 - colour ? 2 : 1;
      if (MODE == 2) {
        const uint32_t t = (uint32_t)compact[dense] << 3;
        if (n_here == 2) { out_keys[pos] = t | (uint32_t)colour; out_vals[pos++] = j; }
        out_keys[pos] = t | 4u;
        out_vals[pos++] = j;
      }
      c += n_here;
    }
  if (MODE == 1) count[j] = c;
}

__global__ void k_tile_dense_of(const int* __restrict__ flags, const int* __restrict__ compact, int64_t ndense,
                                int* __restrict__ dense_of) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndense && flags[d]) dense_of[compact[d]] = (int)d;
}

__device__ __forceinline__ int lower_bound_u32(const uint32_t* __restrict__ a, int n, uint32_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Per tile: the category starts, the end and the neighbouring slices'
// staged ranges; stats[0] = most computed rows of a tile, stats[1] = most
// staged values (all three slices).
__global__ void k_tile_desc(const uint32_t* __restrict__ skeys, int E, int tiles, const int* __restrict__ dense_of,
                            const int* __restrict__ flags, const int* __restrict__ compact, int64_t per_slice,
                            int64_t ndense, int* __restrict__ desc, int* __restrict__ stats) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  int* o = desc + TILE_DESC_INTS * t;
  for (int c = 0; c < 6; ++c) o[c] = lower_bound_u32(skeys, E, ((uint32_t)t << 3) | (uint32_t)c);
  const int64_t d = dense_of[t];
  int staged = o[5] - o[4];
  for (int side = 0; side < 2; ++side) {  // slice z + 1, then z - 1
    const int64_t nd = side == 0 ? d + per_slice : d - per_slice;
    int lo = 0, hi = 0;
    if (nd >= 0 && nd < ndense && flags[nd]) {
      const uint32_t tn = (uint32_t)compact[nd] << 3;
      lo = lower_bound_u32(skeys, E, tn | 4u);
      hi = lower_bound_u32(skeys, E, tn | 5u);
    }
    o[6 + 2 * side] = lo;
    o[7 + 2 * side] = hi;
    staged += hi - lo;
  }
  o[10] = o[11] = 0;
  atomicMax(&stats[0], o[4] - o[0]);
  atomicMax(&stats[1], staged);
}

__global__ void k_tile_entries(const uint32_t* __restrict__ skeys, const int* __restrict__ svals, int E,
                               const int* __restrict__ dense_of, const int* __restrict__ order,
                               const int64_t* __restrict__ keys, int ts, int nt, int2* __restrict__ entries) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int t = (int)(skeys[e] >> 3), cat = (int)(skeys[e] & 7u), j = svals[e];
  const int d = dense_of[t];
  const int ox = (d % nt) * ts - 4, oy = ((d / nt) % nt) * ts - 4, g = ts + 8;
  const int64_t k = keys[order[j]];
  const int lx = key_x(k) - ox, ly = key_y(k) - oy;
  const int interior = (cat < 4 && lx >= 4 && lx < 4 + ts && ly >= 4 && ly < 4 + ts) ? 1 : 0;
  (void)g;
  entries[e] = make_int2(j, lx | (ly << 8) | (interior << 16));
}

// Tile tables from a level's dense z-major index, tile by tile: the
// workgroup of tile t scans its (ts + 8)^2 window of its slice in window
// order (lx-major) and files each row it finds under its categories -- the
// computed colour (ring <= 3 - colour) and the staged values (every row of
// the window).  A count pass gives the per-(tile, category) sizes, one scan
// their offsets, and an emit pass writes the entries in window order, so no
// sort is needed; rows of one category are independent, so their order
// within the category does not change any result.
constexpr int TILE_SCAN_THREADS = 256;

struct TileCell {
  int j, packed, cat;  // schedule position, lx | ly << 8 | interior << 16, computed colour or -1
  bool present;
};

template <bool EMIT>
__device__ __forceinline__ TileCell tile_cell(const int* __restrict__ dense, int dim, int n, int z, int ox, int oy, int ts,
                                              int g, int w, const int* __restrict__ inv) {
  TileCell c{0, 0, -1, false};
  if (w >= g * g) return c;
  const int lx = w / g, ly = w - (w / g) * g;
  const int x = ox + lx, y = oy + ly;
  if (x < 0 || y < 0 || x > n || y > n) return c;
  const int i = dense[((int64_t)z * dim + x) * dim + y];
  if (i < 0) return c;
  c.present = true;
  const int r = max(ring_of(x, ox + 4, ts), ring_of(y, oy + 4, ts));
  const int colour = (x & 1) + 2 * (y & 1);
  c.cat = r <= 3 - colour ? colour : -1;
  if (EMIT) {
    c.j = inv[i];
    c.packed = lx | (ly << 8);
  }
  return c;
}

__global__ void __launch_bounds__(TILE_SCAN_THREADS) k_tile_scan_count(const int* __restrict__ dense, int dim, int n,
                                                                        const int* __restrict__ dense_of, int nt, int ts,
                                                                        int* __restrict__ counts) {
  __shared__ int total[5];
  const int t = blockIdx.x, tid = threadIdx.x;
  if (tid < 5) total[tid] = 0;
  __syncthreads();
  const int d = dense_of[t];
  const int tx = d % nt, ty = (d / nt) % nt, z = d / (nt * nt);
  const int ox = tx * ts - 4, oy = ty * ts - 4, g = ts + 8;
  int cnt[5] = {0, 0, 0, 0, 0};
  for (int w = tid; w < g * g; w += TILE_SCAN_THREADS) {
    const TileCell c = tile_cell<false>(dense, dim, n, z, ox, oy, ts, g, w, nullptr);
    if (!c.present) continue;
    if (c.cat >= 0) ++cnt[c.cat];
    ++cnt[4];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (cnt[k]) atomicAdd(&total[k], cnt[k]);
  __syncthreads();
  if (tid < 5) counts[5 * (int64_t)t + tid] = total[tid];
}

__global__ void __launch_bounds__(TILE_SCAN_THREADS) k_tile_scan_emit(const int* __restrict__ dense, int dim, int n,
                                                                       const int* __restrict__ dense_of, int nt, int ts,
                                                                       const int* __restrict__ inv, const int* __restrict__ offs,
                                                                       int2* __restrict__ entries) {
  constexpr int WAVES = TILE_SCAN_THREADS / 64;
  __shared__ int base[5];
  __shared__ int wave_total[5][WAVES];
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < 5) base[tid] = offs[5 * (int64_t)t + tid];
  const int d = dense_of[t];
  const int tx = d % nt, ty = (d / nt) % nt, z = d / (nt * nt);
  const int ox = tx * ts - 4, oy = ty * ts - 4, g = ts + 8;
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  __syncthreads();
  for (int w0 = 0; w0 < g * g; w0 += TILE_SCAN_THREADS) {
    const TileCell c = tile_cell<true>(dense, dim, n, z, ox, oy, ts, g, w0 + tid, inv);
    int pre[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const bool f = k < 4 ? c.cat == k : c.present;
      const unsigned long long m = __ballot(f);
      pre[k] = __popcll(m & below);
      if (lane == 0) wave_total[k][wave] = __popcll(m);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const bool f = k < 4 ? c.cat == k : c.present;
      if (!f) continue;
      int off = base[k] + pre[k];
      for (int v = 0; v < wave; ++v) off += wave_total[k][v];
      const int lx = c.packed & 255, ly = c.packed >> 8;
      const int interior = (k < 4 && lx >= 4 && lx < 4 + ts && ly >= 4 && ly < 4 + ts) ? 1 : 0;
      entries[off] = make_int2(c.j, c.packed | (interior << 16));
    }
    __syncthreads();
    if (tid < 5) {
      int sum = 0;
      for (int v = 0; v < WAVES; ++v) sum += wave_total[tid][v];
      base[tid] += sum;
    }
    __syncthreads();
  }
}

// desc of tile t from the category offsets (offs[5 t + k], offs[5 tiles] = E)
__global__ void k_tile_desc_offs(const int* __restrict__ offs, int tiles, const int* __restrict__ dense_of,
                                 const int* __restrict__ flags, const int* __restrict__ compact, int64_t per_slice,
                                 int64_t ndense, int* __restrict__ desc, int* __restrict__ stats) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  int* o = desc + TILE_DESC_INTS * t;
  for (int c = 0; c < 6; ++c) o[c] = offs[5 * (int64_t)t + c];
  const int64_t d = dense_of[t];
  int staged = o[5] - o[4];
  for (int side = 0; side < 2; ++side) {  // slice z + 1, then z - 1
    const int64_t nd = side == 0 ? d + per_slice : d - per_slice;
    int lo = 0, hi = 0;
    if (nd >= 0 && nd < ndense && flags[nd]) {
      const int64_t tn = compact[nd];
      lo = offs[5 * tn + 4];
      hi = offs[5 * tn + 5];
    }
    o[6 + 2 * side] = lo;
    o[7 + 2 * side] = hi;
    staged += hi - lo;
  }
  o[10] = o[11] = 0;
  atomicMax(&stats[0], o[4] - o[0]);
  atomicMax(&stats[1], staged);
}

// first tile of each slice z (z = 0..n + 1) from the exclusive scan of the flags
__global__ void k_tile_slice_start(const int* __restrict__ compact, int n, int per_slice, int* __restrict__ out) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z <= n + 1) out[z] = compact[(int64_t)z * per_slice];
}

template <class T, class V> struct TiledArgs {
  const int2* entries;
  const int* desc;
  const RowRecord<T, V>* records;  // by schedule position
  const T* b;                      // right-hand side by schedule position
  T* x0;
  T* x1;
  int g;  // grid side, ts + 8
};

// The slices of one wave: tiles [start[s], end[s]) of slice z0 + 2s; bit s of
// flip selects buffer 1 as the slice's read buffer.
struct WaveTiles {
  int start[RELAX_MAX_SLICES], end[RELAX_MAX_SLICES];
  unsigned flip;
};

// LDS grid layout: per slice layer, four parity planes ((y & 1) * 2 + (x & 1),
// the row colours) of gh * gh cells, x-major inside a plane: rows of one
// colour with consecutive y -- consecutive threads -- read and write
// consecutive words, free of bank conflicts; a neighbour (dx, dy) of a row
// of known colour lies at a constant offset in a known plane.  The tile side
// is even, so grid parities equal key parities.
__device__ __forceinline__ int grid_index(int lx, int ly, int gh) {
  return (((ly & 1) << 1) | (lx & 1)) * gh * gh + (lx >> 1) * gh + (ly >> 1);
}

// One computed row: its record, right-hand side and position; base =
// (lx >> 1) * gh + (ly >> 1), the row's offset inside its plane.
template <class T, class V> struct TiledRow {
  int j, base, interior;
  RowRecord<T, V> op;
  T b;
};

template <class T, class V>
__device__ __forceinline__ void tiled_place(int2 en, int gh, TiledRow<T, V>& w) {
  const int lx = en.y & 255, ly = (en.y >> 8) & 255;
  w.j = en.x;
  w.base = (lx >> 1) * gh + (ly >> 1);
  w.interior = (en.y >> 16) & 1;
}

// Row w.j's record and right-hand side.
template <class T, class V>
__device__ __forceinline__ void tiled_load(const TiledArgs<T, V>& a, TiledRow<T, V>& w) {
  RowRecordWords<T, V> u;
  const record_word* src = (const record_word*)(a.records + (unsigned)w.j);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(u.q) / sizeof(record_word)); ++i) u.q[i] = src[i];
  w.op = u.r;
  w.b = a.b[(unsigned)w.j];
}

// The update of a row of colour C (x parity C & 1, y parity C >> 1):
// neighbour (dx, dy, dz) sits in layer dz + 1, plane of parities
// ((C & 1) + dx, (C >> 1) + dy), at base + ((px + dx) >> 1) * gh + ((py + dy) >> 1).
template <int C, class T, class V>
__device__ __forceinline__ void tiled_update(T* grid, int gh, T* Q, const TiledRow<T, V>& w) {
  constexpr int px = C & 1, py = C >> 1;
  const int pl = gh * gh, layer = 4 * pl;
  T r = w.b;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
    const int qx = px + dx + 2, qy = py + dy + 2;  // shifted to stay nonnegative
    const int plane = ((qy & 1) << 1) | (qx & 1);
    const int off = (dz + 1) * layer + plane * pl + ((qx >> 1) - 1) * gh + ((qy >> 1) - 1);
    r = r - grid[off + w.base] * T(w.op.v[k]);
  }
  const int self = layer + (py * 2 + px) * pl + w.base;
  const T xi = grid[self];
  const T nv = w.op.has ? xi + r * w.op.inv : xi;
  grid[self] = nv;
  if (w.interior) Q[w.j] = nv;
}

// BS threads per tile: thread i computes the tile's i-th row (rows of colour
// c are consecutive, so a colour's rows sit on consecutive threads) and
// stages grid values i, i + BS, ...  Loads go out in dependency steps:
// descriptor; entries and row records; staged values.  Tiles with more than
// BS computed rows take further passes for the rest.
constexpr int TILED_STAGE_UNROLL = 4;

template <class T, class V>
__device__ __forceinline__ void tiled_phase(int phase, T* grid, int gh, T* Q, const TiledRow<T, V>& w) {
  switch (phase) {
    case 0: tiled_update<0>(grid, gh, Q, w); break;
    case 1: tiled_update<1>(grid, gh, Q, w); break;
    case 2: tiled_update<2>(grid, gh, Q, w); break;
    default: tiled_update<3>(grid, gh, Q, w); break;
  }
}

// One tile of slice s of a wave: the bx-th tile of the slice's range.
template <class T, class V, int BS>
__device__ __forceinline__ void relax_tile(const TiledArgs<T, V>& a, const WaveTiles& wv, int bx, int s, T* grid) {
  const int tile = wv.start[s] + bx;
  if (tile >= wv.end[s]) return;
  const bool flip = (wv.flip >> s) & 1u;
  const T* P = flip ? a.x1 : a.x0;
  T* Q = flip ? a.x0 : a.x1;
  const int4 d0 = *(const int4*)(a.desc + TILE_DESC_INTS * tile);
  const int4 d1 = *(const int4*)(a.desc + TILE_DESC_INTS * tile + 4);
  const int4 d2 = *(const int4*)(a.desc + TILE_DESC_INTS * tile + 8);
  const int c0 = d0.x, c1 = d0.y, c2 = d0.z, c3 = d0.w, stage0 = d1.x;
  if (c0 == stage0) return;  // no interior rows: staged data for the neighbouring slices only
  const int tid = (int)threadIdx.x, gh = a.g >> 1, layer = 4 * gh * gh;
  // staged values: this slice [d1.x, d1.y) (layer 1), slice z + 1 [d1.z, d1.w)
  // (layer 2, read buffer of the write side) and slice z - 1 [d2.x, d2.y)
  // (layer 0), as one index range
  const int n_own = d1.y - d1.x, n_up = d1.w - d1.z, n_all = n_own + n_up + (d2.y - d2.x);
  const int e = c0 + tid;
  const bool have = e < stage0;
  const int colour = e < c1 ? 0 : (e < c2 ? 1 : (e < c3 ? 2 : 3));
  // Loads are issued unconditionally, at clamped indices (the tile has
  // computed rows and staged values of its own slice, so c0 and d1.x are
  // entries of it), and only the stores are predicated: with no branch
  // around them, each dependency step waits once for all of its loads.
  const int ec = have ? e : c0;
  TiledRow<T, V> w;
  int2 sen[TILED_STAGE_UNROLL];
  int slay[TILED_STAGE_UNROLL];
  T sval[TILED_STAGE_UNROLL];
  auto stage_entries = [&](int first) {
#pragma unroll
    for (int u = 0; u < TILED_STAGE_UNROLL; ++u) {
      const int i = first + u * BS;
      const int ic = i < n_all ? i : 0;
      const int lay = ic < n_own ? 1 : (ic < n_own + n_up ? 2 : 0);
      const int idx = lay == 1 ? d1.x + ic : (lay == 2 ? d1.z + (ic - n_own) : d2.x + (ic - n_own - n_up));
      slay[u] = i < n_all ? lay : -1;
      sen[u] = a.entries[idx];
    }
  };
  auto stage_values = [&]() {
#pragma unroll
    for (int u = 0; u < TILED_STAGE_UNROLL; ++u) sval[u] = (slay[u] == 2 ? (const T*)Q : P)[(unsigned)sen[u].x];
  };
  auto stage_store = [&]() {
#pragma unroll
    for (int u = 0; u < TILED_STAGE_UNROLL; ++u)
      if (slay[u] >= 0) grid[slay[u] * layer + grid_index(sen[u].y & 255, (sen[u].y >> 8) & 255, gh)] = sval[u];
  };
  stage_entries(tid);
  const int2 en = a.entries[ec];
  // a thread's second row (tiles of more than BS computed rows): its entry
  // now, so its record loads need no further step
  const int2 en2 = a.entries[e + BS < stage0 ? e + BS : c0];
  stage_values();
  tiled_place(en, gh, w);
  tiled_load(a, w);
  for (int i = tid; i < 3 * layer; i += BS) grid[i] = T(0);
  __syncthreads();  // grid cleared
  stage_store();
  for (int first = tid + TILED_STAGE_UNROLL * BS; first - tid < n_all; first += TILED_STAGE_UNROLL * BS) {
    stage_entries(first);
    stage_values();
    stage_store();
  }
  __syncthreads();  // grid staged
  // Further rows of a thread come after its first row in entry order, so
  // their colours are at least its colour: they reuse the first row's
  // registers once it is done.
  const int cstart[5] = {c0, c1, c2, c3, stage0};
  for (int phase = 0; phase < 4; ++phase) {
    if (have && colour == phase) tiled_phase(phase, grid, gh, Q, w);
    for (int e2 = e + BS; e2 < stage0; e2 += BS) {
      if (e2 < cstart[phase] || e2 >= cstart[phase + 1]) continue;
      tiled_place(e2 == e + BS ? en2 : a.entries[e2], gh, w);
      tiled_load(a, w);
      tiled_phase(phase, grid, gh, Q, w);
    }
    __syncthreads();
  }
}

template <class T, class V, int BS>
__global__ void __launch_bounds__(BS) k_relax_tiled(TiledArgs<T, V> a, WaveTiles wv) {
  extern __shared__ unsigned char lds_raw[];
  relax_tile<T, V, BS>(a, wv, (int)blockIdx.x, (int)blockIdx.y, (T*)lds_raw);
}

// Several reconstructions' tiles of the same wave in one launch: members'
// blocks are consecutive ranges of x, [first[m], first[m + 1]); the
// reconstructions share no data, so each tile's updates are those of its
// own reconstruction's launch.
constexpr int RELAX_GROUP_MAX = 8;

template <class T, class V> struct TiledGroup {
  int members;
  int first[RELAX_GROUP_MAX + 1];
  TiledArgs<T, V> a[RELAX_GROUP_MAX];
  WaveTiles wv[RELAX_GROUP_MAX];
};

template <class T, class V, int BS>
__global__ void __launch_bounds__(BS) k_relax_tiled_group(TiledGroup<T, V> g) {
  int m = 0;
  for (int k = 1; k < g.members; ++k) m = (int)blockIdx.x >= g.first[k] ? k : m;
  extern __shared__ unsigned char lds_raw[];
  relax_tile<T, V, BS>(g.a[m], g.wv[m], (int)blockIdx.x - g.first[m], (int)blockIdx.y, (T*)lds_raw);
}

inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- host side

template <class T> struct Schedule {
  DArray<int> order, offsets;
  std::vector<int> host_offsets;
  int max_group = 0;
};

// host: also copy the group offsets to the host (host_offsets, max_group),
// which the launch-per-group relaxations size their grids by; the
// single-workgroup relaxations read them on the device only.
template <class T> static Schedule<T> make_schedule(const DArray<int64_t>& keys, int R, int depth, bool sliced, hipStream_t s,
                                                     bool host = true) {
  Schedule<T> sc;
  const int groups = sliced ? 4 * ((1 << depth) + 1) : 8;
  DArray<uint64_t> g(R, s);
  sc.order.reset(R, s);
  if (R) {
    hipLaunchKernelGGL(k_groups_of, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int64_t*)keys.data(), R, sliced ? 1 : 0,
                       g.data(), sc.order.data());
    KERNEL_CHECK();
    sort_pairs_u64(g.data(), sc.order.data(), R, bits_for((uint64_t)groups), s);
  }
  sc.offsets.reset(groups + 1, s);
  hipLaunchKernelGGL(k_group_offsets, dim3(blocks_for(groups + 1, 256)), dim3(256), 0, s, (const uint64_t*)g.data(), R, groups,
                     sc.offsets.data());
  KERNEL_CHECK();
  if (!host) return sc;
  sc.host_offsets.resize(groups + 1);
  read_bytes(sc.host_offsets.data(), sc.offsets.data(), sizeof(int) * (groups + 1), s);
  for (int t = 0; t < groups; ++t) sc.max_group = std::max(sc.max_group, sc.host_offsets[t + 1] - sc.host_offsets[t]);
  return sc;
}

template <class T, class V = T> struct GsOperator {
  int R = 0;
  DArray<int> cols;
  DArray<V> vals;
  DArray<T> b, inv_diag;
  DArray<unsigned char> has_diag;
  // row of x updated by schedule position j: update_order[j], or j itself
  // when x is kept in schedule order (nullptr)
  const int* update_order = nullptr;
};

template <class T>
static GsOperator<T> gather_operator(const Schedule<T>& sc, const DArray<int>& cols, const DArray<T>& vals, const T* b, int R,
                                     hipStream_t s) {
  GsOperator<T> op;
  op.R = R;
  op.update_order = sc.order.data();
  op.cols.reset(27 * (size_t)R, s);
  op.vals.reset(27 * (size_t)R, s);
  op.b.reset(R, s);
  op.inv_diag.reset(R, s);
  op.has_diag.reset(R, s);
  if (R) {
    hipLaunchKernelGGL(k_gather_operator<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int*)sc.order.data(), R,
                       (const int*)cols.data(), (const T*)vals.data(), b, op.cols.data(), op.vals.data(), op.b.data(),
                       op.inv_diag.data(), op.has_diag.data());
    KERNEL_CHECK();
  }
  return op;
}

template <class T> static void refresh_rhs(GsOperator<T>& op, const Schedule<T>& sc, const T* b, hipStream_t s);

template <class T>
__global__ void k_gather_b(const int* __restrict__ order, int R, const T* __restrict__ b, T* __restrict__ sb) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < R) sb[j] = b[order[j]];
}

template <class T> static void refresh_rhs(GsOperator<T>& op, const Schedule<T>& sc, const T* b, hipStream_t s) {
  if (!op.R) return;
  hipLaunchKernelGGL(k_gather_b<T>, dim3(blocks_for(op.R, 256)), dim3(256), 0, s, (const int*)sc.order.data(), op.R, b,
                     op.b.data());
  KERNEL_CHECK();
}

__global__ void __launch_bounds__(256) k_count_present(const int* __restrict__ cols, int64_t n,
                                                       unsigned long long* __restrict__ out) {
  __shared__ unsigned wave_counts[4];
  unsigned count = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    count += cols[i] >= 0 ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) count += __shfl_down(count, o, 64);
  if ((threadIdx.x & 63) == 0) wave_counts[threadIdx.x >> 6] = count;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(out, (unsigned long long)(wave_counts[0] + wave_counts[1] + wave_counts[2] + wave_counts[3]));
}

// relaxation.py:128-143 (wavefront_groups) + :189-210; returns the launch count
template <class T, class V>
int relax_sliced(const GsOperator<T, V>& op, const Schedule<T>& sc, T* x, int depth, int iters, hipStream_t s) {
  const int n = 1 << depth;
  const unsigned bx = (unsigned)std::max(1, (sc.max_group + 255) / 256);
  int launches = 0;
  for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
    const int first = std::max(0, floordiv(q - n + 1, 2));
    const int last = std::min(iters - 1, q / 2);
    if (last < first) continue;
    const int z0 = n - q + 2 * first;
    const unsigned count = (unsigned)(last - first + 1);
    for (int color = 0; color < 4; ++color) {
      hipLaunchKernelGGL((k_relax_wave<T, V>), dim3(bx, count), dim3(256), 0, s, op.update_order,
                         (const int*)sc.offsets.data(), (const int*)op.cols.data(), (const V*)op.vals.data(),
                         (const T*)op.b.data(), (const T*)op.inv_diag.data(), (const unsigned char*)op.has_diag.data(), x,
                         op.R, z0, color);
      ++launches;
    }
  }
  KERNEL_CHECK();
  return launches;
}

// The launch sequence of relax_sliced as a graph of kernel nodes, cached per
// (precision, depth, iters, row capacity, device, caller stream).  Each call
// waits for the previous launch of the graph to finish (normally long done),
// writes its buffers and slice ranges into every node's arguments
// (hipGraphExecKernelNodeSetParams) and launches the graph on the caller's
// stream.
struct RelaxGraph {
  hipGraph_t graph = nullptr;  // kept: the node handles below belong to it
  hipGraphExec_t exec = nullptr;
  std::vector<hipGraphNode_t> nodes;
  std::vector<int> first_z, slices, color;  // per node
  hipKernelNodeParams params{};
  hipEvent_t done = nullptr;  // completion of the last launch
  std::mutex mu;              // one call at a time writes the arguments and launches
};

// 64-row workgroups spread one launch's rows over the most CUs (a launch holds
// at most a few thousand rows per slice; occupancy does not bound it).
constexpr int RELAX_GRAPH_BS = 64;

template <class T, class V>
static std::unique_ptr<RelaxGraph> build_relax_graph(int n, int iters, unsigned bx) {
  auto g = std::make_unique<RelaxGraph>();
  HIP_CHECK(hipGraphCreate(&g->graph, 0));
  hipGraph_t graph = g->graph;
  RelaxArgs<T, V> d{};
  SliceRanges rg{};
  void* args[2] = {&d, &rg};
  hipKernelNodeParams& kp = g->params;
  kp.func = reinterpret_cast<void*>(&k_relax_wave_args<T, V, RELAX_GRAPH_BS>);
  kp.blockDim = dim3(RELAX_GRAPH_BS);
  kp.sharedMemBytes = 0;
  kp.extra = nullptr;
  hipGraphNode_t prev = nullptr;
  for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
    const int first = std::max(0, floordiv(q - n + 1, 2));
    const int last = std::min(iters - 1, q / 2);
    if (last < first) continue;
    const int z0 = n - q + 2 * first;
    for (int color = 0; color < 4; ++color) {
      kp.gridDim = dim3(bx, (unsigned)(last - first + 1));
      kp.kernelParams = args;
      hipGraphNode_t node = nullptr;
      HIP_CHECK(hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
      g->nodes.push_back(node);
      g->first_z.push_back(z0);
      g->slices.push_back(last - first + 1);
      g->color.push_back(color);
      prev = node;
    }
  }
  HIP_CHECK(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
  HIP_CHECK(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
  return g;
}

template <class T, class V>
int relax_sliced_graph(const GsOperator<T, V>& op, const Schedule<T>& sc, T* x, int depth, int iters, hipStream_t s) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, unsigned, int, hipStream_t>, std::unique_ptr<RelaxGraph>> cache;
  const int n = 1 << depth;
  unsigned bx = 1;
  while (bx * (unsigned)RELAX_GRAPH_BS < (unsigned)std::max(1, sc.max_group)) bx <<= 1;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  RelaxGraph* g = nullptr;
  {
    std::lock_guard<std::mutex> guard(mu);
    std::unique_ptr<RelaxGraph>& slot = cache[std::make_tuple((int)sizeof(V), depth, iters, bx, dev, s)];
    if (!slot) slot = build_relax_graph<T, V>(n, iters, bx);
    g = slot.get();
  }
  std::lock_guard<std::mutex> guard(g->mu);
  HIP_CHECK(hipEventSynchronize(g->done));
  RelaxArgs<T, V> d{op.update_order, op.cols.data(), op.vals.data(), op.b.data(), op.inv_diag.data(), op.has_diag.data(), x, op.R};
  SliceRanges rg;
  void* args[2] = {&d, &rg};
  hipKernelNodeParams kp = g->params;
  kp.kernelParams = args;
  const std::vector<int>& off = sc.host_offsets;
  for (size_t i = 0; i < g->nodes.size(); ++i) {
    for (int t = 0; t < RELAX_MAX_SLICES; ++t) rg.start[t] = rg.end[t] = 0;
    for (int t = 0; t < g->slices[i]; ++t) {
      const int group = 4 * (g->first_z[i] + 2 * t) + g->color[i];
      rg.start[t] = off[group];
      rg.end[t] = off[group + 1];
    }
    kp.gridDim = dim3(bx, (unsigned)g->slices[i]);
    HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec, g->nodes[i], &kp));
  }
  HIP_CHECK(hipGraphLaunch(g->exec, s));
  HIP_CHECK(hipEventRecord(g->done, s));
  return (int)g->nodes.size();
}

// PSR_RELAX=wave selects direct colour-per-launch launches, PSR_RELAX=graph
// their captured graph; the default is the tiled form (one launch per wave)
// where a level has tile tables.  All give the same bits.
static int relax_mode() {
  static int mode = -1;
  if (mode < 0) {
    const char* e = std::getenv("PSR_RELAX");
    mode = (e && std::string(e) == "wave") ? 1 : (e && std::string(e) == "graph") ? 2 : 0;
  }
  return mode;
}

template <class T, class V>
int relax_sliced_any(const GsOperator<T, V>& op, const Schedule<T>& sc, T* x, int depth, int iters, hipStream_t s) {
  switch (relax_mode()) {
    case 1: return relax_sliced(op, sc, x, depth, iters, s);
    case 2:
    default:
      if (iters > RELAX_MAX_SLICES) return relax_sliced(op, sc, x, depth, iters, s);
      return relax_sliced_graph(op, sc, x, depth, iters, s);
  }
}

// ---------------------------------------------------------------- tiled sliced relaxation (host)
// The tile tables of one level (see k_relax_tiled).
struct TileTable {
  int ts = 0, nt = 0, tiles = 0, entries_count = 0;
  DArray<int2> entries;
  DArray<int> desc;
  std::vector<int> slice_start;  // first tile of slice z, z = 0..2^d + 1
  int max_tiles = 0;             // most tiles of one slice
  int max_rows = 0;              // most computed rows of one tile
  int max_staged = 0;            // most staged values of one tile
};

// Tile side for a level: the run's relax_tile (8..40, which keeps the three
// double grids within 64 KiB of LDS) or, for 0, the library default by depth
// (12 up to depth 6, 16 for depths 7-8, 18 beyond); -1 (or
// the PSR_RELAX environment choice of a colour-per-launch form) selects no
// tiles.
constexpr int RELAX_TILE_MAX = 40;
static int relax_tile_size(int requested, int depth) {
  if (requested < 0 || relax_mode() != 0) return 0;
  if (requested >= 8 && requested <= RELAX_TILE_MAX && requested % 2 == 0) return requested;
  // measured on the 800k-point depth-9 workload: the small levels gain from
  // smaller tiles (more workgroups per launch despite more recomputed ring
  // rows), the finest level from 18
  return depth <= 6 ? 12 : depth <= 8 ? 16 : 18;
}

// With the level's dense z-major index (level_dense, side dim) the entries
// come tile by tile from the index (k_tile_scan_*); without it, row by row
// and a sort of the (tile, category) keys.
static void build_tile_table(const int64_t* keys, const int* order, int R, int depth, int ts, const int* level_dense,
                             int dim, TileTable& tt, hipStream_t s) {
  const int n = 1 << depth;
  tt.ts = ts;
  tt.nt = (n + 1 + ts - 1) / ts;
  const int64_t per_slice = (int64_t)tt.nt * tt.nt;
  const int64_t ndense = per_slice * (n + 1);
  if (ndense + 1 > INT32_MAX) throw PsrError(1, "tiled relaxation: too many tiles");
  DArray<int> flags(ndense + 1, s), compact(ndense + 1, s), total(1, s);
  flags.zero();
  hipLaunchKernelGGL(k_tile_members<0>, dim3(blocks_for(R, 256)), dim3(256), 0, s, order, keys, R, ts, tt.nt, flags.data(),
                     (const int*)nullptr, (int*)nullptr, (uint32_t*)nullptr, (int*)nullptr);
  KERNEL_CHECK();
  scan_exclusive_i32(flags.data(), compact.data(), ndense + 1, total.data(), s);
  if (level_dense && dim == n + 1) {
    tt.tiles = read_int(total.data(), s);
    const int tiles = tt.tiles;
    DArray<int> dense_of(std::max(tiles, 1), s);
    hipLaunchKernelGGL(k_tile_dense_of, dim3(blocks_for(ndense, 256)), dim3(256), 0, s, (const int*)flags.data(),
                       (const int*)compact.data(), ndense, dense_of.data());
    KERNEL_CHECK();
    DArray<int> inv(std::max(R, 1), s), offs(5 * (size_t)tiles + 1, s), etotal(1, s);
    hipLaunchKernelGGL(k_inverse_permutation, dim3(blocks_for(R, 256)), dim3(256), 0, s, order, R, inv.data());
    KERNEL_CHECK();
    HIP_CHECK(hipMemsetAsync(offs.data() + 5 * (size_t)tiles, 0, sizeof(int), s));
    if (tiles) {
      hipLaunchKernelGGL(k_tile_scan_count, dim3(tiles), dim3(TILE_SCAN_THREADS), 0, s, level_dense, dim, n,
                         (const int*)dense_of.data(), tt.nt, ts, offs.data());
      KERNEL_CHECK();
    }
    scan_exclusive_i32(offs.data(), offs.data(), 5 * (int64_t)tiles + 1, etotal.data(), s);
    const int E = read_int(etotal.data(), s);
    tt.entries_count = E;
    tt.entries.reset(std::max(E, 1), s);
    DArray<int> small(n + 2 + 2, s);
    HIP_CHECK(hipMemsetAsync(small.data() + n + 2, 0, 2 * sizeof(int), s));
    tt.desc.reset(TILE_DESC_INTS * (size_t)std::max(tiles, 1), s);
    if (tiles) {
      hipLaunchKernelGGL(k_tile_scan_emit, dim3(tiles), dim3(TILE_SCAN_THREADS), 0, s, level_dense, dim, n,
                         (const int*)dense_of.data(), tt.nt, ts, (const int*)inv.data(), (const int*)offs.data(),
                         tt.entries.data());
      KERNEL_CHECK();
      hipLaunchKernelGGL(k_tile_desc_offs, dim3(blocks_for(tiles, 256)), dim3(256), 0, s, (const int*)offs.data(), tiles,
                         (const int*)dense_of.data(), (const int*)flags.data(), (const int*)compact.data(), per_slice,
                         ndense, tt.desc.data(), small.data() + n + 2);
      KERNEL_CHECK();
    }
    hipLaunchKernelGGL(k_tile_slice_start, dim3(blocks_for(n + 2, 256)), dim3(256), 0, s, (const int*)compact.data(), n,
                       (int)per_slice, small.data());
    KERNEL_CHECK();
    std::vector<int> h(n + 4);
    read_bytes(h.data(), small.data(), sizeof(int) * (n + 4), s);
    tt.slice_start.assign(h.begin(), h.begin() + n + 2);
    tt.max_rows = h[n + 2];
    tt.max_staged = h[n + 3];
    tt.max_tiles = 0;
    for (int z = 0; z <= n; ++z) tt.max_tiles = std::max(tt.max_tiles, tt.slice_start[z + 1] - tt.slice_start[z]);
    return;
  }
  DArray<int> count(R + 1, s), pos(R + 1, s), etotal(1, s);
  hipLaunchKernelGGL(k_tile_members<1>, dim3(blocks_for(R, 256)), dim3(256), 0, s, order, keys, R, ts, tt.nt, flags.data(),
                     (const int*)compact.data(), count.data(), (uint32_t*)nullptr, (int*)nullptr);
  KERNEL_CHECK();
  HIP_CHECK(hipMemsetAsync(count.data() + R, 0, sizeof(int), s));
  scan_exclusive_i32(count.data(), pos.data(), R + 1, etotal.data(), s);
  int sizes[2];
  HIP_CHECK(hipMemcpyAsync(&sizes[0], total.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&sizes[1], etotal.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  tt.tiles = sizes[0];
  const int E = sizes[1];
  tt.entries_count = E;
  DArray<uint32_t> skeys(std::max(E, 1), s);
  DArray<int> svals(std::max(E, 1), s);
  hipLaunchKernelGGL(k_tile_members<2>, dim3(blocks_for(R, 256)), dim3(256), 0, s, order, keys, R, ts, tt.nt, flags.data(),
                     (const int*)compact.data(), pos.data(), skeys.data(), svals.data());
  KERNEL_CHECK();
  sort_pairs_u32(skeys.data(), svals.data(), E, bits_for((uint64_t)tt.tiles * 8), s);
  DArray<int> dense_of(std::max(tt.tiles, 1), s);
  hipLaunchKernelGGL(k_tile_dense_of, dim3(blocks_for(ndense, 256)), dim3(256), 0, s, (const int*)flags.data(),
                     (const int*)compact.data(), ndense, dense_of.data());
  KERNEL_CHECK();
  // slice starts and stats in one buffer, read once
  DArray<int> small(n + 2 + 2, s);
  HIP_CHECK(hipMemsetAsync(small.data() + n + 2, 0, 2 * sizeof(int), s));
  tt.desc.reset(TILE_DESC_INTS * (size_t)std::max(tt.tiles, 1), s);
  hipLaunchKernelGGL(k_tile_desc, dim3(blocks_for(tt.tiles, 256)), dim3(256), 0, s, (const uint32_t*)skeys.data(), E, tt.tiles,
                     (const int*)dense_of.data(), (const int*)flags.data(), (const int*)compact.data(), per_slice, ndense,
                     tt.desc.data(), small.data() + n + 2);
  KERNEL_CHECK();
  tt.entries.reset(std::max(E, 1), s);
  hipLaunchKernelGGL(k_tile_entries, dim3(blocks_for(E, 256)), dim3(256), 0, s, (const uint32_t*)skeys.data(),
                     (const int*)svals.data(), E, (const int*)dense_of.data(), order, keys, ts, tt.nt, tt.entries.data());
  KERNEL_CHECK();
  hipLaunchKernelGGL(k_tile_slice_start, dim3(blocks_for(n + 2, 256)), dim3(256), 0, s, (const int*)compact.data(), n,
                     (int)per_slice, small.data());
  KERNEL_CHECK();
  std::vector<int> h(n + 4);
  read_bytes(h.data(), small.data(), sizeof(int) * (n + 4), s);
  tt.slice_start.assign(h.begin(), h.begin() + n + 2);
  tt.max_rows = h[n + 2];
  tt.max_staged = h[n + 3];
  tt.max_tiles = 0;
  for (int z = 0; z <= n; ++z) tt.max_tiles = std::max(tt.max_tiles, tt.slice_start[z + 1] - tt.slice_start[z]);
}

// Threads per tile: enough for most tiles' computed rows in one pass (256 at
// most; larger tiles take further passes).
static int tiled_width(const TileTable& tt) { return tt.max_rows <= 64 ? 64 : tt.max_rows <= 128 ? 128 : 256; }

// The wave sequence of the tiled relaxation as a graph of kernel nodes (one
// per wave), cached per (precision, depth, iters, width, grid x, device,
// caller stream); each call writes its buffers and the waves' tile ranges
// into the node arguments, as relax_sliced_graph does.
struct TiledGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipGraphNode_t> nodes;
  std::vector<int> first_z, slices, first_t;  // per node
  hipKernelNodeParams params{};
  hipEvent_t done = nullptr;
  std::mutex mu;
};

template <class T, class V, int BS> static void* tiled_kernel() {
  return reinterpret_cast<void*>(&k_relax_tiled<T, V, BS>);
}

template <class T, class V, int BS> static void* tiled_group_kernel() {
  return reinterpret_cast<void*>(&k_relax_tiled_group<T, V, BS>);
}

// One kernel node per wave of a level of side n with iters sweeps, chained;
// args points at placeholder arguments of the kernel func.
static std::unique_ptr<TiledGraph> build_wave_graph(int n, int iters, void* func, int width, unsigned bx, size_t lds,
                                                     void** args) {
  auto g = std::make_unique<TiledGraph>();
  HIP_CHECK(hipGraphCreate(&g->graph, 0));
  hipKernelNodeParams& kp = g->params;
  kp.func = func;
  kp.blockDim = dim3(width);
  kp.sharedMemBytes = (unsigned)lds;
  kp.extra = nullptr;
  hipGraphNode_t prev = nullptr;
  for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
    const int first = std::max(0, floordiv(q - n + 1, 2));
    const int last = std::min(iters - 1, q / 2);
    if (last < first) continue;
    kp.gridDim = dim3(bx, (unsigned)(last - first + 1));
    kp.kernelParams = args;
    hipGraphNode_t node = nullptr;
    HIP_CHECK(hipGraphAddKernelNode(&node, g->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
    g->nodes.push_back(node);
    g->first_z.push_back(n - q + 2 * first);
    g->slices.push_back(last - first + 1);
    g->first_t.push_back(first);
    prev = node;
  }
  HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  HIP_CHECK(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
  return g;
}

template <class T, class V>
static std::unique_ptr<TiledGraph> build_tiled_graph(int n, int iters, int width, unsigned bx, size_t lds) {
  TiledArgs<T, V> a{};
  WaveTiles wv{};
  void* args[2] = {&a, &wv};
  void* func = width == 64 ? tiled_kernel<T, V, 64>() : width == 128 ? tiled_kernel<T, V, 128>() : tiled_kernel<T, V, 256>();
  return build_wave_graph(n, iters, func, width, bx, lds, args);
}

// A cached tiled graph with this call's node arguments written, held until
// tiled_launch.
struct TiledLaunch {
  TiledGraph* g = nullptr;
  std::unique_lock<std::mutex> lk;
  int launches = 0;
};

// relax_sliced on the tile tables, in two steps so that a level's node
// arguments can be written on the host while the GPU still relaxes the level
// before: tiled_prepare writes them (x in schedule order holds the start
// values and receives the result; x1 is a second buffer of R values; b is the
// right-hand side by schedule position), tiled_launch copies the start values
// where the first visit reads them and launches.  The operator's absent
// slots must hold +0 (k_level_operator stores them so).
template <class T, class V>
static TiledLaunch tiled_prepare(const TileTable& tt, const RowRecord<T, V>* records, const T* b, T* x, T* x1, int R, int depth,
                                 int iters, hipStream_t s) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, unsigned, int, hipStream_t>, std::unique_ptr<TiledGraph>> cache;
  TiledLaunch t;
  const int n = 1 << depth;
  if (iters > RELAX_MAX_SLICES) throw PsrError(1, "tiled relaxation supports at most 16 sweeps");
  if (R <= 0 || tt.tiles <= 0) return t;
  const int width = tiled_width(tt);
  unsigned bx = 1;
  while (bx < (unsigned)std::max(1, tt.max_tiles)) bx <<= 1;
  const int gside = tt.ts + 8;  // even: the parity planes are (gside / 2)^2 cells
  const size_t lds = sizeof(T) * 3 * (size_t)gside * gside;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  TiledGraph* g = nullptr;
  {
    std::lock_guard<std::mutex> guard(mu);
    std::unique_ptr<TiledGraph>& slot = cache[std::make_tuple((int)sizeof(V), depth, iters, width, bx, dev, s)];
    if (!slot) slot = build_tiled_graph<T, V>(n, iters, width, bx, lds);
    g = slot.get();
  }
  t.lk = std::unique_lock<std::mutex>(g->mu);
  HIP_CHECK(hipEventSynchronize(g->done));
  TiledArgs<T, V> a{tt.entries.data(), tt.desc.data(), records, b, x, x1, gside};
  WaveTiles wv;
  void* args[2] = {&a, &wv};
  hipKernelNodeParams kp = g->params;
  kp.kernelParams = args;
  kp.sharedMemBytes = (unsigned)lds;
  for (size_t i = 0; i < g->nodes.size(); ++i) {
    for (int q = 0; q < RELAX_MAX_SLICES; ++q) wv.start[q] = wv.end[q] = 0;
    wv.flip = 0;
    for (int k = 0; k < g->slices[i]; ++k) {
      const int z = g->first_z[i] + 2 * k, q = g->first_t[i] + k;
      wv.start[k] = tt.slice_start[z];
      wv.end[k] = tt.slice_start[z + 1];
      if ((iters - q) % 2) wv.flip |= 1u << k;
    }
    kp.gridDim = dim3(bx, (unsigned)g->slices[i]);
    HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec, g->nodes[i], &kp));
  }
  t.g = g;
  t.launches = (int)g->nodes.size();
  return t;
}

template <class T> static int tiled_launch(TiledLaunch& t, const T* x, T* x1, int R, int iters, hipStream_t s) {
  if (!t.g) return 0;
  // the first visit reads buffer iters % 2
  if (iters % 2) HIP_CHECK(hipMemcpyAsync(x1, x, sizeof(T) * R, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipGraphLaunch(t.g->exec, s));
  HIP_CHECK(hipEventRecord(t.g->done, s));
  t.lk.unlock();
  t.g = nullptr;
  return t.launches;
}

template <class T, class V>
int relax_tiled_graph(const GsOperator<T, V>& op, const TileTable& tt, const RowRecord<T, V>* records, T* x, T* x1,
                      int depth, int iters, hipStream_t s) {
  TiledLaunch t = tiled_prepare<T, V>(tt, records, op.b.data(), x, x1, op.R, depth, iters, s);
  return tiled_launch<T>(t, x, x1, op.R, iters, s);
}

// ---------------------------------------------------------------- grouped relaxation
// Reconstructions of a batch run in groups; at each finer level every member
// offers its tiled relaxation and one of them launches a single graph whose
// wave launches cover all members' tiles (k_relax_tiled_group), so a group
// of m reconstructions issues the launches of one.  The levels run in
// lockstep: a level's graph starts once every member that will relax it has
// offered it.

// What a member offers at a level: its arguments, tile table and the stream
// (with an event) its inputs are ready on; valid is false for a level it
// relaxes on its own.
struct GroupOffer {
  bool valid = false;
  int value_bytes = 0, vector_bytes = 0, gside = 0, width = 0, iters = 0;
  const void* entries = nullptr;
  const void* desc = nullptr;
  const void* records = nullptr;
  const void* b = nullptr;
  void* x0 = nullptr;
  void* x1 = nullptr;
  const TileTable* tt = nullptr;
  hipEvent_t ready = nullptr;
};

struct RelaxGroup {
  explicit RelaxGroup(int n) : members(n), first(n, -1), arrived(n, -1), finished(n, 0) {}
  ~RelaxGroup() {
    for (auto& kv : levels) {
      for (const GroupOffer& o : kv.second.offers)
        if (o.ready) (void)hipEventDestroy(o.ready);
      for (hipEvent_t e : kv.second.events) (void)hipEventDestroy(e);
    }
  }
  struct Level {
    std::vector<GroupOffer> offers;  // by member
    std::vector<hipEvent_t> done;    // by member: the launch it waits for, or null
    std::vector<hipEvent_t> events;  // owned
    bool launching = false, launched = false;
    int launches = 0;
    std::string error;
  };
  // every member that will relax level L has offered it
  bool complete(int L) const {
    for (int m = 0; m < members; ++m)
      if (!finished[m] && arrived[m] < L && !(first[m] >= 0 && first[m] > L)) return false;
    return true;
  }
  std::mutex mu;
  std::condition_variable cv;
  const int members;
  std::vector<int> first;    // first finer level, -1 until declared
  std::vector<int> arrived;  // last level offered
  std::vector<char> finished;
  std::map<int, Level> levels;
};

static thread_local RelaxGroup* t_group = nullptr;
static thread_local int t_member = -1;

RelaxGroup* relax_group_create(int members) {
  if (members < 1 || members > RELAX_GROUP_MAX) throw PsrError(2, "a relaxation group holds 1 to 8 reconstructions");
  return new RelaxGroup(members);
}
void relax_group_destroy(RelaxGroup* g) { delete g; }
void relax_group_bind(RelaxGroup* g, int member) {
  t_group = g;
  t_member = member;
}

// The bound member relaxes no further level (its cascade ended or failed).
static void group_finish() {
  if (!t_group) return;
  {
    std::lock_guard<std::mutex> lk(t_group->mu);
    t_group->finished[t_member] = 1;
  }
  t_group->cv.notify_all();
}

void relax_group_release() {
  group_finish();
  t_group = nullptr;
  t_member = -1;
}

static void group_declare(int first_level) {
  if (!t_group) return;
  {
    std::lock_guard<std::mutex> lk(t_group->mu);
    t_group->first[t_member] = first_level;
  }
  t_group->cv.notify_all();
}

// One graph over the members ms (same precision, tile side and sweeps) on s;
// returns the event recorded after it and the launch count.
template <class T, class V>
static std::pair<hipEvent_t, int> launch_tiled_group(const std::vector<const GroupOffer*>& ms, int depth, hipStream_t s) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int, int, hipStream_t>, std::unique_ptr<TiledGraph>> cache;
  const int n = 1 << depth, M = (int)ms.size(), iters = ms[0]->iters, gside = ms[0]->gside;
  int width = 64;
  for (const GroupOffer* o : ms) width = std::max(width, o->width);
  const size_t lds = sizeof(T) * 3 * (size_t)gside * gside;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  TiledGraph* g = nullptr;
  {
    std::lock_guard<std::mutex> guard(mu);
    std::unique_ptr<TiledGraph>& slot = cache[std::make_tuple((int)sizeof(V), depth, iters, width, M, dev, s)];
    if (!slot) {
      TiledGroup<T, V> tg{};
      void* args[1] = {&tg};
      void* func = width == 64    ? tiled_group_kernel<T, V, 64>()
                   : width == 128 ? tiled_group_kernel<T, V, 128>()
                                  : tiled_group_kernel<T, V, 256>();
      slot = build_wave_graph(n, iters, func, width, 1, lds, args);
    }
    g = slot.get();
  }
  for (const GroupOffer* o : ms) HIP_CHECK(hipStreamWaitEvent(s, o->ready, 0));
  std::lock_guard<std::mutex> guard(g->mu);
  HIP_CHECK(hipEventSynchronize(g->done));
  TiledGroup<T, V> tg{};
  tg.members = M;
  for (int m = 0; m < M; ++m) {
    const GroupOffer& o = *ms[m];
    tg.a[m] = TiledArgs<T, V>{(const int2*)o.entries, (const int*)o.desc, (const RowRecord<T, V>*)o.records, (const T*)o.b,
                              (T*)o.x0, (T*)o.x1, gside};
  }
  void* args[1] = {&tg};
  hipKernelNodeParams kp = g->params;
  kp.kernelParams = args;
  kp.sharedMemBytes = (unsigned)lds;
  for (size_t i = 0; i < g->nodes.size(); ++i) {
    int total = 0;
    for (int m = 0; m < M; ++m) {
      const TileTable& tt = *ms[m]->tt;
      WaveTiles& wv = tg.wv[m];
      for (int t = 0; t < RELAX_MAX_SLICES; ++t) wv.start[t] = wv.end[t] = 0;
      wv.flip = 0;
      int blocks = 0;
      for (int k = 0; k < g->slices[i]; ++k) {
        const int z = g->first_z[i] + 2 * k, t = g->first_t[i] + k;
        wv.start[k] = tt.slice_start[z];
        wv.end[k] = tt.slice_start[z + 1];
        blocks = std::max(blocks, wv.end[k] - wv.start[k]);
        if ((iters - t) % 2) wv.flip |= 1u << k;
      }
      tg.first[m] = total;
      total += blocks;
    }
    tg.first[M] = total;
    kp.gridDim = dim3((unsigned)std::max(1, total), (unsigned)g->slices[i]);
    HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec, g->nodes[i], &kp));
  }
  HIP_CHECK(hipGraphLaunch(g->exec, s));
  HIP_CHECK(hipEventRecord(g->done, s));
  hipEvent_t ev = nullptr;
  HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(ev, s));
  return {ev, (int)g->nodes.size()};
}

// Launches level L for every offer, one graph per kind of offer, on s.
static void group_launch_level(RelaxGroup::Level& lv, int L, hipStream_t s) {
  std::vector<char> taken(lv.offers.size(), 0);
  for (size_t i = 0; i < lv.offers.size(); ++i) {
    const GroupOffer& oi = lv.offers[i];
    if (!oi.valid || taken[i]) continue;
    std::vector<const GroupOffer*> ms;
    std::vector<size_t> who;
    for (size_t j = i; j < lv.offers.size(); ++j) {
      const GroupOffer& oj = lv.offers[j];
      if (!oj.valid || taken[j] || oj.value_bytes != oi.value_bytes || oj.vector_bytes != oi.vector_bytes ||
          oj.gside != oi.gside || oj.iters != oi.iters)
        continue;
      taken[j] = 1;
      ms.push_back(&oj);
      who.push_back(j);
    }
    std::pair<hipEvent_t, int> r;
    if (oi.vector_bytes == 8 && oi.value_bytes == 4) r = launch_tiled_group<double, float>(ms, L, s);
    else if (oi.vector_bytes == 8) r = launch_tiled_group<double, double>(ms, L, s);
    else r = launch_tiled_group<float, float>(ms, L, s);
    lv.events.push_back(r.first);
    for (size_t j : who) lv.done[j] = r.first;
    lv.launches = r.second;
  }
}

// The bound member's offer at level L; returns the launch count when its
// relaxation ran in a group launch (s then waits for it), else 0.
static int group_relax(int L, GroupOffer offer, hipStream_t s) {
  RelaxGroup* g = t_group;
  const int me = t_member;
  if (offer.valid) {
    HIP_CHECK(hipEventCreateWithFlags(&offer.ready, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(offer.ready, s));
  }
  std::unique_lock<std::mutex> lk(g->mu);
  RelaxGroup::Level& lv = g->levels[L];
  if (lv.offers.empty()) {
    lv.offers.resize(g->members);
    lv.done.assign(g->members, nullptr);
  }
  lv.offers[me] = offer;
  g->arrived[me] = L;
  g->cv.notify_all();
  while (!lv.launched) {
    if (!lv.launching && g->complete(L)) {
      lv.launching = true;
      lk.unlock();
      std::string error;
      try {
        group_launch_level(lv, L, s);
      } catch (const std::exception& e) {
        error = e.what();
      }
      lk.lock();
      lv.error = error;
      lv.launched = true;
      g->cv.notify_all();
      break;
    }
    g->cv.wait(lk);
  }
  const hipEvent_t done = lv.done[me];
  const std::string error = lv.error;
  const int launches = lv.launches;
  lk.unlock();
  if (!error.empty()) throw PsrError(1, "grouped relaxation at depth " + std::to_string(L) + ": " + error);
  if (!done) return 0;
  HIP_CHECK(hipStreamWaitEvent(s, done, 0));
  return launches;
}

// Levels of at most 1,024 rows (one per thread): each thread loads its
// schedule position's 27 columns and values, b, 1/A_ii and x position once
// and keeps them in registers for every sweep, so a colour phase is only LDS
// gathers, the row's update and a barrier.  Same updates in the same order
// as k_relax_block_lds.
constexpr int RELAX_BLOCK_REGS_ROWS = 1024;

template <class T>
__global__ void __launch_bounds__(RELAX_BLOCK_REGS_ROWS) k_relax_block_regs(const int* __restrict__ order, const int* __restrict__ goff,
                                                                            const int* __restrict__ cols, const T* __restrict__ vals,
                                                                            const T* __restrict__ b, const T* __restrict__ inv_diag,
                                                                            const unsigned char* __restrict__ has_diag,
                                                                            T* __restrict__ x, int R, int iters, int forward) {
  extern __shared__ unsigned char lds_raw[];
  T* xs = (T*)lds_raw;
  if (R <= 0) return;
  const int j = (int)threadIdx.x;
  if (j < R) xs[j] = x[j];
  const int jj = j < R ? j : 0;  // clamped loads (R >= 1)
  int cc[27];
  T v[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    cc[k] = cols[(int64_t)k * R + jj];
    v[k] = vals[(int64_t)k * R + jj];
  }
  const T bj = b[jj], invj = inv_diag[jj];
  const int oj = order[jj];
  const bool has = j < R && has_diag[jj] != 0;
  int group = -1;
#pragma unroll
  for (int g = 0; g < 8; ++g) group = j >= goff[g] && j < goff[g + 1] ? g : group;
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < 8; ++c) {
      if (group == (forward ? c : 7 - c)) {
        T r = bj;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          const T t = r - xs[cc[k] >= 0 ? cc[k] : 0] * v[k];
          r = cc[k] >= 0 ? t : r;
        }
        if (has) xs[oj] = xs[oj] + r * invj;
      }
      __syncthreads();
    }
  }
  if (j < R) x[j] = xs[j];
}

// x of up to 8,192 double (16,384 float) rows stays in LDS (64 KiB)
constexpr size_t RELAX_BLOCK_LDS_BYTES = 65536;

template <class T>
void relax_colored(const GsOperator<T>& op, const Schedule<T>& sc, T* x, int iters, bool forward, hipStream_t s) {
  if (!op.R) return;
  const size_t lds = sizeof(T) * (size_t)op.R;
  if (op.update_order && op.R <= RELAX_BLOCK_REGS_ROWS) {
    hipLaunchKernelGGL(k_relax_block_regs<T>, dim3(1), dim3(RELAX_BLOCK_REGS_ROWS), lds, s, op.update_order,
                       (const int*)sc.offsets.data(), (const int*)op.cols.data(), (const T*)op.vals.data(),
                       (const T*)op.b.data(), (const T*)op.inv_diag.data(), (const unsigned char*)op.has_diag.data(), x,
                       op.R, iters, forward ? 1 : 0);
    KERNEL_CHECK();
    return;
  }
  if (op.update_order && lds <= RELAX_BLOCK_LDS_BYTES) {
    hipLaunchKernelGGL(k_relax_block_lds<T>, dim3(1), dim3(1024), lds, s, op.update_order, (const int*)sc.offsets.data(),
                       (const int*)op.cols.data(), (const T*)op.vals.data(), (const T*)op.b.data(),
                       (const T*)op.inv_diag.data(), (const unsigned char*)op.has_diag.data(), x, op.R, iters,
                       forward ? 1 : 0);
    KERNEL_CHECK();
    return;
  }
  hipLaunchKernelGGL(k_relax_block<T>, dim3(1), dim3(1024), 0, s, op.update_order, (const int*)sc.offsets.data(),
                     (const int*)op.cols.data(), (const T*)op.vals.data(), (const T*)op.b.data(), (const T*)op.inv_diag.data(),
                     (const unsigned char*)op.has_diag.data(), x, op.R, iters, forward ? 1 : 0);
  KERNEL_CHECK();
}

// row_order (optional, device): output position j holds row row_order[j]
template <class T> static Stencil<double> assemble_stiffness(const LevelState<T>& rowsL, const LevelState<T>& colsL, int depth,
                                                             int trial_depth, hipStream_t s, const int* row_order = nullptr) {
  Stencil<double> m;
  const int R = rowsL.basis_count;
  m.rows = R;
  m.cols_count = colsL.basis_count;
  m.cols.reset(27 * (size_t)R, s);
  m.vals.reset(27 * (size_t)R, s);
  const int res = 1 << depth;
  DArray<D2> table((res + 1) * 3, s);
  hipLaunchKernelGGL(k_hat_table, dim3(blocks_for((res + 1) * 3, 256)), dim3(256), 0, s, depth, trial_depth, table.data());
  KERNEL_CHECK();
  if (R) {
    hipLaunchKernelGGL(k_stiffness, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int64_t*)rowsL.basis_keys.data(), R,
                       row_order, (const int64_t*)colsL.basis_keys.data(), (int64_t)colsL.basis_count, colsL.columns.view(),
                       colsL.dense_dim ? (const int*)colsL.dense.data() : (const int*)nullptr, colsL.dense_dim, depth, trial_depth,
                       (const D2*)table.data(), m.cols.data(), m.vals.data());
    KERNEL_CHECK();
  }
  return m;
}

template <class T> struct PointSet {
  int P = 0;
  DArray<T> pos, w;
  DArray<uint64_t> bins;
  DArray<int> ids;
  ColumnIndex columns;  // (bx, by) runs of the sorted bins
  DArray<int> cell_start, cell_end;  // dense z-major cell ranges (depth <= 9)
};

// join: order s after the serial normal average (the weights use it); a
// caller on another stream orders that stream itself.
template <class T> static PointSet<T> point_constraints(Result<T>& r, int d, hipStream_t s, bool join = true) {
  PointSet<T> ps;
  if (join) r.join_average();
  LevelState<T>& L = r.levels[d];
  ps.P = (int)L.agg_keys.size();
  ps.pos.reset(3 * (size_t)ps.P, s);
  ps.w.reset(ps.P, s);
  ps.bins.reset(ps.P, s);
  ps.ids.reset(ps.P, s);
  if (ps.P) {
    hipLaunchKernelGGL(k_points<T>, dim3(blocks_for(ps.P, 256)), dim3(256), 0, s, (const int64_t*)L.agg_keys.data(),
                       (const T*)L.agg_vals.data(), ps.P, (const int64_t*)L.tree_keys.data(), (int64_t)L.tree_keys.size(),
                       (const T*)r.normal_average.data(), T(r.prm.point_weight), d, ps.pos.data(), ps.w.data(),
                       ps.bins.data(), ps.ids.data());
    KERNEL_CHECK();
    sort_pairs_u64(ps.bins.data(), ps.ids.data(), ps.P, 3 * d, s);
    const int n = 1 << d;
    const int64_t cells = (int64_t)n * n * n;
    if (d >= 5 && cells <= dense_index_max_cells()) {
      ps.cell_start.reset((size_t)cells, s);
      ps.cell_end.reset((size_t)cells, s);
      ps.cell_start.zero();
      ps.cell_end.zero();
      hipLaunchKernelGGL(k_bin_cells, dim3(blocks_for(ps.P, 256)), dim3(256), 0, s, (const uint64_t*)ps.bins.data(), ps.P, d,
                         ps.cell_start.data(), ps.cell_end.data());
      KERNEL_CHECK();
    } else if (n <= column_table_max_dim()) {
      ps.columns.start.reset((size_t)n * n, s);
      ps.columns.end.reset((size_t)n * n, s);
      ps.columns.start.zero();
      ps.columns.end.zero();
      hipLaunchKernelGGL(k_bin_columns, dim3(blocks_for(ps.P, 256)), dim3(256), 0, s, (const uint64_t*)ps.bins.data(), ps.P, d,
                         ps.columns.start.data(), ps.columns.end.data());
      KERNEL_CHECK();
      ps.columns.dim = n;
    }
  }
  return ps;
}

template <class T> static PointBins<T> point_bins(const PointSet<T>& ps) {
  return PointBins<T>{(const uint64_t*)ps.bins.data(), (const int*)ps.ids.data(), ps.P, ps.columns.view(),
                      ps.cell_start.empty() ? (const int*)nullptr : (const int*)ps.cell_start.data(),
                      ps.cell_end.empty() ? (const int*)nullptr : (const int*)ps.cell_end.data(), (const T*)ps.pos.data(),
                      (const T*)ps.w.data()};
}

template <class T>
static void add_screening(Stencil<double>& m, const DArray<int64_t>& rows, const PointSet<T>& ps, int depth, int trial_depth,
                          double* load, hipStream_t s, const int* row_order = nullptr) {
  if (!m.rows) return;
  hipLaunchKernelGGL(k_screening<T>, dim3(blocks_for(m.rows, 128)), dim3(128), 0, s, (const int64_t*)rows.data(), m.rows,
                     row_order, depth, trial_depth, point_bins(ps), (const int*)m.cols.data(), m.vals.data(), load);
  KERNEL_CHECK();
}

template <class T> static Stencil<T> cast_stencil(Stencil<double>&& m, hipStream_t s) {
  Stencil<T> out;
  out.rows = m.rows;
  out.cols_count = m.cols_count;
  out.cols = std::move(m.cols);
  out.vals.reset(m.vals.size(), s);
  if (m.vals.size()) {
    hipLaunchKernelGGL(k_cast<T>, dim3(blocks_for(m.vals.size(), 256)), dim3(256), 0, s, (const double*)m.vals.data(),
                       out.vals.data(), (int64_t)m.vals.size());
    KERNEL_CHECK();
  }
  return out;
}

// A sliced finer level's operator (colsL = rowsL, trial_depth = depth; op and
// present given) or cross operator (colsL one level coarser) in one pass of
// k_level_operator, rows in schedule order ord (inv = its inverse) or, with ord
// null, in canonical order.
template <class T, class V>
static void level_operator(const LevelState<T>& rowsL, const LevelState<T>& colsL, int depth, int trial_depth,
                           const PointSet<T>* pts, double* load, const int* ord, const int* inv, Stencil<V>& out,
                           GsOperator<T, V>* op, RowRecord<T, V>* records, unsigned long long* present, hipStream_t s) {
  const int R = rowsL.basis_count;
  out.rows = R;
  out.cols_count = colsL.basis_count;
  out.cols.reset(27 * (size_t)R, s);
  out.vals.reset(27 * (size_t)R, s);
  if (op) {
    if (inv) op->cols.reset(27 * (size_t)R, s);
    op->inv_diag.reset(R, s);
    op->has_diag.reset(R, s);
  }
  if (!R) return;
  const int res = 1 << depth;
  DArray<D2> table((res + 1) * 3, s);
  hipLaunchKernelGGL(k_hat_table, dim3(blocks_for((res + 1) * 3, 256)), dim3(256), 0, s, depth, trial_depth, table.data());
  KERNEL_CHECK();
  const PointBins<T> pb = pts ? point_bins(*pts) : PointBins<T>{};
  const int* dense = colsL.dense_dim ? (const int*)colsL.dense.data() : (const int*)nullptr;
  int* sched = op && inv ? op->cols.data() : nullptr;
  T* inv_diag = op ? op->inv_diag.data() : nullptr;
  unsigned char* has_diag = op ? op->has_diag.data() : nullptr;
  if (pts)
    hipLaunchKernelGGL((k_level_operator<T, V, true>), dim3(blocks_for(R, 256)), dim3(256), 0, s,
                       (const int64_t*)rowsL.basis_keys.data(), R, ord, (const int64_t*)colsL.basis_keys.data(),
                       (int64_t)colsL.basis_count, colsL.columns.view(), dense, colsL.dense_dim, depth, trial_depth,
                       (const D2*)table.data(), pb, load, out.cols.data(), out.vals.data(), inv, sched, inv_diag, has_diag,
                       records, present);
  else
    hipLaunchKernelGGL((k_level_operator<T, V, false>), dim3(blocks_for(R, 256)), dim3(256), 0, s,
                       (const int64_t*)rowsL.basis_keys.data(), R, ord, (const int64_t*)colsL.basis_keys.data(),
                       (int64_t)colsL.basis_count, colsL.columns.view(), dense, colsL.dense_dim, depth, trial_depth,
                       (const D2*)table.data(), pb, load, out.cols.data(), out.vals.data(), inv, sched, inv_diag, has_diag,
                       records, present);
  KERNEL_CHECK();
}

template <class T>
static void normal_stencil_load(const LevelState<T>& testL, int test_depth, const int64_t* nkeys, int64_t ncount, ColumnTable ct,
                                DenseIndex dn, const T* nvals, int normal_depth, double* out, hipStream_t s) {
  const int R = testL.basis_count;
  if (!R) return;
  if (!ncount) {
    HIP_CHECK(hipMemsetAsync(out, 0, sizeof(double) * R, s));
    return;
  }
  const int width = test_depth < normal_depth ? 6 : 4;
  const int first = test_depth < normal_depth ? -3 : -2;
  const int rows = (1 << test_depth) + 1;
  DArray<D2> table((size_t)rows * width, s);
  hipLaunchKernelGGL(k_normal_table, dim3(blocks_for(rows * width, 256)), dim3(256), 0, s, test_depth, normal_depth, first, width,
                     table.data());
  KERNEL_CHECK();
  if (dn.idx && width == 4)
    hipLaunchKernelGGL((k_normal_load_dense<T, 4>), dim3(blocks_for(R, 128)), dim3(128), 0, s,
                       (const int64_t*)testL.basis_keys.data(), R, dn, nvals, 3, test_depth, normal_depth, first,
                       (const D2*)table.data(), out);
  else if (dn.idx && width == 6)
    hipLaunchKernelGGL((k_normal_load_dense<T, 6>), dim3(blocks_for(R, 128)), dim3(128), 0, s,
                       (const int64_t*)testL.basis_keys.data(), R, dn, nvals, 3, test_depth, normal_depth, first,
                       (const D2*)table.data(), out);
  else
    hipLaunchKernelGGL(k_normal_load<T>, dim3(blocks_for(R, 128)), dim3(128), 0, s, (const int64_t*)testL.basis_keys.data(), R,
                       nkeys, ncount, ct, dn, nvals, 3, test_depth, normal_depth, first, width, (const D2*)table.data(), out);
  KERNEL_CHECK();
}

// ---------------------------------------------------------------- ordered normal loads
// ordered_normal_loads.py:263-337: float32 runs with base depth 0 sum their
// normal loads in T in a pinned serial order (normal_loads.py:84-92).  One
// thread per row, each following the reference's loop order.

// ordered_normal_loads.py:86-94 (normal_stencil) and :137-152 (contribution)
template <class T>
__device__ T ordered_contribution(int td, int tx, int ty, int tz, int nd, int nx, int ny, int nz, const int64_t* __restrict__ keys,
                                  int64_t count, const T* __restrict__ vals) {
  T dot = T(0);
  const int lim = 1 << nd;
  if (imin(imin(nx, ny), nz) >= 0 && imax(imax(nx, ny), nz) < lim) {
    const int64_t j = find_key(keys, count, basis_key(nd, nx, ny, nz));
    if (j >= 0) {
      const D2 x = normal_integral(td, tx, nd, nx), y = normal_integral(td, ty, nd, ny), z = normal_integral(td, tz, nd, nz);
      const double st[3] = {x.v1 * y.v0 * z.v0, x.v0 * y.v1 * z.v0, x.v0 * y.v0 * z.v1};
#pragma unroll
      for (int a = 0; a < 3; ++a) dot = dot + vals[3 * j + a] * T(st[a]);
    }
  }
  return dot;
}

// :156-171 (same): finer + the 64 same-level records, x-major offsets -2..1
template <class T>
__global__ void k_ol_same(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ nkeys, int64_t ncount,
                          const T* __restrict__ nvals, const T* __restrict__ finer, int depth, T* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  T value = finer[i];
  for (int k = 0; k < 64; ++k)
    value = value + ordered_contribution<T>(depth, cx, cy, cz, depth, cx + k / 16 - 2, cy + (k / 4) % 4 - 2, cz + k % 4 - 2, nkeys,
                                            ncount, nvals);
  out[i] = value;
}

// :97-103 (morton_xy)
__device__ __forceinline__ uint64_t morton_xy(int x, int y) {
  uint64_t key = 0;
  for (int bit = 0; bit < 15; ++bit) {
    key |= (uint64_t)((x >> bit) & 1) << (2 * bit);
    key |= (uint64_t)((y >> bit) & 1) << (2 * bit + 1);
  }
  return key;
}

// :105-122 (parent_order, segmented sort) and :173-188 (parent): the 216
// child-level records of a coarse row, each z layer of 36 in (morton_xy,
// offset index) order; offsets with a negative coordinate sort last
template <class T>
__global__ void k_ol_parent(const int64_t* __restrict__ keys, int C, const int64_t* __restrict__ nkeys, int64_t ncount,
                            const T* __restrict__ nvals, int depth, T* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const int64_t key = keys[i];
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  T value = T(0);
  for (int zl = 0; zl < 6; ++zl) {
    uint64_t order[36];
    for (int m = 0; m < 36; ++m) {
      const int k = 36 * zl + m;
      const int x = 2 * cx + k % 6 - 3, y = 2 * cy + (k / 6) % 6 - 3, z = 2 * cz + k / 36 - 3;
      const uint64_t o = imin(imin(x, y), z) >= 0 ? morton_xy(x, y) : (uint64_t)2147483647;
      order[m] = (o << 8) | (uint64_t)m;
    }
    for (int a = 1; a < 36; ++a) {  // insertion sort: stable by construction of the keys
      const uint64_t v = order[a];
      int b = a - 1;
      while (b >= 0 && order[b] > v) {
        order[b + 1] = order[b];
        --b;
      }
      order[b + 1] = v;
    }
    for (int m = 0; m < 36; ++m) {
      const int k = 36 * zl + (int)(order[m] & 255);
      value = value + ordered_contribution<T>(depth, cx, cy, cz, depth + 1, 2 * cx + k % 6 - 3, 2 * cy + (k / 6) % 6 - 3,
                                              2 * cz + k / 36 - 3, nkeys, ncount, nvals);
    }
  }
  out[i] = value;
}

// :190-212 (restrict): out += the 27 finer values with weights 1/2 per
// nonzero offset axis
template <class T>
__global__ void k_ol_restrict(const int64_t* __restrict__ fine_keys, int64_t F, const T* __restrict__ fine,
                              const int64_t* __restrict__ coarse, int C, int depth, T* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const int64_t key = coarse[i];
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  T value = out[i];
  for (int k = 0; k < 27; ++k) {
    const int ox = k / 9 - 1, oy = (k / 3) % 3 - 1, oz = k % 3 - 1;
    const int nx = 2 * cx + ox, ny = 2 * cy + oy, nz = 2 * cz + oz;
    if (imin(imin(nx, ny), nz) >= 0 && imax(imax(nx, ny), nz) <= (1 << depth)) {
      const int64_t j = find_key(fine_keys, F, basis_key(depth, nx, ny, nz));
      if (j >= 0) {
        T weight = T(1);
        if (ox != 0) weight = weight * T(0.5);
        if (oy != 0) weight = weight * T(0.5);
        if (oz != 0) weight = weight * T(0.5);
        value = value + fine[j] * weight;
      }
    }
  }
  out[i] = value;
}

// :214-238 (prolong): own records plus the 8 coarser cumulative values
// (corner bits x, y, z = 4, 2, 1)
template <class T>
__global__ void k_ol_prolong(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ own_keys, int64_t own_count,
                             const T* __restrict__ own, const int64_t* __restrict__ prev_keys, int64_t prev_count,
                             const T* __restrict__ prev, int depth, T* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int c[3] = {key_x(key), key_y(key), key_z(key)};
  T v[3] = {T(0), T(0), T(0)};
  const int64_t j0 = find_key(own_keys, own_count, key);
  if (j0 >= 0)
    for (int a = 0; a < 3; ++a) v[a] = own[3 * j0 + a];
  if (depth > 0) {
    for (int k = 0; k < 8; ++k) {
      const int p[3] = {c[0] / 2 + (k >> 2), c[1] / 2 + ((k >> 1) & 1), c[2] / 2 + (k & 1)};
      T weight = T(1);
      for (int a = 0; a < 3; ++a) weight = weight * (T(1) - fabs(T(c[a]) * T(0.5) - T(p[a])));
      if (weight > T(0)) {
        const int64_t j = find_key(prev_keys, prev_count, basis_key(depth - 1, p[0], p[1], p[2]));
        if (j >= 0)
          for (int a = 0; a < 3; ++a) v[a] = v[a] + prev[3 * j + a] * weight;
      }
    }
  }
  for (int a = 0; a < 3; ++a) out[3 * (int64_t)i + a] = v[a];
}

// :240-258 (from_coarse): loads += the 64 coarser cumulative values around
// the parent, x-major offsets -2..1
template <class T>
__global__ void k_ol_from_coarse(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ ckeys, int64_t ccount,
                                 const T* __restrict__ cum, int depth, T* __restrict__ loads) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int64_t key = keys[i];
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  T value = T(0);
  for (int k = 0; k < 64; ++k)
    value = value + ordered_contribution<T>(depth, cx, cy, cz, depth - 1, cx / 2 + k / 16 - 2, cy / 2 + (k / 4) % 4 - 2,
                                            cz / 2 + k % 4 - 2, ckeys, ccount, cum);
  loads[i] = loads[i] + value;
}

template <class T>
__global__ void k_widen(const T* __restrict__ in, double* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = double(in[i]);
}

template <class T> static void ordered_normal_loads(Result<T>& r, hipStream_t s) {
  const int D = r.prm.depth;
  auto own_keys = [&](int d) { return (const int64_t*)r.normal_keys.data() + r.normal_offsets[d]; };
  auto own_count = [&](int d) { return (int64_t)(r.normal_offsets[d + 1] - r.normal_offsets[d]); };
  auto own_vals = [&](int d) { return (const T*)r.normal_vals.data() + 3 * (int64_t)r.normal_offsets[d]; };
  std::vector<DArray<T>> loads(D + 1);
  DArray<T> finer(r.levels[D].basis_count, s);
  finer.zero();
  for (int d = D; d >= 0; --d) {
    LevelState<T>& L = r.levels[d];
    const int R = L.basis_count;
    loads[d].reset(R, s);
    if (R) {
      hipLaunchKernelGGL(k_ol_same<T>, dim3(blocks_for(R, 128)), dim3(128), 0, s, (const int64_t*)L.basis_keys.data(), R,
                         own_keys(d), own_count(d), own_vals(d), (const T*)finer.data(), d, loads[d].data());
      KERNEL_CHECK();
    }
    if (d > 0) {
      LevelState<T>& Cl = r.levels[d - 1];
      const int C = Cl.basis_count;
      DArray<T> next(C, s);
      next.zero();
      if (C && own_count(d)) {
        hipLaunchKernelGGL(k_ol_parent<T>, dim3(blocks_for(C, 64)), dim3(64), 0, s, (const int64_t*)Cl.basis_keys.data(), C,
                           own_keys(d), own_count(d), own_vals(d), d - 1, next.data());
        KERNEL_CHECK();
      }
      if (C && d < D) {
        hipLaunchKernelGGL(k_ol_restrict<T>, dim3(blocks_for(C, 128)), dim3(128), 0, s, (const int64_t*)L.basis_keys.data(),
                           (int64_t)R, (const T*)finer.data(), (const int64_t*)Cl.basis_keys.data(), C, d, next.data());
        KERNEL_CHECK();
      }
      finer = std::move(next);
    }
  }
  DArray<T> previous;
  for (int d = 0; d < D; ++d) {
    LevelState<T>& L = r.levels[d];
    const int R = L.basis_count;
    DArray<T> cumulative(3 * (size_t)R, s);
    if (R) {
      const LevelState<T>* P = d > 0 ? &r.levels[d - 1] : nullptr;
      hipLaunchKernelGGL(k_ol_prolong<T>, dim3(blocks_for(R, 128)), dim3(128), 0, s, (const int64_t*)L.basis_keys.data(), R,
                         own_keys(d), own_count(d), own_vals(d), P ? (const int64_t*)P->basis_keys.data() : (const int64_t*)nullptr,
                         (int64_t)(P ? P->basis_count : 0), (const T*)previous.data(), d, cumulative.data());
      KERNEL_CHECK();
    }
    const int F = r.levels[d + 1].basis_count;
    if (F) {
      hipLaunchKernelGGL(k_ol_from_coarse<T>, dim3(blocks_for(F, 128)), dim3(128), 0, s,
                         (const int64_t*)r.levels[d + 1].basis_keys.data(), F, (const int64_t*)L.basis_keys.data(), (int64_t)R,
                         (const T*)cumulative.data(), d + 1, loads[d + 1].data());
      KERNEL_CHECK();
    }
    previous = std::move(cumulative);
  }
  for (int d = 0; d <= D; ++d) {
    LevelState<T>& L = r.levels[d];
    L.normal_load.reset(L.basis_count, s);
    if (L.basis_count) {
      hipLaunchKernelGGL(k_widen<T>, dim3(blocks_for(L.basis_count, 256)), dim3(256), 0, s, (const T*)loads[d].data(),
                         L.normal_load.data(), (int64_t)L.basis_count);
      KERNEL_CHECK();
    }
  }
}

// normal_loads.py:66-190 (stencil path), split by what the base solve needs.
// The base level's load takes the from-child loads of every finer level,
// restricted down level by level: that part runs first, on the solve stream
// (normal_loads_base).  Each finer level's same-level and from-parent loads,
// and the prolonged cumulative normal field behind the latter, feed only that
// level's right-hand side: normal_loads_level computes them level by level,
// on the assembly stream ahead of the level's operators.  Every array is
// summed in the order of a single pass over the levels.  Float32 runs with
// base depth 0 take the ordered path (normal_loads.py:84-92) in one go.
template <class T> struct NormalLoadPlan {
  bool active = false;  // false: every level's load is already set
  int base = 0, depth = 0;
  std::vector<ColumnIndex> records;         // per level, the normal records' column index
  std::vector<DArray<int>> rdense;          // and their dense (x, y, z)-major index
  std::vector<DenseIndex> rdn;
  std::vector<DArray<double>> finer;        // loads restricted to level d from the finer levels
  std::vector<DArray<double>> from_parent;  // from-parent loads, filled by the chain
  DArray<T> previous;                       // the chain's last cumulative coefficients
  int chain = 0;                            // next level of the chain
  bool has_cumulative = false;
  int next_level = 0;                       // next level whose load is still to be set
  std::vector<DArray<double>> from_child;   // normal_loads_base_levels' per-level loads
  DArray<double> same_base;
  bool levels_done = false;
  hipEvent_t ready = nullptr;               // loads stream, after normal_loads_base
  NormalLoadPlan() = default;
  NormalLoadPlan(const NormalLoadPlan&) = delete;
  NormalLoadPlan& operator=(const NormalLoadPlan&) = delete;
  ~NormalLoadPlan() {
    if (ready) (void)hipEventDestroy(ready);
  }
};

// The per-level part of normal_loads_base: each level's records indexed and
// the loads of the tests one level coarser (from_child) and of the base level
// (same level), which need only the tree; s may run it before the
// restriction gathers exist.
template <class T> static void normal_loads_base_levels(Result<T>& r, NormalLoadPlan<T>& plan, hipStream_t s) {
  const int D = r.prm.depth, base = r.prm.base_depth;
  plan.base = base;
  plan.depth = D;
  if (std::is_same<T, float>::value && base == 0) return;  // the ordered loads (normal_loads_base)
  plan.active = true;
  plan.records.resize(D + 1);
  plan.rdense.resize(D + 1);
  plan.rdn.assign(D + 1, DenseIndex{});
  plan.finer.resize(D + 1);
  plan.from_parent.resize(D + 1);
  plan.from_child.resize(D + 1);
  for (int d = 0; d <= D; ++d) {
    const int64_t r0 = r.normal_offsets[d], r1 = r.normal_offsets[d + 1];
    const int64_t* nk = r.normal_keys.data() + r0;
    const T* nv = r.normal_vals.data() + 3 * r0;
    build_column_index(nk, r1 - r0, (1 << d) + 1, plan.records[d], s);
    // dense (x, y, z)-major index of the level's records for the tests taken
    // in canonical order (consecutive z)
    const int64_t rdim = (1 << d) + 1;
    if (r1 > r0 && d >= 5 && rdim * rdim * rdim <= dense_index_max_cells()) {
      plan.rdense[d].reset((size_t)(rdim * rdim * rdim), s);
      HIP_CHECK(hipMemsetAsync(plan.rdense[d].data(), 0xff, sizeof(int) * (size_t)(rdim * rdim * rdim), s));
      hipLaunchKernelGGL(k_dense_fill_xyz, dim3(blocks_for(r1 - r0, 256)), dim3(256), 0, s, nk, r1 - r0, (int)rdim,
                         plan.rdense[d].data());
      KERNEL_CHECK();
      plan.rdn[d].idx = plan.rdense[d].data();
      plan.rdn[d].dim = (int)rdim;
    }
    if (d == base) {
      plan.same_base.reset(r.levels[d].basis_count, s);
      normal_stencil_load(r.levels[d], d, nk, r1 - r0, plan.records[d].view(), plan.rdn[d], nv, d, plan.same_base.data(), s);
    }
    if (d > base) {
      plan.from_child[d - 1].reset(r.levels[d - 1].basis_count, s);
      normal_stencil_load(r.levels[d - 1], d - 1, nk, r1 - r0, plan.records[d].view(), plan.rdn[d], nv, d,
                          plan.from_child[d - 1].data(), s);
    }
  }
  plan.levels_done = true;
}

// The loads the base level needs: normal_loads_base_levels (unless it ran
// already), then every finer level's loads restricted down to the base.
template <class T>
static void normal_loads_base(Result<T>& r, std::vector<DArray<int>>& restr, NormalLoadPlan<T>& plan, hipStream_t s) {
  const int D = r.prm.depth, base = r.prm.base_depth;
  plan.base = base;
  plan.depth = D;
  if (std::is_same<T, float>::value && base == 0) {
    ordered_normal_loads(r, s);
    plan.active = false;
    return;
  }
  if (!plan.levels_done) normal_loads_base_levels(r, plan, s);
  std::vector<DArray<double>>& from_child = plan.from_child;
  DArray<double> same_base = std::move(plan.same_base);
  DArray<double> finer(r.levels[D].basis_count, s);
  finer.zero();
  for (int d = D; d > base; --d) {
    DArray<double> next = std::move(from_child[d - 1]);
    const int C = r.levels[d - 1].basis_count;
    if (C) {
      hipLaunchKernelGGL(k_restrict_apply<double>, dim3(blocks_for(C, 256)), dim3(256), 0, s, (const int*)restr[d - 1].data(),
                         C, (const double*)finer.data(), next.data(), 1);
      KERNEL_CHECK();
    }
    plan.finer[d] = std::move(finer);
    finer = std::move(next);
  }
  LevelState<T>& B = r.levels[base];
  if (B.basis_count) {  // no from-parent load reaches the base level
    hipLaunchKernelGGL(k_add_f64, dim3(blocks_for(B.basis_count, 256)), dim3(256), 0, s, same_base.data(),
                       (const double*)finer.data(), B.basis_count);
    KERNEL_CHECK();
  }
  B.normal_load = std::move(same_base);
  plan.finer[base] = std::move(finer);
  plan.next_level = base + 1;
  HIP_CHECK(hipEventCreateWithFlags(&plan.ready, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(plan.ready, s));
}

// Level d's load (d = plan.next_level): the chain through level d - 1, then
// same-level + restricted finer + from-parent, in that order.  Runs on s,
// which must be ordered after plan.ready.
template <class T> static void normal_loads_level(Result<T>& r, NormalLoadPlan<T>& plan, int d, hipStream_t s) {
  if (!plan.active || d != plan.next_level || d > plan.depth) return;
  while (plan.chain <= d - 1) {
    const int c = plan.chain;
    const int64_t r0 = r.normal_offsets[c], r1 = r.normal_offsets[c + 1];
    const int64_t* nk = r.normal_keys.data() + r0;
    const T* nv = r.normal_vals.data() + 3 * r0;
    plan.has_cumulative = plan.has_cumulative || (r1 > r0);
    LevelState<T>& L = r.levels[c];
    DArray<T> cumulative(3 * (size_t)L.basis_count, s);
    if (L.basis_count) {
      const LevelState<T>* P = c > 0 ? &r.levels[c - 1] : nullptr;
      hipLaunchKernelGGL(k_normal_prolong<T>, dim3(blocks_for(L.basis_count, 256)), dim3(256), 0, s,
                         (const int64_t*)L.basis_keys.data(), L.basis_count, nk, r1 - r0, plan.records[c].view(), nv,
                         P ? (const int64_t*)P->basis_keys.data() : (const int64_t*)nullptr, (int64_t)(P ? P->basis_count : 0),
                         P ? P->columns.view() : ColumnTable{}, (const T*)plan.previous.data(), c, cumulative.data());
      KERNEL_CHECK();
    }
    if (c >= plan.base && plan.has_cumulative) {
      plan.from_parent[c + 1].reset(r.levels[c + 1].basis_count, s);
      DenseIndex bdn;
      if (L.dense_dim) {
        bdn.idx = L.dense.data();
        bdn.dim = L.dense_dim;
        bdn.zmajor = 1;
      }
      normal_stencil_load(r.levels[c + 1], c + 1, (const int64_t*)L.basis_keys.data(), (int64_t)L.basis_count,
                          L.columns.view(), bdn, (const T*)cumulative.data(), c, plan.from_parent[c + 1].data(), s);
    }
    plan.previous = std::move(cumulative);
    ++plan.chain;
  }
  const int64_t r0 = r.normal_offsets[d], r1 = r.normal_offsets[d + 1];
  LevelState<T>& L = r.levels[d];
  const int R = L.basis_count;
  DArray<double> load(R, s);
  normal_stencil_load(L, d, r.normal_keys.data() + r0, r1 - r0, plan.records[d].view(), plan.rdn[d],
                      r.normal_vals.data() + 3 * r0, d, load.data(), s);
  if (R) {
    hipLaunchKernelGGL(k_add_f64, dim3(blocks_for(R, 256)), dim3(256), 0, s, load.data(), (const double*)plan.finer[d].data(), R);
    KERNEL_CHECK();
    if (!plan.from_parent[d].empty()) {
      hipLaunchKernelGGL(k_add_f64, dim3(blocks_for(R, 256)), dim3(256), 0, s, load.data(),
                         (const double*)plan.from_parent[d].data(), R);
      KERNEL_CHECK();
    }
  }
  L.normal_load = std::move(load);
  ++plan.next_level;
}

// Every level's load not set yet, on s (ordered after plan.ready).
template <class T> static void normal_loads_rest(Result<T>& r, NormalLoadPlan<T>& plan, hipStream_t s) {
  while (plan.active && plan.next_level <= plan.depth) normal_loads_level(r, plan, plan.next_level, s);
}

// solver.py:241-317 (regular_multigrid)
// rhs_ready (optional) runs on the host once the levels' operators and
// schedules are enqueued, before the base right-hand side is read: the
// caller makes it ready there (it may depend on work of another stream).
template <class T>
static void regular_multigrid(Result<T>& r, std::vector<DArray<int>>& restr, hipStream_t s,
                              const std::function<void()>& rhs_ready = {}) {
  const int base = r.prm.base_depth;
  LevelState<T>& B = r.levels[base];
  std::vector<DArray<int>> mcols(base + 1);
  std::vector<DArray<T>> mvals(base + 1);
  std::vector<int> Rn(base + 1);
  for (int d = 0; d <= base; ++d) Rn[d] = r.levels[d].basis_count;
  mcols[base].reset(27 * (size_t)Rn[base], s);
  mvals[base].reset(27 * (size_t)Rn[base], s);
  if (Rn[base]) {
    HIP_CHECK(hipMemcpyAsync(mcols[base].data(), B.A.cols.data(), sizeof(int) * 27 * Rn[base], hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(mvals[base].data(), B.A.vals.data(), sizeof(T) * 27 * Rn[base], hipMemcpyDeviceToDevice, s));
  }
  for (int d = base; d > 0; --d) {
    const int C = Rn[d - 1], F = Rn[d];
    mcols[d - 1].reset(27 * (size_t)C, s);
    mvals[d - 1].reset(27 * (size_t)C, s);
    if (C) {
      hipLaunchKernelGGL(k_galerkin<T>, dim3(C), dim3(128), 0, s, (const int*)restr[d - 1].data(), C,
                         (const int*)mcols[d].data(), (const T*)mvals[d].data(), F, (const int*)r.levels[d].P.cols.data(),
                         (const T*)r.levels[d].P.vals.data(), mcols[d - 1].data(), mvals[d - 1].data(),
                         (const int64_t*)r.levels[d - 1].basis_keys.data());
      KERNEL_CHECK();
    }
  }
  std::vector<DArray<T>> rhs(base + 1), x(base + 1);
  for (int d = 0; d <= base; ++d) {
    rhs[d].reset(Rn[d], s);
    rhs[d].zero();
    x[d].reset(Rn[d], s);
    x[d].zero();
  }
  std::vector<Schedule<T>> sched;
  std::vector<GsOperator<T>> ops;
  for (int d = 0; d <= base; ++d) {  // (the gathered b is replaced before each relaxation)
    sched.push_back(make_schedule<T>(r.levels[d].basis_keys, Rn[d], d, false, s, false));
    ops.push_back(gather_operator<T>(sched[d], mcols[d], mvals[d], rhs[d].data(), Rn[d], s));
  }
  if (rhs_ready) rhs_ready();
  if (Rn[base]) HIP_CHECK(hipMemcpyAsync(rhs[base].data(), B.rhs.data(), sizeof(T) * Rn[base], hipMemcpyDeviceToDevice, s));
  const int dc = (base == 0 && r.prm.point_weight == 0.0) ? 1 : 0;
  DArray<T> pad_x(64, s), work(256, s), residual;
  pad_x.zero();
  r.cg_ran.reset((size_t)std::max(0, r.prm.base_v_cycles), s);
  const int iters = r.prm.iters;
  for (int cycle = 0; cycle < r.prm.base_v_cycles; ++cycle) {
    for (int d = base; d > 0; --d) {
      refresh_rhs(ops[d], sched[d], rhs[d].data(), s);
      relax_colored(ops[d], sched[d], x[d].data(), iters, true, s);
      residual.reset(Rn[d], s);
      hipLaunchKernelGGL(k_spmv<T>, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, (const int*)mcols[d].data(),
                         (const T*)mvals[d].data(), Rn[d], 27, (const T*)x[d].data(), residual.data(), T(-1), T(1),
                         (const T*)rhs[d].data());
      KERNEL_CHECK();
      if (Rn[d - 1]) {
        hipLaunchKernelGGL(k_restrict_apply<T>, dim3(blocks_for(Rn[d - 1], 256)), dim3(256), 0, s,
                           (const int*)restr[d - 1].data(), Rn[d - 1], (const T*)residual.data(), rhs[d - 1].data(), 0);
        KERNEL_CHECK();
      }
      x[d - 1].zero();
    }
    {  // the root solve's inputs, for the reverse pass
      DArray<T> rb(std::max(Rn[0], 1), s), rx(64, s);
      if (Rn[0]) HIP_CHECK(hipMemcpyAsync(rb.data(), rhs[0].data(), sizeof(T) * Rn[0], hipMemcpyDeviceToDevice, s));
      HIP_CHECK(hipMemcpyAsync(rx.data(), dc ? pad_x.data() : x[0].data(), sizeof(T) * (dc ? 64 : Rn[0]), hipMemcpyDeviceToDevice, s));
      r.root_b.push_back(std::move(rb));
      r.root_x.push_back(std::move(rx));
    }
    hipLaunchKernelGGL(k_root_solve<T>, dim3(1), dim3(64), 0, s, (const int*)mcols[0].data(), (const T*)mvals[0].data(), Rn[0],
                       (const int64_t*)r.levels[0].basis_keys.data(), (const T*)rhs[0].data(), x[0].data(), pad_x.data(),
                       work.data(), dc, base, cycle, r.prm.cg_accuracy, Rn[0], r.cg_ran.data() + cycle);
    KERNEL_CHECK();
    for (int d = 1; d <= base; ++d) {
      hipLaunchKernelGGL(k_spmv<T>, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, (const int*)r.levels[d].P.cols.data(),
                         (const T*)r.levels[d].P.vals.data(), Rn[d], 8, (const T*)x[d - 1].data(), x[d].data(), T(1), T(1),
                         (const T*)nullptr);
      KERNEL_CHECK();
      x[d - 1].zero();
      relax_colored(ops[d], sched[d], x[d].data(), iters, false, s);
    }
  }
  if (dc && Rn[0]) {
    hipLaunchKernelGGL(k_root_extract<T>, dim3(1), dim3(64), 0, s, (const int64_t*)r.levels[0].basis_keys.data(), Rn[0],
                       (const T*)pad_x.data(), x[0].data());
    KERNEL_CHECK();
  }
  B.correction = std::move(x[base]);
}

// Everything of one finer level that does not depend on the coarser levels'
// solution: the sliced schedule, the operator and the cross operator with
// their screening terms, stored with value type V (T, or float in mixed
// runs) in schedule order (row j = order[j]), the point load, the diagonal
// factors and the operator's columns as schedule positions.
template <class T, class V> struct FinerLevel {
  Schedule<T> sc;
  GsOperator<T, V> op;  // cols (schedule positions), vals, inv_diag, has_diag; b is set by the solve
  DArray<unsigned long long> nnz;  // stored nonzeros, read after the run
  TileTable tiles;                 // tiled relaxation (empty when the colour-per-launch form runs)
  DArray<RowRecord<T, V>> records;  // the tiled relaxation's row records
};

template <class T, class V>
static void prepare_finer(Result<T>& r, int d, Stencil<V>& Aop, Stencil<V>& Cop, bool screened, FinerLevel<T, V>& f,
                          hipStream_t s) {
  LevelState<T>& L = r.levels[d];
  const int R = L.basis_count;
  f.sc = make_schedule<T>(L.basis_keys, R, d, true, s);
  const int* ord = f.sc.order.data();
  const int ts = relax_tile_size(r.prm.relax_tile, d);
  if (ts && R && r.prm.iters <= RELAX_MAX_SLICES)
    build_tile_table(L.basis_keys.data(), ord, R, d, ts, L.dense_dim ? (const int*)L.dense.data() : (const int*)nullptr,
                     L.dense_dim, f.tiles, s);
  L.point_load.reset(R, s);
  if (!screened) L.point_load.zero();
  PointSet<T> pts;
  if (screened) pts = point_constraints(r, d, s, false);
  // x is relaxed in schedule order: for the colour-per-launch form columns
  // become schedule positions, so a phase writes contiguous ranges and
  // gathers from nearby ranges (the tiled form reads no columns)
  DArray<int> inv;
  if (!f.tiles.tiles) {
    inv.reset(R, s);
    hipLaunchKernelGGL(k_inverse_permutation, dim3(blocks_for(R, 256)), dim3(256), 0, s, ord, R, inv.data());
    KERNEL_CHECK();
  }
  f.nnz.reset(1, s);
  f.nnz.zero();
  GsOperator<T, V>& op = f.op;
  op.R = R;
  if (f.tiles.tiles > 0) f.records.reset(R, s);
  level_operator<T, V>(L, L, d, d, screened ? &pts : nullptr, L.point_load.data(), ord, inv.data(), Aop, &op,
                       f.records.data(), f.nnz.data(), s);
  op.vals = std::move(Aop.vals);
  // the cross operator keeps canonical rows: its product with the coarser
  // cumulative solution gathers from neighbouring parents in key order
  level_operator<T, V>(L, r.levels[d - 1], d, d - 1, screened ? &pts : nullptr, (double*)nullptr, (const int*)nullptr,
                       (const int*)nullptr, Cop, (GsOperator<T, V>*)nullptr, (RowRecord<T, V>*)nullptr,
                       (unsigned long long*)nullptr, s);
}

// One finer level's solve (solver.py:343-360): right-hand side, b -= cross *
// cumulative, sliced Gauss-Seidel from zero, then the cumulative solution is
// prolonged and accumulated.
// A finer level's right-hand side and x buffers, and (tiled form, single
// reconstructions) its graph's node arguments: prepared on the host while
// the GPU still works on the level before, so the level's relaxation
// launches as soon as its right-hand side is enqueued.
template <class T> struct PreparedLevel {
  bool ready = false;
  DArray<T> rhs, xs, xs1;
  TiledLaunch launch;
};

template <class T, class V>
static PreparedLevel<T> finer_prepare(Result<T>& r, int d, FinerLevel<T, V>& f, hipStream_t s) {
  PreparedLevel<T> p;
  const int R = r.levels[d].basis_count;
  const bool tiled = f.tiles.tiles > 0;
  p.rhs.reset(R, s);
  p.xs.reset(R, s);
  if (tiled) p.xs1.reset(R, s);
  if (tiled && !t_group)
    p.launch = tiled_prepare<T, V>(f.tiles, (const RowRecord<T, V>*)f.records.data(), p.rhs.data(), p.xs.data(),
                                   p.xs1.data(), R, d, r.prm.iters, s);
  p.ready = true;
  return p;
}

template <class T, class V>
static void solve_finer(Result<T>& r, int d, Stencil<V>& Aop, Stencil<V>& Cop, FinerLevel<T, V>& f, DArray<T>& cumulative,
                        PreparedLevel<T>& prep, hipStream_t s) {
  LevelState<T>& L = r.levels[d];
  const int R = L.basis_count;
  const int* ord = f.sc.order.data();
  GsOperator<T, V>& op = f.op;
  L.rhs = std::move(prep.rhs);
  DArray<T> xs = std::move(prep.xs);
  {  // b = rhs - cross * cumulative(d-1): the canonical-row sum, then k_spmv's
     // alpha = -1, beta = 1 combination at each schedule position; the
     // level's correction starts at zero (solver.py:343)
    DArray<T> coupling(R, s);
    hipLaunchKernelGGL((k_spmv<T, V>), dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int*)Cop.cols.data(),
                       (const V*)Cop.vals.data(), R, 27, (const T*)cumulative.data(), coupling.data(), T(1), T(0),
                       (const T*)nullptr);
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_level_rhs<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const double*)L.normal_load.data(),
                       (const double*)L.point_load.data(), ord, R, (const T*)coupling.data(), L.rhs.data(), xs.data());
    KERNEL_CHECK();
  }
  op.b = std::move(L.rhs);
  r.relax_nnz.push_back(std::move(f.nnz));
  typename Result<T>::RelaxStat st{d, R, 0, 0, nullptr, nullptr, 0.f};
  const bool tiled = f.tiles.tiles > 0;
  DArray<T> xs1 = std::move(prep.xs1);
  if (tiled) {
    st.tile = f.tiles.ts;
    st.width = tiled_width(f.tiles);
    st.tiles = f.tiles.tiles;
    st.entries = f.tiles.entries_count;
    st.max_colour_rows = f.tiles.max_rows;
  }
  HIP_CHECK(hipEventCreate(&st.start));
  HIP_CHECK(hipEventCreate(&st.end));
  HIP_CHECK(hipEventRecord(st.start, s));
  int grouped = 0;
  if (t_group) {  // a batch member: the level's tiles join its group's launch
    GroupOffer o;
    if (tiled) {
      if (r.prm.iters % 2) HIP_CHECK(hipMemcpyAsync(xs1.data(), xs.data(), sizeof(T) * R, hipMemcpyDeviceToDevice, s));
      o.valid = true;
      o.value_bytes = (int)sizeof(V);
      o.vector_bytes = (int)sizeof(T);
      o.gside = f.tiles.ts + 8;
      o.width = tiled_width(f.tiles);
      o.iters = r.prm.iters;
      o.entries = f.tiles.entries.data();
      o.desc = f.tiles.desc.data();
      o.records = f.records.data();
      o.b = op.b.data();
      o.x0 = xs.data();
      o.x1 = xs1.data();
      o.tt = &f.tiles;
    }
    grouped = group_relax(d, o, s);
  }
  if (grouped)
    st.launches = grouped;
  else if (tiled && prep.launch.g)
    st.launches = tiled_launch<T>(prep.launch, xs.data(), xs1.data(), R, r.prm.iters, s);
  else if (tiled)
    st.launches = relax_tiled_graph(op, f.tiles, (const RowRecord<T, V>*)f.records.data(), xs.data(), xs1.data(), d,
                                    r.prm.iters, s);
  else
    st.launches = relax_sliced_any(op, f.sc, xs.data(), d, r.prm.iters, s);
  HIP_CHECK(hipEventRecord(st.end, s));
  r.relax_stats.push_back(st);
  hipLaunchKernelGGL(k_scatter_rows<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, ord, R, (const T*)xs.data(),
                     L.correction.data());
  KERNEL_CHECK();
  // keep the schedule-ordered operator (canonical columns) and right-hand side
  Aop.vals = std::move(op.vals);
  L.rhs = std::move(op.b);
  L.order = std::move(f.sc.order);
  L.scheduled = true;
  DArray<T> next(R, s);
  hipLaunchKernelGGL(k_spmv<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int*)L.P.cols.data(),
                     (const T*)L.P.vals.data(), R, 8, (const T*)cumulative.data(), next.data(), T(1), T(1),
                     (const T*)L.correction.data());
  KERNEL_CHECK();
  cumulative = std::move(next);
}

// The finer levels' operators: A and cross, or their float32 copies in
// mixed runs.
template <class V, class T> static Stencil<V>& finer_operator(LevelState<T>& L, bool cross) {
  if constexpr (std::is_same<V, T>::value)
    return cross ? L.cross : L.A;
  else
    return cross ? L.cross32 : L.A32;
}

// One assembly stream per (device, solve stream), kept for the process: its
// pool blocks and the host thread's launches are reused by every call.
static hipStream_t assembly_stream(int dev, hipStream_t s) { return companion_stream(dev, s, 1); }

// Builds the finer levels' operators on the assembly stream from a host
// thread while the solve stream runs the base solve and the coarser levels'
// relaxation (which leave most of the device idle).  Level d's arrays are
// allocated on the assembly stream, used by the solve stream after it waits
// on level d's event, and returned to the assembly stream's pool only after
// the thread has finished (the destructor), so the thread never reuses a
// block the solve stream still reads.
template <class T, class V> class FinerAssembly {
 public:
  FinerAssembly(Result<T>& r, int first, int last, bool screened, hipStream_t s)
      : r_(r), first_(first), last_(last), s_(s), levels_(std::max(0, last - first + 1)), events_(levels_.size()) {
    if (levels_.empty()) return;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    as_ = assembly_stream(dev, s);
    for (auto& e : events_) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipEvent_t fork;
    HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(fork, s));
    HIP_CHECK(hipStreamWaitEvent(as_, fork, 0));
    HIP_CHECK(hipEventDestroy(fork));
    // the point constraints' weights use the serial normal average
    if (r.average_pending) HIP_CHECK(hipStreamWaitEvent(as_, r.average_done, 0));
    HIP_CHECK(hipEventCreateWithFlags(&gate_, hipEventDisableTiming));
    thread_ = std::thread([this, dev, screened] { run(dev, screened); });
  }
  FinerAssembly(const FinerAssembly&) = delete;
  FinerAssembly& operator=(const FinerAssembly&) = delete;
  ~FinerAssembly() {
    if (!thread_.joinable()) return;
    {
      std::lock_guard<std::mutex> g(mu_);
      cancel_ = true;
      cv_.notify_all();
    }
    thread_.join();
    hipEvent_t end;
    if (hipEventCreateWithFlags(&end, hipEventDisableTiming) == hipSuccess) {
      (void)hipEventRecord(end, as_);
      (void)hipStreamWaitEvent(s_, end, 0);
      (void)hipEventDestroy(end);
    }
    for (auto& e : events_) (void)hipEventDestroy(e);
    (void)hipEventDestroy(gate_);
  }
  // Levels after the first two are assembled only once the solve stream has
  // passed this point (the base solve's single-workgroup kernels slow down
  // several times when they share compute units with the assembly).
  void open_gate() {
    if (!thread_.joinable()) return;
    HIP_CHECK(hipEventRecord(gate_, s_));
    std::lock_guard<std::mutex> g(mu_);
    gate_open_ = true;
    cv_.notify_all();
  }
  // The finer levels' loads: each is set on the assembly stream before its
  // level's event, after the solve stream's plan.ready.
  void set_loads(NormalLoadPlan<T>* plan) {
    if (!thread_.joinable()) return;
    std::lock_guard<std::mutex> g(mu_);
    loads_ = plan;
    cv_.notify_all();
  }
  // Level d's operators and load, once the solve stream is ordered after them.
  FinerLevel<T, V>& wait(int d) {
    const int k = d - first_;
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [&] { return ready_ > k || error_ || done_; });
    if (ready_ <= k) {
      if (error_) std::rethrow_exception(error_);
      throw PsrError(1, "finer level operators were not assembled");
    }
    g.unlock();
    HIP_CHECK(hipStreamWaitEvent(s_, events_[k], 0));
    return levels_[k];
  }

 private:
  void run(int dev, bool screened) {
    try {
      HIP_CHECK(hipSetDevice(dev));
      for (int d = first_; d <= last_; ++d) {
        {
          std::lock_guard<std::mutex> g(mu_);
          if (cancel_) break;
        }
        if (d == first_ + 2) {
          std::unique_lock<std::mutex> g(mu_);
          cv_.wait(g, [&] { return gate_open_ || cancel_; });
          if (cancel_) break;
          g.unlock();
          HIP_CHECK(hipStreamWaitEvent(as_, gate_, 0));
        }
        LevelState<T>& L = r_.levels[d];
        if (!L.basis_count) break;
        prepare_finer<T, V>(r_, d, finer_operator<V>(L, false), finer_operator<V>(L, true), screened,
                            levels_[d - first_], as_);
        NormalLoadPlan<T>* plan = nullptr;
        {
          std::unique_lock<std::mutex> g(mu_);
          cv_.wait(g, [&] { return loads_ != nullptr || cancel_; });
          if (cancel_) break;
          plan = loads_;
        }
        if (!loads_waited_) {
          if (plan->ready) HIP_CHECK(hipStreamWaitEvent(as_, plan->ready, 0));
          loads_waited_ = true;
        }
        normal_loads_level(r_, *plan, d, as_);
        HIP_CHECK(hipEventRecord(events_[d - first_], as_));
        std::lock_guard<std::mutex> g(mu_);
        ready_ = d - first_ + 1;
        cv_.notify_all();
      }
    } catch (...) {
      std::lock_guard<std::mutex> g(mu_);
      error_ = std::current_exception();
    }
    std::lock_guard<std::mutex> g(mu_);
    done_ = true;
    cv_.notify_all();
  }

  Result<T>& r_;
  int first_, last_;
  hipStream_t s_, as_ = nullptr;
  hipEvent_t gate_ = nullptr;
  std::vector<FinerLevel<T, V>> levels_;
  std::vector<hipEvent_t> events_;
  std::thread thread_;
  std::mutex mu_;
  std::condition_variable cv_;
  int ready_ = 0;
  bool cancel_ = false, done_ = false, gate_open_ = false;
  NormalLoadPlan<T>* loads_ = nullptr;
  bool loads_waited_ = false;  // assembly thread only
  std::exception_ptr error_;
};

// The base solve, then each finer level once its operators and load are ready.
template <class T, class V>
static void solve_cascade(Result<T>& r, std::vector<DArray<int>>& restr, bool screened, FinerAssembly<T, V>& finer,
                          hipEvent_t loads_ready, hipStream_t s) {
  const int base = r.prm.base_depth;
  DArray<T> cumulative;
  PreparedLevel<T> next;  // the coming finer level's buffers and graph
  for (int d = base; d <= r.prm.solve_depth; ++d) {
    LevelState<T>& L = r.levels[d];
    const int R = L.basis_count;
    if (!R) break;
    if (d == base) {
      L.point_load.reset(R, s);
      L.point_load.zero();
      L.rhs.reset(R, s);
      Stencil<double> A = assemble_stiffness(L, L, d, d, s);
      if (screened) {
        PointSet<T> pts = point_constraints(r, d, s);
        add_screening(A, L.basis_keys, pts, d, d, L.point_load.data(), s);
      }
      L.A = cast_stencil<T>(std::move(A), s);
      // the base operators are set up while the normal loads finish on
      // their own stream; the right-hand side waits for them
      regular_multigrid(r, restr, s, [&] {
        HIP_CHECK(hipStreamWaitEvent(s, loads_ready, 0));
        hipLaunchKernelGGL(k_rhs<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const double*)L.normal_load.data(),
                           (const double*)L.point_load.data(), R, L.rhs.data());
        KERNEL_CHECK();
      });
      cumulative.reset(R, s);
      HIP_CHECK(hipMemcpyAsync(cumulative.data(), L.correction.data(), sizeof(T) * R, hipMemcpyDeviceToDevice, s));
      finer.open_gate();
    } else {
      FinerLevel<T, V>& f = finer.wait(d);
      PreparedLevel<T> cur = next.ready ? std::move(next) : finer_prepare<T, V>(r, d, f, s);
      next = PreparedLevel<T>();
      solve_finer<T, V>(r, d, finer_operator<V>(L, false), finer_operator<V>(L, true), f, cumulative, cur, s);
      // the next level's node arguments are written while this one relaxes
      if (d + 1 <= r.prm.solve_depth && r.levels[d + 1].basis_count)
        next = finer_prepare<T, V>(r, d + 1, finer.wait(d + 1), s);
    }
    L.solved = true;
  }
}

// Normal loads, then the cascadic solve from the base level up
// (solver.py:332-360); the finer levels' operators and loads are assembled
// alongside.
template <class T, class V>
static void solve_levels(Result<T>& r, std::vector<DArray<int>>& restr, bool screened, NormalLoadPlan<T>& plan,
                         hipStream_t s) {
  const int base = r.prm.base_depth;
  group_declare(base + 1);
  // The normal loads the base level needs run on a loads stream beside the
  // base operators' setup on s; loads_ready orders their readers after them.
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  const hipStream_t h = companion_stream(dev, s, 2);
  hipEvent_t start = nullptr, loads_ready = nullptr;
  HIP_CHECK(hipEventCreateWithFlags(&start, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&loads_ready, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(start, s));
  HIP_CHECK(hipStreamWaitEvent(h, start, 0));
  {
    FinerAssembly<T, V> finer(r, base + 1, r.prm.solve_depth, screened, s);
    normal_loads_base(r, restr, plan, h);
    HIP_CHECK(hipEventRecord(loads_ready, h));
    finer.set_loads(&plan);
    solve_cascade<T, V>(r, restr, screened, finer, loads_ready, s);
  }  // the solve stream is ordered after the assembly stream here
  group_finish();
  HIP_CHECK(hipStreamWaitEvent(s, loads_ready, 0));  // also when the cascade stopped before the base level
  (void)hipEventDestroy(start);
  (void)hipEventDestroy(loads_ready);
  normal_loads_rest(r, plan, s);
}

template <class T> void stage_solve(Result<T>& r) {
  hipStream_t s = r.s;
  const int D = r.prm.depth;
  // The base level's per-level normal loads need only the tree: they start
  // on the loads stream now, beside the index, prolongation and restriction
  // set-up below.  The plan outlives the assembly, whose stream reads it.
  NormalLoadPlan<T> plan;
  {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    const hipStream_t h = companion_stream(dev, s, 2);
    hipEvent_t tree_done = nullptr;
    HIP_CHECK(hipEventCreateWithFlags(&tree_done, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(tree_done, s));
    HIP_CHECK(hipStreamWaitEvent(h, tree_done, 0));
    (void)hipEventDestroy(tree_done);
    normal_loads_base_levels(r, plan, h);
  }
  for (int d = 0; d <= D; ++d) {
    LevelState<T>& L = r.levels[d];
    build_column_index((const int64_t*)L.basis_keys.data(), L.basis_count, (1 << d) + 1, L.columns, s);
    const int64_t dim = (1 << d) + 1;
    if (L.basis_count && d >= 5 && dim * dim * dim <= dense_index_max_cells()) {
      L.dense.reset((size_t)(dim * dim * dim), s);
      HIP_CHECK(hipMemsetAsync(L.dense.data(), 0xff, sizeof(int) * (size_t)(dim * dim * dim), s));  // -1
      hipLaunchKernelGGL(k_dense_fill, dim3(blocks_for(L.basis_count, 256)), dim3(256), 0, s, (const int64_t*)L.basis_keys.data(),
                         L.basis_count, (int)dim, L.dense.data());
      KERNEL_CHECK();
      L.dense_dim = (int)dim;
    }
  }
  // prolongation P_d and restriction gathers (solver.py:340-342)
  std::vector<DArray<int>> restr(D + 1);
  for (int d = 1; d <= D; ++d) {
    LevelState<T>& F = r.levels[d];
    LevelState<T>& C = r.levels[d - 1];
    F.P.rows = F.basis_count;
    F.P.cols_count = C.basis_count;
    F.P.cols.reset(8 * (size_t)F.basis_count, s);
    F.P.vals.reset(8 * (size_t)F.basis_count, s);
    if (F.basis_count) {
      hipLaunchKernelGGL(k_prolongation<T>, dim3(blocks_for(F.basis_count, 256)), dim3(256), 0, s,
                         (const int64_t*)F.basis_keys.data(), F.basis_count, (const int64_t*)C.basis_keys.data(),
                         (int64_t)C.basis_count, C.columns.view(), d, F.P.cols.data(), F.P.vals.data());
      KERNEL_CHECK();
    }
    restr[d - 1].reset(27 * (size_t)C.basis_count, s);
    if (C.basis_count) {
      hipLaunchKernelGGL(k_restriction, dim3(blocks_for(C.basis_count, 256)), dim3(256), 0, s,
                         (const int64_t*)C.basis_keys.data(), C.basis_count, (const int64_t*)F.basis_keys.data(),
                         (int64_t)F.basis_count, F.columns.view(), d, restr[d - 1].data());
      KERNEL_CHECK();
    }
  }
  for (int d = 0; d <= D; ++d) {
    LevelState<T>& L = r.levels[d];
    L.correction.reset(L.basis_count, s);
    L.correction.zero();
  }
  const bool screened = r.prm.point_weight > 0 && !r.prm.exact_interpolation;
  if (r.prm.float_operators)
    solve_levels<T, float>(r, restr, screened, plan, s);
  else
    solve_levels<T, T>(r, restr, screened, plan, s);
  for (int d = 0; d < D; ++d) r.levels[d].restriction = std::move(restr[d]);
  for (int d = 0; d <= D; ++d) {
    r.levels[d].dense.free();
    r.levels[d].dense_dim = 0;
  }
}

// CSR -> slot layout in entry order (GS only iterates present slots in order)
template <class T>
__global__ void k_csr_to_slots(const int* __restrict__ offsets, const int* __restrict__ columns, const T* __restrict__ values,
                               int R, int* __restrict__ cols, T* __restrict__ vals, int* __restrict__ too_long) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int o0 = offsets[i], o1 = offsets[i + 1];
  if (o1 - o0 > 27) *too_long = 1;
  for (int k = 0; k < 27; ++k) {
    const int e = o0 + k;
    cols[(int64_t)k * R + i] = e < o1 ? columns[e] : -1;
    vals[(int64_t)k * R + i] = e < o1 ? values[e] : T(0);
  }
}

template <class T>
void relax_sliced_csr(const int64_t* keys, int R, const int* offsets, const int* columns, const T* values, const T* b, T* x,
                      int depth, int iters, hipStream_t s) {
  if (R <= 0) return;
  DArray<int64_t> k(R, s);
  HIP_CHECK(hipMemcpyAsync(k.data(), keys, sizeof(int64_t) * R, hipMemcpyDeviceToDevice, s));
  DArray<int> cols(27 * (size_t)R, s), flag(1, s);
  DArray<T> vals(27 * (size_t)R, s);
  flag.zero();
  hipLaunchKernelGGL(k_csr_to_slots<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, offsets, columns, values, R, cols.data(),
                     vals.data(), flag.data());
  KERNEL_CHECK();
  if (read_int(flag.data(), s)) throw PsrError(2, "relaxation rows hold at most 27 entries");
  Schedule<T> sc = make_schedule<T>(k, R, depth, true, s);
  GsOperator<T> op = gather_operator<T>(sc, cols, vals, b, R, s);
  (void)relax_sliced_any(op, sc, x, depth, iters, s);
}

// relaxation.py:261-271 (relax_scheduled) on caller CSR arrays: z < 0 runs
// iters 8-colour sweeps of row_schedule(keys, depth) (colours 0..7, or 7..0);
// z >= 0 relaxes slice z of the sliced schedule, taking the colours whose
// bit 2 equals z & 1 in that order (group z*4 + (colour & 3)).  Small
// operators run in one workgroup; larger ones launch each colour group.
template <class T>
void relax_scheduled_csr(const int64_t* keys, int R, const int* offsets, const int* columns, const T* values, const T* b, T* x,
                         int depth, int iters, int forward, int z, hipStream_t s) {
  if (z > (1 << depth)) throw PsrError(2, "slice z must lie in [0, 2**depth] (or be negative for 8-colour sweeps)");
  if (R <= 0) return;
  DArray<int64_t> k(R, s);
  HIP_CHECK(hipMemcpyAsync(k.data(), keys, sizeof(int64_t) * R, hipMemcpyDeviceToDevice, s));
  DArray<int> cols(27 * (size_t)R, s), flag(1, s);
  DArray<T> vals(27 * (size_t)R, s);
  flag.zero();
  hipLaunchKernelGGL(k_csr_to_slots<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, offsets, columns, values, R, cols.data(),
                     vals.data(), flag.data());
  KERNEL_CHECK();
  if (read_int(flag.data(), s)) throw PsrError(2, "relaxation rows hold at most 27 entries");
  Schedule<T> sc = make_schedule<T>(k, R, depth, z >= 0, s);
  GsOperator<T> op = gather_operator<T>(sc, cols, vals, b, R, s);
  if (z < 0 && R <= 65536) {
    relax_colored(op, sc, x, iters, forward != 0, s);
    return;
  }
  const unsigned bx = (unsigned)std::max(1, (sc.max_group + 255) / 256);
  auto group = [&](int g) {
    hipLaunchKernelGGL((k_relax_wave<T, T>), dim3(bx, 1), dim3(256), 0, s, op.update_order, (const int*)sc.offsets.data(),
                       (const int*)op.cols.data(), (const T*)op.vals.data(), (const T*)op.b.data(), (const T*)op.inv_diag.data(),
                       (const unsigned char*)op.has_diag.data(), x, op.R, g >> 2, g & 3);
  };
  for (int it = 0; it < iters; ++it)
    for (int k8 = 0; k8 < 8; ++k8) {
      const int color = forward ? k8 : 7 - k8;
      if (z < 0) group(color);
      else if ((color >> 2) == (z & 1)) group(4 * z + (color & 3));
    }
  KERNEL_CHECK();
}

// ================================================================ adjoint
// Reverse of the cascadic solve and of the normal loads with respect to the
// normal-record values, holding the octree, matrices, schedules and the root
// solve's branch fixed (cascadic_adjoint.py:71-104, relaxation_adjoint.py,
// multigrid_adjoint.py, normal_load_adjoint.py:127-161).  Gradients are
// float64 throughout; scatters to shared neighbours use float64 atomics, as
// the reference's reverse kernels do (relaxation_adjoint.py:47-66).

// Reverse of one Gauss-Seidel launch (relaxation_adjoint.py:47-66 without the
// matrix-value terms): rows of the launch's groups in reverse of the forward
// update, x_i += (b_i - sum_k A_ik x_k) / A_ii.  x is canonical (row
// order[j] at schedule position j); rows of one launch are uncoupled.
template <class T, class V>
__global__ void __launch_bounds__(256) k_gs_reverse(const int* __restrict__ order, const int* __restrict__ goff, int first_group,
                                                    int group_stride, const int* __restrict__ cols, const V* __restrict__ vals,
                                                    const T* __restrict__ inv_diag, const unsigned char* __restrict__ has_diag,
                                                    int R, double* __restrict__ gx, double* __restrict__ gb) {
  const int group = first_group + (int)blockIdx.y * group_stride;
  const int j = goff[group] + (int)blockIdx.x * blockDim.x + (int)threadIdx.x;
  if (j >= goff[group + 1] || !has_diag[j]) return;
  const int i = order[j];
  const double rbar = gx[i] * double(inv_diag[j]);
  if (rbar == 0.0) return;
  gb[i] += rbar;
  for (int k = 0; k < 27; ++k) {
    const int c = cols[(int64_t)k * R + j];
    if (c >= 0) atomicAdd(&gx[c], -rbar * double(vals[(int64_t)k * R + j]));
  }
}

// ---- matrix-value gradients (relaxation_adjoint.py:29-73, multigrid_adjoint.py,
// assembly_adjoint.py:44-91): the reverse sweeps need the x each row saw, so
// the forward relaxations are replayed once with x snapshots.

// colour of a row in the sliced order: (x&1) + 2(y&1); in the 8-colour order
// also + 4(z&1)
__device__ __forceinline__ int sliced_colour(int64_t key) { return (key_x(key) & 1) + 2 * (key_y(key) & 1); }
__device__ __forceinline__ int colour8(int64_t key) { return (key_x(key) & 1) + 2 * (key_y(key) & 1) + 4 * (key_z(key) & 1); }

// One launch of relax_sliced that also writes each updated row's new x to
// snap[t * R + row] (t = visit of the slice: first_visit + slice index).
template <class T, class V>
__global__ void __launch_bounds__(256) k_relax_wave_snap(const int* __restrict__ order, const int* __restrict__ goff,
                                                         const int* __restrict__ cols, const V* __restrict__ vals,
                                                         const T* __restrict__ b, const T* __restrict__ inv_diag,
                                                         const unsigned char* __restrict__ has_diag, T* __restrict__ x, int R,
                                                         int first_z, int color, int first_visit, T* __restrict__ snap) {
  const int group = 4 * (first_z + 2 * (int)blockIdx.y) + color;
  const int j = goff[group] + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= goff[group + 1]) return;
  gs_update(j, R, order, cols, vals, b, inv_diag, has_diag, x);
  const int i = order[j];
  snap[(int64_t)(first_visit + (int)blockIdx.y) * R + i] = x[i];
}

// Reverse of one sliced launch with the matrix terms: row i's update at visit
// t saw x_c from visit t for neighbours earlier in the visit (the slice above,
// or an earlier colour of its own slice) and from visit t-1 otherwise; the
// diagonal entry also carries -rbar * (x_new - x_old) through 1/A_ii.
// gvals is in the operator's (schedule-position) layout.
template <class T, class V>
__global__ void __launch_bounds__(256) k_gs_reverse_sliced_m(const int* __restrict__ order, const int* __restrict__ goff,
                                                             int first_z, int color, int first_visit, const int* __restrict__ cols,
                                                             const V* __restrict__ vals, const T* __restrict__ inv_diag,
                                                             const unsigned char* __restrict__ has_diag, int R,
                                                             const int64_t* __restrict__ keys, const T* __restrict__ snap,
                                                             double* __restrict__ gx, double* __restrict__ gb,
                                                             double* __restrict__ gvals) {
  const int group = 4 * (first_z + 2 * (int)blockIdx.y) + color;
  const int j = goff[group] + (int)blockIdx.x * blockDim.x + (int)threadIdx.x;
  if (j >= goff[group + 1] || !has_diag[j]) return;
  const int t = first_visit + (int)blockIdx.y;
  const int i = order[j];
  const double rbar = gx[i] * double(inv_diag[j]);
  if (rbar == 0.0) return;
  gb[i] += rbar;
  const int64_t ki = keys[i];
  const int zi = key_z(ki), ci = sliced_colour(ki);
  const T* now = snap + (int64_t)t * R;
  const T* before = t > 0 ? snap + (int64_t)(t - 1) * R : nullptr;
  for (int k = 0; k < 27; ++k) {
    const int c = cols[(int64_t)k * R + j];
    if (c < 0) continue;
    atomicAdd(&gx[c], -rbar * double(vals[(int64_t)k * R + j]));
    const int64_t kc = keys[c];
    const int zc = key_z(kc);
    const bool earlier = zc > zi || (zc == zi && sliced_colour(kc) < ci);
    double seen = earlier ? double(now[c]) : (before ? double(before[c]) : 0.0);
    if (c == i) seen = double(now[i]);  // x_old + (x_new - x_old)
    gvals[(int64_t)k * R + j] -= rbar * seen;
  }
}

// relax_colored on one workgroup with x in global memory, writing x before
// the call and after every sweep to snap[(sweep + 1) * R + i] (snap[0..R):
// the input x).
template <class T>
__global__ void __launch_bounds__(1024) k_relax_block_snap(const int* __restrict__ order, const int* __restrict__ goff,
                                                           const int* __restrict__ cols, const T* __restrict__ vals,
                                                           const T* __restrict__ b, const T* __restrict__ inv_diag,
                                                           const unsigned char* __restrict__ has_diag, T* __restrict__ x, int R,
                                                           int iters, int forward, T* __restrict__ snap) {
  for (int i = threadIdx.x; i < R; i += blockDim.x) snap[i] = x[i];
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    for (int c = 0; c < 8; ++c) {
      const int g = forward ? c : 7 - c;
      for (int j = goff[g] + threadIdx.x; j < goff[g + 1]; j += blockDim.x)
        gs_update(j, R, order, cols, vals, b, inv_diag, has_diag, x);
      __syncthreads();
    }
    for (int i = threadIdx.x; i < R; i += blockDim.x) snap[(int64_t)(it + 1) * R + i] = x[i];
    __syncthreads();
  }
}

// Reverse of one colour group of an 8-colour sweep with the matrix terms
// (canonical layout: gvals[k * R + order[j]]).
template <class T>
__global__ void __launch_bounds__(256) k_gs_reverse_colored_m(const int* __restrict__ order, const int* __restrict__ goff, int g,
                                                              int forward, const int* __restrict__ cols, const T* __restrict__ vals,
                                                              const T* __restrict__ inv_diag,
                                                              const unsigned char* __restrict__ has_diag, int R,
                                                              const int64_t* __restrict__ keys, const T* __restrict__ before,
                                                              const T* __restrict__ after, double* __restrict__ gx,
                                                              double* __restrict__ gb, double* __restrict__ gvals) {
  const int j = goff[g] + (int)blockIdx.x * blockDim.x + (int)threadIdx.x;
  if (j >= goff[g + 1] || !has_diag[j]) return;
  const int i = order[j];
  const double rbar = gx[i] * double(inv_diag[j]);
  if (rbar == 0.0) return;
  gb[i] += rbar;
  const int pos = forward ? g : 7 - g;
  for (int k = 0; k < 27; ++k) {
    const int c = cols[(int64_t)k * R + j];
    if (c < 0) continue;
    atomicAdd(&gx[c], -rbar * double(vals[(int64_t)k * R + j]));
    const int gc = colour8(keys[c]);
    const bool earlier = (forward ? gc : 7 - gc) < pos;
    double seen = earlier ? double(after[c]) : double(before[c]);
    if (c == i) seen = double(after[i]);
    gvals[(int64_t)k * R + i] -= rbar * seen;
  }
}

// out[j] = 0.5 * g[order[j]] (or g[j])
__global__ void k_half_gather(const int* __restrict__ order, int R, const double* __restrict__ g, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < R) out[j] = 0.5 * g[order ? order[j] : j];
}

// gvals[k * R + j] += alpha * gy[row] * x[col] for a stencil's present slots
// (the matrix gradient of y += alpha * A x), rows at positions j of canonical
// row order[j] (or j)
template <class T>
__global__ void k_stencil_outer(const int* __restrict__ order, const int* __restrict__ cols, int R, int slots,
                                const double* __restrict__ gy, const T* __restrict__ x, double alpha, double* __restrict__ gvals) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const double g = gy[order ? order[j] : j] * alpha;
  if (g == 0.0) return;
  for (int k = 0; k < slots; ++k) {
    const int c = cols[(int64_t)k * R + j];
    if (c >= 0) gvals[(int64_t)k * R + j] += g * double(x[c]);
  }
}

// Reverse of k_galerkin (one workgroup per coarse row r): the coarse slot
// gradients reach the fine operator entries through the same R and P weights.
template <class T>
__global__ void __launch_bounds__(128) k_galerkin_vjp(const int* __restrict__ rcols, int C, const int* __restrict__ acols, int F,
                                                      const int* __restrict__ pcols, const T* __restrict__ pvals,
                                                      const int64_t* __restrict__ ckeys, const double* __restrict__ gcoarse,
                                                      double* __restrict__ gfine) {
  const int r = blockIdx.x;
  const int t = threadIdx.x;
  if (t >= 125) return;
  const int64_t rkey = ckeys[r];
  const int rx = key_x(rkey), ry = key_y(rkey), rz = key_z(rkey);
  const int tx = t / 25, ty = (t / 5) % 5, tz = t % 5;
  int col = -1;
  for (int kr = 0; kr < 27; ++kr) {
    const int k = rcols[(int64_t)kr * C + r];
    if (k < 0) continue;
    const int jx = tx - kr / 9, jy = ty - (kr / 3) % 3, jz = tz - kr % 3;
    if ((unsigned)jx > 2u || (unsigned)jy > 2u || (unsigned)jz > 2u) continue;
    const int c = acols[(int64_t)(jx * 9 + jy * 3 + jz) * F + k];
    if (c >= 0) col = c;
  }
  if (col < 0) return;
  double gra = 0.0;  // gradient of (R A)_{r,t}
  for (int kp = 0; kp < 8; ++kp) {
    const int pc = pcols[(int64_t)kp * F + col];
    if (pc < 0) continue;
    const int64_t pkey = ckeys[pc];
    const int ox = key_x(pkey) - rx + 1, oy = key_y(pkey) - ry + 1, oz = key_z(pkey) - rz + 1;
    if ((unsigned)ox > 2u || (unsigned)oy > 2u || (unsigned)oz > 2u) continue;
    gra += gcoarse[(int64_t)(ox * 9 + oy * 3 + oz) * C + r] * double(pvals[(int64_t)kp * F + col]);
  }
  if (gra == 0.0) return;
  for (int kr = 0; kr < 27; ++kr) {
    const int k = rcols[(int64_t)kr * C + r];
    if (k < 0) continue;
    const int jx = tx - kr / 9, jy = ty - (kr / 3) % 3, jz = tz - kr % 3;
    if ((unsigned)jx > 2u || (unsigned)jy > 2u || (unsigned)jz > 2u) continue;
    const int j = jx * 9 + jy * 3 + jz;
    if (acols[(int64_t)j * F + k] < 0) continue;
    atomicAdd(&gfine[(int64_t)j * F + k], restriction_weight(kr) * gra);
  }
}

// d/dp of max(0, 1 - |p s - c|): -sign(t) s for 0 < |t| < 1, zero at the
// kink t = 0 and where clamped (assembly_adjoint.py:38-47)
template <class T> __device__ __forceinline__ double hat_deriv(T p, T s, T c) {
  const T t = p * s - c;
  if (!(fabs(t) < T(1)) || t == T(0)) return 0.0;
  return t < T(0) ? double(s) : -double(s);
}

// screening_stencils.py:34-70 reversed at one level: row j's slot gradients
// of the operator (gA, trial depth = depth) and of the cross operator (gC,
// trial depth = depth - 1) and its point-load gradient (gload) reach the
// constraint points' positions and weights; each of the three may be null.
template <class T>
__global__ void __launch_bounds__(128) k_screening_vjp(const int64_t* __restrict__ rows, int R, const int* __restrict__ row_order,
                                                       int depth, PointBins<T> pb, const double* __restrict__ gA,
                                                       const double* __restrict__ gC, const double* __restrict__ gload,
                                                       double* __restrict__ gpos, double* __restrict__ gw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const int i = row_order ? row_order[j] : j;
  const int64_t key = rows[i];
  const int cx = key_x(key), cy = key_y(key), cz = key_z(key);
  const int n = 1 << depth;
  double GA[27], GC[27];
  bool any = false;
  for (int k = 0; k < 27; ++k) {
    GA[k] = gA ? gA[(int64_t)k * R + j] : 0.0;
    GC[k] = gC ? gC[(int64_t)k * R + j] : 0.0;
    any = any || GA[k] != 0.0 || GC[k] != 0.0;
  }
  const double GL = gload ? gload[j] : 0.0;
  if (!any && GL == 0.0) return;
  const T s = T(1 << depth);
  for (int a = 0; a < 2; ++a)
    for (int bb = 0; bb < 2; ++bb)
      for (int kz = 0; kz < 2; ++kz) {
        const int bx = cx - a, by = cy - bb, bz = cz - kz;
        if (bx < 0 || by < 0 || bz < 0 || bx >= n || by >= n || bz >= n) continue;
        int lo, hi;
        if (pb.cell_start) {
          const int64_t c = ((int64_t)bz * n + bx) * n + by;
          lo = pb.cell_start[c];
          hi = pb.cell_end[c];
        } else {
          const uint64_t want = ((uint64_t)bx << (2 * depth)) | ((uint64_t)by << depth) | (uint64_t)bz;
          int l = 0, h = pb.P;
          if (pb.bin_cols.start) {
            l = pb.bin_cols.start[bx * n + by];
            h = pb.bin_cols.end[bx * n + by];
          }
          while (l < h) {
            const int mid = (l + h) >> 1;
            if (pb.bin_keys[mid] < want) l = mid + 1; else h = mid;
          }
          lo = l;
          hi = l;
          while (hi < pb.P && pb.bin_keys[hi] == want) ++hi;
        }
        for (int jj = lo; jj < hi; ++jj) {
          const int index = pb.bin_ids[jj];
          const T weight = pb.w[index];
          if (weight == T(0)) continue;
          const T p[3] = {pb.pos[3 * index], pb.pos[3 * index + 1], pb.pos[3 * index + 2]};
          const T cc[3] = {T(cx), T(cy), T(cz)};
          double h[3], dh[3];
          for (int ax = 0; ax < 3; ++ax) {
            const T v = T(1) - fabs(p[ax] * s - cc[ax]);
            h[ax] = v > T(0) ? double(v) : 0.0;
            dh[ax] = hat_deriv(p[ax], s, cc[ax]);
          }
          const double phi = h[0] * h[1] * h[2];
          const double dphi[3] = {dh[0] * h[1] * h[2], h[0] * dh[1] * h[2], h[0] * h[1] * dh[2]};
          double S = GL, Sp[3] = {0.0, 0.0, 0.0};  // sum_k G_k trial_k and its derivatives (load: trial = 1)
          for (int which = 0; which < 2; ++which) {
            const double* G = which == 0 ? GA : GC;
            if ((which == 0 && !gA) || (which == 1 && !gC)) continue;
            const int td = which == 0 ? depth : depth - 1;
            const int shift = depth - td;
            const int px = (cx >> shift) - 1, py = (cy >> shift) - 1, pz = (cz >> shift) - 1;
            const T st = T(1 << td);
            const T pc[3] = {T(px), T(py), T(pz)};
            double hx[3][3], dx[3][3];
            for (int ax = 0; ax < 3; ++ax)
              for (int o = 0; o < 3; ++o) {
                const T v = T(1) - fabs(p[ax] * st - (pc[ax] + T(o)));
                hx[ax][o] = v > T(0) ? double(v) : 0.0;
                dx[ax][o] = hat_deriv(p[ax], st, pc[ax] + T(o));
              }
            for (int k = 0; k < 27; ++k) {
              if (G[k] == 0.0) continue;
              const int ox = k / 9, oy = (k / 3) % 3, oz = k % 3;
              S += G[k] * hx[0][ox] * hx[1][oy] * hx[2][oz];
              Sp[0] += G[k] * dx[0][ox] * hx[1][oy] * hx[2][oz];
              Sp[1] += G[k] * hx[0][ox] * dx[1][oy] * hx[2][oz];
              Sp[2] += G[k] * hx[0][ox] * hx[1][oy] * dx[2][oz];
            }
          }
          const double w = double(weight);
          atomicAdd(&gw[index], phi * S);
          for (int ax = 0; ax < 3; ++ax) atomicAdd(&gpos[3 * index + ax], w * (dphi[ax] * S + phi * Sp[ax]));
        }
      }
}

// constraints.py:55-84 reversed: centroid p = v[0:3] / v[3] and weight
// v[3] 2^d (point_weight avg[1]) (zero where the forward zeroed it).
template <class T>
__global__ void __launch_bounds__(256) k_points_vjp(const T* __restrict__ avals, const T* __restrict__ w, int P, int depth,
                                                    double point_weight, const double* __restrict__ gpos,
                                                    const double* __restrict__ gw, double* __restrict__ gagg,
                                                    double* __restrict__ gavg1) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double part = 0.0;
  if (a < P && w[a] != T(0)) {
    const double v3 = double(avals[4 * (int64_t)a + 3]);
    for (int c = 0; c < 3; ++c) gagg[3 * (int64_t)a + c] += gpos[3 * (int64_t)a + c] / v3;
    part = gw[a] * v3 * double(1 << depth) * point_weight;
  }
  __shared__ double red[256];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0.0) atomicAdd(gavg1, red[0]);
}

// gx[col] += alpha * A[row, col] * gy[row] (A^T gy), rows at positions j of
// canonical row order[j] (or j)
template <class V>
__global__ void k_stencil_scatter(const int* __restrict__ order, const int* __restrict__ cols, const V* __restrict__ vals, int R,
                                  int slots, const double* __restrict__ gy, double alpha, double* __restrict__ gx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R) return;
  const double g = gy[order ? order[j] : j] * alpha;
  if (g == 0.0) return;
  for (int k = 0; k < slots; ++k) {
    const int c = cols[(int64_t)k * R + j];
    if (c >= 0) atomicAdd(&gx[c], g * double(vals[(int64_t)k * R + j]));
  }
}

// transpose of k_restrict_apply: gx[fine] += w_k * gy[coarse]
__global__ void k_restrict_scatter(const int* __restrict__ cols, int C, const double* __restrict__ gy, double* __restrict__ gx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C) return;
  const double g = gy[r];
  if (g == 0.0) return;
  for (int k = 0; k < 27; ++k) {
    const int c = cols[(int64_t)k * C + r];
    if (c >= 0) atomicAdd(&gx[c], restriction_weight(k) * g);
  }
}

__global__ void k_axpy_d(double* __restrict__ y, const double* __restrict__ x, double a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// transpose of k_normal_load: record gradients from the load gradient
__global__ void k_normal_load_vjp(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ nkeys, int64_t ncount,
                                  ColumnTable ct, int test_depth, int normal_depth, int first, int width,
                                  const D2* __restrict__ table, const double* __restrict__ gout, double* __restrict__ gvals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const double g = gout[i];
  if (g == 0.0) return;
  const int64_t key = keys[i];
  const int c0 = key_x(key), c1 = key_y(key), c2 = key_z(key);
  const int e0 = normal_center(c0, test_depth, normal_depth), e1 = normal_center(c1, test_depth, normal_depth),
            e2 = normal_center(c2, test_depth, normal_depth);
  const int lim = 1 << normal_depth;
  const int zlo = imax(e2 + first, 0), zhi = imin(e2 + first + width - 1, lim - 1);
  for (int x = 0; x < width; ++x) {
    const int nx = e0 + x + first;
    if (nx < 0 || nx >= lim) continue;
    for (int y = 0; y < width; ++y) {
      const int ny = e1 + y + first;
      if (ny < 0 || ny >= lim || zlo > zhi) continue;
      int64_t found[6];
      find_run_table<6>(nkeys, ncount, ct, normal_depth, nx, ny, zlo, zhi - zlo + 1, found);
      for (int z = 0; z < width; ++z) {
        const int nz = e2 + z + first;
        if (nz < zlo || nz > zhi) continue;
        const int64_t jn = found[nz - zlo];
        if (jn < 0) continue;
        const D2 ix = table[c0 * width + x], iy = table[c1 * width + y], iz = table[c2 * width + z];
        atomicAdd(&gvals[3 * jn], g * ix.v1 * iy.v0 * iz.v0);
        atomicAdd(&gvals[3 * jn + 1], g * ix.v0 * iy.v1 * iz.v0);
        atomicAdd(&gvals[3 * jn + 2], g * ix.v0 * iy.v0 * iz.v1);
      }
    }
  }
}

// transpose of k_normal_prolong: own records and the previous level's
// cumulative normal coefficients from this level's
template <class T>
__global__ void k_normal_prolong_vjp(const int64_t* __restrict__ keys, int R, const int64_t* __restrict__ own_keys,
                                     int64_t own_count, ColumnTable own_ct, const int64_t* __restrict__ prev_keys,
                                     int64_t prev_count, ColumnTable ct, int depth, const double* __restrict__ gcum,
                                     double* __restrict__ gown, double* __restrict__ gprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const double g0 = gcum[3 * i], g1 = gcum[3 * i + 1], g2 = gcum[3 * i + 2];
  if (g0 == 0.0 && g1 == 0.0 && g2 == 0.0) return;
  const int64_t key = keys[i];
  const int c[3] = {key_x(key), key_y(key), key_z(key)};
  int64_t own;
  find_run_table<1>(own_keys, own_count, own_ct, depth, c[0], c[1], c[2], 1, &own);
  if (own >= 0) {  // one row per record key: no other writer in this launch
    gown[3 * own] += g0; gown[3 * own + 1] += g1; gown[3 * own + 2] += g2;
  }
  if (depth == 0) return;
  int64_t idx[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) find_run_table<2>(prev_keys, prev_count, ct, depth - 1, c[0] / 2 + a, c[1] / 2 + b, c[2] / 2, 2, idx[a][b]);
  for (int corner = 0; corner < 8; ++corner) {
    int p[3] = {c[0] / 2, c[1] / 2, c[2] / 2};
    T weight = T(1);
    for (int axis = 0; axis < 3; ++axis) {
      if ((corner >> axis) & 1) p[axis] += 1;
      weight = weight * (T(1) - fabs(T(c[axis]) * T(0.5) - T(p[axis])));
    }
    if (weight > T(0)) {
      const int64_t j = idx[corner & 1][(corner >> 1) & 1][(corner >> 2) & 1];
      if (j >= 0) {
        const double w = double(weight);
        atomicAdd(&gprev[3 * j], w * g0);
        atomicAdd(&gprev[3 * j + 1], w * g1);
        atomicAdd(&gprev[3 * j + 2], w * g2);
      }
    }
  }
}

namespace {

DArray<double> zeros_d(size_t n, hipStream_t s) {
  DArray<double> a(n, s);
  a.zero();
  return a;
}

// reverse of relax_sliced: the launch program backwards, colours 3..0
template <class T, class V>
void gs_sliced_reverse(const Schedule<T>& sc, const int* cols, const V* vals, const T* inv_diag, const unsigned char* has_diag,
                       int R, int depth, int iters, double* gx, double* gb, hipStream_t s) {
  const int n = 1 << depth;
  const unsigned bx = (unsigned)std::max(1, (sc.max_group + 255) / 256);
  std::vector<std::pair<int, int>> program;  // (first slice, slices)
  for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
    const int first = std::max(0, floordiv(q - n + 1, 2));
    const int last = std::min(iters - 1, q / 2);
    if (last < first) continue;
    program.emplace_back(n - q + 2 * first, last - first + 1);
  }
  for (auto it = program.rbegin(); it != program.rend(); ++it)
    for (int color = 3; color >= 0; --color) {
      hipLaunchKernelGGL((k_gs_reverse<T, V>), dim3(bx, (unsigned)it->second), dim3(256), 0, s, (const int*)sc.order.data(),
                         (const int*)sc.offsets.data(), 4 * it->first + color, 8, cols, vals, inv_diag, has_diag, R, gx, gb);
    }
  KERNEL_CHECK();
}

// reverse of relax_colored (8-colour sweeps, forward 0..7 or reverse 7..0)
template <class T>
void gs_colored_reverse(const GsOperator<T>& op, const Schedule<T>& sc, int iters, bool forward, double* gx, double* gb,
                        hipStream_t s) {
  if (!op.R) return;
  const unsigned bx = (unsigned)std::max(1, (sc.max_group + 255) / 256);
  for (int it = 0; it < iters; ++it)
    for (int c = 0; c < 8; ++c) {
      const int g = forward ? 7 - c : c;
      hipLaunchKernelGGL((k_gs_reverse<T, T>), dim3(bx, 1), dim3(256), 0, s, op.update_order, (const int*)sc.offsets.data(), g, 0,
                         (const int*)op.cols.data(), (const T*)op.vals.data(), (const T*)op.inv_diag.data(),
                         (const unsigned char*)op.has_diag.data(), op.R, gx, gb);
    }
  KERNEL_CHECK();
}

template <class V>
void stencil_scatter(const int* order, const Stencil<V>& m, int slots, const double* gy, double alpha, double* gx, hipStream_t s) {
  if (!m.rows) return;
  hipLaunchKernelGGL(k_stencil_scatter<V>, dim3(blocks_for(m.rows, 256)), dim3(256), 0, s, order, (const int*)m.cols.data(),
                     (const V*)m.vals.data(), m.rows, slots, gy, alpha, gx);
  KERNEL_CHECK();
}

// One finer level backwards: gx (w.r.t. its correction, consumed) gives the
// right-hand-side gradient grhs and adds to gcum_prev (w.r.t. the cumulative
// solution of the level below).
template <class T, class V>
void finer_backward(Result<T>& r, int d, const Stencil<V>& Aop, const Stencil<V>& Cop, DArray<double>& gx, DArray<double>& grhs,
                    DArray<double>& gcum_prev, hipStream_t s, DArray<double>* gA = nullptr, DArray<double>* gC = nullptr,
                    const T* cum_prev = nullptr) {
  LevelState<T>& L = r.levels[d];
  const int R = L.basis_count;
  Schedule<T> sc = make_schedule<T>(L.basis_keys, R, d, true, s);
  DArray<T> inv_diag(R, s);
  DArray<unsigned char> has_diag(R, s);
  hipLaunchKernelGGL((k_diag_sched<T, V>), dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int*)sc.order.data(), R,
                     (const int*)Aop.cols.data(), (const V*)Aop.vals.data(), inv_diag.data(), has_diag.data());
  KERNEL_CHECK();
  grhs = zeros_d(R, s);
  if (!gA) {
    gs_sliced_reverse<T, V>(sc, (const int*)Aop.cols.data(), (const V*)Aop.vals.data(), (const T*)inv_diag.data(),
                            (const unsigned char*)has_diag.data(), R, d, r.prm.iters, gx.data(), grhs.data(), s);
  } else {
    // replay the level's relaxation from zero (relaxation_adjoint.py:29-46:
    // the x each visit saw), then the reverse with the matrix terms
    const int n = 1 << d, iters = r.prm.iters;
    const unsigned bx = (unsigned)std::max(1, (sc.max_group + 255) / 256);
    struct Step { int z0, slices, first; };
    std::vector<Step> program;
    for (int q = 0; q <= n + 2 * (iters - 1); ++q) {
      const int first = std::max(0, floordiv(q - n + 1, 2));
      const int last = std::min(iters - 1, q / 2);
      if (last < first) continue;
      program.push_back({n - q + 2 * first, last - first + 1, first});
    }
    DArray<T> x(R, s), snap((size_t)iters * R, s);
    x.zero();
    for (const Step& st : program)
      for (int color = 0; color < 4; ++color) {
        hipLaunchKernelGGL((k_relax_wave_snap<T, V>), dim3(bx, (unsigned)st.slices), dim3(256), 0, s, (const int*)sc.order.data(),
                           (const int*)sc.offsets.data(), (const int*)Aop.cols.data(), (const V*)Aop.vals.data(),
                           (const T*)L.rhs.data(), (const T*)inv_diag.data(), (const unsigned char*)has_diag.data(), x.data(), R,
                           st.z0, color, st.first, snap.data());
      }
    KERNEL_CHECK();
    gA->reset(27 * (size_t)R, s);
    gA->zero();
    for (auto it = program.rbegin(); it != program.rend(); ++it)
      for (int color = 3; color >= 0; --color) {
        hipLaunchKernelGGL((k_gs_reverse_sliced_m<T, V>), dim3(bx, (unsigned)it->slices), dim3(256), 0, s,
                           (const int*)sc.order.data(), (const int*)sc.offsets.data(), it->z0, color, it->first,
                           (const int*)Aop.cols.data(), (const V*)Aop.vals.data(), (const T*)inv_diag.data(),
                           (const unsigned char*)has_diag.data(), R, (const int64_t*)L.basis_keys.data(), (const T*)snap.data(),
                           gx.data(), grhs.data(), gA->data());
      }
    KERNEL_CHECK();
  }
  // b = rhs - cross * cumulative(d-1); the cross operator has canonical rows
  stencil_scatter<V>((const int*)nullptr, Cop, 27, (const double*)grhs.data(), -1.0, gcum_prev.data(), s);
  if (gC && R) {
    gC->reset(27 * (size_t)R, s);
    gC->zero();
    hipLaunchKernelGGL(k_stencil_outer<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int*)nullptr,
                       (const int*)Cop.cols.data(), R, 27, (const double*)grhs.data(), cum_prev, -1.0, gC->data());
    KERNEL_CHECK();
  }
}

// Reverse of relax_colored with the matrix terms: sweeps from the last, each
// sweep's colours in reverse, from the replay's snapshots (snap[0]: x before
// the call, snap[s + 1]: after sweep s); gvals in the canonical layout.
template <class T>
void gs_colored_reverse_m(const GsOperator<T>& op, const Schedule<T>& sc, int iters, bool forward, const int64_t* keys,
                          const T* snap, double* gx, double* gb, double* gvals, hipStream_t s) {
  if (!op.R) return;
  const unsigned bx = (unsigned)std::max(1, (sc.max_group + 255) / 256);
  for (int it = iters - 1; it >= 0; --it)
    for (int c = 7; c >= 0; --c) {
      const int g = forward ? c : 7 - c;
      hipLaunchKernelGGL(k_gs_reverse_colored_m<T>, dim3(bx), dim3(256), 0, s, op.update_order, (const int*)sc.offsets.data(), g,
                         forward ? 1 : 0, (const int*)op.cols.data(), (const T*)op.vals.data(), (const T*)op.inv_diag.data(),
                         (const unsigned char*)op.has_diag.data(), op.R, keys, snap + (size_t)it * op.R,
                         snap + (size_t)(it + 1) * op.R, gx, gb, gvals);
    }
  KERNEL_CHECK();
}

template <class T>
void relax_colored_snap(const GsOperator<T>& op, const Schedule<T>& sc, T* x, int iters, bool forward, T* snap, hipStream_t s) {
  if (!op.R) return;
  hipLaunchKernelGGL(k_relax_block_snap<T>, dim3(1), dim3(1024), 0, s, op.update_order, (const int*)sc.offsets.data(),
                     (const int*)op.cols.data(), (const T*)op.vals.data(), (const T*)op.b.data(), (const T*)op.inv_diag.data(),
                     (const unsigned char*)op.has_diag.data(), x, op.R, iters, forward ? 1 : 0, snap);
  KERNEL_CHECK();
}

}  // namespace

// transpose of normal_stencil_load
template <class T>
void normal_load_vjp(const LevelState<T>& testL, int test_depth, const int64_t* nkeys, int64_t ncount, ColumnTable ct,
                     int normal_depth, const double* gout, double* gvals, hipStream_t s) {
  const int R = testL.basis_count;
  if (!R || !ncount) return;
  const int width = test_depth < normal_depth ? 6 : 4;
  const int first = test_depth < normal_depth ? -3 : -2;
  const int rows = (1 << test_depth) + 1;
  DArray<D2> table((size_t)rows * width, s);
  hipLaunchKernelGGL(k_normal_table, dim3(blocks_for(rows * width, 256)), dim3(256), 0, s, test_depth, normal_depth, first, width,
                     table.data());
  KERNEL_CHECK();
  hipLaunchKernelGGL(k_normal_load_vjp, dim3(blocks_for(R, 128)), dim3(128), 0, s, (const int64_t*)testL.basis_keys.data(), R,
                     nkeys, ncount, ct, test_depth, normal_depth, first, width, (const D2*)table.data(), gout, gvals);
  KERNEL_CHECK();
}

// Reverse of normal_loads (normal_loads.py:66-190, normal_load_adjoint.py:127-161):
// gload[d] (w.r.t. the level's load, canonical; empty = zero) gives the
// gradient w.r.t. every normal record's three values (grecords, accumulated).
template <class T>
void normal_loads_backward(Result<T>& r, std::vector<DArray<double>>& gload, int top, double* grecords, hipStream_t s) {
  const int D = r.prm.depth, base = r.prm.base_depth;
  auto has = [&](int d) { return d <= top && d < (int)gload.size() && !gload[d].empty(); };
  // finer(d-1) = R finer(d) + from_child(d-1): G_d = gload_d + R^T G_{d-1}
  std::vector<DArray<double>> G(D + 1);
  for (int d = base; d <= D; ++d) {
    const int R = r.levels[d].basis_count;
    G[d] = zeros_d(R, s);
    if (has(d) && R) {
      hipLaunchKernelGGL(k_axpy_d, dim3(blocks_for(R, 256)), dim3(256), 0, s, G[d].data(), (const double*)gload[d].data(), 1.0,
                         (int64_t)R);
      KERNEL_CHECK();
    }
    const int C = d > base ? r.levels[d - 1].basis_count : 0;
    if (C && R) {
      hipLaunchKernelGGL(k_restrict_scatter, dim3(blocks_for(C, 256)), dim3(256), 0, s,
                         (const int*)r.levels[d - 1].restriction.data(), C, (const double*)G[d - 1].data(), G[d].data());
      KERNEL_CHECK();
    }
  }
  std::vector<DArray<double>> gcumn(D + 1);
  bool has_cumulative = false;
  for (int d = 0; d <= D; ++d) {
    const int64_t r0 = r.normal_offsets[d], r1 = r.normal_offsets[d + 1];
    const int64_t* nk = r.normal_keys.data() + r0;
    double* gv = grecords + 3 * r0;
    has_cumulative = has_cumulative || (r1 > r0);
    LevelState<T>& L = r.levels[d];
    ColumnIndex records;
    build_column_index(nk, r1 - r0, (1 << d) + 1, records, s);
    if (d >= base) {
      if (has(d)) normal_load_vjp(L, d, nk, r1 - r0, records.view(), d, (const double*)gload[d].data(), gv, s);
      if (d > base) normal_load_vjp(r.levels[d - 1], d - 1, nk, r1 - r0, records.view(), d, (const double*)G[d - 1].data(), gv, s);
    }
    if (d < D) {
      gcumn[d] = zeros_d(3 * (size_t)L.basis_count, s);
      if (d >= base && has_cumulative && has(d + 1))
        normal_load_vjp(r.levels[d + 1], d + 1, (const int64_t*)L.basis_keys.data(), (int64_t)L.basis_count, L.columns.view(), d,
                        (const double*)gload[d + 1].data(), gcumn[d].data(), s);
    }
  }
  // cumulative normal coefficients, finest first (normal_loads.py:27-63)
  for (int d = D - 1; d >= 0; --d) {
    LevelState<T>& L = r.levels[d];
    if (!L.basis_count) continue;
    const int64_t r0 = r.normal_offsets[d], r1 = r.normal_offsets[d + 1];
    ColumnIndex records;
    build_column_index(r.normal_keys.data() + r0, r1 - r0, (1 << d) + 1, records, s);
    const LevelState<T>* P = d > 0 ? &r.levels[d - 1] : nullptr;
    hipLaunchKernelGGL(k_normal_prolong_vjp<T>, dim3(blocks_for(L.basis_count, 256)), dim3(256), 0, s,
                       (const int64_t*)L.basis_keys.data(), L.basis_count, (const int64_t*)(r.normal_keys.data() + r0), r1 - r0,
                       records.view(), P ? (const int64_t*)P->basis_keys.data() : (const int64_t*)nullptr,
                       (int64_t)(P ? P->basis_count : 0), P ? P->columns.view() : ColumnTable{}, d,
                       (const double*)gcumn[d].data(), grecords + 3 * r0, d > 0 ? gcumn[d - 1].data() : (double*)nullptr);
    KERNEL_CHECK();
  }
}

template <class T> void solve_backward(Result<T>& r, const double* gcoef, double* grecords, hipStream_t s, SolveGrads* mg) {
  const int D = r.prm.depth, base = r.prm.base_depth;
  int top = base - 1;
  for (int d = base; d <= r.prm.solve_depth && d <= D; ++d)
    if (r.levels[d].solved) top = d;
  if (top < base) return;
  std::vector<DArray<double>> grhs(D + 1);
  std::vector<DArray<T>> cum;
  if (mg) {
    mg->top = top;
    mg->gA.clear();
    mg->gC.clear();
    mg->gA.resize(D + 1);
    mg->gC.resize(D + 1);
    // the forward's cumulative solutions: cum(base) = x_base, cum(d) = x_d + P_d cum(d-1)
    cum.resize(D + 1);
    for (int d = base; d <= top; ++d) {
      const int R = r.levels[d].basis_count;
      cum[d].reset(R, s);
      if (R) HIP_CHECK(hipMemcpyAsync(cum[d].data(), r.levels[d].correction.data(), sizeof(T) * R, hipMemcpyDeviceToDevice, s));
      if (d > base && R) {
        hipLaunchKernelGGL(k_spmv<T>, dim3(blocks_for(R, 256)), dim3(256), 0, s, (const int*)r.levels[d].P.cols.data(),
                           (const T*)r.levels[d].P.vals.data(), R, 8, (const T*)cum[d - 1].data(), cum[d].data(), T(1), T(1),
                           (const T*)nullptr);
        KERNEL_CHECK();
      }
    }
  }
  // finer levels: cumulative(d) = P_d cumulative(d-1) + x_d
  DArray<double> gcum;
  for (int d = top; d > base; --d) {
    LevelState<T>& L = r.levels[d];
    const int R = L.basis_count, C = r.levels[d - 1].basis_count;
    DArray<double> gx(R, s);
    HIP_CHECK(hipMemcpyAsync(gx.data(), gcoef + r.level_offsets[d], sizeof(double) * R, hipMemcpyDeviceToDevice, s));
    if (!gcum.empty()) {
      hipLaunchKernelGGL(k_axpy_d, dim3(blocks_for(R, 256)), dim3(256), 0, s, gx.data(), (const double*)gcum.data(), 1.0,
                         (int64_t)R);
      KERNEL_CHECK();
    }
    DArray<double> gprev = zeros_d(C, s);
    if (!gcum.empty()) stencil_scatter<T>(nullptr, L.P, 8, (const double*)gcum.data(), 1.0, gprev.data(), s);
    DArray<double>* gA = mg ? &mg->gA[d] : nullptr;
    DArray<double>* gC = mg ? &mg->gC[d] : nullptr;
    const T* cprev = mg ? (const T*)cum[d - 1].data() : nullptr;
    if (r.prm.float_operators) finer_backward<T, float>(r, d, L.A32, L.cross32, gx, grhs[d], gprev, s, gA, gC, cprev);
    else finer_backward<T, T>(r, d, L.A, L.cross, gx, grhs[d], gprev, s, gA, gC, cprev);
    gcum = std::move(gprev);
  }
  // base level: regular multigrid
  {
    std::vector<int> Rn(base + 1);
    for (int d = 0; d <= base; ++d) Rn[d] = r.levels[d].basis_count;
    DArray<double> gxb(Rn[base], s);
    HIP_CHECK(hipMemcpyAsync(gxb.data(), gcoef + r.level_offsets[base], sizeof(double) * Rn[base], hipMemcpyDeviceToDevice, s));
    if (!gcum.empty() && Rn[base]) {
      hipLaunchKernelGGL(k_axpy_d, dim3(blocks_for(Rn[base], 256)), dim3(256), 0, s, gxb.data(), (const double*)gcum.data(), 1.0,
                         (int64_t)Rn[base]);
      KERNEL_CHECK();
    }
    grhs[base] = zeros_d(Rn[base], s);
    if (mg) {
      mg->gA[base].reset(27 * (size_t)Rn[base], s);
      mg->gA[base].zero();
    }
    auto ran = [&](int cycle) { return std::find(r.cg_runs.begin(), r.cg_runs.end(), cycle) != r.cg_runs.end(); };
    const int maxiter = Rn[0];
    DArray<double> cg_store((size_t)(maxiter + 1) * 3 * 64 + 4 * (size_t)(maxiter + 1) + 64, s);
    auto root_reverse = [&](int cycle, const int* cols0, const T* vals0, double* gx0, double* grhs0, int dc, double* gvals0) {
      hipLaunchKernelGGL(k_root_cg_vjp<T>, dim3(1), dim3(64), 0, s, cols0, vals0, Rn[0], (const int64_t*)r.levels[0].basis_keys.data(),
                         (const T*)r.root_b[cycle].data(), (const T*)r.root_x[cycle].data(), dc, base, cycle, r.prm.cg_accuracy,
                         maxiter, cg_store.data(), gx0, grhs0, gvals0);
      KERNEL_CHECK();
    };
    if (base == 0 && Rn[0] && !r.cg_runs.empty()) {
      // the whole base solve is the root solve, once per cycle, x carried
      const int dc = r.prm.point_weight == 0.0 ? 1 : 0;
      DArray<double> gpad(64, s);
      double* g = gxb.data();
      if (dc) {
        hipLaunchKernelGGL(k_root_pad_grad, dim3(1), dim3(64), 0, s, (const int64_t*)r.levels[0].basis_keys.data(), Rn[0],
                           (const double*)gxb.data(), gpad.data());
        KERNEL_CHECK();
        g = gpad.data();
      }
      for (int cycle = r.prm.base_v_cycles - 1; cycle >= 0; --cycle)
        if (ran(cycle)) root_reverse(cycle, (const int*)r.levels[0].A.cols.data(), (const T*)r.levels[0].A.vals.data(), g,
                                     grhs[base].data(), dc, mg ? mg->gA[0].data() : (double*)nullptr);
    }
    if (base > 0) {
      // rebuild the Galerkin hierarchy, schedules and operators
      std::vector<DArray<int>> mcols(base + 1);
      std::vector<DArray<T>> mvals(base + 1);
      LevelState<T>& B = r.levels[base];
      mcols[base].reset(27 * (size_t)Rn[base], s);
      mvals[base].reset(27 * (size_t)Rn[base], s);
      HIP_CHECK(hipMemcpyAsync(mcols[base].data(), B.A.cols.data(), sizeof(int) * 27 * Rn[base], hipMemcpyDeviceToDevice, s));
      HIP_CHECK(hipMemcpyAsync(mvals[base].data(), B.A.vals.data(), sizeof(T) * 27 * Rn[base], hipMemcpyDeviceToDevice, s));
      for (int d = base; d > 0; --d) {
        const int C = Rn[d - 1], F = Rn[d];
        mcols[d - 1].reset(27 * (size_t)C, s);
        mvals[d - 1].reset(27 * (size_t)C, s);
        if (C) {
          hipLaunchKernelGGL(k_galerkin<T>, dim3(C), dim3(128), 0, s, (const int*)r.levels[d - 1].restriction.data(),
                             C, (const int*)mcols[d].data(), (const T*)mvals[d].data(), F, (const int*)r.levels[d].P.cols.data(),
                             (const T*)r.levels[d].P.vals.data(), mcols[d - 1].data(), mvals[d - 1].data(),
                             (const int64_t*)r.levels[d - 1].basis_keys.data());
          KERNEL_CHECK();
        }
      }
      // gather_operator reads b over each level's rows
      DArray<T> zero_b(std::max(1, *std::max_element(Rn.begin(), Rn.end())), s);
      zero_b.zero();
      std::vector<Schedule<T>> sched;
      std::vector<GsOperator<T>> ops;
      for (int d = 0; d <= base; ++d) {
        sched.push_back(make_schedule<T>(r.levels[d].basis_keys, Rn[d], d, false, s, false));
        ops.push_back(gather_operator<T>(sched[d], mcols[d], mvals[d], zero_b.data(), Rn[d], s));
      }
      const int iters = r.prm.iters, cycles = r.prm.base_v_cycles;
      // replay of the V-cycles (regular_multigrid) with x snapshots
      std::vector<std::vector<DArray<T>>> snap_down(cycles), snap_up(cycles);
      std::vector<DArray<double>> gM(base + 1);
      if (mg) {
        std::vector<DArray<T>> rhs(base + 1), x(base + 1);
        for (int d = 0; d <= base; ++d) {
          rhs[d].reset(Rn[d], s);
          rhs[d].zero();
          x[d].reset(Rn[d], s);
          x[d].zero();
          gM[d].reset(27 * (size_t)Rn[d], s);
          gM[d].zero();
        }
        if (Rn[base]) HIP_CHECK(hipMemcpyAsync(rhs[base].data(), B.rhs.data(), sizeof(T) * Rn[base], hipMemcpyDeviceToDevice, s));
        DArray<T> pad_x(64, s), work(256, s), residual;
        pad_x.zero();
        DArray<int> flag(1, s);
        for (int cycle = 0; cycle < cycles; ++cycle) {
          snap_down[cycle].resize(base + 1);
          snap_up[cycle].resize(base + 1);
          for (int d = base; d > 0; --d) {
            refresh_rhs(ops[d], sched[d], rhs[d].data(), s);
            snap_down[cycle][d].reset((size_t)(iters + 1) * Rn[d], s);
            relax_colored_snap(ops[d], sched[d], x[d].data(), iters, true, snap_down[cycle][d].data(), s);
            residual.reset(Rn[d], s);
            hipLaunchKernelGGL(k_spmv<T>, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, (const int*)mcols[d].data(),
                               (const T*)mvals[d].data(), Rn[d], 27, (const T*)x[d].data(), residual.data(), T(-1), T(1),
                               (const T*)rhs[d].data());
            KERNEL_CHECK();
            if (Rn[d - 1]) {
              hipLaunchKernelGGL(k_restrict_apply<T>, dim3(blocks_for(Rn[d - 1], 256)), dim3(256), 0, s,
                                 (const int*)r.levels[d - 1].restriction.data(), Rn[d - 1], (const T*)residual.data(),
                                 rhs[d - 1].data(), 0);
              KERNEL_CHECK();
            }
            x[d - 1].zero();
          }
          hipLaunchKernelGGL(k_root_solve<T>, dim3(1), dim3(64), 0, s, (const int*)mcols[0].data(), (const T*)mvals[0].data(), Rn[0],
                             (const int64_t*)r.levels[0].basis_keys.data(), (const T*)rhs[0].data(), x[0].data(), pad_x.data(),
                             work.data(), 0, base, cycle, r.prm.cg_accuracy, Rn[0], flag.data());
          KERNEL_CHECK();
          for (int d = 1; d <= base; ++d) {
            hipLaunchKernelGGL(k_spmv<T>, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, (const int*)r.levels[d].P.cols.data(),
                               (const T*)r.levels[d].P.vals.data(), Rn[d], 8, (const T*)x[d - 1].data(), x[d].data(), T(1), T(1),
                               (const T*)nullptr);
            KERNEL_CHECK();
            x[d - 1].zero();
            snap_up[cycle][d].reset((size_t)(iters + 1) * Rn[d], s);
            relax_colored_snap(ops[d], sched[d], x[d].data(), iters, false, snap_up[cycle][d].data(), s);
          }
        }
      }
      std::vector<DArray<double>> gxl(base + 1), grl(base + 1);
      for (int d = 0; d < base; ++d) {
        gxl[d].reset(Rn[d], s);
        grl[d].reset(Rn[d], s);
      }
      gxl[base] = std::move(gxb);
      grl[base].reset(Rn[base], s);
      for (int cycle = cycles - 1; cycle >= 0; --cycle) {
        for (int d = 0; d < base; ++d) {
          gxl[d].zero();
          grl[d].zero();
        }
        grl[base].zero();
        // up: x_d += P_d x_{d-1}, then reverse-colour sweeps
        for (int d = base; d >= 1; --d) {
          if (mg)
            gs_colored_reverse_m<T>(ops[d], sched[d], iters, false, (const int64_t*)r.levels[d].basis_keys.data(),
                                    (const T*)snap_up[cycle][d].data(), gxl[d].data(), grl[d].data(), gM[d].data(), s);
          else
            gs_colored_reverse<T>(ops[d], sched[d], iters, false, gxl[d].data(), grl[d].data(), s);
          stencil_scatter<T>(nullptr, r.levels[d].P, 8, (const double*)gxl[d].data(), 1.0, gxl[d - 1].data(), s);
        }
        // root: when skipped x_0 stays zero; when its CG ran, reverse it
        if (ran(cycle) && Rn[0]) root_reverse(cycle, (const int*)mcols[0].data(), (const T*)mvals[0].data(), gxl[0].data(),
                                              grl[0].data(), 0, mg ? gM[0].data() : (double*)nullptr);
        // down: forward sweeps, residual, restriction
        for (int d = 1; d <= base; ++d) {
          DArray<double> gres = zeros_d(Rn[d], s);
          if (Rn[d - 1]) {
            hipLaunchKernelGGL(k_restrict_scatter, dim3(blocks_for(Rn[d - 1], 256)), dim3(256), 0, s,
                               (const int*)r.levels[d - 1].restriction.data(), Rn[d - 1], (const double*)grl[d - 1].data(),
                               gres.data());
            KERNEL_CHECK();
          }
          if (Rn[d]) {
            hipLaunchKernelGGL(k_axpy_d, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, grl[d].data(), (const double*)gres.data(),
                               1.0, (int64_t)Rn[d]);
            KERNEL_CHECK();
            // A_d^T over the level's canonical stencil
            hipLaunchKernelGGL(k_stencil_scatter<T>, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, (const int*)nullptr,
                               (const int*)mcols[d].data(), (const T*)mvals[d].data(), Rn[d], 27, (const double*)gres.data(), -1.0,
                               gxl[d].data());
            KERNEL_CHECK();
            if (mg) {  // residual = rhs - A x_after_down
              hipLaunchKernelGGL(k_stencil_outer<T>, dim3(blocks_for(Rn[d], 256)), dim3(256), 0, s, (const int*)nullptr,
                                 (const int*)mcols[d].data(), Rn[d], 27, (const double*)gres.data(),
                                 (const T*)snap_down[cycle][d].data() + (size_t)iters * Rn[d], -1.0, gM[d].data());
              KERNEL_CHECK();
            }
          }
          if (mg)
            gs_colored_reverse_m<T>(ops[d], sched[d], iters, true, (const int64_t*)r.levels[d].basis_keys.data(),
                                    (const T*)snap_down[cycle][d].data(), gxl[d].data(), grl[d].data(), gM[d].data(), s);
          else
            gs_colored_reverse<T>(ops[d], sched[d], iters, true, gxl[d].data(), grl[d].data(), s);
        }
        if (Rn[base]) {
          hipLaunchKernelGGL(k_axpy_d, dim3(blocks_for(Rn[base], 256)), dim3(256), 0, s, grhs[base].data(),
                             (const double*)grl[base].data(), 1.0, (int64_t)Rn[base]);
          KERNEL_CHECK();
        }
      }
      if (mg) {
        // Galerkin products A_{d-1} = R A_d P, coarse to fine
        for (int d = 1; d <= base; ++d) {
          const int C = Rn[d - 1], F = Rn[d];
          if (!C || !F) continue;
          hipLaunchKernelGGL(k_galerkin_vjp<T>, dim3(C), dim3(128), 0, s, (const int*)r.levels[d - 1].restriction.data(), C,
                             (const int*)mcols[d].data(), F, (const int*)r.levels[d].P.cols.data(),
                             (const T*)r.levels[d].P.vals.data(), (const int64_t*)r.levels[d - 1].basis_keys.data(),
                             (const double*)gM[d - 1].data(), gM[d].data());
          KERNEL_CHECK();
        }
        mg->gA[base] = std::move(gM[base]);
      }
    }
  }
  // right-hand side = fp(load) + fp(0.5) * fp(point_load) (solver.py:119-126)
  normal_loads_backward(r, grhs, top, grecords, s);
  if (mg) mg->grhs = std::move(grhs);
}

// Screening reversed per level (see the declaration).
template <class T>
void screening_backward(Result<T>& r, const SolveGrads& mg, std::vector<DArray<double>>& gagg, double* gavg1, hipStream_t s) {
  const int base = r.prm.base_depth;
  gagg.resize(r.prm.depth + 1);
  for (int d = base; d <= mg.top; ++d) {
    LevelState<T>& L = r.levels[d];
    const int R = L.basis_count;
    if (!R || mg.gA[d].empty() || mg.grhs[d].empty()) continue;
    PointSet<T> pts = point_constraints(r, d, s, false);
    if (!pts.P) continue;
    const int* ord = d > base ? (const int*)L.order.data() : (const int*)nullptr;
    // the point load enters b = T(load) + T(0.5) T(point_load), in the load's layout
    DArray<double> gload(R, s);
    hipLaunchKernelGGL(k_half_gather, dim3(blocks_for(R, 256)), dim3(256), 0, s, ord, R, (const double*)mg.grhs[d].data(),
                       gload.data());
    KERNEL_CHECK();
    DArray<double> gpos = zeros_d(3 * (size_t)pts.P, s), gw = zeros_d(pts.P, s);
    // the operator and point load at schedule positions (finer levels), the
    // cross operator at canonical rows
    hipLaunchKernelGGL(k_screening_vjp<T>, dim3(blocks_for(R, 128)), dim3(128), 0, s, (const int64_t*)L.basis_keys.data(), R, ord, d,
                       point_bins(pts), (const double*)mg.gA[d].data(), (const double*)nullptr, (const double*)gload.data(),
                       gpos.data(), gw.data());
    KERNEL_CHECK();
    if (d > base && !mg.gC[d].empty()) {
      hipLaunchKernelGGL(k_screening_vjp<T>, dim3(blocks_for(R, 128)), dim3(128), 0, s, (const int64_t*)L.basis_keys.data(), R,
                         (const int*)nullptr, d, point_bins(pts), (const double*)nullptr, (const double*)mg.gC[d].data(),
                         (const double*)nullptr, gpos.data(), gw.data());
      KERNEL_CHECK();
    }
    if (gagg[d].empty()) gagg[d] = zeros_d(3 * (size_t)pts.P, s);
    hipLaunchKernelGGL(k_points_vjp<T>, dim3(blocks_for(pts.P, 256)), dim3(256), 0, s, (const T*)L.agg_vals.data(),
                       (const T*)pts.w.data(), pts.P, d, r.prm.point_weight, (const double*)gpos.data(), (const double*)gw.data(),
                       gagg[d].data(), gavg1);
    KERNEL_CHECK();
  }
}

template void relax_sliced_csr<float>(const int64_t*, int, const int*, const int*, const float*, const float*, float*, int, int,
                                      hipStream_t);
template void relax_scheduled_csr<float>(const int64_t*, int, const int*, const int*, const float*, const float*, float*, int,
                                          int, int, int, hipStream_t);
template void relax_scheduled_csr<double>(const int64_t*, int, const int*, const int*, const double*, const double*, double*,
                                           int, int, int, int, hipStream_t);
template void relax_sliced_csr<double>(const int64_t*, int, const int*, const int*, const double*, const double*, double*, int,
                                       int, hipStream_t);

template void stage_solve<float>(Result<float>&);
template void stage_solve<double>(Result<double>&);
template void solve_backward<float>(Result<float>&, const double*, double*, hipStream_t, SolveGrads*);
template void solve_backward<double>(Result<double>&, const double*, double*, hipStream_t, SolveGrads*);
template void screening_backward<float>(Result<float>&, const SolveGrads&, std::vector<DArray<double>>&, double*, hipStream_t);
template void screening_backward<double>(Result<double>&, const SolveGrads&, std::vector<DArray<double>>&, double*, hipStream_t);

}  // namespace psr
