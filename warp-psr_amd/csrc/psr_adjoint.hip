// Reverse mode of the reconstruction with respect to the input normals, with
// the octree, sample grouping, splat depths and solver branches held fixed
// (src/warp_psr/reconstruction_adjoint.py:61-185): field queries
// (field.py:214-286), the isovalue (field_adjoint.py:134-142), cumulative
// coefficients (field_adjoint.py:20-36), the cascadic solve and normal loads
// (solve_backward), the normal splats (normals.py:46-104) and the unit
// normals of the samples (samples.py:90-134).  Gradients are float64;
// scatters to shared entries use float64 atomics, as the reference's reverse
// kernels do, so their last bits depend on the atomic order.
#include "psr_field_index.h"
#include "psr_splat.h"

#include <cmath>

namespace psr {

namespace {

template <class T> PSR_HD int clamp_i(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// field.py:214-286 (query_vjp, reference evaluation) at one cube point p:
// coefficient / cumulative gradients (atomics) and the cube-space query
// gradient.  Lookup decisions are the forward's (same T arithmetic).
template <class T>
__device__ void query_vjp(const FieldIndex& f, int depth, const T pin[3], double gv, const double gg[3], const T* coef,
                          const T* cum, double* gcoef, double* gcum, double qp[3]) {
  qp[0] = qp[1] = qp[2] = 0.0;
  if (gv == 0.0 && gg[0] == 0.0 && gg[1] == 0.0 && gg[2] == 0.0) return;
  T p[3] = {pin[0], pin[1], pin[2]};
  const T lo = tmin(tmin(p[0], p[1]), p[2]), hi = tmax(tmax(p[0], p[1]), p[2]);
  if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && lo >= T(0) && hi <= T(1))) return;
  int leaf = 0;
  for (int d = depth; d >= 0; --d) {
    const int n = 1 << d;
    const int c0 = clamp_i<T>((int)ceil(p[0] * T(n)) - 1, 0, n - 1);
    const int c1 = clamp_i<T>((int)ceil(p[1] * T(n)) - 1, 0, n - 1);
    const int c2 = clamp_i<T>((int)ceil(p[2] * T(n)) - 1, 0, n - 1);
    int64_t idx;
    level_run<1>(f, d, c0, c1, c2, 1, &idx);
    if (idx >= 0) { leaf = d; break; }
  }
  for (int a = 0; a < 3; ++a) {
    if (p[a] == T(0)) p[a] = T(1e-6);
    else if (p[a] == T(1)) p[a] = T(1.0 - 1e-6);
  }
  const int begin = leaf > 0 ? leaf - 1 : 0, end = (leaf < depth ? leaf + 1 : depth) + 1;
  for (int d = begin; d < end; ++d) {
    const int n = 1 << d;
    int cell[3], base[3];
    double fr[3];
    for (int a = 0; a < 3; ++a) {
      const T t = p[a] * T(n);
      cell[a] = clamp_i<T>((int)ceil(t) - 1, 0, n - 1);
      base[a] = (int)floor(t);
      fr[a] = double(t - T(base[a]));
    }
    const bool cumulative_level = d == leaf - 1;
    for (int corner = 0; corner < 8; ++corner) {
      const int node[3] = {cell[0] + corner / 4, cell[1] + (corner / 2) % 2, cell[2] + corner % 2};
      if (imin(imin(node[0] - base[0], node[1] - base[1]), node[2] - base[2]) < 0) continue;
      int64_t index;
      level_run<1>(f, d, node[0], node[1], node[2], 1, &index);
      if (index < 0) continue;
      double h[3], sg[3];
      for (int a = 0; a < 3; ++a) {
        h[a] = 1.0 - fr[a];
        sg[a] = -1.0;
        if (node[a] > base[a]) { h[a] = fr[a]; sg[a] = 1.0; }
      }
      const double bg[3] = {double(n) * sg[0] * h[1] * h[2], double(n) * h[0] * sg[1] * h[2], double(n) * h[0] * h[1] * sg[2]};
      const double gc = gv * h[0] * h[1] * h[2] + gg[0] * bg[0] + gg[1] * bg[1] + gg[2] * bg[2];
      double c;
      if (cumulative_level) {
        c = double(cum[index]);
        if (gcum) atomicAdd(&gcum[index], gc);
      } else {
        c = double(coef[index]);
        if (gcoef) atomicAdd(&gcoef[index], gc);
      }
      const double nn = double(n) * double(n);
      const double xy = c * nn * sg[0] * sg[1] * h[2], xz = c * nn * sg[0] * h[1] * sg[2], yz = c * nn * h[0] * sg[1] * sg[2];
      qp[0] += gv * c * bg[0] + (gg[1] * xy + gg[2] * xz);
      qp[1] += gv * c * bg[1] + (gg[0] * xy + gg[2] * yz);
      qp[2] += gv * c * bg[2] + (gg[0] * xz + gg[1] * yz);
    }
  }
  for (int a = 0; a < 3; ++a)
    if (pin[a] == T(0) || pin[a] == T(1)) qp[a] = 0.0;  // the nudged coordinate is a constant
}

// Model-space queries (cube = L q + t; the returned gradient is L^T of the
// cube gradient): value/gradient seeds -> coefficient, cumulative and query
// gradients (field_adjoint.py:62-104, 144-154).
template <class T>
__global__ void k_query_vjp(FieldIndex f, int depth, const T* __restrict__ coef, const T* __restrict__ cum,
                            const T* __restrict__ queries, int64_t m, Affine<T> A, const T* __restrict__ gvalue,
                            const T* __restrict__ ggrad, double* __restrict__ gcoef, double* __restrict__ gcum,
                            T* __restrict__ gquery) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T q0 = queries[3 * i], q1 = queries[3 * i + 1], q2 = queries[3 * i + 2];
  T p[3];
  for (int r = 0; r < 3; ++r) p[r] = A.L[3 * r] * q0 + A.L[3 * r + 1] * q1 + A.L[3 * r + 2] * q2 + A.t[r];
  const double gv = gvalue ? double(gvalue[i]) : 0.0;
  double gg[3] = {0.0, 0.0, 0.0};
  if (ggrad) {  // model gradient = L^T cube gradient, so the cube seed is L * seed
    const double s0 = double(ggrad[3 * i]), s1 = double(ggrad[3 * i + 1]), s2 = double(ggrad[3 * i + 2]);
    for (int r = 0; r < 3; ++r) gg[r] = double(A.L[3 * r]) * s0 + double(A.L[3 * r + 1]) * s1 + double(A.L[3 * r + 2]) * s2;
  }
  double qp[3];
  query_vjp<T>(f, depth, p, gv, gg, coef, cum, gcoef, gcum, qp);
  if (gquery)
    for (int c = 0; c < 3; ++c)
      gquery[3 * i + c] += T(double(A.L[c]) * qp[0] + double(A.L[3 + c]) * qp[1] + double(A.L[6 + c]) * qp[2]);
}

// Isovalue = sum w v(centroid) / sum w (reconstruction.py:101-116): each
// centroid value receives seed * w / mass (field_adjoint.py:134-142); the
// centroids are sums / w in cube space.
template <class T>
__global__ void k_iso_vjp(FieldIndex f, int depth, const T* __restrict__ coef, const T* __restrict__ cum,
                          const T* __restrict__ sums, const T* __restrict__ weights, int G, double scale,
                          double* __restrict__ gcoef, double* __restrict__ gcum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const T w = weights[i];
  const T p[3] = {sums[3 * i] / w, sums[3 * i + 1] / w, sums[3 * i + 2] / w};
  const double gg[3] = {0.0, 0.0, 0.0};
  double qp[3];
  query_vjp<T>(f, depth, p, scale * double(w), gg, coef, cum, gcoef, gcum, qp);
}

__global__ void k_sum_weights(const void* __restrict__ weights, int is_double, int G, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double total = 0.0;
  for (int i = 0; i < G; ++i) total += is_double ? ((const double*)weights)[i] : double(((const float*)weights)[i]);
  *out = total;
}

// field_adjoint.py:20-36 (cumulative_backward), as a gather: coarse node j at
// d-1 collects weight * work of its fine nodes 2j + (-1, 0, 1) per axis.
__global__ void k_cumulative_vjp(FieldIndex f, int d, double* __restrict__ work) {
  const int64_t j = f.off[d - 1] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= f.off[d]) return;
  const int64_t key = f.keys[j];
  const int x = key_x(key), y = key_y(key), z = key_z(key);
  double acc = 0.0;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b) {
      const int fx = 2 * x + a, fy = 2 * y + b;
      if (fx < 0 || fy < 0) continue;
      int64_t found[3];
      const int z0 = 2 * z - 1;
      if (z0 >= 0) {
        level_run<3>(f, d, fx, fy, z0, 3, found);
      } else {
        found[0] = -1;
        level_run<2>(f, d, fx, fy, 0, 2, found + 1);
      }
      for (int c = 0; c < 3; ++c) {
        if (found[c] < 0) continue;
        const double w = (a ? 0.5 : 1.0) * (b ? 0.5 : 1.0) * (c != 1 ? 0.5 : 1.0);
        acc += w * work[found[c]];
      }
    }
  work[j] += acc;
}

__global__ void k_add_d(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

// normals.py:46-104 reversed for the normal sums: every valid splat of
// aggregate src adds scale_k * (-record gradient) (records are negated sums)
// to d(value)/d(n'), n' = nsums/w / |nsums/w| * w, then through the
// normalisation to nsums.
template <class T>
__global__ void k_normal_splat_vjp(int G, const T* __restrict__ sums, const T* __restrict__ nsums, const T* __restrict__ weights,
                                   const T* __restrict__ dw, int minimum, int maximum, const int64_t* __restrict__ nkeys,
                                   const int64_t* __restrict__ noff, const double* __restrict__ grec,
                                   double* __restrict__ gnsums) {
  const int src = blockIdx.x * blockDim.x + threadIdx.x;
  if (src >= G) return;
  double gn[3] = {0.0, 0.0, 0.0};
  bool any = false;
  for (int k = 0; k < 54; ++k) {
    const NormalRecord<T> rec = normal_record<T>(src, k, sums, nsums, weights, dw, minimum, maximum, true);
    if (!rec.valid) continue;
    const int64_t lo = noff[rec.top], hi = noff[rec.top + 1];
    const int64_t idx = find_key(nkeys + lo, hi - lo, basis_key(rec.top, rec.x, rec.y, rec.z));
    if (idx < 0) continue;
    const double* g = grec + 3 * (lo + idx);
    gn[0] -= rec.scale * g[0];
    gn[1] -= rec.scale * g[1];
    gn[2] -= rec.scale * g[2];
    any = true;
  }
  if (!any) return;
  const double w = double(weights[src]);
  const double m0 = double(nsums[3 * src]) / w, m1 = double(nsums[3 * src + 1]) / w, m2 = double(nsums[3 * src + 2]) / w;
  const double len = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
  if (!(len > 0.0)) return;
  const double dot = m0 * gn[0] + m1 * gn[1] + m2 * gn[2];
  const double l3 = len * len * len;
  // n'_c = m_c / len * w and m = nsums / w: d n'_c / d nsums_e = (delta_ce / len - m_c m_e / len^3)
  gnsums[3 * src] += gn[0] / len - m0 * dot / l3;
  gnsums[3 * src + 1] += gn[1] / len - m1 * dot / l3;
  gnsums[3 * src + 2] += gn[2] / len - m2 * dot / l3;
}

// samples.py:90-134 reversed for the normals: nsums[g] = sum of unit normals
// u = n / |n| with n = (N1, N2, N0) * scale over the group's accepted inputs
template <class T>
__global__ void k_unit_normal_vjp(const T* __restrict__ nrm, int64_t n, T scale, const int* __restrict__ group,
                                  const int* __restrict__ status, const double* __restrict__ gnsums, T* __restrict__ gnormals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = group[i];
  if (g < 0 || status[i] != 0) return;
  const double gu[3] = {gnsums[3 * g], gnsums[3 * g + 1], gnsums[3 * g + 2]};
  const double s = double(scale);
  const double v[3] = {double(nrm[3 * i + 1]) * s, double(nrm[3 * i + 2]) * s, double(nrm[3 * i]) * s};
  const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (!(l > 0.0)) return;
  const double u[3] = {v[0] / l, v[1] / l, v[2] / l};
  const double ug = u[0] * gu[0] + u[1] * gu[1] + u[2] * gu[2];
  double gv[3];
  for (int c = 0; c < 3; ++c) gv[c] = (gu[c] - u[c] * ug) / l;
  gnormals[3 * i + 1] += T(gv[0] * s);
  gnormals[3 * i + 2] += T(gv[1] * s);
  gnormals[3 * i] += T(gv[2] * s);
}

}  // namespace

template <class T>
void field_vjp(Result<T>& r, const T* queries, int64_t m, const T* gvalue, const T* ggrad, double* gcoef, double* gcum,
               T* gquery, hipStream_t s) {
  if (m <= 0) return;
  Affine<T> A;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) A.L[3 * i + j] = r.model_to_cube[4 * i + j];
    A.t[i] = r.model_to_cube[4 * i + 3];
  }
  hipLaunchKernelGGL(k_query_vjp<T>, dim3(blocks_for(m, 256)), dim3(256), 0, s, field_index(r), r.prm.depth,
                     (const T*)r.coefficients.data(), (const T*)r.cumulative.data(), queries, m, A, gvalue, ggrad, gcoef, gcum,
                     gquery);
  KERNEL_CHECK();
}

template <class T>
void backward_normals(Result<T>& r, const double* gcoef_in, const double* gcum_in, double iso_seed, T* gnormals,
                      const T* normals, int64_t n, hipStream_t s) {
  const int64_t K = r.level_offsets.back();
  const int D = r.prm.depth;
  // coefficient gradient = direct seeds + cumulative chain (+ isovalue)
  DArray<double> gcoef(K, s), work(K, s);
  gcoef.zero();
  work.zero();
  if (gcoef_in) HIP_CHECK(hipMemcpyAsync(gcoef.data(), gcoef_in, sizeof(double) * K, hipMemcpyDeviceToDevice, s));
  if (gcum_in) HIP_CHECK(hipMemcpyAsync(work.data(), gcum_in, sizeof(double) * K, hipMemcpyDeviceToDevice, s));
  if (iso_seed != 0.0 && r.G) {
    DArray<double> mass(1, s);
    hipLaunchKernelGGL(k_sum_weights, dim3(1), dim3(64), 0, s, (const void*)r.weights.data(), sizeof(T) == 8 ? 1 : 0, r.G,
                       mass.data());
    KERNEL_CHECK();
    double mass_h = 0.0;
    read_bytes(&mass_h, mass.data(), sizeof(double), s);
    hipLaunchKernelGGL(k_iso_vjp<T>, dim3(blocks_for(r.G, 256)), dim3(256), 0, s, field_index(r), D, (const T*)r.coefficients.data(),
                       (const T*)r.cumulative.data(), (const T*)r.position_sums.data(), (const T*)r.weights.data(), r.G,
                       iso_seed / mass_h, gcoef.data(), work.data());
    KERNEL_CHECK();
  }
  const FieldIndex f = field_index(r);
  for (int d = D; d >= 1; --d) {
    const int cnt = r.level_offsets[d] - r.level_offsets[d - 1];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_cumulative_vjp, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, f, d, work.data());
    KERNEL_CHECK();
  }
  if (K) {
    hipLaunchKernelGGL(k_add_d, dim3(blocks_for(K, 256)), dim3(256), 0, s, gcoef.data(), (const double*)work.data(), K);
    KERNEL_CHECK();
  }
  work.free();
  // solve and normal loads -> normal records
  const int64_t records = (int64_t)r.normal_keys.size();
  DArray<double> grec(3 * (size_t)records, s);
  grec.zero();
  solve_backward(r, (const double*)gcoef.data(), grec.data(), s);
  // normal splats -> normal sums -> input normals
  DArray<int64_t> noff(r.normal_offsets.size(), s);
  std::vector<int64_t> ho(r.normal_offsets.begin(), r.normal_offsets.end());
  HIP_CHECK(hipMemcpyAsync(noff.data(), ho.data(), sizeof(int64_t) * ho.size(), hipMemcpyHostToDevice, s));
  DArray<double> gns(3 * (size_t)r.G, s);
  gns.zero();
  hipLaunchKernelGGL(k_normal_splat_vjp<T>, dim3(blocks_for(r.G, 256)), dim3(256), 0, s, r.G, (const T*)r.position_sums.data(),
                     (const T*)r.normal_sums.data(), (const T*)r.weights.data(), (const T*)r.dw.data(), r.prm.base_depth, D,
                     (const int64_t*)r.normal_keys.data(), (const int64_t*)noff.data(), (const double*)grec.data(), gns.data());
  KERNEL_CHECK();
  hipLaunchKernelGGL(k_unit_normal_vjp<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, normals, n, r.transform[4],
                     (const int*)r.input_group.data(), (const int*)r.input_status.data(), (const double*)gns.data(), gnormals);
  KERNEL_CHECK();
  HIP_CHECK(hipStreamSynchronize(s));  // the temporaries above return to the pool
}

template void field_vjp<float>(Result<float>&, const float*, int64_t, const float*, const float*, double*, double*, float*,
                               hipStream_t);
template void field_vjp<double>(Result<double>&, const double*, int64_t, const double*, const double*, double*, double*, double*,
                                hipStream_t);
template void backward_normals<float>(Result<float>&, const double*, const double*, double, float*, const float*, int64_t,
                                      hipStream_t);
template void backward_normals<double>(Result<double>&, const double*, const double*, double, double*, const double*, int64_t,
                                       hipStream_t);

}  // namespace psr
