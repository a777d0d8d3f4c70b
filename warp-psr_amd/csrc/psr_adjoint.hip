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
                          const T* cum, double* gcoef, double* gcum, double qp[3], double fg[3] = nullptr) {
  qp[0] = qp[1] = qp[2] = 0.0;
  if (fg) fg[0] = fg[1] = fg[2] = 0.0;
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
      if (fg)
        for (int a = 0; a < 3; ++a) fg[a] += c * bg[a];
    }
  }
  for (int a = 0; a < 3; ++a)
    if (pin[a] == T(0) || pin[a] == T(1)) qp[a] = 0.0;  // the nudged coordinate is a constant
}

// Model-space queries (cube = L q + t; the returned gradient is L^T of the
// cube gradient): value/gradient seeds -> coefficient, cumulative and query
// gradients (field_adjoint.py:62-104, 144-154).
// With gtransform, the loss's dependence on the normalisation through the
// query mapping is reduced into gtransform[4] = d/d(inv, t1, t2, t3): the
// cube point is (inv q1 + t1, inv q2 + t2, inv q0 + t3) (the model-to-cube
// map, transforms.py:66-79) and the model gradient is L^T grad f.
template <class T>
__global__ void __launch_bounds__(256) k_query_vjp(FieldIndex f, int depth, const T* __restrict__ coef, const T* __restrict__ cum,
                                                   const T* __restrict__ queries, int64_t m, Affine<T> A,
                                                   const T* __restrict__ gvalue, const T* __restrict__ ggrad,
                                                   double* __restrict__ gcoef, double* __restrict__ gcum,
                                                   T* __restrict__ gquery, double* __restrict__ gtransform) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < m) {
    const T q0 = queries[3 * i], q1 = queries[3 * i + 1], q2 = queries[3 * i + 2];
    T p[3];
    for (int r = 0; r < 3; ++r) p[r] = A.L[3 * r] * q0 + A.L[3 * r + 1] * q1 + A.L[3 * r + 2] * q2 + A.t[r];
    const double gv = gvalue ? double(gvalue[i]) : 0.0;
    double gg[3] = {0.0, 0.0, 0.0}, sm[3] = {0.0, 0.0, 0.0};
    if (ggrad) {  // model gradient = L^T cube gradient, so the cube seed is L * seed
      for (int c = 0; c < 3; ++c) sm[c] = double(ggrad[3 * i + c]);
      for (int r = 0; r < 3; ++r) gg[r] = double(A.L[3 * r]) * sm[0] + double(A.L[3 * r + 1]) * sm[1] + double(A.L[3 * r + 2]) * sm[2];
    }
    double qp[3], fg[3];
    query_vjp<T>(f, depth, p, gv, gg, coef, cum, gcoef, gcum, qp, gtransform ? fg : nullptr);
    if (gquery)
      for (int c = 0; c < 3; ++c)
        gquery[3 * i + c] += T(double(A.L[c]) * qp[0] + double(A.L[3 + c]) * qp[1] + double(A.L[6 + c]) * qp[2]);
    if (gtransform) {
      part[0] = qp[0] * double(q1) + qp[1] * double(q2) + qp[2] * double(q0) + sm[1] * fg[0] + sm[2] * fg[1] + sm[0] * fg[2];
      part[1] = qp[0];
      part[2] = qp[1];
      part[3] = qp[2];
    }
  }
  if (!gtransform) return;
  __shared__ double red[4][256];
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = part[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 4) atomicAdd(&gtransform[threadIdx.x], red[threadIdx.x][0]);
}

// Isovalue = sum w v(centroid) / sum w (reconstruction.py:101-116): each
// centroid value receives seed * w / mass (field_adjoint.py:134-142); the
// centroids are sums / w in cube space.
// With gsums, the centroid's own gradient (seed * w / mass * grad f) reaches
// the position sums through p = sums / w.
template <class T>
__global__ void k_iso_vjp(FieldIndex f, int depth, const T* __restrict__ coef, const T* __restrict__ cum,
                          const T* __restrict__ sums, const T* __restrict__ weights, int G, double scale,
                          double* __restrict__ gcoef, double* __restrict__ gcum, double* __restrict__ gsums) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const T w = weights[i];
  const T p[3] = {sums[3 * i] / w, sums[3 * i + 1] / w, sums[3 * i + 2] / w};
  const double gg[3] = {0.0, 0.0, 0.0};
  double qp[3];
  query_vjp<T>(f, depth, p, scale * double(w), gg, coef, cum, gcoef, gcum, qp);
  if (gsums)
    for (int a = 0; a < 3; ++a) gsums[3 * i + a] += qp[a] / double(w);
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

// normals.py:46-104 reversed: every valid splat of aggregate src adds
// scale_k * (-record gradient) (records are negated sums) to d(value)/d(n'),
// n' = nsums/w / |nsums/w| * w, then through the normalisation to nsums.
// With gsums / gdw, the record scale's derivatives (centroid, splat depth,
// weight) reach the position sums and the density's (depth, weight).
template <class T>
__global__ void k_normal_splat_vjp(int G, const T* __restrict__ sums, const T* __restrict__ nsums, const T* __restrict__ weights,
                                   const T* __restrict__ dw, int minimum, int maximum, const int64_t* __restrict__ nkeys,
                                   const int64_t* __restrict__ noff, const double* __restrict__ grec,
                                   double* __restrict__ gnsums, double* __restrict__ gsums, double* __restrict__ gdw) {
  const int src = blockIdx.x * blockDim.x + threadIdx.x;
  if (src >= G) return;
  double gn[3] = {0.0, 0.0, 0.0}, gp[3] = {0.0, 0.0, 0.0}, gdepth = 0.0, gwgt = 0.0;
  bool any = false;
  for (int k = 0; k < 54; ++k) {
    const RecordJacobian J = normal_record_jacobian<T>(src, k, sums, nsums, weights, dw, minimum, maximum);
    if (!J.valid) continue;
    const int64_t lo = noff[J.top], hi = noff[J.top + 1];
    const int64_t idx = find_key(nkeys + lo, hi - lo, basis_key(J.top, J.x, J.y, J.z));
    if (idx < 0) continue;
    const double* g = grec + 3 * (lo + idx);
    const double gv[3] = {-g[0], -g[1], -g[2]};  // gradient of the record's un-negated value
    gn[0] += J.s * gv[0];
    gn[1] += J.s * gv[1];
    gn[2] += J.s * gv[2];
    const double dot = J.n[0] * gv[0] + J.n[1] * gv[1] + J.n[2] * gv[2];
    for (int a = 0; a < 3; ++a) gp[a] += dot * J.ds_dp[a];
    gdepth += dot * J.ds_ddepth;
    gwgt += dot * J.ds_dwgt;
    any = true;
  }
  if (!any) return;
  const double w = double(weights[src]);
  if (gsums)
    for (int a = 0; a < 3; ++a) gsums[3 * src + a] += gp[a] / w;
  if (gdw) {
    gdw[2 * src] += gdepth;
    gdw[2 * src + 1] += gwgt;
  }
  if (!gnsums) return;
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

// density.py:152-182 (_depths) reversed: the selected level d is the
// forward's (same T comparisons); raw depth and weight h^2 / density feed
// back to density[d] and, below the target, density[d + 1].
template <class T>
__global__ void k_depths_vjp(const T* __restrict__ dens, int G, int kd, T target, const double* __restrict__ gdw,
                             double* __restrict__ gdens) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const int stride = kd + 1;
  const double graw = gdw[2 * i], gw = gdw[2 * i + 1];
  if (graw == 0.0 && gw == 0.0) return;
  int d = kd;
  T density = dens[(int64_t)i * stride + d];
  if (density >= target) {
    const double D = double(density), h = 1.0 / double(1 << d);
    gdens[(int64_t)i * stride + d] += graw / (D * LN2_D * 2.0) - gw * h * h / (D * D);
    return;
  }
  T fine = density;
  while (density < target && d > 0) {
    fine = density;
    d -= 1;
    density = dens[(int64_t)i * stride + d];
  }
  const double D = double(density), F = double(fine), h = 1.0 / double(1 << d);
  const double L1 = log(D / double(target)), L2 = log(D / F);
  double gD = -gw * h * h / (D * D);
  if (L2 != 0.0 && d + 1 <= kd) {
    gD += graw * (L2 - L1) / (D * L2 * L2);
    gdens[(int64_t)i * stride + d + 1] += graw * L1 / (L2 * L2 * F);
  }
  gdens[(int64_t)i * stride + d] += gD;
}

// density.py:108-150 reversed at one level d: the centroid's B-spline
// derivatives (to the position sums) and each record's weight in the value
// (to the density splat records, float64 atomics).
template <class T>
__global__ void k_density_eval_vjp(const int64_t* __restrict__ lkeys, const T* __restrict__ lvals, int64_t lcount,
                                   const T* __restrict__ sums, const T* __restrict__ weights, int G, int d, int stride,
                                   const double* __restrict__ gdens, double* __restrict__ gsums, double* __restrict__ glvals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const double g = gdens[(int64_t)i * stride + d];
  if (g == 0.0) return;
  const T w = weights[i];
  const T p[3] = {sums[3 * i] / w, sums[3 * i + 1] / w, sums[3 * i + 2] / w};
  const int n = 1 << d;
  const T h = T(1) / T(n);
  int c[3];
  double bv[3][3], db[3][3];
  for (int a = 0; a < 3; ++a) {
    c[a] = cell_index(p[a], n);
    const double t = double((p[a] - T(c[a]) * h) / h);
    const Quad q = quadratic(t);
    for (int o = 0; o < 3; ++o) {
      bv[a][o] = quad_at(q, o);
      db[a][o] = quad_deriv(t, o) / double(h);
    }
  }
  double gp[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 27; ++k) {
    const int x = k / 9, y = k / 3 % 3, z = k % 3;
    const int tx = c[0] + x - 1, ty = c[1] + y - 1, tz = c[2] + z - 1;
    if (imin(imin(tx, ty), tz) < 0 || imax(imax(tx, ty), tz) >= n) continue;
    const int64_t idx = find_key(lkeys, lcount, basis_key(d, tx, ty, tz));
    if (idx < 0) continue;
    const double v = double(lvals[idx]);
    gp[0] += db[0][x] * bv[1][y] * bv[2][z] * v;
    gp[1] += bv[0][x] * db[1][y] * bv[2][z] * v;
    gp[2] += bv[0][x] * bv[1][y] * db[2][z] * v;
    atomicAdd(&glvals[idx], g * bv[0][x] * bv[1][y] * bv[2][z]);
  }
  for (int a = 0; a < 3; ++a) gsums[3 * i + a] += g * gp[a] / double(w);
}

// density.py:74-106 (_splat) reversed at one level d: aggregate a's 27
// records weigh v[3] * scale * basis(p - node), p = v[0:3] / v[3]; the record
// gradients reach the aggregate's position sums.
template <class T>
__global__ void k_density_splat_vjp(const int64_t* __restrict__ akeys, const T* __restrict__ avals, int64_t m, int d, T scale,
                                    const int64_t* __restrict__ lkeys, int64_t lcount, const double* __restrict__ glvals,
                                    double* __restrict__ gagg) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= m) return;
  const int64_t key = akeys[a];
  const int c[3] = {key_x(key), key_y(key), key_z(key)};
  const T* v = avals + 4 * a;
  const T p[3] = {v[0] / v[3], v[1] / v[3], v[2] / v[3]};
  const int n = 1 << d;
  const T h = T(1) / T(n);
  double bv[3][3], db[3][3];
  for (int ax = 0; ax < 3; ++ax) {
    const double t = double((p[ax] - T(c[ax]) * h) / h);
    const Quad q = quadratic(t);
    for (int o = 0; o < 3; ++o) {
      bv[ax][o] = quad_at(q, o);
      db[ax][o] = quad_deriv(t, o) / double(h);
    }
  }
  const double weight = double(v[3] * scale);
  double gp[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 27; ++k) {
    const int x = k / 9, y = k / 3 % 3, z = k % 3;
    const int tx = c[0] + x - 1, ty = c[1] + y - 1, tz = c[2] + z - 1;
    if (imin(imin(tx, ty), tz) < 0 || imax(imax(tx, ty), tz) >= n) continue;
    const int64_t idx = find_key(lkeys, lcount, basis_key(d, tx, ty, tz));
    if (idx < 0) continue;
    const double g = glvals[idx] * weight;
    gp[0] += g * db[0][x] * bv[1][y] * bv[2][z];
    gp[1] += g * bv[0][x] * db[1][y] * bv[2][z];
    gp[2] += g * bv[0][x] * bv[1][y] * db[2][z];
  }
  const double inv = 1.0 / double(v[3]);
  for (int ax = 0; ax < 3; ++ax) gagg[3 * a + ax] = gp[ax] * inv;
}

// Aggregates are sums over their leaf groups (density.py:228-240): group g's
// position sums receive the gradient of its ancestor at level d.
__global__ void k_aggregate_to_groups(const int64_t* __restrict__ gkeys, int G, int depth, const int64_t* __restrict__ akeys,
                                      int64_t m, int d, const double* __restrict__ gagg, double* __restrict__ gsums) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int64_t key = gkeys[g];
  const int s = depth - d;
  const int64_t idx = find_key(akeys, m, basis_key(d, key_x(key) >> s, key_y(key) >> s, key_z(key) >> s));
  if (idx < 0) return;
  for (int a = 0; a < 3; ++a) gsums[3 * g + a] += gagg[3 * idx + a];
}

// normals.py:111-163 reversed for the weight component: average[1] =
// sum(dw[2s+1] w_s) / sum(w_s) over the groups with a nonzero mean normal,
// so its gradient reaches each such group's splat weight as w_s / sum(w_s).
template <class T>
__global__ void k_average_vjp(int G, const T* __restrict__ nsums, const T* __restrict__ weights, const T* __restrict__ average,
                              const double* __restrict__ gavg1, double* __restrict__ gdw) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const double ga = gavg1[0];
  const T total = average[2];
  if (ga == 0.0 || !(total > T(0))) return;
  const T w = weights[g];
  const T n0 = nsums[3 * g] / w, n1 = nsums[3 * g + 1] / w, n2 = nsums[3 * g + 2] / w;
  if (!(sqrt(n0 * n0 + n1 * n1 + n2 * n2) > T(0))) return;
  gdw[2 * g + 1] += ga * double(w) / double(total);
}

// samples.py:90-134 reversed for the positions: q = (P1, P2, P0) * inv +
// (t1, t2, t3) summed per group.  The transform's own gradients (inv, t1..t3)
// are reduced into acc[4] for the bounds' extrema owners.
template <class T>
__global__ void __launch_bounds__(256) k_sample_position_vjp(const T* __restrict__ pts, int64_t n, double inv,
                                                             const int* __restrict__ group, const int* __restrict__ status,
                                                             const double* __restrict__ gsums, T* __restrict__ gpoints,
                                                             double* __restrict__ acc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const int g = group[i];
    if (g >= 0 && status[i] == 0) {
      const double gq[3] = {gsums[3 * g], gsums[3 * g + 1], gsums[3 * g + 2]};
      gpoints[3 * i] += T(gq[2] * inv);
      gpoints[3 * i + 1] += T(gq[0] * inv);
      gpoints[3 * i + 2] += T(gq[1] * inv);
      part[0] = gq[0] * double(pts[3 * i + 1]) + gq[1] * double(pts[3 * i + 2]) + gq[2] * double(pts[3 * i]);
      part[1] = gq[0];
      part[2] = gq[1];
      part[3] = gq[2];
    }
  }
  __shared__ double red[4][256];
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = part[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 4) atomicAdd(&acc[threadIdx.x], red[threadIdx.x][0]);
}

// samples.py bounds owners (sample_adjoint.py:16-31): the lowest index whose
// coordinate equals each of the six bounds.
template <class T>
__global__ void k_extrema_owners(const T* __restrict__ pts, int64_t n, const T* __restrict__ bounds, int* __restrict__ owners) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; ++a) {
    const T v = pts[3 * i + a];
    if (v == bounds[a]) atomicMin(&owners[a], (int)i);
    if (v == bounds[a + 3]) atomicMin(&owners[a + 3], (int)i);
  }
}

template <class T>
__global__ void k_add_bound_grads(const int* __restrict__ owners, const double* __restrict__ gb, T* __restrict__ gpoints) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int b = 0; b < 6; ++b)
    if (owners[b] >= 0 && owners[b] < 0x7fffffff) gpoints[3 * (int64_t)owners[b] + b % 3] += T(gb[b]);
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

__global__ void k_fill_owners(int* __restrict__ owners) {
  if (threadIdx.x < 6 && blockIdx.x == 0) owners[threadIdx.x] = 0x7fffffff;
}

}  // namespace

template <class T>
void field_vjp(Result<T>& r, const T* queries, int64_t m, const T* gvalue, const T* ggrad, double* gcoef, double* gcum,
               T* gquery, double* gtransform, hipStream_t s) {
  if (m <= 0) return;
  Affine<T> A;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) A.L[3 * i + j] = r.model_to_cube[4 * i + j];
    A.t[i] = r.model_to_cube[4 * i + 3];
  }
  hipLaunchKernelGGL(k_query_vjp<T>, dim3(blocks_for(m, 256)), dim3(256), 0, s, field_index(r), r.prm.depth,
                     (const T*)r.coefficients.data(), (const T*)r.cumulative.data(), queries, m, A, gvalue, ggrad, gcoef, gcum,
                     gquery, gtransform);
  KERNEL_CHECK();
}

// samples.py:54-88 (_transform) reversed on the host: gradients of (inv, t1,
// t2, t3) -> the six bounds, with Warp's max adjoint (the first operand when
// it is strictly larger).  The normal scale t4 does not reach the unit
// normals, which are invariant to it.
template <class T> static void transform_vjp(const T* b, double scale, const double acc[4], double gb[6]) {
  for (int k = 0; k < 6; ++k) gb[k] = 0.0;
  const double e[3] = {double(b[3] - b[0]), double(b[4] - b[1]), double(b[5] - b[2])};
  const T e1 = b[4] - b[1], e2 = b[5] - b[2], e0 = b[3] - b[0];
  const T inner = wmax(e1, e2);
  const T width_t = wmax(e0, inner) * T(scale);
  const double width = double(width_t), inv = 1.0 / width;
  double ginv = acc[0], gwidth = 0.0;
  for (int a = 0; a < 3; ++a) {
    const int axis = (a + 1) % 3;
    const double origin = (double(b[axis + 3]) + double(b[axis])) / 2.0 - width / 2.0;
    const double gt = acc[1 + a];
    const double gorigin = -gt * inv;
    ginv += -gt * origin;
    gb[axis + 3] += gorigin / 2.0;
    gb[axis] += gorigin / 2.0;
    gwidth += -gorigin / 2.0;
  }
  gwidth += -ginv * inv * inv;
  const double gext = gwidth * scale;
  int arg = 0;
  if (!(e0 > inner)) arg = (e1 > e2) ? 1 : 2;
  (void)e;
  gb[arg + 3] += gext;
  gb[arg] -= gext;
}

// Gradients of a field loss with respect to the input normals and, for runs
// without screening, the input positions (reconstruction_adjoint.py:89-177):
// coefficients and cumulative coefficients (+ isovalue) -> cascadic solve
// and normal loads -> normal records -> splats -> (normal sums, centroids,
// splat depth and weight) -> density levels and aggregates -> samples.
template <class T>
void backward_inputs(Result<T>& r, const double* gcoef_in, const double* gcum_in, double iso_seed, const double* gtransform,
                     const T* points, const T* normals, int64_t n, T* gpoints, T* gnormals, hipStream_t s) {
  const int64_t K = r.level_offsets.back();
  const int D = r.prm.depth, G = r.G;
  const bool positions = gpoints != nullptr;
  // coefficient gradient = direct seeds + cumulative chain (+ isovalue)
  DArray<double> gcoef(K, s), work(K, s), gsums;
  gcoef.zero();
  work.zero();
  if (positions) {
    gsums.reset(3 * (size_t)G, s);
    gsums.zero();
  }
  if (gcoef_in) HIP_CHECK(hipMemcpyAsync(gcoef.data(), gcoef_in, sizeof(double) * K, hipMemcpyDeviceToDevice, s));
  if (gcum_in) HIP_CHECK(hipMemcpyAsync(work.data(), gcum_in, sizeof(double) * K, hipMemcpyDeviceToDevice, s));
  if (iso_seed != 0.0 && G) {
    DArray<double> mass(1, s);
    hipLaunchKernelGGL(k_sum_weights, dim3(1), dim3(64), 0, s, (const void*)r.weights.data(), sizeof(T) == 8 ? 1 : 0, G,
                       mass.data());
    KERNEL_CHECK();
    double mass_h = 0.0;
    read_bytes(&mass_h, mass.data(), sizeof(double), s);
    hipLaunchKernelGGL(k_iso_vjp<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, field_index(r), D, (const T*)r.coefficients.data(),
                       (const T*)r.cumulative.data(), (const T*)r.position_sums.data(), (const T*)r.weights.data(), G,
                       iso_seed / mass_h, gcoef.data(), work.data(), positions ? gsums.data() : (double*)nullptr);
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
  // screened runs: the positions also reach the operators through the
  // constraint points (screening_stencils.py:34-70, constraints.py:55-84)
  const bool screened = r.prm.point_weight > 0 && !r.prm.exact_interpolation;
  SolveGrads mg;
  solve_backward(r, (const double*)gcoef.data(), grec.data(), s, positions && screened ? &mg : (SolveGrads*)nullptr);
  // normal splats -> normal sums (-> input normals), centroids and (depth, weight)
  DArray<int64_t> noff(r.normal_offsets.size(), s);
  std::vector<int64_t> ho(r.normal_offsets.begin(), r.normal_offsets.end());
  HIP_CHECK(hipMemcpyAsync(noff.data(), ho.data(), sizeof(int64_t) * ho.size(), hipMemcpyHostToDevice, s));
  DArray<double> gns, gdw;
  if (gnormals) {
    gns.reset(3 * (size_t)G, s);
    gns.zero();
  }
  if (positions) {
    gdw.reset(2 * (size_t)G, s);
    gdw.zero();
  }
  hipLaunchKernelGGL(k_normal_splat_vjp<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, G, (const T*)r.position_sums.data(),
                     (const T*)r.normal_sums.data(), (const T*)r.weights.data(), (const T*)r.dw.data(), r.prm.base_depth, D,
                     (const int64_t*)r.normal_keys.data(), (const int64_t*)noff.data(), (const double*)grec.data(),
                     gnormals ? gns.data() : (double*)nullptr, positions ? gsums.data() : (double*)nullptr,
                     positions ? gdw.data() : (double*)nullptr);
  KERNEL_CHECK();
  if (gnormals) {
    hipLaunchKernelGGL(k_unit_normal_vjp<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, normals, n, r.transform[4],
                       (const int*)r.input_group.data(), (const int*)r.input_status.data(), (const double*)gns.data(), gnormals);
    KERNEL_CHECK();
  }
  if (positions && screened) {
    std::vector<DArray<double>> gagg;
    DArray<double> gavg1(1, s);
    gavg1.zero();
    screening_backward(r, mg, gagg, gavg1.data(), s);
    for (int d = r.prm.base_depth; d <= mg.top && d < (int)gagg.size(); ++d) {
      const int64_t m = (int64_t)r.levels[d].agg_keys.size();
      if (gagg[d].empty() || !m) continue;
      hipLaunchKernelGGL(k_aggregate_to_groups, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const int64_t*)r.sample_keys.data(),
                         G, D, (const int64_t*)r.levels[d].agg_keys.data(), m, d, (const double*)gagg[d].data(), gsums.data());
      KERNEL_CHECK();
    }
    hipLaunchKernelGGL(k_average_vjp<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, G, (const T*)r.normal_sums.data(),
                       (const T*)r.weights.data(), (const T*)r.normal_average.data(), (const double*)gavg1.data(), gdw.data());
    KERNEL_CHECK();
  }
  if (positions) {
    // splat depth and weight -> densities (density.py:152-182)
    const int kd = r.prm.kernel_depth, stride = kd + 1;
    DArray<double> gdens((size_t)G * stride, s);
    gdens.zero();
    hipLaunchKernelGGL(k_depths_vjp<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const T*)r.densities.data(), G, kd,
                       T(r.prm.samples_per_node), (const double*)gdw.data(), gdens.data());
    KERNEL_CHECK();
    // densities -> centroids and density splat records -> aggregates -> groups
    for (int d = 0; d <= kd; ++d) {
      LevelState<T>& L = r.levels[d];
      const int64_t lcount = (int64_t)L.dens_keys.size(), m = (int64_t)L.agg_keys.size();
      if (!lcount || !m) continue;
      DArray<double> glvals(lcount, s), gagg(3 * (size_t)m, s);
      glvals.zero();
      hipLaunchKernelGGL(k_density_eval_vjp<T>, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const int64_t*)L.dens_keys.data(),
                         (const T*)L.dens_vals.data(), lcount, (const T*)r.position_sums.data(), (const T*)r.weights.data(), G,
                         d, stride, (const double*)gdens.data(), gsums.data(), glvals.data());
      KERNEL_CHECK();
      hipLaunchKernelGGL(k_density_splat_vjp<T>, dim3(blocks_for(m, 256)), dim3(256), 0, s, (const int64_t*)L.agg_keys.data(),
                         (const T*)L.agg_vals.data(), m, d, r.density_scale, (const int64_t*)L.dens_keys.data(), lcount,
                         (const double*)glvals.data(), gagg.data());
      KERNEL_CHECK();
      hipLaunchKernelGGL(k_aggregate_to_groups, dim3(blocks_for(G, 256)), dim3(256), 0, s, (const int64_t*)r.sample_keys.data(),
                         G, D, (const int64_t*)L.agg_keys.data(), m, d, (const double*)gagg.data(), gsums.data());
      KERNEL_CHECK();
    }
    // group position sums -> input positions and the transform (samples.py:54-134)
    DArray<double> acc(4, s);
    acc.zero();
    if (gtransform) HIP_CHECK(hipMemcpyAsync(acc.data(), gtransform, sizeof(double) * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_sample_position_vjp<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, points, n, double(r.transform[0]),
                       (const int*)r.input_group.data(), (const int*)r.input_status.data(), (const double*)gsums.data(),
                       gpoints, acc.data());
    KERNEL_CHECK();
    double acc_h[4], gb[6];
    read_bytes(acc_h, acc.data(), sizeof(acc_h), s);
    transform_vjp<T>(r.bounds, r.prm.scale, acc_h, gb);
    DArray<T> bounds(6, s);
    DArray<int> owners(6, s);
    DArray<double> gbd(6, s);
    HIP_CHECK(hipMemcpyAsync(bounds.data(), r.bounds, sizeof(T) * 6, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(gbd.data(), gb, sizeof(gb), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_owners, dim3(1), dim3(64), 0, s, owners.data());
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_extrema_owners<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, points, n, (const T*)bounds.data(),
                       owners.data());
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_add_bound_grads<T>, dim3(1), dim3(64), 0, s, (const int*)owners.data(), (const double*)gbd.data(),
                       gpoints);
    KERNEL_CHECK();
  }
  HIP_CHECK(hipStreamSynchronize(s));  // the temporaries above (and host staging) return to the pool
}

template <class T>
void backward_normals(Result<T>& r, const double* gcoef_in, const double* gcum_in, double iso_seed, T* gnormals,
                      const T* normals, int64_t n, hipStream_t s) {
  backward_inputs<T>(r, gcoef_in, gcum_in, iso_seed, (const double*)nullptr, (const T*)nullptr, normals, n, (T*)nullptr,
                     gnormals, s);
}

template void field_vjp<float>(Result<float>&, const float*, int64_t, const float*, const float*, double*, double*, float*,
                               double*, hipStream_t);
template void field_vjp<double>(Result<double>&, const double*, int64_t, const double*, const double*, double*, double*, double*,
                                double*, hipStream_t);
template void backward_normals<float>(Result<float>&, const double*, const double*, double, float*, const float*, int64_t,
                                      hipStream_t);
template void backward_normals<double>(Result<double>&, const double*, const double*, double, double*, const double*, int64_t,
                                       hipStream_t);
template void backward_inputs<float>(Result<float>&, const double*, const double*, double, const double*, const float*,
                                     const float*, int64_t, float*, float*, hipStream_t);
template void backward_inputs<double>(Result<double>&, const double*, const double*, double, const double*, const double*,
                                      const double*, int64_t, double*, double*, hipStream_t);

}  // namespace psr
