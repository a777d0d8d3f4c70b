// Marching cubes on a uniform grid of model-space nodes (meshing.py:150-280):
// slab-wise cell classification into a one-bit mask, compaction of the
// crossing cells, one vertex per crossing grid edge (sorted by edge key) and
// triangles from a case table generated here from the marching-cubes method
// (corner order (000,100,110,010,001,101,111,011), as meshing.py:29-33).
//
// Case table.  Corners with value >= 0 are positive (an edge exactly on the
// threshold keeps both endpoints nonnegative).  On each cube face the
// crossing edges are joined by segments; on an ambiguous face (two diagonal
// positive corners) each positive corner is cut off by its own segment, so a
// face shared by two cells is cut the same way in both and the surface has no
// cracks.  The segments close into loops; a loop is oriented so that its
// normal points toward its edges' positive endpoints (faces point toward
// increasing values) and is fanned into triangles from its first edge.
#include "psr_runtime.h"
#include "psr_mesh.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace psr {

namespace {

constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// edges as (lower, upper) corner along their axis
constexpr int kEdge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6}, {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
constexpr int kFace[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

int edge_of(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((kEdge[e][0] == a && kEdge[e][1] == b) || (kEdge[e][0] == b && kEdge[e][1] == a)) return e;
  return -1;
}

McTable build_table() {
  McTable t{};
  for (int code = 0; code < 256; ++code) {
    auto positive = [&](int c) { return ((code >> c) & 1) != 0; };
    std::vector<int> adj[12];
    for (int f = 0; f < 6; ++f) {
      const int* q = kFace[f];
      int e[4];
      for (int k = 0; k < 4; ++k) e[k] = edge_of(q[k], q[(k + 1) % 4]);  // e[k] joins q[k] and q[k+1]
      std::vector<int> cross;
      for (int k = 0; k < 4; ++k)
        if (positive(q[k]) != positive(q[(k + 1) % 4])) cross.push_back(e[k]);
      auto join = [&](int a, int b) {
        adj[a].push_back(b);
        adj[b].push_back(a);
      };
      if (cross.size() == 2) {
        join(cross[0], cross[1]);
      } else if (cross.size() == 4) {
        if (positive(q[0])) {  // q0, q2 positive: cut each off
          join(e[3], e[0]);
          join(e[1], e[2]);
        } else {  // q1, q3 positive
          join(e[0], e[1]);
          join(e[2], e[3]);
        }
      }
    }
    bool seen[12] = {};
    int ntri = 0;
    for (int start = 0; start < 12; ++start) {
      if (seen[start] || adj[start].empty()) continue;
      if (adj[start].size() != 2) throw PsrError(1, "marching-cubes table: open loop");
      std::vector<int> loop;
      int prev = -1, cur = start;
      do {
        loop.push_back(cur);
        seen[cur] = true;
        const int next = adj[cur][0] != prev ? adj[cur][0] : adj[cur][1];
        prev = cur;
        cur = next;
      } while (cur != start && loop.size() <= 12);
      if (cur != start) throw PsrError(1, "marching-cubes table: loop does not close");
      // orientation: Newell normal of the edge midpoints vs the direction
      // from the loop's negative endpoints to its positive ones
      double n[3] = {0, 0, 0}, d[3] = {0, 0, 0};
      const size_t L = loop.size();
      for (size_t i = 0; i < L; ++i) {
        double m0[3], m1[3];
        for (int a = 0; a < 3; ++a) {
          m0[a] = 0.5 * (kCorner[kEdge[loop[i]][0]][a] + kCorner[kEdge[loop[i]][1]][a]);
          m1[a] = 0.5 * (kCorner[kEdge[loop[(i + 1) % L]][0]][a] + kCorner[kEdge[loop[(i + 1) % L]][1]][a]);
        }
        n[0] += m0[1] * m1[2] - m0[2] * m1[1];
        n[1] += m0[2] * m1[0] - m0[0] * m1[2];
        n[2] += m0[0] * m1[1] - m0[1] * m1[0];
        const int lo = kEdge[loop[i]][0], hi = kEdge[loop[i]][1];
        const int pos = positive(lo) ? lo : hi, neg = positive(lo) ? hi : lo;
        for (int a = 0; a < 3; ++a) d[a] += kCorner[pos][a] - kCorner[neg][a];
      }
      const double dot = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
      if (dot == 0.0) throw PsrError(1, "marching-cubes table: degenerate loop");
      if (dot < 0) std::reverse(loop.begin(), loop.end());
      for (size_t i = 1; i + 1 < L; ++i) {
        if (ntri >= MC_MAX_TRIANGLES) throw PsrError(1, "marching-cubes table: too many triangles");
        t.edges[code][3 * ntri] = (int8_t)loop[0];
        t.edges[code][3 * ntri + 1] = (int8_t)loop[i];
        t.edges[code][3 * ntri + 2] = (int8_t)loop[i + 1];
        ++ntri;
      }
    }
    t.count[code] = (int8_t)ntri;
  }
  return t;
}

__constant__ int8_t c_tri_count[256];
__constant__ int8_t c_tri_edges[256][3 * MC_MAX_TRIANGLES];
__constant__ int8_t c_edge[12][2];
__constant__ int8_t c_corner[8][3];

// The table in each device's constant memory, uploaded once per device.
void upload_table() {
  static std::mutex mu;
  static std::vector<int> done;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  if (std::find(done.begin(), done.end(), dev) != done.end()) return;
  const McTable& t = mc_table();
  int8_t edge[12][2], corner[8][3];
  for (int e = 0; e < 12; ++e)
    for (int k = 0; k < 2; ++k) edge[e][k] = (int8_t)kEdge[e][k];
  for (int c = 0; c < 8; ++c)
    for (int a = 0; a < 3; ++a) corner[c][a] = (int8_t)kCorner[c][a];
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_tri_count), t.count, sizeof(t.count)));
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_tri_edges), t.edges, sizeof(t.edges)));
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_edge), edge, sizeof(edge)));
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_corner), corner, sizeof(corner)));
  done.push_back(dev);
}

struct Grid {
  int nx, ny, nz;
  float lo[3], hi[3], step[3];
};

// meshing.py:37-41 and :88-94 (lower + step * node, clamped to the box),
// rounded after the product and after the sum
__device__ __forceinline__ void node_point(const Grid& g, int x, int y, int z, float* out) {
  const int c[3] = {x, y, z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float p = g.lo[a] + g.step[a] * float(c[a]);
    out[a] = fminf(fmaxf(p, g.lo[a]), g.hi[a]);
  }
}

__global__ void k_mc_slab_points(Grid g, int start, int64_t n, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t plane = (int64_t)g.ny * g.nz;
  const int x = start + (int)(i / plane), y = (int)((i / g.nz) % g.ny), z = (int)(i % g.nz);
  node_point(g, x, y, z, out + 3 * i);
}

// meshing.py:44-59 (_mark_cells): cells of planes [start, start + width)
__global__ void k_mc_mark(Grid g, const float* __restrict__ values, int start, int64_t n, uint32_t* __restrict__ mask,
                          int* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t cy = g.ny - 1, cz = g.nz - 1;
  const int x = (int)(i / (cy * cz)), y = (int)((i / cz) % cy), z = (int)(i % cz);
  int code = 0;
  bool finite = true;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int qx = x + c_corner[c][0], qy = y + c_corner[c][1], qz = z + c_corner[c][2];
    const float v = values[((int64_t)qx * g.ny + qy) * g.nz + qz];
    finite = finite && isfinite(v);
    if (v >= 0.0f) code |= 1 << c;
  }
  if (!finite) atomicOr(status, 1);
  if (code != 0 && code != 255) {
    const int64_t index = ((int64_t)(start + x) * cy + y) * cz + z;
    atomicOr(&mask[index >> 5], 1u << (unsigned)(index & 31));
  }
}

__global__ void k_mc_word_counts(const uint32_t* __restrict__ mask, int64_t words, int* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) counts[i] = __popc(mask[i]);
}

// meshing.py:75-85 (_emit_cells): ascending cell index
__global__ void k_mc_emit(const uint32_t* __restrict__ mask, int64_t words, const int* __restrict__ offsets,
                          int* __restrict__ cells) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  uint32_t bits = mask[w];
  int out = offsets[w];
  while (bits) {
    const int b = __ffs(bits) - 1;
    cells[out++] = (int)(w * 32 + b);
    bits &= bits - 1;
  }
}

__device__ __forceinline__ void cell_xyz(const Grid& g, int cell, int& x, int& y, int& z) {
  const int cy = g.ny - 1, cz = g.nz - 1;
  x = cell / (cy * cz);
  y = (cell / cz) % cy;
  z = cell % cz;
}

// meshing.py:88-94 (_cell_points): 8 corners per retained cell
__global__ void k_mc_corner_points(Grid g, const int* __restrict__ cells, int64_t n, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  cell_xyz(g, cells[i >> 3], x, y, z);
  const int c = (int)(i & 7);
  node_point(g, x + c_corner[c][0], y + c_corner[c][1], z + c_corner[c][2], out + 3 * i);
}

__device__ __forceinline__ int cell_code(const float* __restrict__ v) {
  int code = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (v[c] >= 0.0f) code |= 1 << c;
  return code;
}

// crossing edges and triangles per retained cell
__global__ void k_mc_cell_counts(const float* __restrict__ values, int n, int* __restrict__ edges, int* __restrict__ tris) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int code = cell_code(values + 8 * (int64_t)i);
  int ne = 0;
  for (int e = 0; e < 12; ++e) ne += ((code >> c_edge[e][0]) & 1) != ((code >> c_edge[e][1]) & 1);
  edges[i] = ne;
  tris[i] = c_tri_count[code];
}

// grid edge key: lower node index * 3 + axis
__device__ __forceinline__ uint64_t edge_key(const Grid& g, int x, int y, int z, int e) {
  const int lo = c_edge[e][0], hi = c_edge[e][1];
  const int axis = c_corner[hi][0] != c_corner[lo][0] ? 0 : (c_corner[hi][1] != c_corner[lo][1] ? 1 : 2);
  const uint64_t node =
      ((uint64_t)(x + c_corner[lo][0]) * (uint64_t)g.ny + (uint64_t)(y + c_corner[lo][1])) * (uint64_t)g.nz +
      (uint64_t)(z + c_corner[lo][2]);
  return node * 3 + (uint64_t)axis;
}

__global__ void k_mc_edge_keys(Grid g, const int* __restrict__ cells, const float* __restrict__ values, int n,
                               const int* __restrict__ offsets, uint64_t* __restrict__ keys, int* __restrict__ refs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int code = cell_code(values + 8 * (int64_t)i);
  int x, y, z;
  cell_xyz(g, cells[i], x, y, z);
  int out = offsets[i];
  for (int e = 0; e < 12; ++e) {
    if (((code >> c_edge[e][0]) & 1) == ((code >> c_edge[e][1]) & 1)) continue;
    keys[out] = edge_key(g, x, y, z, e);
    refs[out] = i * 12 + e;
    ++out;
  }
}

// one vertex per unique edge: t = v_lo / (v_lo - v_hi) along the edge's
// axis, in grid units, then lower + step * coordinate
__global__ void k_mc_vertices(Grid g, const int* __restrict__ cells, const float* __restrict__ values,
                              const int* __restrict__ refs, const int* __restrict__ head_pos, int V, float* __restrict__ vertices,
                              int* __restrict__ vref) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int ref = refs[head_pos[v]];
  const int cell = ref / 12, e = ref % 12;
  int x, y, z;
  cell_xyz(g, cells[cell], x, y, z);
  const int lo = c_edge[e][0], hi = c_edge[e][1];
  const float va = values[8 * (int64_t)cell + lo], vb = values[8 * (int64_t)cell + hi];
  const float t = va / (va - vb);
  const int base[3] = {x + c_corner[lo][0], y + c_corner[lo][1], z + c_corner[lo][2]};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float gc = float(base[a]);
    if (c_corner[hi][a] != c_corner[lo][a]) gc = gc + t;
    vertices[3 * (int64_t)v + a] = g.lo[a] + g.step[a] * gc;
  }
  vref[2 * v] = 8 * cell + lo;
  vref[2 * v + 1] = 8 * cell + hi;
}

__device__ __forceinline__ int find_edge(const uint64_t* __restrict__ keys, int V, uint64_t key) {
  int l = 0, h = V;
  while (l < h) {
    const int m = (l + h) >> 1;
    if (keys[m] < key) l = m + 1; else h = m;
  }
  return l;
}

__global__ void k_mc_triangles(Grid g, const int* __restrict__ cells, const float* __restrict__ values, int n,
                               const int* __restrict__ tri_offsets, const uint64_t* __restrict__ ukeys, int V,
                               int* __restrict__ indices) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int code = cell_code(values + 8 * (int64_t)i);
  const int nt = c_tri_count[code];
  if (!nt) return;
  int x, y, z;
  cell_xyz(g, cells[i], x, y, z);
  const int64_t out = 3 * (int64_t)tri_offsets[i];
  for (int k = 0; k < 3 * nt; ++k) indices[out + k] = find_edge(ukeys, V, edge_key(g, x, y, z, c_tri_edges[code][k]));
}

__global__ void k_mc_unique_keys(const uint64_t* __restrict__ keys, const int* __restrict__ head_pos, int V,
                                 uint64_t* __restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) out[v] = keys[head_pos[v]];
}

// d vertex / d (v_lo, v_hi): only the edge-axis coordinate moves,
// d t / d v_lo = -v_hi / (v_lo - v_hi)^2, d t / d v_hi = v_lo / (v_lo - v_hi)^2
__global__ void k_mc_vertex_vjp(Grid g, const float* __restrict__ values, const int* __restrict__ vref, int V,
                                const float* __restrict__ gvert, float* __restrict__ gvalues) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int ia = vref[2 * v], ib = vref[2 * v + 1];
  const int lo = ia & 7, hi = ib & 7;
  const int axis = c_corner[hi][0] != c_corner[lo][0] ? 0 : (c_corner[hi][1] != c_corner[lo][1] ? 1 : 2);
  const float va = values[ia], vb = values[ib];
  const float d = va - vb;
  const float gt = gvert[3 * (int64_t)v + axis] * g.step[axis];
  if (gt == 0.0f) return;
  atomicAdd(&gvalues[ia], gt * (-vb / (d * d)));
  atomicAdd(&gvalues[ib], gt * (va / (d * d)));
}

}  // namespace

const McTable& mc_table() {
  static const McTable t = build_table();
  return t;
}

struct McState {
  hipStream_t s;
  Grid g;
  int64_t cells_total = 0, words = 0;
  DArray<uint32_t> mask;
  DArray<int> status;
  DArray<int> cells;
  int retained = -1;
  // mesh
  DArray<float> vertices;
  DArray<int> indices, vref;
  int V = 0, T = 0;
};

McState* mc_create(const int shape[3], const float lower[3], const float upper[3], hipStream_t s) {
  for (int a = 0; a < 3; ++a) {
    if (shape[a] < 2) throw PsrError(2, "resolution must be an integer or three integers >= 2");
    if (!std::isfinite(lower[a]) || !std::isfinite(upper[a]) || !(lower[a] < upper[a]))
      throw PsrError(2, "bounds must be finite with lower < upper on every axis");
  }
  const int64_t cells = (int64_t)(shape[0] - 1) * (shape[1] - 1) * (shape[2] - 1);
  if (cells > INT32_MAX) throw PsrError(2, "extraction cell count exceeds int32 indexing");
  if ((int64_t)shape[0] * shape[1] * shape[2] > INT32_MAX) throw PsrError(2, "grid node count exceeds int32 sampling indices");
  upload_table();
  auto* m = new McState();
  m->s = s;
  Grid& g = m->g;
  g.nx = shape[0];
  g.ny = shape[1];
  g.nz = shape[2];
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = lower[a];
    g.hi[a] = upper[a];
    // meshing.py:142-143: float32 (upper - lower) / (n - 1)
    const float width = upper[a] - lower[a];
    g.step[a] = width / float(shape[a] - 1);
    if (!std::isfinite(g.step[a]) || !(g.step[a] > 0.0f)) {
      delete m;
      throw PsrError(2, "bounds/spacing must be representable in float32");
    }
  }
  m->cells_total = cells;
  m->words = (cells + 31) / 32;
  m->mask.reset(m->words, s);
  m->mask.zero();
  m->status.reset(1, s);
  m->status.zero();
  return m;
}

void mc_free(McState* m) { delete m; }

int64_t mc_slab_nodes(const McState& m, int start, int width) {
  if (start < 0 || width < 1 || start + width > m.g.nx - 1) throw PsrError(2, "slab outside the grid's cells");
  return (int64_t)(width + 1) * m.g.ny * m.g.nz;
}

void mc_slab_points(McState& m, int start, int width, float* out) {
  const int64_t n = mc_slab_nodes(m, start, width);
  hipLaunchKernelGGL(k_mc_slab_points, dim3(blocks_for(n, 256)), dim3(256), 0, m.s, m.g, start, n, out);
  KERNEL_CHECK();
}

void mc_slab_values(McState& m, int start, int width, const float* values) {
  (void)mc_slab_nodes(m, start, width);
  const int64_t n = (int64_t)width * (m.g.ny - 1) * (m.g.nz - 1);
  hipLaunchKernelGGL(k_mc_mark, dim3(blocks_for(n, 256)), dim3(256), 0, m.s, m.g, values, start, n, m.mask.data(),
                     m.status.data());
  KERNEL_CHECK();
}

int mc_select(McState& m, int* invalid) {
  DArray<int> counts(m.words, m.s), offsets(m.words, m.s), total(1, m.s);
  hipLaunchKernelGGL(k_mc_word_counts, dim3(blocks_for(m.words, 256)), dim3(256), 0, m.s, (const uint32_t*)m.mask.data(),
                     m.words, counts.data());
  KERNEL_CHECK();
  scan_exclusive_i32(counts.data(), offsets.data(), m.words, total.data(), m.s);
  int host[2];
  read_bytes(&host[0], total.data(), sizeof(int), m.s);
  read_bytes(&host[1], m.status.data(), sizeof(int), m.s);
  *invalid = host[1];
  const int count = host[0];
  if ((int64_t)count * 12 > INT32_MAX) throw PsrError(2, "retained corner count exceeds int32 indexing");
  m.cells.reset(count, m.s);
  if (count) {
    hipLaunchKernelGGL(k_mc_emit, dim3(blocks_for(m.words, 256)), dim3(256), 0, m.s, (const uint32_t*)m.mask.data(), m.words,
                       (const int*)offsets.data(), m.cells.data());
    KERNEL_CHECK();
  }
  m.retained = count;
  return count;
}

void mc_corner_points(McState& m, float* out) {
  if (m.retained < 0) throw PsrError(2, "select the cells first");
  const int64_t n = 8 * (int64_t)m.retained;
  if (!n) return;
  hipLaunchKernelGGL(k_mc_corner_points, dim3(blocks_for(n, 256)), dim3(256), 0, m.s, m.g, (const int*)m.cells.data(), n, out);
  KERNEL_CHECK();
}

void mc_triangulate(McState& m, const float* values, int* V, int* T) {
  if (m.retained < 0) throw PsrError(2, "select the cells first");
  const int n = m.retained;
  hipStream_t s = m.s;
  m.V = m.T = 0;
  if (n) {
    DArray<int> ecount(n, s), eoff(n, s), tcount(n, s), toff(n, s), etotal(1, s), ttotal(1, s);
    hipLaunchKernelGGL(k_mc_cell_counts, dim3(blocks_for(n, 256)), dim3(256), 0, s, values, n, ecount.data(), tcount.data());
    KERNEL_CHECK();
    scan_exclusive_i32(ecount.data(), eoff.data(), n, etotal.data(), s);
    scan_exclusive_i32(tcount.data(), toff.data(), n, ttotal.data(), s);
    const int E = read_int(etotal.data(), s);
    const int Tn = read_int(ttotal.data(), s);
    DArray<uint64_t> keys(E, s);
    DArray<int> refs(E, s), heads(E, s), prefix(E, s), vtotal(1, s);
    if (E) {
      hipLaunchKernelGGL(k_mc_edge_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, m.g, (const int*)m.cells.data(), values, n,
                         (const int*)eoff.data(), keys.data(), refs.data());
      KERNEL_CHECK();
      int bits = 1;
      while (bits < 64 && (((uint64_t)m.g.nx * m.g.ny * m.g.nz * 3) >> bits)) ++bits;
      sort_pairs_u64(keys.data(), refs.data(), E, bits, s);
      mark_heads_u64(keys.data(), E, ~0ull, heads.data(), s);
      scan_exclusive_i32(heads.data(), prefix.data(), E, vtotal.data(), s);
      const int Vn = read_int(vtotal.data(), s);
      DArray<int> head_pos(Vn, s);
      compact_heads(heads.data(), prefix.data(), E, head_pos.data(), s);
      DArray<uint64_t> ukeys(Vn, s);
      m.vertices.reset(3 * (size_t)Vn, s);
      m.vref.reset(2 * (size_t)Vn, s);
      m.indices.reset(3 * (size_t)Tn, s);
      if (Vn) {
        hipLaunchKernelGGL(k_mc_vertices, dim3(blocks_for(Vn, 256)), dim3(256), 0, s, m.g, (const int*)m.cells.data(), values,
                           (const int*)refs.data(), (const int*)head_pos.data(), Vn, m.vertices.data(), m.vref.data());
        KERNEL_CHECK();
        hipLaunchKernelGGL(k_mc_unique_keys, dim3(blocks_for(Vn, 256)), dim3(256), 0, s, (const uint64_t*)keys.data(),
                           (const int*)head_pos.data(), Vn, ukeys.data());
        KERNEL_CHECK();
      }
      if (Tn) {
        hipLaunchKernelGGL(k_mc_triangles, dim3(blocks_for(n, 256)), dim3(256), 0, s, m.g, (const int*)m.cells.data(), values,
                           n, (const int*)toff.data(), (const uint64_t*)ukeys.data(), Vn, m.indices.data());
        KERNEL_CHECK();
      }
      m.V = Vn;
      m.T = Tn;
    }
  }
  *V = m.V;
  *T = m.T;
}

void mc_copy(const McState& m, float* vertices, int* indices) {
  if (m.V) HIP_CHECK(hipMemcpyAsync(vertices, m.vertices.data(), sizeof(float) * 3 * m.V, hipMemcpyDeviceToDevice, m.s));
  if (m.T) HIP_CHECK(hipMemcpyAsync(indices, m.indices.data(), sizeof(int) * 3 * m.T, hipMemcpyDeviceToDevice, m.s));
}

void mc_vertex_vjp(const McState& m, const float* values, const float* gvert, float* gvalues) {
  if (!m.V) return;
  hipLaunchKernelGGL(k_mc_vertex_vjp, dim3(blocks_for(m.V, 256)), dim3(256), 0, m.s, m.g, values, (const int*)m.vref.data(), m.V,
                     gvert, gvalues);
  KERNEL_CHECK();
}

}  // namespace psr
