// Tree metadata (tree_metadata.py:95-153): every retained brood of every
// level -- the active support nodes plus the broods of the parents of the
// finer level, of the samples, of the density splat records and of the normal
// records -- in global lattice coordinates (local + 2^(d+1)), two padding
// levels above the root, and per node its flags, ghost mark, parent index
// and eight child indices.  Built on demand from a finished reconstruction.
#include "psr_result.h"
#include "psr_meta.h"

namespace psr {

namespace {

constexpr int64_t kInvalid = 0x7fffffffffffffffLL;

__device__ __forceinline__ int floor_half(int v) { return v >> 1; }  // floor(v / 2) (tree.py:26-28)

// tree.py:37-39 (_parent_keys)
__global__ void k_meta_parents(const int64_t* __restrict__ fine, int64_t n, uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = fine[i];
  out[i] = (uint64_t)tree_key(floor_half(tkey_x(k)), floor_half(tkey_y(k)), floor_half(tkey_z(k)));
}

// tree_metadata.py:14-26 (_broods): the brood of each key at depth d; field
// keys (basis-key encoded) of another depth give no brood
__global__ void k_meta_broods(const int64_t* __restrict__ keys, int64_t n, int depth, int field_keys, uint64_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8 * n) return;
  const int64_t i = t >> 3;
  const int k = (int)(t & 7);
  const int64_t key = keys[i];
  int x, y, z;
  bool valid = true;
  if (field_keys) {
    x = key_x(key);
    y = key_y(key);
    z = key_z(key);
    valid = (int)(key >> 48) == depth;
  } else {
    x = tkey_x(key);
    y = tkey_y(key);
    z = tkey_z(key);
  }
  out[t] = (uint64_t)kInvalid;
  if (valid)
    out[t] = (uint64_t)tree_key(floor_half(x) * 2 + (k & 1), floor_half(y) * 2 + ((k >> 1) & 1), floor_half(z) * 2 + ((k >> 2) & 1));
}

__global__ void k_meta_copy_keys(const int64_t* __restrict__ in, int64_t n, uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint64_t)in[i];
}

__global__ void k_meta_compact(const uint64_t* __restrict__ keys, const int* __restrict__ heads, const int* __restrict__ prefix,
                               int64_t n, int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && heads[i]) out[prefix[i]] = (int64_t)keys[i];
}

// tree_metadata.py:29-39 (_global_keys, _padding)
__global__ void k_meta_global(const int64_t* __restrict__ local, int64_t n, int depth, int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = local[i];
  const int off = 1 << (depth + 1);
  out[i] = tree_key(tkey_x(k) + off, tkey_y(k) + off, tkey_z(k) + off);
}

__global__ void k_meta_padding(int64_t n, int64_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = tree_key(i / 4, (i / 2) % 2, i % 2);
}

// tree_metadata.py:42-74 (_links)
__global__ void k_meta_links(const int64_t* __restrict__ keys, int64_t n, const int64_t* __restrict__ parent_keys, int64_t np,
                             const int64_t* __restrict__ child_keys, int64_t nc, const int64_t* __restrict__ active, int64_t na,
                             const int64_t* __restrict__ active_child, int64_t nac, int depth, uint8_t* __restrict__ flags,
                             uint8_t* __restrict__ ghost, int* __restrict__ parents, int* __restrict__ children) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t key = keys[i];
  const int c[3] = {tkey_x(key), tkey_y(key), tkey_z(key)};
  parents[i] = (int)find_key(parent_keys, np, tree_key(floor_half(c[0]), floor_half(c[1]), floor_half(c[2])));
  bool is_ghost = depth == -2;
  int bits = 0;
  if (depth >= 0) {
    const int off = 1 << (depth + 1);
    const int l[3] = {c[0] - off, c[1] - off, c[2] - off};
    is_ghost = find_key(active, na, tree_key(l[0], l[1], l[2])) < 0;
    if (!is_ghost && imin(imin(l[0], l[1]), l[2]) >= 0) {
      const int hi = imax(imax(l[0], l[1]), l[2]);
      if (hi < (1 << depth)) bits |= 1 | 4;  // SPACE and quadratic Dirichlet FEM validity
      if (hi <= (1 << depth)) bits |= 2;     // degree-one Neumann FEM validity
    }
  }
  bool has_child = false, child_is_active = depth < -1;
  for (int k = 0; k < 8; ++k) {
    const int q[3] = {2 * c[0] + (k & 1), 2 * c[1] + ((k >> 1) & 1), 2 * c[2] + ((k >> 2) & 1)};
    const int64_t child = find_key(child_keys, nc, tree_key(q[0], q[1], q[2]));
    children[8 * i + k] = (int)child;
    if (child >= 0) {
      has_child = true;
      if (depth >= -1) {
        const int off = 1 << (depth + 2);
        child_is_active = child_is_active || find_key(active_child, nac, tree_key(q[0] - off, q[1] - off, q[2] - off)) >= 0;
      }
    }
  }
  if (has_child && !child_is_active) bits |= 64;  // the children's ghost flag, stored on the parent
  flags[i] = (uint8_t)bits;
  ghost[i] = (uint8_t)is_ghost;
}

// keyed.py:157-178 (unique_keys) over uint64 candidates; kInvalid is dropped
DArray<int64_t> unique_u64(DArray<uint64_t>& cand, int64_t n, hipStream_t s) {
  DArray<int64_t> out;
  if (!n) return out;
  DArray<int> ids(n, s), heads(n, s), prefix(n, s), total(1, s);
  fill_iota(ids.data(), n, s);
  sort_pairs_u64(cand.data(), ids.data(), n, 63, s);
  mark_heads_u64(cand.data(), n, (uint64_t)kInvalid, heads.data(), s);
  scan_exclusive_i32(heads.data(), prefix.data(), n, total.data(), s);
  const int count = read_int(total.data(), s);
  out.reset(count, s);
  if (count) {
    hipLaunchKernelGGL(k_meta_compact, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const uint64_t*)cand.data(),
                       (const int*)heads.data(), (const int*)prefix.data(), n, out.data());
    KERNEL_CHECK();
  }
  return out;
}

}  // namespace

template <class T> void build_tree_metadata(Result<T>& r, TreeMetadata& meta, hipStream_t s) {
  const int D = r.prm.depth;
  std::vector<DArray<int64_t>> raw(D + 1);
  for (int d = D; d >= 0; --d) {
    LevelState<T>& L = r.levels[d];
    // parents of the finer level's nodes
    DArray<int64_t> parents;
    if (d < D && raw[d + 1].size()) {
      const int64_t nf = (int64_t)raw[d + 1].size();
      DArray<uint64_t> pk(nf, s);
      hipLaunchKernelGGL(k_meta_parents, dim3(blocks_for(nf, 256)), dim3(256), 0, s, (const int64_t*)raw[d + 1].data(), nf,
                         pk.data());
      KERNEL_CHECK();
      parents = unique_u64(pk, nf, s);
    }
    struct Source { const int64_t* keys; int64_t n; };
    std::vector<Source> sources;
    if (d == D) sources.push_back({(const int64_t*)r.sample_keys.data(), (int64_t)r.sample_keys.size()});
    if (L.dens_keys.size()) sources.push_back({(const int64_t*)L.dens_keys.data(), (int64_t)L.dens_keys.size()});
    sources.push_back({(const int64_t*)r.normal_keys.data() + r.normal_offsets[d],
                       (int64_t)(r.normal_offsets[d + 1] - r.normal_offsets[d])});
    const int64_t na = (int64_t)L.tree_keys.size();
    int64_t total = na + 8 * (int64_t)parents.size();
    for (const Source& src : sources) total += 8 * src.n;
    DArray<uint64_t> cand(total, s);
    int64_t offset = 0;
    if (na) {
      hipLaunchKernelGGL(k_meta_copy_keys, dim3(blocks_for(na, 256)), dim3(256), 0, s, (const int64_t*)L.tree_keys.data(), na,
                         cand.data());
      KERNEL_CHECK();
    }
    offset += na;
    if (parents.size()) {
      const int64_t np = (int64_t)parents.size();
      hipLaunchKernelGGL(k_meta_broods, dim3(blocks_for(8 * np, 256)), dim3(256), 0, s, (const int64_t*)parents.data(), np, d, 0,
                         cand.data() + offset);
      KERNEL_CHECK();
      offset += 8 * np;
    }
    for (const Source& src : sources) {
      if (!src.n) continue;
      hipLaunchKernelGGL(k_meta_broods, dim3(blocks_for(8 * src.n, 256)), dim3(256), 0, s, src.keys, src.n, d, 1,
                         cand.data() + offset);
      KERNEL_CHECK();
      offset += 8 * src.n;
    }
    raw[d] = unique_u64(cand, total, s);
  }
  // levels: two padding levels, then every depth in global coordinates
  meta.levels.clear();
  meta.levels.resize(D + 3);
  for (int g = 0; g < D + 3; ++g) {
    TreeMetadata::Level& lv = meta.levels[g];
    lv.local_depth = g - 2;
    if (g < 2) {
      const int64_t n = g == 0 ? 1 : 8;
      lv.keys.reset(n, s);
      hipLaunchKernelGGL(k_meta_padding, dim3(1), dim3(64), 0, s, n, lv.keys.data());
      KERNEL_CHECK();
    } else {
      const int d = g - 2;
      const int64_t n = (int64_t)raw[d].size();
      lv.keys.reset(n, s);
      if (n) {
        hipLaunchKernelGGL(k_meta_global, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const int64_t*)raw[d].data(), n, d,
                           lv.keys.data());
        KERNEL_CHECK();
      }
    }
  }
  for (int g = 0; g < D + 3; ++g) {
    TreeMetadata::Level& lv = meta.levels[g];
    const int d = lv.local_depth;
    const int64_t n = (int64_t)lv.keys.size();
    lv.flags.reset(n, s);
    lv.ghost.reset(n, s);
    lv.parents.reset(n, s);
    lv.children.reset(8 * (size_t)n, s);
    if (!n) continue;
    const DArray<int64_t>* pk = g > 0 ? &meta.levels[g - 1].keys : nullptr;
    const DArray<int64_t>* ck = g + 1 < D + 3 ? &meta.levels[g + 1].keys : nullptr;
    const DArray<int64_t>* act = d >= 0 ? &r.levels[d].tree_keys : nullptr;
    const DArray<int64_t>* act_child = (d >= -1 && d < D) ? &r.levels[d + 1].tree_keys : nullptr;
    hipLaunchKernelGGL(k_meta_links, dim3(blocks_for(n, 128)), dim3(128), 0, s, (const int64_t*)lv.keys.data(), n,
                       pk ? (const int64_t*)pk->data() : (const int64_t*)nullptr, pk ? (int64_t)pk->size() : 0,
                       ck ? (const int64_t*)ck->data() : (const int64_t*)nullptr, ck ? (int64_t)ck->size() : 0,
                       act ? (const int64_t*)act->data() : (const int64_t*)nullptr, act ? (int64_t)act->size() : 0,
                       act_child ? (const int64_t*)act_child->data() : (const int64_t*)nullptr,
                       act_child ? (int64_t)act_child->size() : 0, d, lv.flags.data(), lv.ghost.data(), lv.parents.data(),
                       lv.children.data());
    KERNEL_CHECK();
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

template void build_tree_metadata<float>(Result<float>&, TreeMetadata&, hipStream_t);
template void build_tree_metadata<double>(Result<double>&, TreeMetadata&, hipStream_t);

}  // namespace psr
