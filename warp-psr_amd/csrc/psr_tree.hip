// Stage 4: level-by-level brood tree (src/warp_psr/tree.py:152-201).
//
// Candidate broods of level d are level d-1 cells with coordinates in
// [-2^d, 2^d); they are packed as three (d+1)-bit fields (offset by 2^d), so the
// unique-sort costs ceil((3d+4)/8) radix passes instead of the 60-bit tree key.
// Children of the sorted unique broods are placed at their lexicographic rank
// directly (no second sort).
#include "psr_result.h"

namespace psr {

namespace {

struct BroodPack {
  int d;
  PSR_HD uint64_t pack(int x, int y, int z) const {
    const int off = 1 << d, b = d + 1;
    return ((uint64_t)(x + off) << (2 * b)) | ((uint64_t)(y + off) << b) | (uint64_t)(z + off);
  }
  PSR_HD uint64_t sentinel() const { return (uint64_t)1 << (3 * (d + 1)); }
  PSR_HD int64_t to_tree_key(uint64_t p) const {
    const int off = 1 << d, b = d + 1;
    const uint64_t m = ((uint64_t)1 << b) - 1;
    return tree_key((int)(p >> (2 * b)) - off, (int)((p >> b) & m) - off, (int)(p & m) - off);
  }
};

PSR_HD bool in_support(int x, int y, int z, int lim) {
  return imin(imin(x, y), z) >= -lim && imax(imax(x, y), z) < lim;
}

// tree.py:41-50 (_seed_broods)
template <class T>
__global__ void k_seed_broods(const int64_t* __restrict__ nkeys, const T* __restrict__ nvals, int64_t count, int d,
                              BroodPack bp, uint64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t key = nkeys[i];
  const int nd = key_d(key);
  const T* v = nvals + 3 * i;
  const T w = -T(0);  // the negated fourth record component
  const T len2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + w * w;
  uint64_t packed = bp.sentinel();
  if (nd >= d && double(len2) > 0.0) {
    const int sh = nd - d;
    packed = bp.pack((key_x(key) >> sh) >> 1, (key_y(key) >> sh) >> 1, (key_z(key) >> sh) >> 1);
  }
  out[i] = packed;
}

// tree.py:30-34 (_coarse, lo = -1)
__global__ void k_coarse_broods(int cc, BroodPack bp, uint64_t* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cc * cc * cc) return;
  out[i] = bp.pack(i / (cc * cc) - 1, (i / cc) % cc - 1, i % cc - 1);
}

// tree.py:52-64 (_support_broods) for radius 1 (27 per request) or 0
__global__ void k_support_broods(const int64_t* __restrict__ src, int64_t count, int radius, int lim, BroodPack bp,
                                 uint64_t* __restrict__ out) {
  const int width = 2 * radius + 1, w3 = width * width * width;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= count * w3) return;
  const int64_t i = idx / w3;
  const int j = (int)(idx % w3);
  const int64_t key = src[i];
  const int px = tkey_x(key) >> 1, py = tkey_y(key) >> 1, pz = tkey_z(key) >> 1;
  const int x = px + j / (width * width) - radius, y = py + (j / width) % width - radius, z = pz + j % width - radius;
  out[idx] = in_support(x, y, z, lim) ? bp.pack(x, y, z) : bp.sentinel();
}

__global__ void k_parent_tree_keys(const int64_t* __restrict__ src, int64_t count, BroodPack bp, uint64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t key = src[i];
  out[i] = bp.pack(tkey_x(key) >> 1, tkey_y(key) >> 1, tkey_z(key) >> 1);
}

__global__ void k_unpack_unique(const uint64_t* __restrict__ keys, const int* __restrict__ heads, const int* __restrict__ prefix,
                                int64_t n, BroodPack bp, int64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && heads[i]) out[prefix[i]] = bp.to_tree_key(keys[i]);
}

// tree.py:78-96: lexicographic rank of each child of the sorted unique broods.
__global__ void k_expand_broods(const int64_t* __restrict__ broods, int64_t nb, int64_t* __restrict__ children) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int64_t key = broods[i];
  const int64_t xp = (key >> 40) << 40, xyp = (key >> 20) << 20;
  const int64_t x0 = lower_bound(broods, nb, xp), x1 = lower_bound(broods, nb, xp + ((int64_t)1 << 40));
  const int64_t xy0 = lower_bound(broods, nb, xyp), xy1 = lower_bound(broods, nb, xyp + ((int64_t)1 << 20));
  const int cx = 2 * tkey_x(key), cy = 2 * tkey_y(key), cz = 2 * tkey_z(key);
  for (int k = 0; k < 8; ++k) {
    const int ox = (k >> 2) & 1, oy = (k >> 1) & 1, oz = k & 1;
    // children before this one: all of smaller x runs, then the x half, the
    // smaller y runs inside this x run, the y half, earlier z, then z half
    const int64_t dest = 8 * x0 + 4 * ox * (x1 - x0) + 4 * (xy0 - x0) + 2 * oy * (xy1 - xy0) + 2 * (i - xy0) + oz;
    children[dest] = tree_key(cx + ox, cy + oy, cz + oz);
  }
}

// tree.py:105-123 (basis marks/fill)
__global__ void k_basis_marks(const int64_t* __restrict__ keys, int64_t n, int d, int* __restrict__ marks) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = keys[i];
  const int x = tkey_x(k), y = tkey_y(k), z = tkey_z(k);
  const int lim = 1 << d;
  marks[i] = (imin(imin(x, y), z) >= 0 && imax(imax(x, y), z) <= lim) ? 1 : 0;
}

__global__ void k_basis_fill(const int64_t* __restrict__ keys, const int* __restrict__ marks, const int* __restrict__ prefix,
                             int64_t n, int d, int64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !marks[i]) return;
  const int64_t k = keys[i];
  out[prefix[i]] = basis_key(d, tkey_x(k), tkey_y(k), tkey_z(k));
}

// Sorted unique tree keys of packed candidates (keyed.py:157-178).
DArray<int64_t> unique_packed(DArray<uint64_t>& cand, int64_t total, BroodPack bp, hipStream_t s) {
  DArray<int64_t> out;
  if (!total) return out;
  DArray<int> ids(total, s), heads(total, s), prefix(total, s), cnt(1, s);
  fill_iota(ids.data(), total, s);
  sort_pairs_u64(cand.data(), ids.data(), total, bits_for(bp.sentinel()), s);
  mark_heads_u64(cand.data(), total, bp.sentinel(), heads.data(), s);
  scan_exclusive_i32(heads.data(), prefix.data(), total, cnt.data(), s);
  const int count = read_int(cnt.data(), s);
  out.reset(count, s);
  if (count) {
    hipLaunchKernelGGL(k_unpack_unique, dim3(blocks_for(total, 256)), dim3(256), 0, s, (const uint64_t*)cand.data(),
                       (const int*)heads.data(), (const int*)prefix.data(), total, bp, out.data());
    KERNEL_CHECK();
  }
  return out;
}

}  // namespace

template <class T> void stage_tree(Result<T>& r) {
  hipStream_t s = r.s;
  const int depth = r.prm.depth, full = r.prm.full_depth, base = r.prm.base_depth;
  r.levels.resize(depth + 1);
  DArray<int64_t> broods, requests;
  bool have_requests = false;
  // per-level basis counts stay on the device until the loop ends (one read):
  // each level's basis keys are allocated for all 8 * broods candidates
  DArray<int> basis_counts(depth + 1, s);
  basis_counts.zero();
  for (int d = depth; d >= 0; --d) {
    const int64_t r0 = r.normal_offsets[d], r1 = r.normal_offsets[d + 1];
    const int64_t count = r1 - r0;
    // parents = previous broods (level-d cells); kept for this iteration
    DArray<int64_t> parents = std::move(broods);
    BroodPack bp{d};
    if (d > base) {
      requests.reset(parents.size(), s);
      if (parents.size())
        HIP_CHECK(hipMemcpyAsync(requests.data(), parents.data(), parents.size() * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    } else if (have_requests) {
      // requests move up one level: unique parents (tree.py:176-180)
      const int64_t m = (int64_t)requests.size();
      DArray<uint64_t> rk(m, s);
      BroodPack up{d + 1};  // parents of level-(d+1) cells are level-d cells; pack range of level d+1 broods
      if (m) {
        hipLaunchKernelGGL(k_parent_tree_keys, dim3(blocks_for(m, 256)), dim3(256), 0, s, (const int64_t*)requests.data(), m,
                           up, rk.data());
        KERNEL_CHECK();
      }
      requests = unique_packed(rk, m, up, s);
    } else {
      requests.reset(parents.size(), s);
      if (parents.size())
        HIP_CHECK(hipMemcpyAsync(requests.data(), parents.data(), parents.size() * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    }
    have_requests = true;
    const int cc = d <= full ? (d == 0 ? 2 : (1 << (d - 1)) + 2) : 0;
    const int64_t nc = (int64_t)cc * cc * cc;
    const int64_t ns = (int64_t)requests.size() * 27;
    const int64_t np = (int64_t)parents.size();
    const int64_t total = count + nc + ns + np;
    DArray<uint64_t> cand(total, s);
    if (count) {
      hipLaunchKernelGGL(k_seed_broods<T>, dim3(blocks_for(count, 256)), dim3(256), 0, s,
                         (const int64_t*)r.normal_keys.data() + r0, (const T*)r.normal_vals.data() + 3 * r0, count, d, bp,
                         cand.data());
      KERNEL_CHECK();
    }
    if (nc) {
      hipLaunchKernelGGL(k_coarse_broods, dim3(blocks_for(nc, 256)), dim3(256), 0, s, cc, bp, cand.data() + count);
      KERNEL_CHECK();
    }
    if (ns) {
      hipLaunchKernelGGL(k_support_broods, dim3(blocks_for(ns, 256)), dim3(256), 0, s, (const int64_t*)requests.data(),
                         (int64_t)requests.size(), 1, 1 << d, bp, cand.data() + count + nc);
      KERNEL_CHECK();
    }
    if (np) {
      hipLaunchKernelGGL(k_support_broods, dim3(blocks_for(np, 256)), dim3(256), 0, s, (const int64_t*)parents.data(), np, 0,
                         1 << d, bp, cand.data() + count + nc + ns);
      KERNEL_CHECK();
    }
    broods = unique_packed(cand, total, bp, s);
    LevelState<T>& L = r.levels[d];
    const int64_t nb = (int64_t)broods.size();
    L.tree_keys.reset(8 * nb, s);
    if (nb) {
      hipLaunchKernelGGL(k_expand_broods, dim3(blocks_for(nb, 256)), dim3(256), 0, s, (const int64_t*)broods.data(), nb,
                         L.tree_keys.data());
      KERNEL_CHECK();
    }
    // compact_basis (tree.py:126-143)
    const int64_t nt = 8 * nb;
    L.basis_keys.reset(nt, s);
    if (nt) {
      DArray<int> marks(nt, s), prefix(nt, s);
      hipLaunchKernelGGL(k_basis_marks, dim3(blocks_for(nt, 256)), dim3(256), 0, s, (const int64_t*)L.tree_keys.data(), nt, d,
                         marks.data());
      KERNEL_CHECK();
      scan_exclusive_i32(marks.data(), prefix.data(), nt, basis_counts.data() + d, s);
      hipLaunchKernelGGL(k_basis_fill, dim3(blocks_for(nt, 256)), dim3(256), 0, s, (const int64_t*)L.tree_keys.data(),
                         (const int*)marks.data(), (const int*)prefix.data(), nt, d, L.basis_keys.data());
      KERNEL_CHECK();
    }
  }
  std::vector<int> counts(depth + 1);
  read_bytes(counts.data(), basis_counts.data(), sizeof(int) * (depth + 1), s);
  for (int d = 0; d <= depth; ++d) {
    r.levels[d].basis_count = counts[d];
    r.levels[d].basis_keys.shrink((size_t)counts[d]);
  }
}

template void stage_tree<float>(Result<float>&);
template void stage_tree<double>(Result<double>&);

}  // namespace psr
