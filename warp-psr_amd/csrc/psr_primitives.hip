// Device-wide primitives: exclusive scan and a stable LSD radix sort.
//
// The sort ranks each 8-bit digit inside a 4096-item tile with 64-lane
// ballots (peer masks over the digit bits), so equal digits keep input order:
// the pass is stable, which is what gives every serial group sum of the
// pipeline its reference summation order.
#include "psr_runtime.h"

#include <cstring>

namespace psr {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (256 threads, 4 waves).
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wave_sums[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = wave_inclusive_scan(v);
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  int offset = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    int s = wave_sums[w];
    if (w < wave) offset += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return offset + inc - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums(const int* __restrict__ in, int64_t n, int* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int s = 0;
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    int64_t i = base + r * SCAN_THREADS + threadIdx.x;
    if (i < n) s += in[i];
  }
  int total;
  block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const int* __restrict__ in, int* __restrict__ out, int64_t n,
                                                             const int* __restrict__ tile_offsets) {
  __shared__ int tile[SCAN_TILE];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    int64_t i = base + r * SCAN_THREADS + threadIdx.x;
    tile[r * SCAN_THREADS + threadIdx.x] = i < n ? in[i] : 0;
  }
  __syncthreads();
  int local[SCAN_ITEMS];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    local[k] = s;
    s += tile[threadIdx.x * SCAN_ITEMS + k];
  }
  int total;
  int prefix = block_exclusive_scan(s, &total);
  const int off = tile_offsets ? tile_offsets[blockIdx.x] : 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) tile[threadIdx.x * SCAN_ITEMS + k] = local[k] + prefix + off;
  __syncthreads();
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    int64_t i = base + r * SCAN_THREADS + threadIdx.x;
    if (i < n) out[i] = tile[r * SCAN_THREADS + threadIdx.x];
  }
}

// total = exclusive prefix of the last element + the last input value
__global__ void k_scan_total(const int* __restrict__ last_input, const int* __restrict__ out, int64_t n, int* total) {
  *total = out[n - 1] + *last_input;
}

// ------------------------------------------------------------ radix sort
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_radix_count(const K* __restrict__ keys, int64_t n, int shift,
                                                             int* __restrict__ counts, int nblocks) {
  __shared__ int hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    int64_t i = base + r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&hist[(int)((keys[i] >> shift) & 255)], 1);
  }
  __syncthreads();
  counts[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const K* __restrict__ keys, const int* __restrict__ vals,
                                                               K* __restrict__ keys_out, int* __restrict__ vals_out,
                                                               int64_t n, int shift, const int* __restrict__ offsets,
                                                               int nblocks) {
  __shared__ int base_off[256];
  __shared__ int running[256];
  __shared__ int wave_cnt[RS_WAVES][256];
  base_off[threadIdx.x] = offsets[threadIdx.x * nblocks + blockIdx.x];
  running[threadIdx.x] = 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) wave_cnt[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = base + r * RS_THREADS + threadIdx.x;
    const bool valid = i < n;
    K key = valid ? keys[i] : K(0);
    int val = valid ? vals[i] : 0;
    const int digit = (int)((key >> shift) & 255);
    uint64_t peers = __ballot(valid);
    if (!valid) peers = ~peers;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (digit >> b) & 1;
      uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & lt_mask);
    if (valid && rank == 0) wave_cnt[wave][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = base_off[digit] + running[digit] + rank;
      for (int w = 0; w < wave; ++w) pos += wave_cnt[w][digit];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) add += wave_cnt[w][threadIdx.x];
    running[threadIdx.x] += add;
    // the next round's clear of wave_cnt happens after this barrier
    __syncthreads();
  }
}

template <class K> void sort_pairs(K* keys, int* vals, int64_t n, int bits, hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  const int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
  DArray<K> kalt(n, s);
  DArray<int> valt(n, s);
  DArray<int> counts((size_t)256 * nblocks, s);
  K* kin = keys;
  int* vin = vals;
  K* kout = kalt.data();
  int* vout = valt.data();
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_count<K>, dim3(nblocks), dim3(RS_THREADS), 0, s, kin, n, shift, counts.data(), nblocks);
    KERNEL_CHECK();
    scan_exclusive_i32(counts.data(), counts.data(), (int64_t)256 * nblocks, nullptr, s);
    hipLaunchKernelGGL(k_radix_scatter<K>, dim3(nblocks), dim3(RS_THREADS), 0, s, kin, vin, kout, vout, n, shift,
                       counts.data(), nblocks);
    KERNEL_CHECK();
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (kin != keys) {
    HIP_CHECK(hipMemcpyAsync(keys, kin, n * sizeof(K), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(vals, vin, n * sizeof(int), hipMemcpyDeviceToDevice, s));
  }
}

template <class K>
__global__ void k_mark_heads(const K* __restrict__ keys, int64_t n, K sentinel, int* __restrict__ heads) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  K k = keys[i];
  heads[i] = (k != sentinel && (i == 0 || k != keys[i - 1])) ? 1 : 0;
}

__global__ void k_iota(int* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int)i;
}

__global__ void k_compact_heads(const int* __restrict__ heads, const int* __restrict__ prefix, int64_t n,
                                int* __restrict__ head_pos) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && heads[i]) head_pos[prefix[i]] = (int)i;
}

}  // namespace

void scan_exclusive_i32(const int* in, int* out, int64_t n, int* total_dev, hipStream_t s) {
  if (n <= 0) {
    if (total_dev) HIP_CHECK(hipMemsetAsync(total_dev, 0, sizeof(int), s));
    return;
  }
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  DArray<int> last;
  if (total_dev) {  // keep the last input: out may alias in
    last.reset(1, s);
    HIP_CHECK(hipMemcpyAsync(last.data(), in + (n - 1), sizeof(int), hipMemcpyDeviceToDevice, s));
  }
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, n, (const int*)nullptr);
    KERNEL_CHECK();
  } else {
    DArray<int> sums(tiles, s);
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, n, sums.data());
    KERNEL_CHECK();
    scan_exclusive_i32(sums.data(), sums.data(), tiles, nullptr, s);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, out, n,
                       (const int*)sums.data());
    KERNEL_CHECK();
  }
  if (total_dev) {
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, s, (const int*)last.data(), (const int*)out, n, total_dev);
    KERNEL_CHECK();
  }
}

void sort_pairs_u32(uint32_t* keys, int* vals, int64_t n, int bits, hipStream_t s) { sort_pairs(keys, vals, n, bits, s); }
void sort_pairs_u64(uint64_t* keys, int* vals, int64_t n, int bits, hipStream_t s) { sort_pairs(keys, vals, n, bits, s); }

void mark_heads_u32(const uint32_t* keys, int64_t n, uint32_t sentinel, int* heads, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mark_heads<uint32_t>, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, sentinel, heads);
  KERNEL_CHECK();
}
void mark_heads_u64(const uint64_t* keys, int64_t n, uint64_t sentinel, int* heads, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mark_heads<uint64_t>, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, sentinel, heads);
  KERNEL_CHECK();
}

void fill_iota(int* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, n);
  KERNEL_CHECK();
}

void compact_heads(const int* heads, const int* prefix, int64_t n, int* head_pos, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_compact_heads, dim3(blocks_for(n, 256)), dim3(256), 0, s, heads, prefix, n, head_pos);
  KERNEL_CHECK();
}

namespace {
struct PinnedScratch {
  void* p = nullptr;
  PinnedScratch() { (void)hipHostMalloc(&p, 1 << 16, hipHostMallocDefault); }
  ~PinnedScratch() { if (p) (void)hipHostFree(p); }
};
thread_local PinnedScratch g_pinned;
}  // namespace

namespace {
__device__ __forceinline__ int column_of(int64_t k, int dim) { return key_x(k) * dim + key_y(k); }

__global__ void k_column_index(const int64_t* __restrict__ keys, int64_t n, int dim, int* __restrict__ start,
                               int* __restrict__ end) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = column_of(keys[i], dim);
  if (i == 0 || column_of(keys[i - 1], dim) != c) start[c] = (int)i;
  if (i == n - 1 || column_of(keys[i + 1], dim) != c) end[c] = (int)(i + 1);
}
}  // namespace

void build_column_index(const int64_t* keys, int64_t n, int dim, ColumnIndex& out, hipStream_t s) {
  out.dim = 0;
  if (dim <= 0 || dim > COLUMN_TABLE_MAX_DIM || n <= 0) {
    out.start.free();
    out.end.free();
    return;
  }
  const size_t cells = (size_t)dim * dim;
  out.start.reset(cells, s);
  out.end.reset(cells, s);
  out.start.zero();
  out.end.zero();
  hipLaunchKernelGGL(k_column_index, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, dim, out.start.data(),
                     out.end.data());
  KERNEL_CHECK();
  out.dim = dim;
}

void read_bytes(void* host, const void* dev, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  if (bytes <= (1 << 16) && g_pinned.p) {
    HIP_CHECK(hipMemcpyAsync(g_pinned.p, dev, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(host, g_pinned.p, bytes);
  } else {
    HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

int read_int(const int* dev, hipStream_t s) {
  int v = 0;
  read_bytes(&v, dev, sizeof(int), s);
  return v;
}

}  // namespace psr
