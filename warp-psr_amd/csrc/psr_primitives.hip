// Device-wide primitives: exclusive scan and a stable LSD radix sort.
//
// The sort ranks each 8-bit digit inside a 4096-item tile with 64-lane
// ballots (peer masks over the digit bits), so equal digits keep input order:
// the pass is stable, which is what gives every serial group sum of the
// pipeline its reference summation order.
#include "psr_runtime.h"

#include <cstdlib>

#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

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

// Exclusive scan of one 4,096-value tile plus its offset; the last tile also
// writes the total (its offset plus its sum) when total is non-null.  The
// tile is read into LDS before anything is written, so out may alias in.
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const int* __restrict__ in, int* __restrict__ out, int64_t n,
                                                             const int* __restrict__ tile_offsets, int* __restrict__ total_out) {
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
  if (total_out && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total_out = off + total;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) tile[threadIdx.x * SCAN_ITEMS + k] = local[k] + prefix + off;
  __syncthreads();
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    int64_t i = base + r * SCAN_THREADS + threadIdx.x;
    if (i < n) out[i] = tile[r * SCAN_THREADS + threadIdx.x];
  }
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

constexpr int RS_CHUNK = RS_TILE / RS_WAVES;  // items one wave ranks: RS_ITEMS rounds of 64

// Lanes of the wave holding the same 8-bit digit (and the same validity).
__device__ __forceinline__ uint64_t digit_peers(int digit, bool valid) {
  uint64_t peers = __ballot(valid);
  if (!valid) peers = ~peers;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (digit >> b) & 1;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Stable rank of RS_ITEMS keys per lane inside the wave's chunk.  Each wave
// owns one row of hist: the lanes of a wave execute each LDS instruction
// together, so every lane reads the running count before the group's first
// lane advances it, and no barrier is needed inside the chunk.
template <class K>
__device__ __forceinline__ void wave_rank(const K (&key)[RS_ITEMS], int64_t first, int64_t n, int shift, int* hist_row,
                                          int (&pos)[RS_ITEMS]) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const bool valid = first + r * 64 < n;
    const int digit = (int)((key[r] >> shift) & 255);
    const uint64_t peers = digit_peers(digit, valid);
    const int rank = __popcll(peers & lt_mask);
    const int before = hist_row[digit];
    pos[r] = before + rank;
    if (valid && rank == 0) hist_row[digit] = before + __popcll(peers);
  }
}

// One pass over a tile of RS_TILE items: wave w ranks the contiguous chunk
// [w*RS_CHUNK, (w+1)*RS_CHUNK) of the tile, the per-wave digit counts become
// the tile's offsets (global digit offset + earlier waves), and each item is
// written to offset + rank.  Items keep input order within a digit.  KO may
// be narrower or wider than KI: the first and last passes of a sort of
// 64-bit keys whose bits fit 32 convert on the way through.
template <class KI, class KO>
__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const KI* __restrict__ keys, const int* __restrict__ vals,
                                                               KO* __restrict__ keys_out, int* __restrict__ vals_out,
                                                               int64_t n, int shift, const int* __restrict__ offsets,
                                                               int nblocks) {
  __shared__ int hist[RS_WAVES][256];
#pragma unroll
  for (int w = 0; w < RS_WAVES; ++w) hist[w][threadIdx.x] = 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t first = (int64_t)blockIdx.x * RS_TILE + wave * RS_CHUNK + lane;
  KI key[RS_ITEMS];
  int val[RS_ITEMS], pos[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = first + r * 64;
    key[r] = i < n ? keys[i] : KI(0);
    val[r] = i < n ? vals[i] : 0;
  }
  __syncthreads();
  wave_rank(key, first, n, shift, hist[wave], pos);
  __syncthreads();
  {
    int run = offsets[threadIdx.x * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      const int c = hist[w][threadIdx.x];
      hist[w][threadIdx.x] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    if (first + r * 64 < n) {
      const int p = hist[wave][(int)((key[r] >> shift) & 255)] + pos[r];
      keys_out[p] = (KO)key[r];
      vals_out[p] = val[r];
    }
  }
}

// Whole sort of n <= RS_TILE pairs in one workgroup: every pass ranks and
// scatters inside LDS, with the same per-digit order as the tiled passes.
template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_sort_tile(K* __restrict__ keys, int* __restrict__ vals, int n, int bits) {
  __shared__ K skey[RS_TILE];
  __shared__ int sval[RS_TILE];
  __shared__ int hist[RS_WAVES][256];
  for (int i = threadIdx.x; i < n; i += RS_THREADS) {
    skey[i] = keys[i];
    sval[i] = vals[i];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int first = wave * RS_CHUNK + lane;
  for (int shift = 0; shift < bits; shift += 8) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) hist[w][threadIdx.x] = 0;
    __syncthreads();
    K key[RS_ITEMS];
    int val[RS_ITEMS], pos[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const int i = first + r * 64;
      key[r] = i < n ? skey[i] : K(0);
      val[r] = i < n ? sval[i] : 0;
    }
    wave_rank(key, first, n, shift, hist[wave], pos);
    __syncthreads();
    int c[RS_WAVES], count = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      c[w] = hist[w][threadIdx.x];
      count += c[w];
    }
    int total;
    int run = block_exclusive_scan(count, &total);
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      hist[w][threadIdx.x] = run;
      run += c[w];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      if (first + r * 64 < n) {
        const int p = hist[wave][(int)((key[r] >> shift) & 255)] + pos[r];
        skey[p] = key[r];
        sval[p] = val[r];
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += RS_THREADS) {
    keys[i] = skey[i];
    vals[i] = sval[i];
  }
}

// One counting + scan + scatter pass from (kin, vin) to (kout, vout).
template <class KI, class KO>
void radix_pass(const KI* kin, const int* vin, KO* kout, int* vout, int64_t n, int shift, int* counts, int nblocks,
                hipStream_t s) {
  hipLaunchKernelGGL(k_radix_count<KI>, dim3(nblocks), dim3(RS_THREADS), 0, s, kin, n, shift, counts, nblocks);
  KERNEL_CHECK();
  scan_exclusive_i32(counts, counts, (int64_t)256 * nblocks, nullptr, s);
  hipLaunchKernelGGL((k_radix_scatter<KI, KO>), dim3(nblocks), dim3(RS_THREADS), 0, s, kin, vin, kout, vout, n, shift,
                     (const int*)counts, nblocks);
  KERNEL_CHECK();
}

template <class K> void radix_sort_pairs(K* keys, int* vals, int64_t n, int bits, hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  if (n <= RS_TILE) {
    hipLaunchKernelGGL(k_sort_tile<K>, dim3(1), dim3(RS_THREADS), 0, s, keys, vals, (int)n, bits);
    KERNEL_CHECK();
    return;
  }
  const int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
  const int passes = (bits + 7) / 8;
  DArray<int> valt(n, s);
  DArray<int> counts((size_t)256 * nblocks, s);
  int* vin = vals;
  int* vout = valt.data();
  if (sizeof(K) == 8 && passes >= 2 && bits <= 32) {
    // 64-bit keys whose bits fit 32: the middle passes move 32-bit keys
    DArray<uint32_t> ka(n, s), kb(n, s);
    radix_pass(keys, vin, ka.data(), vout, n, 0, counts.data(), nblocks, s);
    std::swap(vin, vout);
    uint32_t* kin = ka.data();
    uint32_t* kout = kb.data();
    for (int p = 1; p + 1 < passes; ++p) {
      radix_pass(kin, vin, kout, vout, n, 8 * p, counts.data(), nblocks, s);
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
    radix_pass(kin, vin, keys, vout, n, 8 * (passes - 1), counts.data(), nblocks, s);
    if (vout != vals) HIP_CHECK(hipMemcpyAsync(vals, vout, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    return;
  }
  DArray<K> kalt(n, s);
  K* kin = keys;
  K* kout = kalt.data();
  for (int p = 0; p < passes; ++p) {
    radix_pass(kin, vin, kout, vout, n, 8 * p, counts.data(), nblocks, s);
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

// Exclusive scan of up to SCAN_SMALL_TILES tiles in one workgroup, tile after
// tile with a running carry (one launch where the tiled form takes three).
constexpr int SCAN_SMALL_TILES = 4;

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_small(const int* __restrict__ in, int* __restrict__ out, int64_t n,
                                                             int* __restrict__ total_out) {
  __shared__ int tile[SCAN_TILE];
  int carry = 0;
  for (int64_t base = 0; base < n; base += SCAN_TILE) {
    for (int r = 0; r < SCAN_ITEMS; ++r) {
      const int64_t i = base + r * SCAN_THREADS + threadIdx.x;
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
    const int prefix = block_exclusive_scan(s, &total);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) tile[threadIdx.x * SCAN_ITEMS + k] = local[k] + prefix + carry;
    __syncthreads();
    for (int r = 0; r < SCAN_ITEMS; ++r) {
      const int64_t i = base + r * SCAN_THREADS + threadIdx.x;
      if (i < n) out[i] = tile[r * SCAN_THREADS + threadIdx.x];
    }
    carry += total;
    __syncthreads();
  }
  if (total_out && threadIdx.x == 0) *total_out = carry;
}

void scan_exclusive_i32(const int* in, int* out, int64_t n, int* total_dev, hipStream_t s) {
  if (n <= 0) {
    if (total_dev) HIP_CHECK(hipMemsetAsync(total_dev, 0, sizeof(int), s));
    return;
  }
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, n, (const int*)nullptr, total_dev);
    KERNEL_CHECK();
  } else if (tiles <= SCAN_SMALL_TILES) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, n, total_dev);
    KERNEL_CHECK();
  } else {
    DArray<int> sums(tiles, s);
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, n, sums.data());
    KERNEL_CHECK();
    scan_exclusive_i32(sums.data(), sums.data(), tiles, nullptr, s);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, out, n,
                       (const int*)sums.data(), total_dev);
    KERNEL_CHECK();
  }
}

void sort_pairs_u32(uint32_t* keys, int* vals, int64_t n, int bits, hipStream_t s) { radix_sort_pairs(keys, vals, n, bits, s); }
void sort_pairs_u64(uint64_t* keys, int* vals, int64_t n, int bits, hipStream_t s) { radix_sort_pairs(keys, vals, n, bits, s); }

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
// 64 KiB pinned staging buffers for small device-to-host reads, one per host
// thread.  A thread returns its buffer to a process-wide free list when it
// ends; the buffers and the list are never freed, so no HIP call runs from a
// thread-local or static destructor (those can run after the HIP runtime has
// shut down at process exit).
struct PinnedList {
  std::mutex mu;
  std::vector<void*> free;
};
PinnedList& pinned_list() {
  static PinnedList* list = new PinnedList();  // leaked on purpose
  return *list;
}
struct PinnedScratch {
  void* p = nullptr;
  PinnedScratch() {
    PinnedList& l = pinned_list();
    {
      std::lock_guard<std::mutex> g(l.mu);
      if (!l.free.empty()) {
        p = l.free.back();
        l.free.pop_back();
        return;
      }
    }
    if (hipHostMalloc(&p, 1 << 16, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
  }
  ~PinnedScratch() {
    if (!p) return;
    PinnedList& l = pinned_list();
    std::lock_guard<std::mutex> g(l.mu);
    l.free.push_back(p);
  }
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

static int64_t env_limit(const char* name, int64_t fallback, int64_t lo, int64_t hi) {
  const char* v = std::getenv(name);
  if (!v || !*v) return fallback;
  char* end = nullptr;
  const long long x = std::strtoll(v, &end, 10);
  if (*end) return fallback;
  return x < lo ? lo : (x > hi ? hi : (int64_t)x);
}

int64_t dense_index_max_cells() { return env_limit("PSR_DENSE_INDEX_MAX_CELLS", (int64_t)1 << 28, 0, (int64_t)1 << 28); }
int column_table_max_dim() { return (int)env_limit("PSR_COLUMN_TABLE_MAX_DIM", 4097, 0, 4097); }
int field_table_max_depth() { return (int)env_limit("PSR_FIELD_TABLE_MAX_DEPTH", 11, -1, 11); }

// both arrays of a column index cleared in one launch
__global__ void k_zero_pair(int* __restrict__ a, int* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = 0;
    b[i] = 0;
  }
}

void build_column_index(const int64_t* keys, int64_t n, int dim, ColumnIndex& out, hipStream_t s) {
  out.dim = 0;
  if (dim <= 0 || dim > column_table_max_dim() || n <= 0) {
    out.start.free();
    out.end.free();
    return;
  }
  const size_t cells = (size_t)dim * dim;
  out.start.reset(cells, s);
  out.end.reset(cells, s);
  hipLaunchKernelGGL(k_zero_pair, dim3(blocks_for((int64_t)cells, 256)), dim3(256), 0, s, out.start.data(), out.end.data(),
                     (int64_t)cells);
  KERNEL_CHECK();
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

hipStream_t companion_stream(int dev, hipStream_t s, int role) {
  static std::mutex mu;
  static std::map<std::tuple<int, hipStream_t, int>, hipStream_t> streams;
  std::lock_guard<std::mutex> g(mu);
  hipStream_t& c = streams[std::make_tuple(dev, s, role)];
  if (!c) HIP_CHECK(hipStreamCreateWithFlags(&c, hipStreamNonBlocking));
  return c;
}

int read_int(const int* dev, hipStream_t s) {
  int v = 0;
  read_bytes(&v, dev, sizeof(int), s);
  return v;
}

}  // namespace psr
