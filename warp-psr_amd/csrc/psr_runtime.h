// Device memory and launch plumbing: a stream-ordered caching allocator and a
// move-only device array.  All work of one reconstruction is issued on one
// stream, so a block returned to the cache may be handed out again on the same
// stream without synchronisation (stream order protects it).
#pragma once
#include "psr_common.h"

#include <map>
#include <memory>

namespace psr {

class Pool {
 public:
  static Pool& instance() {
    static Pool pool;
    return pool;
  }
  void* allocate(size_t bytes, hipStream_t stream) {
    size_t cls = size_class(bytes);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto& lst = free_[{stream, cls}];
      if (!lst.empty()) {
        void* p = lst.back();
        lst.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
      // release cached blocks of every stream and retry once
      (void)hipGetLastError();
      release_all();
      HIP_CHECK(hipMalloc(&p, cls));
    }
    std::lock_guard<std::mutex> g(mu_);
    sizes_[p] = cls;
    total_ += cls;
    return p;
  }
  void release(void* p, hipStream_t stream) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    free_[{stream, sizes_.at(p)}].push_back(p);
  }
  // Frees every cached block (after synchronising the device).
  void release_all() {
    std::lock_guard<std::mutex> g(mu_);
    (void)hipDeviceSynchronize();
    for (auto& kv : free_)
      for (void* p : kv.second) {
        total_ -= sizes_.at(p);
        sizes_.erase(p);
        (void)hipFree(p);
      }
    free_.clear();
  }
  size_t reserved_bytes() {
    std::lock_guard<std::mutex> g(mu_);
    return total_;
  }

 private:
  static size_t size_class(size_t bytes) {
    if (bytes < 256) bytes = 256;
    // powers of two up to 1 MiB, then 1/8-octave steps (bounded waste)
    size_t c = 256;
    while (c < bytes && c < (1u << 20)) c <<= 1;
    if (c >= bytes) return c;
    size_t step = 1;
    while ((step << 4) <= bytes) step <<= 1;  // step = highest power of two / 8
    return (bytes + step - 1) / step * step;
  }
  std::mutex mu_;
  std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> free_;
  std::unordered_map<void*, size_t> sizes_;
  size_t total_ = 0;
};

template <class T> class DArray {
 public:
  DArray() = default;
  DArray(size_t n, hipStream_t s) { reset(n, s); }
  ~DArray() { free(); }
  DArray(const DArray&) = delete;
  DArray& operator=(const DArray&) = delete;
  DArray(DArray&& o) noexcept : p_(o.p_), n_(o.n_), s_(o.s_) { o.p_ = nullptr; o.n_ = 0; }
  DArray& operator=(DArray&& o) noexcept {
    if (this != &o) {
      free();
      p_ = o.p_; n_ = o.n_; s_ = o.s_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  void reset(size_t n, hipStream_t s) {
    free();
    s_ = s;
    n_ = n;
    p_ = n ? (T*)Pool::instance().allocate(n * sizeof(T), s) : nullptr;
  }
  void free() {
    if (p_) Pool::instance().release(p_, s_);
    p_ = nullptr;
    n_ = 0;
  }
  // keep only the first n elements (n <= size()); the block is unchanged and
  // returns to the pool whole
  void shrink(size_t n) {
    if (n < n_) n_ = n;
  }
  T* data() const { return p_; }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  hipStream_t stream() const { return s_; }
  void zero() {
    if (n_) HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(T), s_));
  }
  std::vector<T> to_host() const {
    std::vector<T> h(n_);
    if (n_) {
      HIP_CHECK(hipMemcpyAsync(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost, s_));
      HIP_CHECK(hipStreamSynchronize(s_));
    }
    return h;
  }
  void from_host(const T* h, size_t n, hipStream_t s) {
    reset(n, s);
    if (n) HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
  hipStream_t s_ = nullptr;
};

// Small pinned host staging buffer for scalar reads that size the next stage.
class HostScalars {
 public:
  explicit HostScalars(size_t bytes = 4096) { HIP_CHECK(hipHostMalloc(&p_, bytes, hipHostMallocDefault)); }
  ~HostScalars() { if (p_) (void)hipHostFree(p_); }
  template <class T> T* as() { return (T*)p_; }

 private:
  void* p_ = nullptr;
};

inline unsigned blocks_for(size_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

// ------------------------------------------------------------------ primitives
// Implemented in psr_primitives.hip.
// Exclusive scan of n int32 values (out may alias in); returns nothing, total
// (sum of all inputs) is written to *total_dev when non-null.
void scan_exclusive_i32(const int* in, int* out, int64_t n, int* total_dev, hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits.
// keys/vals are sorted in place (scratch from the pool).
void sort_pairs_u32(uint32_t* keys, int* vals, int64_t n, int bits, hipStream_t s);
void sort_pairs_u64(uint64_t* keys, int* vals, int64_t n, int bits, hipStream_t s);
// heads[i] = 1 where sorted keys start a new run and are not the sentinel.
void mark_heads_u32(const uint32_t* keys, int64_t n, uint32_t sentinel, int* heads, hipStream_t s);
void mark_heads_u64(const uint64_t* keys, int64_t n, uint64_t sentinel, int* heads, hipStream_t s);
inline void sort_pairs(uint32_t* keys, int* vals, int64_t n, int bits, hipStream_t s) { sort_pairs_u32(keys, vals, n, bits, s); }
inline void sort_pairs(uint64_t* keys, int* vals, int64_t n, int bits, hipStream_t s) { sort_pairs_u64(keys, vals, n, bits, s); }
inline void mark_heads(const uint32_t* keys, int64_t n, uint32_t sentinel, int* heads, hipStream_t s) {
  mark_heads_u32(keys, n, sentinel, heads, s);
}
inline void mark_heads(const uint64_t* keys, int64_t n, uint64_t sentinel, int* heads, hipStream_t s) {
  mark_heads_u64(keys, n, sentinel, heads, s);
}
// A companion stream of caller stream s on device dev for one role (0: the
// serial normal average, 1: finer-level operator assembly): non-blocking,
// default priority, created on first use and kept for the process, so its
// pool blocks and queues are reused from call to call.
hipStream_t companion_stream(int dev, hipStream_t s, int role);
// Reads one int from device memory (synchronises the stream).
int read_int(const int* dev, hipStream_t s);
void read_bytes(void* host, const void* dev, size_t bytes, hipStream_t s);
// iota
void fill_iota(int* out, int64_t n, hipStream_t s);
// After heads + exclusive scan: compacted positions of heads (head_pos[g] = i).
void compact_heads(const int* heads, const int* prefix, int64_t n, int* head_pos, hipStream_t s);
// (x, y) column index over one level's sorted, same-depth keys whose x and y
// lie in [0, dim): start/end of each column's run.  Levels whose dim exceeds
// max_dim get no table (view() is then null and lookups search the array).
struct ColumnIndex {
  DArray<int> start, end;
  int dim = 0;
  ColumnTable view() const {
    ColumnTable t;
    if (dim) { t.start = start.data(); t.end = end.data(); t.dim = dim; }
    return t;
  }
};
void build_column_index(const int64_t* keys, int64_t n, int dim, ColumnIndex& out, hipStream_t s);
// Size limits of the lookup tables.  A level beyond a limit takes the next
// method -- dense index, then (x, y) column table, then a search of the
// sorted keys -- with the same results, slower and in less memory.  The
// defaults (2^28 dense cells: depth 9; column tables up to 4097 per side:
// depth 12; the field's column table up to depth 11) can be lowered with
// PSR_DENSE_INDEX_MAX_CELLS, PSR_COLUMN_TABLE_MAX_DIM and
// PSR_FIELD_TABLE_MAX_DEPTH (read at each use), e.g. to cap memory.
int64_t dense_index_max_cells();
int column_table_max_dim();
int field_table_max_depth();

inline int bits_for(uint64_t maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

}  // namespace psr
