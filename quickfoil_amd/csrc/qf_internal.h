// Internal declarations shared by the HIP translation units of libquickfoil_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "quickfoil_amd.h"

namespace qf {

// ------------------------------------------------------------------------------------------
// Errors: internal code throws qf::Error; every extern "C" entry point catches it and returns
// a status, so no exception crosses the C ABI.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define QF_CHECK_HIP(expr)                                                            \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      throw ::qf::Error(QF_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define QF_REQUIRE(cond, msg)                                        \
  do {                                                               \
    if (!(cond)) throw ::qf::Error(QF_ERR_INVALID, std::string(msg)); \
  } while (0)

// ------------------------------------------------------------------------------------------
// Device memory pool: size-classed free lists over hipMalloc.  Work is ordered on one stream,
// so a block released after enqueueing its last use can be handed to later work on the same
// stream without a synchronisation.
class Pool {
 public:
  explicit Pool(int device) : device_(device) {}
  ~Pool() { release_all(); }
  void* get(size_t bytes);
  void put(void* p);
  void release_all();
  // Blocks released so far: work on another stream that writes a block taken after a release
  // must first wait for the work queued before it.
  uint64_t releases() const { return releases_; }

 private:
  static size_t size_class(size_t bytes);
  int device_;
  uint64_t releases_ = 0;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> live_;
};

struct Context;

// Scoped device buffer from the context pool.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(Context* ctx, size_t bytes);
  ~DevBuf();
  DevBuf(DevBuf&& o) noexcept : ctx_(o.ctx_), ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <class T> T* as() const { return static_cast<T*>(ptr_); }
  void* get() const { return ptr_; }
  size_t bytes() const { return bytes_; }

 private:
  Context* ctx_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// A relation partitioned on one column: all columns in partition order (SoA) plus offsets.
struct PartitionedTable {
  int radix_bits = 0;
  int num_columns = 0;
  int64_t num_rows = 0;
  std::vector<DevBuf> columns;      // num_columns (+1 when tids are carried)
  DevBuf offsets;                   // 2^radix_bits + 1 int64
  std::vector<int64_t> h_offsets;   // host copy
  bool has_tid = false;
};

// Accumulated event timings of one named phase.
struct PhaseTimer {
  std::string name;
  double ms = 0;
  int64_t launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  std::unique_ptr<Pool> pool;
  int num_cus = 256;
  // Background partition cache keyed by (column pointer, rows, column, radix bits).
  std::map<std::tuple<const void*, int64_t, int, int>, std::shared_ptr<PartitionedTable>> bg_cache;
  // Evaluation cache (partitioned relations and their HBM tables), owned by evaluate.hip.
  std::shared_ptr<void> eval_cache;
  // Captured evaluation step (graph and the buffers it uses), owned by evaluate.hip.
  std::shared_ptr<void> graph_cache;
  // Autotuner choices per workload shape, owned by evaluate.hip.
  std::shared_ptr<void> tune_cache;
  // Pinned host staging for a plan's uploads and the counts of uncaptured steps, owned by
  // evaluate.hip.
  std::shared_ptr<void> host_stage;
  // Private non-blocking stream for captured graphs (the caller's stream may be the legacy
  // default stream, which cannot be captured); ordered after the caller's stream by an event.
  hipStream_t graph_stream = nullptr;
  hipEvent_t graph_order_event = nullptr;
  // Streams (default priority) and events for independent work forked from `stream` and joined
  // back to it (the binding table's partitions per join variable); made on first use.
  std::vector<hipStream_t> aux_streams;
  std::vector<hipEvent_t> aux_events;
  // While a step is being captured, released pool blocks move here instead of the pool.
  std::vector<DevBuf>* capture_hold = nullptr;
  bool timing = false;
  std::vector<PhaseTimer> phases;
  std::vector<hipEvent_t> event_pool;
};

// Records events around a phase on the context stream when timing is enabled.
class PhaseScope {
 public:
  PhaseScope(Context* ctx, const char* name);
  ~PhaseScope();

 private:
  Context* ctx_;
  int index_ = -1;
  hipEvent_t start_ = nullptr;
};
// Resolves pending events (synchronises) into the accumulated phase times.
void collect_phase_times(Context* ctx);
void release_events(Context* ctx);

// ------------------------------------------------------------------------------------------
// Shared device helpers.

// murmur3 finaliser: a bijection on uint32 used for internal partitioning and slots.
__host__ __device__ inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// Reference hashing (utility/Hash.hpp:20-28): identity for int32, boost-style combine.
__host__ __device__ inline uint32_t ref_hash(int32_t v) { return (uint32_t)v; }
__host__ __device__ inline uint32_t ref_hash_combine(uint32_t seed, int32_t v) {
  return seed ^ (ref_hash(v) + 0x9e3779b9u + (seed << 6) + (seed >> 2));
}
__host__ __device__ inline uint32_t next_pow2_u32(uint32_t v) {
  v--;
  v |= v >> 1;
  v |= v >> 2;
  v |= v >> 4;
  v |= v >> 8;
  v |= v >> 16;
  return v + 1;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------------------------------
// A fixed device region that an operation carves its temporaries from, in order, instead of
// taking them from the context pool (for entry points that allocate nothing).
struct Scratch {
  char* base = nullptr;
  size_t capacity = 0;
  size_t used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at + bytes > capacity) throw Error(QF_ERR_INVALID, "scratch region too small");
    used = at + bytes;
    return base + at;
  }
};

// ------------------------------------------------------------------------------------------
// Primitives implemented in primitives.hip.

// Exclusive prefix sum of n int64 values (in and out may alias).  Returns nothing; the total
// is out[n-1] + in[n-1] and is not computed here.  Temporaries come from `scratch` when given.
void scan_exclusive_i64(Context* ctx, const int64_t* in, int64_t* out, int64_t n, Scratch* scratch = nullptr);
// Bytes scan_exclusive_i64 takes from a Scratch for n values (256-byte aligned pieces).
size_t scan_scratch_bytes(int64_t n);

void fill_i32(Context* ctx, int32_t* p, int32_t v, int64_t n);
void fill_i64(Context* ctx, int64_t* p, int64_t v, int64_t n);
void fill_u64(Context* ctx, uint64_t* p, uint64_t v, int64_t n);

// Digit of a row for partitioning: ((mixed ? mix32(k) : (uint32)k) >> shift) & ((1<<bits)-1).
struct DigitFn {
  int mixed = 0;   // 0: identity hash of the value, 1: mix32
  int shift = 0;
  int bits = 0;
};

// Stable partition of num_rows rows (num_in_cols int32 input columns, optionally followed by a
// generated row-id column) by the digit of column key_col.  total_bits may exceed 8; the work
// is split into stable LSD passes of at most 8 bits.  Output columns are written with element
// stride out_stride (1 = SoA, 2 = interleaved pairs).  When d_offsets is non-null it receives
// 2^total_bits + 1 partition starts.  total digit: mixed ? mix32(k) >> (32 - total_bits)
// : k & ((1 << total_bits) - 1).
void stable_partition(Context* ctx, const int32_t* const* in_cols, int num_in_cols, int key_col,
                      bool gen_tid, int64_t num_rows, int mixed, int total_bits,
                      int32_t* const* out_cols, int out_stride, int64_t* d_offsets, Scratch* scratch = nullptr);
// Bytes stable_partition (nkeys = 1) or stable_partition_multi takes from a Scratch for num_rows
// rows of ncols written columns per key.
size_t partition_scratch_bytes(int64_t num_rows, int ncols, int total_bits, int nkeys = 1);

// The same rows partitioned by each of nkeys <= PT_MAXKEYS key columns (mixed digits as
// stable_partition, SoA outputs): out_cols[k * ncols + c] is written column c of key k, where
// ncols = num_in_cols (+ 1 for the row id), and d_offsets[k] (or d_offsets null) receives key k's
// 2^total_bits + 1 partition starts.  The first pass reads every input tile once for all keys.
constexpr int PT_MAXKEYS = 4;
void stable_partition_multi(Context* ctx, const int32_t* const* in_cols, int num_in_cols, const int* key_cols,
                            int nkeys, bool gen_tid, int64_t num_rows, int mixed, int total_bits,
                            int32_t* const* out_cols, int64_t* const* d_offsets, Scratch* scratch = nullptr);

// Stable sort of (key, payload...) rows by uint32 key values below 2^key_bits (raw digits).
void stable_sort_by_key(Context* ctx, const int32_t* const* in_cols, int num_in_cols, int key_col,
                        bool gen_tid, int64_t num_rows, int key_bits, int32_t* const* out_cols);

// Host helpers.
void copy_d2h(Context* ctx, void* dst, const void* src, size_t bytes);
void copy_h2d(Context* ctx, void* dst, const void* src, size_t bytes);
void sync(Context* ctx);

// Runs f(ctx) with the device set and converts exceptions into a status and last_error.
template <class F>
inline int guarded(qf_context* c, F&& f) {
  Context* ctx = reinterpret_cast<Context*>(c);
  if (!ctx) return QF_ERR_INVALID;
  try {
    QF_CHECK_HIP(hipSetDevice(ctx->device));
    f(ctx);
    return QF_OK;
  } catch (const Error& e) {
    ctx->last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    ctx->last_error = "host allocation failed";
    return QF_ERR_OOM;
  } catch (const std::exception& e) {
    ctx->last_error = e.what();
    return QF_ERR_INVALID;
  }
}

}  // namespace qf
