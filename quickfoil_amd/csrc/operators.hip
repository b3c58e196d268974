// Context lifetime and the operators that keep the reference's exact layouts:
// RadixPartition (PartitionTuple pairs), BuildHashTableOnPartitions and BuildHashTableOnTable
// (FoilHashTable chains).
//
// A chained table built by sequential head insertion (operations/BuildHashTable.cpp:81-89) is
// a pure function of the bucket of every row: a row's link points at the previous row of the
// same bucket and each bucket's head is its last row.  The device build therefore stably sorts
// (bucket, row) pairs with the radix partition and links neighbours, which reproduces the
// reference arrays bit for bit without any ordering between threads.
#include "qf_internal.h"

#include <algorithm>

namespace qf {

void noalloc_scratch_setup(int device);  // noalloc.hip

static int bits_for(uint64_t n) {
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) < n) ++b;
  return b;
}

// ---- BuildHashTableOnPartitions -----------------------------------------------------------

__global__ void k_partition_bucket_keys(const int32_t* pairs, const int64_t* offsets, int radix_bits,
                                        const int64_t* bucket_offsets, int64_t n, int32_t* gkey,
                                        int32_t* local) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t v = pairs[2 * i];
  const uint32_t p = ref_hash(v) & ((1u << radix_bits) - 1u);
  const int64_t np = offsets[p + 1] - offsets[p];
  const uint32_t nb = next_pow2_u32((uint32_t)np);
  const uint32_t mask = (nb - 1u) << radix_bits;
  const uint32_t b = (ref_hash(v) & mask) >> radix_bits;  // HASH_BIT_MODULO
  gkey[i] = (int32_t)(bucket_offsets[p] + b);
  local[i] = (int32_t)(i - offsets[p]);
}

// Sorted (bucket, local index, global index): link consecutive rows of one bucket.
__global__ void k_link_sorted(const int32_t* skey, const int32_t* slocal, const int32_t* sglobal, int64_t n,
                              int32_t* buckets, int32_t* next) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t g = skey[k];
  const bool has_prev = k > 0 && skey[k - 1] == g;
  next[sglobal[k]] = has_prev ? slocal[k - 1] + 1 : 0;
  if (k == n - 1 || skey[k + 1] != g) buckets[g] = slocal[k] + 1;
}

// ---- BuildHashTableOnTable ------------------------------------------------------------------

struct KeyCols {
  const int32_t* c[QF_MAX_ARITY];
  int n;
};

__device__ inline uint32_t ref_hash_keys(const KeyCols& k, int64_t row) {
  uint32_t h = ref_hash(k.c[0][row]);
  for (int i = 1; i < k.n; ++i) h = ref_hash_combine(h, k.c[i][row]);
  return h;
}

__global__ void k_table_bucket_keys(KeyCols keys, int64_t n, uint32_t mask, int32_t* bucket) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bucket[i] = (int32_t)(ref_hash_keys(keys, i) & mask);
}

static void build_table_from_buckets(Context* ctx, const int32_t* bucket, int64_t n, int key_bits,
                                     const int32_t* local_or_null, int32_t* d_buckets, int32_t* d_next) {
  // Stable sort (bucket, local, row) by bucket, then link.
  DevBuf sk(ctx, n * 4), sl(ctx, n * 4), sg(ctx, n * 4);
  if (local_or_null) {
    const int32_t* in[2] = {bucket, local_or_null};
    int32_t* out[3] = {sk.as<int32_t>(), sl.as<int32_t>(), sg.as<int32_t>()};
    stable_sort_by_key(ctx, in, 2, 0, true, n, key_bits, out);
  } else {
    const int32_t* in[1] = {bucket};
    int32_t* out[2] = {sk.as<int32_t>(), sg.as<int32_t>()};
    stable_sort_by_key(ctx, in, 1, 0, true, n, key_bits, out);
  }
  const int32_t* slocal = local_or_null ? sl.as<int32_t>() : sg.as<int32_t>();
  hipLaunchKernelGGL(k_link_sorted, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream,
                     sk.as<int32_t>(), slocal, sg.as<int32_t>(), n, d_buckets, d_next);
  QF_CHECK_HIP(hipGetLastError());
}

void build_hash_table_on_table(Context* ctx, const qf_table* keys, int32_t* d_buckets, int32_t* d_next) {
  QF_REQUIRE(keys && keys->num_columns >= 1 && keys->num_columns <= QF_MAX_ARITY, "1..6 key columns");
  const int64_t n = keys->num_rows;
  QF_REQUIRE(n >= 0 && n < ((int64_t)1 << 31), "row count out of range");
  if (n == 0) return;
  const uint32_t nb = next_pow2_u32((uint32_t)n);
  KeyCols kc{};
  kc.n = keys->num_columns;
  for (int i = 0; i < kc.n; ++i) kc.c[i] = keys->columns[i];
  DevBuf bucket(ctx, n * 4);
  hipLaunchKernelGGL(k_table_bucket_keys, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, kc, n,
                     nb - 1u, bucket.as<int32_t>());
  QF_CHECK_HIP(hipGetLastError());
  fill_i32(ctx, d_buckets, 0, nb);
  build_table_from_buckets(ctx, bucket.as<int32_t>(), n, bits_for(nb), nullptr, d_buckets, d_next);
}

// ---- phase timing ------------------------------------------------------------------------------

static hipEvent_t take_event(Context* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  QF_CHECK_HIP(hipEventCreate(&e));
  return e;
}

PhaseScope::PhaseScope(Context* ctx, const char* name) : ctx_(ctx) {
  if (!ctx->timing) return;
  for (size_t i = 0; i < ctx->phases.size(); ++i)
    if (ctx->phases[i].name == name) index_ = (int)i;
  if (index_ < 0) {
    ctx->phases.emplace_back();
    ctx->phases.back().name = name;
    index_ = (int)ctx->phases.size() - 1;
  }
  start_ = take_event(ctx);
  QF_CHECK_HIP(hipEventRecord(start_, ctx->stream));
}

PhaseScope::~PhaseScope() {
  if (index_ < 0) return;
  hipEvent_t stop = nullptr;
  try {
    stop = take_event(ctx_);
  } catch (...) {
    return;
  }
  (void)hipEventRecord(stop, ctx_->stream);
  ctx_->phases[index_].pending.emplace_back(start_, stop);
  ctx_->phases[index_].launches++;
}

void collect_phase_times(Context* ctx) {
  for (PhaseTimer& p : ctx->phases) {
    for (auto& ev : p.pending) {
      QF_CHECK_HIP(hipEventSynchronize(ev.second));
      float ms = 0;
      QF_CHECK_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
      p.ms += ms;
      ctx->event_pool.push_back(ev.first);
      ctx->event_pool.push_back(ev.second);
    }
    p.pending.clear();
  }
}

void release_events(Context* ctx) {
  for (PhaseTimer& p : ctx->phases)
    for (auto& ev : p.pending) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  ctx->phases.clear();
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->event_pool.clear();
}

}  // namespace qf

// =============================================================================================
// C ABI

using namespace qf;

extern "C" {

int qf_context_create(int device, qf_context** out) {
  if (!out) return QF_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    return QF_ERR_DEVICE;
  }
  Context* ctx = new (std::nothrow) Context();
  if (!ctx) return QF_ERR_OOM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return QF_ERR_DEVICE;
  }
  ctx->own_stream = true;
  ctx->pool.reset(new Pool(device));
  noalloc_scratch_setup(device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  *out = reinterpret_cast<qf_context*>(ctx);
  return QF_OK;
}

void qf_context_destroy(qf_context* c) {
  Context* ctx = reinterpret_cast<Context*>(c);
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->bg_cache.clear();
  ctx->graph_cache.reset();
  ctx->eval_cache.reset();
  release_events(ctx);
  ctx->pool.reset();
  if (ctx->graph_stream) {
    (void)hipStreamSynchronize(ctx->graph_stream);
    (void)hipStreamDestroy(ctx->graph_stream);
  }
  if (ctx->graph_order_event) (void)hipEventDestroy(ctx->graph_order_event);
  for (hipStream_t st : ctx->aux_streams) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  for (hipEvent_t ev : ctx->aux_events) (void)hipEventDestroy(ev);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int qf_context_set_stream(qf_context* c, void* hip_stream) {
  return guarded(c, [&](Context* ctx) {
    QF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) QF_CHECK_HIP(hipStreamDestroy(ctx->stream));
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    ctx->own_stream = false;
  });
}

int qf_synchronize(qf_context* c) {
  return guarded(c, [&](Context* ctx) { sync(ctx); });
}

const char* qf_last_error(const qf_context* c) {
  const Context* ctx = reinterpret_cast<const Context*>(c);
  return ctx ? ctx->last_error.c_str() : "null context";
}

int qf_context_trim(qf_context* c) {
  return guarded(c, [&](Context* ctx) {
    sync(ctx);
    ctx->bg_cache.clear();
    ctx->graph_cache.reset();
    ctx->eval_cache.reset();
    ctx->pool->release_all();
  });
}

int qf_context_enable_timing(qf_context* c, int enable) {
  return guarded(c, [&](Context* ctx) { ctx->timing = enable != 0; });
}

int qf_context_reset_timing(qf_context* c) {
  return guarded(c, [&](Context* ctx) {
    collect_phase_times(ctx);
    for (PhaseTimer& p : ctx->phases) {
      p.ms = 0;
      p.launches = 0;
    }
  });
}

int qf_context_phase_times(qf_context* c, const char** names, double* ms, int64_t* launches, int max_phases) {
  int count = 0;
  const int st = guarded(c, [&](Context* ctx) {
    collect_phase_times(ctx);
    count = (int)ctx->phases.size();
    for (int i = 0; i < count && i < max_phases; ++i) {
      if (names) names[i] = ctx->phases[i].name.c_str();
      if (ms) ms[i] = ctx->phases[i].ms;
      if (launches) launches[i] = ctx->phases[i].launches;
    }
  });
  return st == QF_OK ? count : -st;
}

int qf_radix_partition(qf_context* c, const int32_t* d_values, int64_t n, int radix_bits, int32_t* d_out_pairs,
                       int64_t* d_offsets) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(radix_bits >= 1 && radix_bits <= 16, "radix_bits must be in [1, 16]");
    QF_REQUIRE(n >= 0 && n < ((int64_t)1 << 31), "row count out of range");
    QF_REQUIRE(n == 0 || (d_values && d_out_pairs), "null column");
    QF_REQUIRE(d_offsets, "null offsets");
    const int32_t* in[1] = {d_values};
    int32_t* out[2] = {d_out_pairs, d_out_pairs + 1};
    stable_partition(ctx, in, 1, 0, true, n, 0, radix_bits, out, 2, d_offsets);
  });
}

int64_t qf_partition_bucket_layout(const int64_t* h_offsets, int num_partitions, int64_t* h_bucket_offsets) {
  if (!h_offsets || num_partitions <= 0) return -QF_ERR_INVALID;
  int64_t total = 0;
  for (int p = 0; p < num_partitions; ++p) {
    if (h_bucket_offsets) h_bucket_offsets[p] = total;
    const int64_t np = h_offsets[p + 1] - h_offsets[p];
    if (np < 0) return -QF_ERR_INVALID;
    if (np > 0) total += next_pow2_u32((uint32_t)np);
  }
  if (h_bucket_offsets) h_bucket_offsets[num_partitions] = total;
  return total;
}

int qf_build_hash_table_on_partitions(qf_context* c, const int32_t* d_pairs, const int64_t* d_offsets,
                                      int radix_bits, const int64_t* d_bucket_offsets, int64_t total_buckets,
                                      int64_t n, int32_t* d_buckets, int32_t* d_next) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(radix_bits >= 1 && radix_bits <= 16, "radix_bits must be in [1, 16]");
    QF_REQUIRE(n >= 0 && n < ((int64_t)1 << 31) && total_buckets >= 0 && total_buckets < ((int64_t)1 << 31),
               "size out of range");
    if (total_buckets > 0) fill_i32(ctx, d_buckets, 0, total_buckets);
    if (n == 0) return;
    DevBuf gkey(ctx, n * 4), local(ctx, n * 4);
    hipLaunchKernelGGL(k_partition_bucket_keys, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream,
                       d_pairs, d_offsets, radix_bits, d_bucket_offsets, n, gkey.as<int32_t>(),
                       local.as<int32_t>());
    QF_CHECK_HIP(hipGetLastError());
    build_table_from_buckets(ctx, gkey.as<int32_t>(), n, bits_for((uint64_t)total_buckets), local.as<int32_t>(),
                             d_buckets, d_next);
  });
}

int qf_build_hash_table_on_table(qf_context* c, const qf_table* keys, int32_t* d_buckets, int32_t* d_next) {
  return guarded(c, [&](Context* ctx) { build_hash_table_on_table(ctx, keys, d_buckets, d_next); });
}

}  // extern "C"
