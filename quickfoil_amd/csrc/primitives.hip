// Device primitives: memory pool, fills, exclusive scan, and the stable radix partition that
// RadixPartition, the chained-table builds and the evaluator's co-partitioning all use.
//
// Stable partition of one tile (4096 rows, 16 waves of 64 lanes x 4 rows) per workgroup:
//   1. histogram kernel: per-tile digit counts in LDS, written bin-major to HBM;
//   2. exclusive scan over (bin, tile) gives each tile's output base per bin;
//   3. scatter kernel: the tile's rows are loaded into registers first (every column's loads in
//      flight together), per-wave LDS histograms and wave-level prefix sums give every row its
//      stable rank (peers found with one ballot per digit bit), rows are staged in LDS in
//      bin order, then written out so consecutive lanes store consecutive addresses of a bin.
// Digits wider than 8 bits are done as stable LSD passes of <= 8 bits.
#include "qf_internal.h"

#include <algorithm>

namespace qf {

// ------------------------------------------------------------------------------------------
// Pool

size_t Pool::size_class(size_t bytes) {
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}

void* Pool::get(size_t bytes) {
  if (bytes == 0) bytes = 1;
  const size_t c = size_class(bytes);
  auto it = free_.find(c);
  void* p = nullptr;
  if (it != free_.end()) {
    p = it->second;
    free_.erase(it);
  } else {
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      // Return cached blocks to HIP and retry once.
      (void)hipGetLastError();
      for (auto& kv : free_) (void)hipFree(kv.second);
      free_.clear();
      e = hipMalloc(&p, c);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(QF_ERR_OOM, "hipMalloc of " + std::to_string(c) + " bytes failed");
      }
    }
  }
  live_[p] = c;
  return p;
}

void Pool::put(void* p) {
  if (!p) return;
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.emplace(it->second, p);
  live_.erase(it);
  ++releases_;
}

void Pool::release_all() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
}

DevBuf::DevBuf(Context* ctx, size_t bytes) : ctx_(ctx), bytes_(bytes) { ptr_ = ctx->pool->get(bytes); }
DevBuf::~DevBuf() {
  if (!ptr_ || !ctx_) return;
  if (ctx_->capture_hold) {
    // Keep the block reserved for the graph being captured.
    DevBuf keep;
    keep.ctx_ = ctx_;
    keep.ptr_ = ptr_;
    keep.bytes_ = bytes_;
    ptr_ = nullptr;
    ctx_->capture_hold->push_back(std::move(keep));
    return;
  }
  ctx_->pool->put(ptr_);
}
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    if (ptr_ && ctx_) ctx_->pool->put(ptr_);
    ctx_ = o.ctx_;
    ptr_ = o.ptr_;
    bytes_ = o.bytes_;
    o.ptr_ = nullptr;
  }
  return *this;
}

void copy_d2h(Context* ctx, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  QF_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  QF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
}
void copy_h2d(Context* ctx, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  QF_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  QF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
}
void sync(Context* ctx) { QF_CHECK_HIP(hipStreamSynchronize(ctx->stream)); }

// ------------------------------------------------------------------------------------------
// Fills (kernels rather than hipMemsetAsync so they can sit inside captured graphs).

template <class T>
__global__ void k_fill(T* p, T v, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <class T>
static void fill_impl(Context* ctx, T* p, T v, int64_t n) {
  if (n <= 0) return;
  const int blocks = (int)std::min<int64_t>(ceil_div(n, 256), 4096);
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, ctx->stream, p, v, n);
  QF_CHECK_HIP(hipGetLastError());
}
void fill_i32(Context* ctx, int32_t* p, int32_t v, int64_t n) { fill_impl(ctx, p, v, n); }
void fill_i64(Context* ctx, int64_t* p, int64_t v, int64_t n) { fill_impl(ctx, p, v, n); }
void fill_u64(Context* ctx, uint64_t* p, uint64_t v, int64_t n) { fill_impl(ctx, p, v, n); }

// ------------------------------------------------------------------------------------------
// Exclusive scan of int64 (three kernels, recursive over block sums).

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ inline int64_t block_exclusive_scan_i64(int64_t v, int64_t* lds, int64_t* total) {
  // Inclusive scan inside each wave by shuffles, then the four waves' totals through LDS.
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const long long y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) lds[w] = incl;
  __syncthreads();
  int64_t before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < SCAN_THREADS / 64; ++i) {
    const int64_t t = lds[i];
    if (i < w) before += t;
    all += t;
  }
  *total = all;
  __syncthreads();  // lds is free again for the caller
  return before + (int64_t)incl - v;
}

__global__ void k_scan_reduce(const int64_t* in, int64_t n, int64_t* block_sums) {
  __shared__ int64_t lds[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int64_t idx = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  int64_t total;
  block_exclusive_scan_i64(s, lds, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void k_scan_down(const int64_t* in, int64_t* out, int64_t n, const int64_t* block_offsets) {
  __shared__ int64_t lds[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t vals[SCAN_ITEMS];
  int64_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int64_t idx = base + i;
    vals[i] = idx < n ? in[idx] : 0;
    s += vals[i];
  }
  int64_t total;
  int64_t run = block_exclusive_scan_i64(s, lds, &total) + (block_offsets ? block_offsets[blockIdx.x] : 0);
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const int64_t idx = base + i;
    if (idx < n) out[idx] = run;
    run += vals[i];
  }
}

static size_t aligned256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t scan_scratch_bytes(int64_t n) {
  size_t total = 0;
  for (int64_t blocks = ceil_div(n, SCAN_TILE); blocks > 1; blocks = ceil_div(blocks, SCAN_TILE))
    total += aligned256((size_t)blocks * sizeof(int64_t));
  return total;
}

void scan_exclusive_i64(Context* ctx, const int64_t* in, int64_t* out, int64_t n, Scratch* scratch) {
  if (n <= 0) return;
  const int64_t blocks = ceil_div(n, SCAN_TILE);
  if (blocks == 1) {
    hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, in, out, n, nullptr);
    QF_CHECK_HIP(hipGetLastError());
    return;
  }
  DevBuf pooled;
  int64_t* sums = nullptr;
  if (scratch) {
    sums = static_cast<int64_t*>(scratch->take(blocks * sizeof(int64_t)));
  } else {
    pooled = DevBuf(ctx, blocks * sizeof(int64_t));
    sums = pooled.as<int64_t>();
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)blocks), dim3(SCAN_THREADS), 0, ctx->stream, in, n, sums);
  QF_CHECK_HIP(hipGetLastError());
  scan_exclusive_i64(ctx, sums, sums, blocks, scratch);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)blocks), dim3(SCAN_THREADS), 0, ctx->stream, in, out, n, sums);
  QF_CHECK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------
// Stable radix partition.
//
// One call partitions the same rows by up to PT_MAXKEYS key columns (the binding table by each
// clause join variable): every key has its own outputs, bases and offsets.  The first pass reads
// each input tile once and scatters it for every key in turn; later passes and the offsets run
// over all keys' tiles in one launch.

constexpr int PT_THREADS = 1024;
constexpr int PT_WAVES = PT_THREADS / 64;
constexpr int PT_IPT = 4;
constexpr int PT_TILE = PT_THREADS * PT_IPT;
constexpr int PT_MAXCOLS = 8;
constexpr int PT_MAXBINS = 256;
static_assert(PT_MAXKEYS == 4, "key slots of PassArgs");

struct PassArgs {
  const int32_t* in[PT_MAXKEYS][PT_MAXCOLS];  // per key: input columns (the same for every key in
                                              // a first pass that reads each tile once)
  int32_t* out[PT_MAXKEYS][PT_MAXCOLS];       // per key: output columns
  int key_col[PT_MAXKEYS];
  int nkeys;
  int ncols;      // columns written (inputs plus the generated row id, if any)
  int tid_col;    // index of the generated row-id column, or -1
  int out_stride;
  int64_t n;          // rows of every key's input
  int64_t num_tiles;  // tiles of every key's input
  int mixed, shift, bits;
};

__device__ inline uint32_t digit_of(int32_t k, int mixed, int shift, int bits) {
  const uint32_t h = mixed ? mix32((uint32_t)k) : (uint32_t)k;
  return bits == 0 ? 0u : (h >> shift) & ((1u << bits) - 1u);
}

// Digit counts of one tile per wave (sixteen tiles per workgroup, key blockIdx.y), written
// key-major, then bin-major (counts[(key * bins + b) * num_tiles + t]), so that one exclusive scan
// gives every (key, tile) its output base per bin, plus key * n.
constexpr int PT_HIST_RPL = PT_TILE / 64;  // rows per lane
__global__ __launch_bounds__(PT_THREADS) void k_part_hist(PassArgs a, int64_t* counts) {
  __shared__ uint32_t hist[PT_WAVES][PT_MAXBINS];
  const int nb = 1 << a.bits;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int key_index = blockIdx.y;
  const int64_t tile = (int64_t)blockIdx.x * PT_WAVES + w;
  if (tile >= a.num_tiles) return;  // whole waves only; no workgroup barrier follows
  for (int b = lane; b < nb; b += 64) hist[w][b] = 0;
  __builtin_amdgcn_wave_barrier();
  const int64_t base = tile * PT_TILE;
  const int32_t* key = a.in[key_index][a.key_col[key_index]] + base;
  const int rows = (int)min((int64_t)PT_TILE, a.n - base);
  int32_t k[PT_HIST_RPL];
  if (rows == PT_TILE) {
#pragma unroll
    for (int i = 0; i < PT_HIST_RPL; ++i) k[i] = key[i * 64 + lane];
#pragma unroll
    for (int i = 0; i < PT_HIST_RPL; ++i) atomicAdd(&hist[w][digit_of(k[i], a.mixed, a.shift, a.bits)], 1u);
  } else {
    for (int r = lane; r < rows; r += 64) atomicAdd(&hist[w][digit_of(key[r], a.mixed, a.shift, a.bits)], 1u);
  }
  __builtin_amdgcn_wave_barrier();
  int64_t* out = counts + (int64_t)key_index * nb * a.num_tiles;
  for (int b = lane; b < nb; b += 64) out[(int64_t)b * a.num_tiles + tile] = hist[w][b];
}

// LDS of the scatter kernel beyond its static arrays: the staged tile (ncols x PT_TILE words)
// and one digit byte per row.
static size_t scatter_lds_bytes(int ncols) { return (size_t)ncols * PT_TILE * 4 + PT_TILE; }

// Tiles of one scatter workgroup.  Workgroups are dealt to the 8 XCDs round robin; when the grid
// is a multiple of 8, XCD x owns the contiguous tiles [x*n/8, (x+1)*n/8) and its workgroups
// stride through them, so the tiles an XCD works on at once are neighbours and the runs they
// write into one bin meet in that XCD's L2.
struct TileWalk {
  int64_t next, end, step;
  __device__ TileWalk(int64_t n) {
    const int64_t G = gridDim.x, b = blockIdx.x;
    if (G % 8 == 0) {
      const int64_t x = b & 7, q = n >> 3, r = n & 7;
      const int64_t start = x * q + (x < r ? x : r);
      next = start + (b >> 3);
      end = start + q + (x < r ? 1 : 0);
      step = G >> 3;
    } else {
      next = b;
      end = n;
      step = G;
    }
  }
};

// Loads one tile of key `ki`'s input (NC columns) into registers: every column's loads are in
// flight together.  Addresses are a wave-uniform base per column plus the lane's offset.
template <int NC>
__device__ inline void load_tile(const PassArgs& a, int ki, int64_t tile, int lane, int w, int32_t (&v)[PT_IPT][NC]) {
  const int64_t wave_base = tile * PT_TILE + (int64_t)w * 64 * PT_IPT;
  if (wave_base + 64 * PT_IPT <= a.n) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c == a.tid_col) {
#pragma unroll
        for (int i = 0; i < PT_IPT; ++i) v[i][c] = (int32_t)(wave_base + i * 64 + lane);
      } else {
        const int32_t* src = a.in[ki][c] + wave_base;
#pragma unroll
        for (int i = 0; i < PT_IPT; ++i) v[i][c] = src[i * 64 + lane];
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < PT_IPT; ++i) {
    const int64_t row = wave_base + (int64_t)i * 64 + lane;
    const bool ok = row < a.n;
#pragma unroll
    for (int c = 0; c < NC; ++c) v[i][c] = c == a.tid_col ? (int32_t)row : (ok ? a.in[ki][c][row] : 0);
  }
}

// Persistent workgroups, software pipelined: a tile's rows are loaded into registers while the
// previous tile is written out of LDS.  A workgroup walks the tiles of all keys' inputs, key after
// key (one key, or tables too wide for k_part_scatter_keys, which takes several keys together).
// NC = a.ncols.  4096-row tiles: at 7-bit digits a bin's run in a tile averages 32 rows (128 B per
// column), which measured 10-16 % faster than 2048-row tiles at the same number of waves per CU.
template <int NC>
__global__ __launch_bounds__(PT_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_part_scatter(
    PassArgs a, const int64_t* bases) {
  extern __shared__ int32_t stage[];                // ncols x PT_TILE rows in bin order, then digits
  // Per-wave counts, then per-wave running starts, 16 bits each (a tile has 4096 rows): wave w
  // uses half (w & 1) of row (w >> 1).
  __shared__ uint32_t wrun[PT_WAVES / 2][PT_MAXBINS];
  __shared__ int64_t gbase[PT_MAXBINS];  // output row of bin b's first tile row, minus its tile start
  __shared__ uint32_t wtot[PT_MAXBINS / 64];
  uint8_t* dig8 = reinterpret_cast<uint8_t*>(stage + a.ncols * PT_TILE);

  const int nb = 1 << a.bits;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int wsh = 16 * (w & 1);
  // Walk position g -> (key of the input, tile of that input).
  const int64_t T = a.num_tiles;
  TileWalk walk(T * a.nkeys);
  int32_t v[PT_IPT][NC];
  if (walk.next < walk.end) load_tile<NC>(a, (int)(walk.next / T), walk.next % T, lane, w, v);

  for (int64_t g = walk.next; g < walk.end; g += walk.step) {
    const int kin = (int)(g / T);
    const int64_t tile = g % T;
    const int k_end = kin + 1;
    // Each wave owns PT_IPT consecutive 64-row groups of the tile: (wave, group, lane) is row order.
    const int64_t tile_base = tile * PT_TILE;
    const int64_t wave_base = tile_base + (int64_t)w * 64 * PT_IPT;
    for (int k = kin; k < k_end; ++k) {
      const int key_col = a.key_col[k];
      for (int b = threadIdx.x; b < PT_WAVES / 2 * PT_MAXBINS; b += PT_THREADS) (&wrun[0][0])[b] = 0;
      uint32_t d[PT_IPT];
#pragma unroll
      for (int i = 0; i < PT_IPT; ++i) {
        int32_t key = v[i][0];
#pragma unroll
        for (int c = 1; c < NC; ++c)
          if (c == key_col) key = v[i][c];
        d[i] = digit_of(key, a.mixed, a.shift, a.bits);
      }
      __syncthreads();  // counts cleared; the previous write-out is done with LDS
#pragma unroll
      for (int i = 0; i < PT_IPT; ++i)
        if (wave_base + (int64_t)i * 64 + lane < a.n) atomicAdd(&wrun[w >> 1][d[i]], 1u << wsh);
      __syncthreads();

      // Bin totals, their exclusive scan (one lane per bin, 64 bins per wave), per-wave starts.
      uint32_t tot = 0, incl = 0;
      if (threadIdx.x < PT_MAXBINS) {
        if ((int)threadIdx.x < nb)
          for (int ww = 0; ww < PT_WAVES / 2; ++ww) {
            const uint32_t pair = wrun[ww][threadIdx.x];
            tot += (pair & 0xffffu) + (pair >> 16);
          }
        incl = tot;
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t y = __shfl_up(incl, off, 64);
          if (lane >= off) incl += y;
        }
        if (lane == 63) wtot[w] = incl;
      }
      __syncthreads();
      if ((int)threadIdx.x < nb) {
        uint32_t excl = incl - tot;
        for (int ww = 0; ww < w; ++ww) excl += wtot[ww];
        uint32_t at = excl;
        for (int ww = 0; ww < PT_WAVES / 2; ++ww) {
          const uint32_t pair = wrun[ww][threadIdx.x];
          const uint32_t lo = at, hi = at + (pair & 0xffffu);
          wrun[ww][threadIdx.x] = lo | (hi << 16);
          at = hi + (pair >> 16);
        }
        // The scan over all keys' counts puts key k's rows after the k * n rows of the keys before.
        gbase[threadIdx.x] = bases[((int64_t)k * nb + threadIdx.x) * T + tile] - (int64_t)k * a.n - (int64_t)excl;
      }
      __syncthreads();

      // Stable ranks within the wave (peers: lanes of the group with the same digit, one ballot per
      // digit bit), rows staged in LDS in bin order.
#pragma unroll
      for (int i = 0; i < PT_IPT; ++i) {
        const int64_t row = wave_base + (int64_t)i * 64 + lane;
        const bool valid = row < a.n;
        const uint64_t vmask = __ballot(valid);
        if (vmask == 0) break;
        uint64_t peers = vmask;
        for (int kb = 0; kb < a.bits; ++kb) {
          const bool bit = (d[i] >> kb) & 1u;
          const uint64_t bk = __ballot(bit);
          peers &= bit ? bk : ~bk;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        uint32_t pos = 0;
        if (valid) pos = ((wrun[w >> 1][d[i]] >> wsh) & 0xffffu) + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
#pragma unroll
          for (int c = 0; c < NC; ++c) stage[c * PT_TILE + pos] = v[i][c];
          dig8[pos] = (uint8_t)d[i];
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) atomicAdd(&wrun[w >> 1][d[i]], (uint32_t)__popcll(peers) << wsh);
        __builtin_amdgcn_wave_barrier();
      }
      __syncthreads();

      // After the last key's staging the next tile's loads fly while this one is written out.
      if (k == k_end - 1 && g + walk.step < walk.end) {
        const int64_t gn = g + walk.step;
        load_tile<NC>(a, (int)(gn / T), gn % T, lane, w, v);
      }

      // Consecutive threads store consecutive rows of a bin.
      const int tile_rows = (int)min((int64_t)PT_TILE, a.n - tile_base);
      int32_t* const* out = a.out[k];
      for (int j = threadIdx.x; j < tile_rows; j += PT_THREADS) {
        const int64_t dest = gbase[dig8[j]] + j;
#pragma unroll
        for (int c = 0; c < NC; ++c) out[c][dest * a.out_stride] = stage[c * PT_TILE + j];
      }
    }
  }
}

// Every key's key column of one tile (rows past the end read as 0).
__device__ inline void load_key_columns(const PassArgs& a, int64_t tile, int lane, int w,
                                        int32_t (&kv)[PT_MAXKEYS][PT_IPT]) {
  const int64_t wave_base = tile * PT_TILE + (int64_t)w * 64 * PT_IPT;
  const bool full = wave_base + 64 * PT_IPT <= a.n;
#pragma unroll
  for (int k = 0; k < PT_MAXKEYS; ++k) {
    if (k >= a.nkeys) break;
    const int32_t* src = a.in[k][a.key_col[k]] + wave_base;
#pragma unroll
    for (int i = 0; i < PT_IPT; ++i) kv[k][i] = full || wave_base + i * 64 + lane < a.n ? src[i * 64 + lane] : 0;
  }
}

// Every key's tile of one tile index together (the binding table by each clause join variable):
// the digit counts of all keys are taken in one phase and scanned in one phase (a thread per key
// and bin), then each key's tile is staged and written out in turn: two barriers per key after
// the first instead of five.  SHARED (a first pass over one input for every key): the tile is
// read once for all keys.  Otherwise the key columns of all keys' tiles (their digits) load while
// the tile before is written out, the first key's rows during the count and scan phases, and each
// further key's rows while the key before is written out.
template <int NC, bool SHARED>
__global__ __launch_bounds__(PT_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_part_scatter_keys(
    PassArgs a, const int64_t* bases) {
  extern __shared__ int32_t stage[];  // ncols x PT_TILE rows in bin order, then digits
  // Per key: per-wave counts, then per-wave running starts, 16 bits each (wave w uses half
  // (w & 1) of row (w >> 1)).
  __shared__ uint32_t wrun[PT_MAXKEYS][PT_WAVES / 2][PT_MAXBINS];
  __shared__ int64_t gbase[PT_MAXKEYS][PT_MAXBINS];  // output row of bin b's first tile row, minus its tile start
  __shared__ uint32_t wtot[PT_WAVES];
  static_assert(PT_MAXKEYS * PT_MAXBINS == PT_THREADS, "one scan thread per key and bin");
  uint8_t* dig8 = reinterpret_cast<uint8_t*>(stage + a.ncols * PT_TILE);

  const int nb = 1 << a.bits;
  const int nk = a.nkeys;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int wsh = 16 * (w & 1);
  const int sk = threadIdx.x / PT_MAXBINS, sb = threadIdx.x % PT_MAXBINS;  // the thread's key and bin in the scan
  const int64_t T = a.num_tiles;
  TileWalk walk(T);
  for (int b = threadIdx.x; b < nk * (PT_WAVES / 2) * PT_MAXBINS; b += PT_THREADS) (&wrun[0][0][0])[b] = 0;
  int32_t v[PT_IPT][NC];
  int32_t kv[SHARED ? 1 : PT_MAXKEYS][PT_IPT];  // !SHARED: every key's key column of the tile
  if (walk.next < walk.end) {
    if constexpr (SHARED)
      load_tile<NC>(a, 0, walk.next, lane, w, v);
    else
      load_key_columns(a, walk.next, lane, w, kv);
  }

  for (int64_t tile = walk.next; tile < walk.end; tile += walk.step) {
    const int64_t tile_base = tile * PT_TILE;
    const int64_t wave_base = tile_base + (int64_t)w * 64 * PT_IPT;
    // Every key's digit of the lane's rows, a byte each.
    uint32_t d4[PT_IPT];
#pragma unroll
    for (int i = 0; i < PT_IPT; ++i) {
      d4[i] = 0;
#pragma unroll
      for (int k = 0; k < PT_MAXKEYS; ++k) {
        if (k >= nk) break;
        int32_t key;
        if constexpr (SHARED) {
          const int kc = a.key_col[k];
          key = v[i][0];
#pragma unroll
          for (int c = 1; c < NC; ++c)
            if (c == kc) key = v[i][c];
        } else {
          key = kv[k][i];
        }
        d4[i] |= digit_of(key, a.mixed, a.shift, a.bits) << (8 * k);
      }
    }
    // The first key's rows load during the count and scan phases.
    if constexpr (!SHARED) load_tile<NC>(a, 0, tile, lane, w, v);
    __syncthreads();  // counts cleared; the previous tile's write-out is done with LDS
#pragma unroll
    for (int i = 0; i < PT_IPT; ++i)
      if (wave_base + (int64_t)i * 64 + lane < a.n)
        for (int k = 0; k < nk; ++k) atomicAdd(&wrun[k][w >> 1][(d4[i] >> (8 * k)) & 0xffu], 1u << wsh);
    __syncthreads();

    // Bin totals of every key, their exclusive scans (64 bins per wave), per-wave starts.
    uint32_t tot = 0;
    if (sk < nk && sb < nb)
      for (int ww = 0; ww < PT_WAVES / 2; ++ww) {
        const uint32_t pair = wrun[sk][ww][sb];
        tot += (pair & 0xffffu) + (pair >> 16);
      }
    uint32_t incl = tot;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    if (sk < nk && sb < nb) {
      uint32_t excl = incl - tot;
      for (int ww = sk * (PT_MAXBINS / 64); ww < w; ++ww) excl += wtot[ww];
      uint32_t at = excl;
      for (int ww = 0; ww < PT_WAVES / 2; ++ww) {
        const uint32_t pair = wrun[sk][ww][sb];
        const uint32_t lo = at, hi = at + (pair & 0xffffu);
        wrun[sk][ww][sb] = lo | (hi << 16);
        at = hi + (pair >> 16);
      }
      // The scan over all keys' counts puts key k's rows after the k * n rows of the keys before.
      gbase[sk][sb] = bases[((int64_t)sk * nb + sb) * T + tile] - (int64_t)sk * a.n - (int64_t)excl;
    }
    __syncthreads();

    const int tile_rows = (int)min((int64_t)PT_TILE, a.n - tile_base);
    for (int k = 0; k < nk; ++k) {
      // Stable ranks within the wave (peers: lanes of the group with the same digit, one ballot per
      // digit bit), rows staged in LDS in bin order.
#pragma unroll
      for (int i = 0; i < PT_IPT; ++i) {
        const int64_t row = wave_base + (int64_t)i * 64 + lane;
        const bool valid = row < a.n;
        const uint64_t vmask = __ballot(valid);
        if (vmask == 0) break;
        const uint32_t d = (d4[i] >> (8 * k)) & 0xffu;
        uint64_t peers = vmask;
        for (int kb = 0; kb < a.bits; ++kb) {
          const bool bit = (d >> kb) & 1u;
          const uint64_t bk = __ballot(bit);
          peers &= bit ? bk : ~bk;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        uint32_t pos = 0;
        if (valid) pos = ((wrun[k][w >> 1][d] >> wsh) & 0xffffu) + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
#pragma unroll
          for (int c = 0; c < NC; ++c) stage[c * PT_TILE + pos] = v[i][c];
          dig8[pos] = (uint8_t)d;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) atomicAdd(&wrun[k][w >> 1][d], (uint32_t)__popcll(peers) << wsh);
        __builtin_amdgcn_wave_barrier();
      }
      __syncthreads();

      if (k == nk - 1) {
        // No rank reads the counts any more: clear them for the next tile, whose loads fly while
        // this one is written out.
        for (int b = threadIdx.x; b < nk * (PT_WAVES / 2) * PT_MAXBINS; b += PT_THREADS) (&wrun[0][0][0])[b] = 0;
        if (tile + walk.step < walk.end) {
          if constexpr (SHARED)
            load_tile<NC>(a, 0, tile + walk.step, lane, w, v);
          else
            load_key_columns(a, tile + walk.step, lane, w, kv);
        }
      } else if (!SHARED) {
        load_tile<NC>(a, k + 1, tile, lane, w, v);  // the staged rows no longer need the registers
      }

      // Consecutive threads store consecutive rows of a bin.
      int32_t* const* out = a.out[k];
      for (int j = threadIdx.x; j < tile_rows; j += PT_THREADS) {
        const int64_t dest = gbase[k][dig8[j]] + j;
#pragma unroll
        for (int c = 0; c < NC; ++c) out[c][dest * a.out_stride] = stage[c * PT_TILE + j];
      }
      if (k < nk - 1) __syncthreads();  // the next key's staging overwrites the tile
    }
  }
}

// offsets[b] = first row whose digit is >= b, for b in [0, num_parts]; rows are sorted by digit.
// Each thread takes PT_OFF_RPT rows, 256 apart (coalesced); key blockIdx.y.
constexpr int PT_OFF_RPT = 8;
struct OffsetArgs {
  const int32_t* key[PT_MAXKEYS];
  int64_t* offsets[PT_MAXKEYS];
};
__global__ void k_part_offsets(OffsetArgs o, int stride, int64_t n, int mixed, int shift, int bits,
                               int64_t num_parts) {
  const int32_t* key = o.key[blockIdx.y];
  int64_t* offsets = o.offsets[blockIdx.y];
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x * PT_OFF_RPT + threadIdx.x;
  if (n == 0) {
    if (i0 == 0)
      for (int64_t b = 0; b <= num_parts; ++b) offsets[b] = 0;
    return;
  }
  // All keys first (the stores below could alias them as far as the compiler knows).
  int32_t k[PT_OFF_RPT], kp[PT_OFF_RPT];
#pragma unroll
  for (int j = 0; j < PT_OFF_RPT; ++j) {
    const int64_t i = i0 + (int64_t)j * blockDim.x;
    k[j] = i < n ? key[i * stride] : 0;
    kp[j] = (i < n && i > 0) ? key[(i - 1) * stride] : 0;
  }
#pragma unroll
  for (int j = 0; j < PT_OFF_RPT; ++j) {
    const int64_t i = i0 + (int64_t)j * blockDim.x;
    if (i >= n) return;
    const int64_t d = digit_of(k[j], mixed, shift, bits);
    const int64_t dp = i == 0 ? -1 : (int64_t)digit_of(kp[j], mixed, shift, bits);
    for (int64_t b = dp + 1; b <= d; ++b) offsets[b] = i;
    if (i == n - 1)
      for (int64_t b = d + 1; b <= num_parts; ++b) offsets[b] = n;
  }
}

template <int NC, bool SHARED>
static void launch_scatter_keys(Context* ctx, const PassArgs& a, const int64_t* bases) {
  const size_t lds = scatter_lds_bytes(NC);
  static bool lds_attr[64] = {};
  int dev = 0;
  QF_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64 || !lds_attr[dev]) {
    QF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_part_scatter_keys<NC, SHARED>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (dev >= 0 && dev < 64) lds_attr[dev] = true;
  }
  int per_cu = 0;
  QF_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_part_scatter_keys<NC, SHARED>, PT_THREADS, lds));
  QF_REQUIRE(per_cu >= 1, "the multi-key scatter does not fit a CU");
  int64_t grid = (int64_t)per_cu * ctx->num_cus;
  if (grid >= 8) grid &= ~int64_t(7);
  grid = std::min<int64_t>(grid, a.num_tiles);
  hipLaunchKernelGGL((k_part_scatter_keys<NC, SHARED>), dim3((unsigned)grid), dim3(PT_THREADS), lds, ctx->stream, a, bases);
}

template <int NC>
static void launch_scatter_one(Context* ctx, const PassArgs& a, const int64_t* bases) {
  const size_t lds = scatter_lds_bytes(NC);
  // Once per device (outside any graph capture in practice: the first pass runs eagerly).
  static bool lds_attr[64] = {};
  int dev = 0;
  QF_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64 || !lds_attr[dev]) {
    QF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_part_scatter<NC>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (dev >= 0 && dev < 64) lds_attr[dev] = true;
  }
  // Persistent grid: as many workgroups as fit the GPU at once (a multiple of 8), at most one per tile.
  int per_cu = 0;
  QF_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_part_scatter<NC>, PT_THREADS, lds));
  int64_t grid = (int64_t)std::max(1, per_cu) * ctx->num_cus;
  if (grid >= 8) grid &= ~int64_t(7);
  grid = std::min<int64_t>(grid, a.num_tiles * a.nkeys);
  hipLaunchKernelGGL((k_part_scatter<NC>), dim3((unsigned)grid), dim3(PT_THREADS), lds, ctx->stream, a, bases);
}

template <int NC, bool SHARED>
static void launch_scatter_nc(Context* ctx, const PassArgs& a, const int64_t* bases) {
  // The multi-key kernel's LDS (staged tile plus every key's counts and bases) fits a CU up to
  // seven columns; wider tables take the per-key kernel, which reads a shared input once per key.
  if (a.nkeys > 1 && NC <= 7)
    launch_scatter_keys<NC, SHARED>(ctx, a, bases);
  else
    launch_scatter_one<NC>(ctx, a, bases);
}

template <bool SHARED>
static void launch_scatter(Context* ctx, const PassArgs& a, const int64_t* bases) {
  switch (a.ncols) {
    case 1: launch_scatter_nc<1, SHARED>(ctx, a, bases); break;
    case 2: launch_scatter_nc<2, SHARED>(ctx, a, bases); break;
    case 3: launch_scatter_nc<3, SHARED>(ctx, a, bases); break;
    case 4: launch_scatter_nc<4, SHARED>(ctx, a, bases); break;
    case 5: launch_scatter_nc<5, SHARED>(ctx, a, bases); break;
    case 6: launch_scatter_nc<6, SHARED>(ctx, a, bases); break;
    case 7: launch_scatter_nc<7, SHARED>(ctx, a, bases); break;
    default: launch_scatter_nc<8, SHARED>(ctx, a, bases); break;
  }
}

// One pass: per-tile counts of every key, their exclusive scan, the scatter.  `shared`: every
// key reads the same input (a.in[0]), each tile read once.
static void run_pass(Context* ctx, PassArgs a, bool shared, Scratch* scratch) {
  a.num_tiles = std::max<int64_t>(1, ceil_div(a.n, PT_TILE));
  const int nb = 1 << a.bits;
  const int64_t m = (int64_t)a.nkeys * nb * a.num_tiles;
  const size_t count_bytes = (size_t)m * sizeof(int64_t);
  DevBuf pooled;
  int64_t* counts = nullptr;
  if (scratch) {
    counts = static_cast<int64_t*>(scratch->take(count_bytes));
  } else {
    pooled = DevBuf(ctx, count_bytes);
    counts = pooled.as<int64_t>();
  }
  hipLaunchKernelGGL(k_part_hist, dim3((unsigned)ceil_div(a.num_tiles, PT_WAVES), (unsigned)a.nkeys), dim3(PT_THREADS),
                     0, ctx->stream, a, counts);
  QF_CHECK_HIP(hipGetLastError());
  scan_exclusive_i64(ctx, counts, counts, m, scratch);
  if (shared)
    launch_scatter<true>(ctx, a, counts);
  else
    launch_scatter<false>(ctx, a, counts);
  QF_CHECK_HIP(hipGetLastError());
}

// Digit bits of each LSD pass: at most 8 per pass, split evenly (low part first).
static std::vector<int> pass_bits(int total_bits) {
  const int passes = std::max(1, (total_bits + 7) / 8);
  std::vector<int> bits(passes, total_bits / passes);
  for (int p = 0; p < total_bits % passes; ++p) ++bits[p];
  return bits;
}

size_t partition_scratch_bytes(int64_t num_rows, int ncols, int total_bits, int nkeys) {
  if (num_rows <= 0) return 0;
  const std::vector<int> bits = pass_bits(total_bits);
  const int passes = (int)bits.size();
  size_t total = 0;
  for (int s = 0; s < 2 && s < passes - 1; ++s)
    total += (size_t)nkeys * ncols * aligned256((size_t)num_rows * sizeof(int32_t));
  const int64_t tiles = std::max<int64_t>(1, ceil_div(num_rows, PT_TILE));
  for (int b : bits) {
    const int64_t m = (int64_t)nkeys * ((int64_t)1 << b) * tiles;
    total += aligned256((size_t)m * sizeof(int64_t)) + scan_scratch_bytes(m);
  }
  return total;
}

// out_cols[k * ncols + c]: output column c of key k; d_offsets[k] (or null): its partition starts.
static void partition_impl(Context* ctx, const int32_t* const* in_cols, int num_in_cols, const int* key_cols,
                           int nkeys, bool gen_tid, int64_t num_rows, int mixed, int total_bits,
                           int32_t* const* out_cols, int out_stride, int64_t* const* d_offsets, Scratch* scratch) {
  const int ncols = num_in_cols + (gen_tid ? 1 : 0);
  QF_REQUIRE(ncols <= PT_MAXCOLS, "too many columns to partition");
  QF_REQUIRE(nkeys >= 1 && nkeys <= PT_MAXKEYS, "bad number of partition keys");
  QF_REQUIRE(total_bits >= 0 && total_bits <= 32, "bad radix bit count");
  for (int k = 0; k < nkeys; ++k) QF_REQUIRE(key_cols[k] >= 0 && key_cols[k] < num_in_cols, "bad key column");
  if (num_rows > 0) {
    const std::vector<int> bits = pass_bits(total_bits);
    const int passes = (int)bits.size();
    // Intermediate columns of the LSD passes (two sets, used alternately), per key.
    std::vector<DevBuf> pooled;
    std::vector<int32_t*> tmp[2];
    if (passes > 1)
      for (int s = 0; s < 2 && s < passes - 1; ++s)
        for (int i = 0; i < nkeys * ncols; ++i) {
          const size_t bytes = (size_t)num_rows * sizeof(int32_t);
          if (scratch) {
            tmp[s].push_back(static_cast<int32_t*>(scratch->take(bytes)));
          } else {
            pooled.emplace_back(ctx, bytes);
            tmp[s].push_back(pooled.back().as<int32_t>());
          }
        }
    int lo = 0;
    for (int p = 0; p < passes; ++p) {
      PassArgs a{};
      a.n = num_rows;
      a.nkeys = nkeys;
      a.ncols = ncols;
      a.mixed = mixed;
      a.bits = bits[p];
      a.shift = mixed ? (32 - total_bits + lo) : lo;
      if (a.bits == 0) a.shift = 0;
      for (int k = 0; k < nkeys; ++k) {
        a.key_col[k] = key_cols[k];
        for (int c = 0; c < ncols; ++c) {
          if (p == 0)
            a.in[k][c] = c < num_in_cols ? in_cols[c] : nullptr;
          else
            a.in[k][c] = tmp[(p - 1) & 1][k * ncols + c];
          a.out[k][c] = p == passes - 1 ? out_cols[k * ncols + c] : tmp[p & 1][k * ncols + c];
        }
      }
      a.tid_col = p == 0 && gen_tid ? num_in_cols : -1;
      a.out_stride = p == passes - 1 ? out_stride : 1;
      run_pass(ctx, a, p == 0, scratch);
      lo += bits[p];
    }
  }
  if (d_offsets) {
    const int64_t parts = (int64_t)1 << total_bits;
    const int full_shift = mixed ? 32 - total_bits : 0;
    const int64_t blocks = std::max<int64_t>(1, ceil_div(num_rows, 256 * PT_OFF_RPT));
    OffsetArgs o{};
    for (int k = 0; k < nkeys; ++k) {
      o.key[k] = out_cols[k * ncols + key_cols[k]];
      o.offsets[k] = d_offsets[k];
    }
    hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)blocks, (unsigned)nkeys), dim3(256), 0, ctx->stream, o,
                       out_stride, num_rows, mixed, total_bits == 0 ? 0 : full_shift, total_bits, parts);
    QF_CHECK_HIP(hipGetLastError());
  }
}

void stable_partition(Context* ctx, const int32_t* const* in_cols, int num_in_cols, int key_col, bool gen_tid,
                      int64_t num_rows, int mixed, int total_bits, int32_t* const* out_cols, int out_stride,
                      int64_t* d_offsets, Scratch* scratch) {
  int64_t* offs[1] = {d_offsets};
  partition_impl(ctx, in_cols, num_in_cols, &key_col, 1, gen_tid, num_rows, mixed, total_bits, out_cols, out_stride,
                 d_offsets ? offs : nullptr, scratch);
}

void stable_partition_multi(Context* ctx, const int32_t* const* in_cols, int num_in_cols, const int* key_cols,
                            int nkeys, bool gen_tid, int64_t num_rows, int mixed, int total_bits,
                            int32_t* const* out_cols, int64_t* const* d_offsets, Scratch* scratch) {
  partition_impl(ctx, in_cols, num_in_cols, key_cols, nkeys, gen_tid, num_rows, mixed, total_bits, out_cols, 1,
                 d_offsets, scratch);
}

void stable_sort_by_key(Context* ctx, const int32_t* const* in_cols, int num_in_cols, int key_col, bool gen_tid,
                        int64_t num_rows, int key_bits, int32_t* const* out_cols) {
  partition_impl(ctx, in_cols, num_in_cols, &key_col, 1, gen_tid, num_rows, 0, key_bits, out_cols, 1, nullptr,
                 nullptr);
}

}  // namespace qf
