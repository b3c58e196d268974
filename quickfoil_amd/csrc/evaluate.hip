// Candidate-literal evaluation on the GPU.
//
// Semantics (operations/CountAggregator.cpp:127-237 over HashJoin + Filter): for literal L with
// bound positions B, match(r, s) <=> P[i][s] == T[v_i][r] for all i in B, and
//   binding_pos(L) = #{(r, s) : match, r < N_pos},   covered_pos(L) = #{r < N_pos : exists s},
// likewise for negatives.  The reference enumerates every matching pair; here every binding
// row r gets cnt(r) = #{s : P[B](s) = T[v_B](r)} from a table of distinct B-tuples of the
// relation with their multiplicities, so binding = sum cnt and covered = sum [cnt > 0].  The
// integers are identical and the work no longer grows with the number of matching pairs
// (skewed keys make that number quadratic).
//
// Layout.  The relation is co-partitioned with the binding table: both are stably partitioned
// by the top R bits of mix32(join value), the relation on its join column with all columns
// carried (SoA), the binding table on the clause join variable with the referenced columns and
// the row id carried.  A work item is (join group, partition q, tile of binding rows of q);
// items are counted per (group, partition) on the device, scanned, and found by binary search,
// so no host loop touches partitions.  An item stages its binding tile in LDS, then for each
// key-column subset the group's literals use it builds an open-addressing count table of the
// partition's relation rows in LDS and probes it for every literal of that subset.  Partitions
// too large for an LDS table ("heavy": where skewed keys land) have their tables built once in
// HBM per (relation, join column, subset) by a multi-block kernel and cached with the
// partitions; their items probe those.
//
// Counts: per-lane u64 accumulators, 64-lane shuffle reductions, one atomic per wave and
// literal.  Covered bitvectors (optional) are set with atomicOr at the binding row id.
#include "qf_internal.h"

#include <algorithm>
#include <map>
#include <set>

namespace qf {

constexpr int EV_THREADS = 256;
constexpr int EV_MAXK = QF_MAX_ARITY;  // key columns of one subset tuple (distinct relation columns)
constexpr int EV_MAXEQ = 4;            // binding-side equalities of one literal (repeated columns)
constexpr int EV_MAXSUB = 32;          // subsets per group
constexpr int EV_MAXTC = 12;           // binding columns staged per group (plus the row id)
constexpr int EV_TBL_WORDS = 12288;    // LDS words for one count table (48 KB)
constexpr int EV_MIN_CAP = 16;

struct DevSubset {
  int nk;
  int pcol[EV_MAXK];
  const int64_t* hbase;  // per heavy ordinal: word offset of its HBM table
  const uint32_t* hbuf;  // HBM tables of this subset
};

struct DevLit {
  int slot;
  int subset;
  int nk;
  int neq;
  int tcol[EV_MAXK];        // staged binding column per key position
  int eq[EV_MAXEQ][2];      // staged binding columns that must be equal in the row
};

struct DevGroup {
  const int32_t* pcols[QF_MAX_ARITY];  // relation, partitioned on the join column
  const int64_t* poff;
  const int32_t* hq;                   // per partition: heavy ordinal or -1
  const int32_t* tcols[EV_MAXTC + 1];  // binding table partitioned on the join variable; [ntc] = row id
  const int64_t* toff;
  int ntc;
  int nsub;
  DevSubset sub[EV_MAXSUB];
  int lit_begin, nlit;
};

// Table layout (cap slots): state[cap] | cnt[cap] | keys[nk][cap].  state = 1 + partition-
// relative row that claimed the slot; empty slots have cnt == 0.
__device__ inline uint32_t tuple_hash(const int32_t* v, int nk) {
  uint32_t h = 0x2545f491u;
  for (int k = 0; k < nk; ++k) h = (h ^ (uint32_t)v[k]) * 0x9e3779b1u + (uint32_t)k;
  return mix32(h);
}

__host__ __device__ inline uint32_t table_cap_lds(int64_t np) {
  uint32_t c = next_pow2_u32((uint32_t)((np * 4 + 2) / 3));
  return c < EV_MIN_CAP ? EV_MIN_CAP : c;
}
__host__ __device__ inline uint32_t table_cap_hbm(int64_t np) {
  uint32_t c = next_pow2_u32((uint32_t)(2 * np));
  return c < EV_MIN_CAP ? EV_MIN_CAP : c;
}

__device__ inline uint32_t table_claim(uint32_t* state, uint32_t cap, const int32_t* vals, int nk,
                                       const int32_t* const* pcols, const int* pcol, int64_t p0, int64_t r) {
  uint32_t slot = tuple_hash(vals, nk) & (cap - 1);
  while (true) {
    const uint32_t prev = atomicCAS(&state[slot], 0u, (uint32_t)(r + 1));
    if (prev == 0) return slot;
    bool eq = true;
    for (int k = 0; k < nk; ++k) eq = eq && (pcols[pcol[k]][p0 + prev - 1] == vals[k]);
    if (eq) return slot;
    slot = (slot + 1) & (cap - 1);
  }
}

__device__ inline uint32_t table_probe(const uint32_t* cnt, const int32_t* keys, uint32_t cap, const int32_t* vals,
                                       int nk) {
  uint32_t slot = tuple_hash(vals, nk) & (cap - 1);
  while (true) {
    const uint32_t c = cnt[slot];
    if (c == 0) return 0;
    bool eq = true;
    for (int k = 0; k < nk; ++k) eq = eq && (keys[(size_t)k * cap + slot] == vals[k]);
    if (eq) return c;
    slot = (slot + 1) & (cap - 1);
  }
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------------------------------------------------
// HBM tables for heavy partitions, per (relation, join column, subset).

struct DevRel {
  const int32_t* pcols[QF_MAX_ARITY];
  const int64_t* poff;
};

struct HeavyJob {
  int q;
  uint32_t cap;
  int64_t base;    // word offset of the table
  int64_t r0, r1;  // relation rows of partition q (relative), or slots
};

__global__ __launch_bounds__(EV_THREADS) void k_heavy_build(DevRel rel, DevSubset sub, const HeavyJob* jobs,
                                                            uint32_t* tables) {
  const HeavyJob J = jobs[blockIdx.x];
  const int64_t p0 = rel.poff[J.q];
  uint32_t* state = tables + J.base;
  uint32_t* cnt = state + J.cap;
  const int lane = threadIdx.x & 63;
  for (int64_t base = J.r0; base < J.r1; base += EV_THREADS) {
    const int64_t r = base + threadIdx.x;
    const bool valid = r < J.r1;
    uint32_t slot = 0;
    if (valid) {
      int32_t vals[EV_MAXK];
      for (int k = 0; k < sub.nk; ++k) vals[k] = rel.pcols[sub.pcol[k]][p0 + r];
      slot = table_claim(state, J.cap, vals, sub.nk, rel.pcols, sub.pcol, p0, r);
    }
    // Skewed partitions repeat one key: the lanes sharing the first lane's slot add at once.
    const uint64_t active = __ballot(valid);
    if (active == 0) continue;
    const int first = __ffsll((unsigned long long)active) - 1;
    const uint32_t lead = __shfl(slot, first, 64);
    const uint64_t peers = __ballot(valid && slot == lead);
    if (valid) {
      if (slot == lead) {
        if (lane == first) atomicAdd(&cnt[slot], (uint32_t)__popcll(peers));
      } else {
        atomicAdd(&cnt[slot], 1u);
      }
    }
  }
}

__global__ __launch_bounds__(EV_THREADS) void k_heavy_finalize(DevRel rel, DevSubset sub, const HeavyJob* jobs,
                                                               uint32_t* tables) {
  const HeavyJob J = jobs[blockIdx.x];
  const int64_t p0 = rel.poff[J.q];
  const uint32_t* state = tables + J.base;
  int32_t* keys = reinterpret_cast<int32_t*>(tables + J.base + 2 * (int64_t)J.cap);
  for (int64_t s = J.r0 + threadIdx.x; s < J.r1; s += EV_THREADS) {
    const uint32_t st = state[s];
    if (st)
      for (int k = 0; k < sub.nk; ++k) keys[(int64_t)k * J.cap + s] = rel.pcols[sub.pcol[k]][p0 + st - 1];
  }
}

// ------------------------------------------------------------------------------------------
// Work items: chunks of binding rows per (group, partition), counted and scanned on the device.

__global__ void k_count_items(const DevGroup* groups, int num_groups, int R, int tile, int64_t* counts) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t parts = (int64_t)1 << R;
  if (id >= (int64_t)num_groups * parts) return;
  const DevGroup& G = groups[id >> R];
  const int64_t q = id & (parts - 1);
  const int64_t np = G.poff[q + 1] - G.poff[q];
  const int64_t nt = G.toff[q + 1] - G.toff[q];
  counts[id] = (np > 0 && nt > 0) ? (nt + tile - 1) / tile : 0;
}

struct EvalArgs {
  const DevGroup* groups;
  const DevLit* lits;
  const int64_t* starts;        // exclusive scan of per-(group, partition) item counts
  const int64_t* counts;
  int64_t num_cells;            // groups x 2^R
  int R;
  int64_t num_positive;
  unsigned long long* counts_out;  // num_slots x 4
  unsigned long long* bits;     // num_slots x bit_words, or null
  int64_t bit_words;
  int tile_rows;
};

__global__ __launch_bounds__(EV_THREADS) void k_eval(EvalArgs a) {
  extern __shared__ uint32_t lds[];
  uint32_t* tbl = lds;                                                // EV_TBL_WORDS
  int32_t* tstage = reinterpret_cast<int32_t*>(lds + EV_TBL_WORDS);  // (ntc + 1) x tile_rows
  const int lane = threadIdx.x & 63;
  const int64_t total = a.starts[a.num_cells - 1] + a.counts[a.num_cells - 1];

  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    // Cell = last (group, partition) whose first item is <= it.
    int64_t lo = 0, hi = a.num_cells - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (a.starts[mid] <= it) lo = mid;
      else hi = mid - 1;
    }
    const int64_t cell = lo;
    const int64_t chunk = it - a.starts[cell];
    const DevGroup& G = a.groups[cell >> a.R];
    const int64_t q = cell & (((int64_t)1 << a.R) - 1);
    const int64_t p0 = G.poff[q];
    const int64_t np = G.poff[q + 1] - p0;
    const int64_t t0 = G.toff[q] + chunk * a.tile_rows;
    const int tn = (int)min((int64_t)a.tile_rows, G.toff[q + 1] - t0);
    const int ntc = G.ntc;
    const int hq = G.hq[q];

    __syncthreads();  // the previous item is done with the tile and the table
    for (int c = 0; c <= ntc; ++c) {
      const int32_t* src = G.tcols[c] + t0;
      int32_t* dst = tstage + c * a.tile_rows;
      for (int i = threadIdx.x; i < tn; i += EV_THREADS) dst[i] = src[i];
    }

    for (int s = 0; s < G.nsub; ++s) {
      const DevSubset& S = G.sub[s];
      const uint32_t* cnt;
      const int32_t* keys;
      uint32_t cap;
      if (hq < 0) {
        cap = table_cap_lds(np);
        uint32_t* state = tbl;
        uint32_t* c_ = tbl + cap;
        __syncthreads();  // previous subset's probes are done with the table
        for (uint32_t i = threadIdx.x; i < 2 * cap; i += EV_THREADS) tbl[i] = 0;
        __syncthreads();
        for (int64_t r = threadIdx.x; r < np; r += EV_THREADS) {
          int32_t vals[EV_MAXK];
          for (int k = 0; k < S.nk; ++k) vals[k] = G.pcols[S.pcol[k]][p0 + r];
          const uint32_t slot = table_claim(state, cap, vals, S.nk, G.pcols, S.pcol, p0, r);
          atomicAdd(&c_[slot], 1u);
        }
        __syncthreads();
        int32_t* k_ = reinterpret_cast<int32_t*>(tbl + 2 * cap);
        for (uint32_t sl = threadIdx.x; sl < cap; sl += EV_THREADS) {
          const uint32_t st = state[sl];
          if (st)
            for (int k = 0; k < S.nk; ++k) k_[(size_t)k * cap + sl] = G.pcols[S.pcol[k]][p0 + st - 1];
        }
        __syncthreads();
        cnt = c_;
        keys = k_;
      } else {
        cap = table_cap_hbm(np);
        const uint32_t* base = S.hbuf + S.hbase[hq];
        cnt = base + cap;
        keys = reinterpret_cast<const int32_t*>(base + 2 * (int64_t)cap);
        __syncthreads();  // tile staging complete
      }

      for (int l = G.lit_begin; l < G.lit_begin + G.nlit; ++l) {
        const DevLit L = a.lits[l];
        if (L.subset != s) continue;
        uint64_t bp = 0, bn = 0;
        uint32_t cp = 0, cn = 0;
        for (int i = threadIdx.x; i < tn; i += EV_THREADS) {
          bool ok = true;
          for (int e = 0; e < L.neq; ++e)
            ok = ok && tstage[L.eq[e][0] * a.tile_rows + i] == tstage[L.eq[e][1] * a.tile_rows + i];
          uint32_t c = 0;
          if (ok) {
            int32_t vals[EV_MAXK];
            for (int k = 0; k < L.nk; ++k) vals[k] = tstage[L.tcol[k] * a.tile_rows + i];
            c = table_probe(cnt, keys, cap, vals, L.nk);
          }
          const int32_t row = tstage[ntc * a.tile_rows + i];
          if (row < a.num_positive) {
            bp += c;
            cp += (c != 0);
          } else {
            bn += c;
            cn += (c != 0);
          }
          if (a.bits && c) atomicOr(&a.bits[(int64_t)L.slot * a.bit_words + (row >> 6)], 1ull << (row & 63));
        }
        bp = wave_sum_u64(bp);
        bn = wave_sum_u64(bn);
        const uint32_t cpn = wave_sum_u32(cp | (cn << 16));  // <= 64 lanes x 16 rows per count
        if (lane == 0) {
          unsigned long long* out = a.counts_out + (int64_t)L.slot * 4;
          if (cpn & 0xffffu) atomicAdd(&out[0], (unsigned long long)(cpn & 0xffffu));
          if (cpn >> 16) atomicAdd(&out[1], (unsigned long long)(cpn >> 16));
          if (bp) atomicAdd(&out[2], (unsigned long long)bp);
          if (bn) atomicAdd(&out[3], (unsigned long long)bn);
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// Host side.

struct LitSpec {
  int slot;
  std::vector<int> pcols;                 // distinct relation columns of the key (sorted)
  std::vector<int> bcols;                 // binding column per key position
  std::vector<std::pair<int, int>> eqs;   // binding columns that must be equal (repeated columns)
};

struct GroupSpec {
  int relation = -1;
  int join_col = -1;
  int join_var = -1;
  std::vector<LitSpec> lits;
};

static int literal_join_key(const qf_literal& l, int arity, const int* types) {
  // FoilLiteral::UpdateJoinKey (schema/FoilLiteral.hpp:133-144): bound position with the
  // largest variable type id, first on ties.
  int jk = -1;
  for (int pos = 0; pos < arity; ++pos) {
    if (l.vars[pos] < 0) continue;
    const int t = types ? types[pos] : 0;
    if (jk == -1 || (types ? types[jk] : 0) < t) jk = pos;
  }
  return jk;
}

// Sorts key positions by relation column; a relation column bound twice (P[c] == T[x] and
// P[c] == T[y]) keeps one key position and adds the row-side equality T[x] == T[y].
static void normalise_key(LitSpec& ls) {
  std::vector<std::pair<int, int>> kv;
  for (size_t i = 0; i < ls.pcols.size(); ++i) kv.emplace_back(ls.pcols[i], ls.bcols[i]);
  std::sort(kv.begin(), kv.end());
  ls.pcols.clear();
  ls.bcols.clear();
  for (auto& p : kv) {
    if (!ls.pcols.empty() && ls.pcols.back() == p.first) {
      if (ls.bcols.back() != p.second) ls.eqs.emplace_back(ls.bcols.back(), p.second);
      continue;
    }
    ls.pcols.push_back(p.first);
    ls.bcols.push_back(p.second);
  }
  QF_REQUIRE((int)ls.eqs.size() <= EV_MAXEQ, "too many repeated columns in one literal");
}

static int auto_radix_bits(int64_t max_rows) {
  // Partitions of about 512 relation rows keep most count tables inside one LDS table.
  int r = 0;
  while (r < 16 && (max_rows >> r) > 512) ++r;
  return r;
}

// HBM count tables of the heavy partitions of one (relation, join column) for one subset.
struct HeavyTables {
  DevBuf buf;
  DevBuf base;  // int64 per heavy ordinal
};

// A relation partitioned on one join column, with its heavy partitions and their tables.
struct RelPart {
  std::shared_ptr<PartitionedTable> pt;
  DevBuf hq;                         // int32 per partition
  std::vector<int64_t> heavy_q;      // heavy partitions
  std::map<std::vector<int>, std::shared_ptr<HeavyTables>> tables;
};

struct RelCache {
  std::map<std::tuple<const void*, int64_t, int, int>, std::shared_ptr<RelPart>> parts;
};

static RelCache& rel_cache(Context* ctx) {
  // The context's generic cache slot holds this translation unit's relation cache.
  if (!ctx->eval_cache) ctx->eval_cache = std::make_shared<RelCache>();
  return *std::static_pointer_cast<RelCache>(ctx->eval_cache);
}

static std::shared_ptr<RelPart> partition_relation(Context* ctx, const qf_table& rel, int join_col, int R,
                                                   bool cache) {
  const auto key = std::make_tuple((const void*)rel.columns[0], rel.num_rows, join_col, R);
  RelCache& rc = rel_cache(ctx);
  if (cache) {
    auto it = rc.parts.find(key);
    if (it != rc.parts.end()) return it->second;
  }
  PhaseScope ps(ctx, "partition_relations");
  auto rp = std::make_shared<RelPart>();
  auto pt = std::make_shared<PartitionedTable>();
  pt->radix_bits = R;
  pt->num_columns = rel.num_columns;
  pt->num_rows = rel.num_rows;
  for (int c = 0; c < rel.num_columns; ++c) pt->columns.emplace_back(ctx, std::max<int64_t>(1, rel.num_rows) * 4);
  pt->offsets = DevBuf(ctx, (((size_t)1 << R) + 1) * 8);
  std::vector<int32_t*> outs;
  for (auto& b : pt->columns) outs.push_back(b.as<int32_t>());
  stable_partition(ctx, rel.columns, rel.num_columns, join_col, false, rel.num_rows, 1, R, outs.data(), 1,
                   pt->offsets.as<int64_t>());
  pt->h_offsets.resize(((size_t)1 << R) + 1);
  copy_d2h(ctx, pt->h_offsets.data(), pt->offsets.get(), pt->h_offsets.size() * 8);
  rp->pt = pt;
  // Heavy: the LDS table of the widest possible key (all columns) does not fit.
  const int64_t parts = (int64_t)1 << R;
  std::vector<int32_t> hq(parts, -1);
  for (int64_t q = 0; q < parts; ++q) {
    const int64_t np = pt->h_offsets[q + 1] - pt->h_offsets[q];
    if (np > 0 && (int64_t)table_cap_lds(np) * (rel.num_columns + 2) > EV_TBL_WORDS) {
      hq[q] = (int32_t)rp->heavy_q.size();
      rp->heavy_q.push_back(q);
    }
  }
  rp->hq = DevBuf(ctx, parts * 4);
  copy_h2d(ctx, rp->hq.get(), hq.data(), parts * 4);
  if (cache) rc.parts[key] = rp;
  return rp;
}

static std::shared_ptr<HeavyTables> heavy_tables(Context* ctx, RelPart& rp, const std::vector<int>& pcols) {
  auto it = rp.tables.find(pcols);
  if (it != rp.tables.end()) return it->second;
  auto ht = std::make_shared<HeavyTables>();
  const int nk = (int)pcols.size();
  std::vector<int64_t> base;
  std::vector<HeavyJob> build, fin;
  int64_t words = 0;
  for (int64_t q : rp.heavy_q) {
    const int64_t np = rp.pt->h_offsets[q + 1] - rp.pt->h_offsets[q];
    const uint32_t cap = table_cap_hbm(np);
    base.push_back(words);
    for (int64_t r = 0; r < np; r += 8192) build.push_back(HeavyJob{(int)q, cap, words, r, std::min(np, r + 8192)});
    for (int64_t r = 0; r < cap; r += 65536)
      fin.push_back(HeavyJob{(int)q, cap, words, r, std::min<int64_t>(cap, r + 65536)});
    words += (int64_t)cap * (nk + 2);
  }
  ht->buf = DevBuf(ctx, std::max<int64_t>(1, words) * 4);
  ht->base = DevBuf(ctx, std::max<size_t>(1, base.size()) * 8);
  if (!base.empty()) {
    PhaseScope ps(ctx, "heavy_tables");
    copy_h2d(ctx, ht->base.get(), base.data(), base.size() * 8);
    DevBuf dbuild(ctx, build.size() * sizeof(HeavyJob)), dfin(ctx, fin.size() * sizeof(HeavyJob));
    copy_h2d(ctx, dbuild.get(), build.data(), build.size() * sizeof(HeavyJob));
    copy_h2d(ctx, dfin.get(), fin.data(), fin.size() * sizeof(HeavyJob));
    fill_i32(ctx, ht->buf.as<int32_t>(), 0, words);
    DevRel rel{};
    for (int c = 0; c < rp.pt->num_columns; ++c) rel.pcols[c] = rp.pt->columns[c].as<int32_t>();
    rel.poff = rp.pt->offsets.as<int64_t>();
    DevSubset sub{};
    sub.nk = nk;
    for (int k = 0; k < nk; ++k) sub.pcol[k] = pcols[k];
    hipLaunchKernelGGL(k_heavy_build, dim3((unsigned)build.size()), dim3(EV_THREADS), 0, ctx->stream, rel, sub,
                       dbuild.as<HeavyJob>(), ht->buf.as<uint32_t>());
    QF_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_heavy_finalize, dim3((unsigned)fin.size()), dim3(EV_THREADS), 0, ctx->stream, rel, sub,
                       dfin.as<HeavyJob>(), ht->buf.as<uint32_t>());
    QF_CHECK_HIP(hipGetLastError());
    sync(ctx);  // job lists are released at the end of this scope
  }
  rp.tables[pcols] = ht;
  return ht;
}

// Binding table partitioned on column `var`, carrying `cols` (must include var) and row ids.
static std::shared_ptr<PartitionedTable> partition_binding(Context* ctx, const qf_table& b, int var,
                                                           const std::vector<int>& cols, int R) {
  PhaseScope ps(ctx, "partition_binding");
  auto pt = std::make_shared<PartitionedTable>();
  pt->radix_bits = R;
  pt->num_columns = (int)cols.size();
  pt->num_rows = b.num_rows;
  pt->has_tid = true;
  for (size_t c = 0; c <= cols.size(); ++c) pt->columns.emplace_back(ctx, std::max<int64_t>(1, b.num_rows) * 4);
  pt->offsets = DevBuf(ctx, (((size_t)1 << R) + 1) * 8);
  // Batches of at most 6 data columns, each with the key column, so every batch partitions the
  // same way (the partition is a deterministic, stable function of the key column).
  std::vector<int> others;
  for (int c : cols)
    if (c != var) others.push_back(c);
  size_t pos = 0;
  bool first = true;
  const int key_index = (int)(std::find(cols.begin(), cols.end(), var) - cols.begin());
  do {
    std::vector<const int32_t*> in{b.columns[var]};
    std::vector<int32_t*> out{pt->columns[key_index].as<int32_t>()};
    std::vector<int> batch;
    while (pos < others.size() && batch.size() < 6) batch.push_back(others[pos++]);
    for (int c : batch) {
      in.push_back(b.columns[c]);
      out.push_back(pt->columns[std::find(cols.begin(), cols.end(), c) - cols.begin()].as<int32_t>());
    }
    if (first) out.push_back(pt->columns[cols.size()].as<int32_t>());
    stable_partition(ctx, in.data(), (int)in.size(), 0, first, b.num_rows, 1, R, out.data(), 1,
                     first ? pt->offsets.as<int64_t>() : nullptr);
    first = false;
  } while (pos < others.size());
  return pt;
}

struct Prepared {
  std::vector<DevGroup> groups;
  std::vector<DevLit> lits;
  int max_tc = 0;
  std::vector<std::shared_ptr<PartitionedTable>> keep_bind;
  std::vector<std::shared_ptr<RelPart>> keep_rel;
  std::vector<std::shared_ptr<HeavyTables>> keep_heavy;
};

static void prepare(Context* ctx, const qf_table& binding, const qf_table* relations,
                    const std::vector<GroupSpec>& specs, int R, bool cache, Prepared& P) {
  std::map<std::pair<int, int>, std::shared_ptr<RelPart>> rel_parts;
  std::map<int, std::set<int>> var_cols;
  for (const GroupSpec& g : specs) {
    auto key = std::make_pair(g.relation, g.join_col);
    if (!rel_parts.count(key)) rel_parts[key] = partition_relation(ctx, relations[g.relation], g.join_col, R, cache);
    var_cols[g.join_var].insert(g.join_var);
    for (const LitSpec& l : g.lits) {
      for (int c : l.bcols) var_cols[g.join_var].insert(c);
      for (auto& e : l.eqs) {
        var_cols[g.join_var].insert(e.first);
        var_cols[g.join_var].insert(e.second);
      }
    }
  }
  std::map<int, std::shared_ptr<PartitionedTable>> bind_parts;
  std::map<int, std::vector<int>> var_col_list;
  for (auto& kv : var_cols) {
    std::vector<int> cols(kv.second.begin(), kv.second.end());
    QF_REQUIRE((int)cols.size() <= EV_MAXTC, "too many binding columns referenced by one join variable");
    var_col_list[kv.first] = cols;
    bind_parts[kv.first] = partition_binding(ctx, binding, kv.first, cols, R);
  }
  for (auto& kv : rel_parts) P.keep_rel.push_back(kv.second);
  for (auto& kv : bind_parts) P.keep_bind.push_back(kv.second);

  for (const GroupSpec& g : specs) {
    RelPart& rp = *rel_parts[std::make_pair(g.relation, g.join_col)];
    const auto& bp = bind_parts[g.join_var];
    const std::vector<int>& cols = var_col_list[g.join_var];
    auto tidx = [&](int c) { return (int)(std::find(cols.begin(), cols.end(), c) - cols.begin()); };
    DevGroup dg{};
    for (int c = 0; c < rp.pt->num_columns; ++c) dg.pcols[c] = rp.pt->columns[c].as<int32_t>();
    dg.poff = rp.pt->offsets.as<int64_t>();
    dg.hq = rp.hq.as<int32_t>();
    dg.ntc = (int)cols.size();
    for (size_t c = 0; c <= cols.size(); ++c) dg.tcols[c] = bp->columns[c].as<int32_t>();
    dg.toff = bp->offsets.as<int64_t>();
    std::vector<std::vector<int>> subsets;
    dg.lit_begin = (int)P.lits.size();
    for (const LitSpec& l : g.lits) {
      int s = (int)(std::find(subsets.begin(), subsets.end(), l.pcols) - subsets.begin());
      if (s == (int)subsets.size()) subsets.push_back(l.pcols);
      QF_REQUIRE((int)subsets.size() <= EV_MAXSUB, "too many key subsets in one join group");
      DevLit dl{};
      dl.slot = l.slot;
      dl.subset = s;
      dl.nk = (int)l.pcols.size();
      for (int k = 0; k < dl.nk; ++k) dl.tcol[k] = tidx(l.bcols[k]);
      dl.neq = (int)l.eqs.size();
      for (int e = 0; e < dl.neq; ++e) {
        dl.eq[e][0] = tidx(l.eqs[e].first);
        dl.eq[e][1] = tidx(l.eqs[e].second);
      }
      P.lits.push_back(dl);
    }
    dg.nlit = (int)P.lits.size() - dg.lit_begin;
    dg.nsub = (int)subsets.size();
    for (int s = 0; s < dg.nsub; ++s) {
      dg.sub[s].nk = (int)subsets[s].size();
      for (int k = 0; k < dg.sub[s].nk; ++k) dg.sub[s].pcol[k] = subsets[s][k];
      if (!rp.heavy_q.empty()) {
        auto ht = heavy_tables(ctx, rp, subsets[s]);
        dg.sub[s].hbuf = ht->buf.as<uint32_t>();
        dg.sub[s].hbase = ht->base.as<int64_t>();
        P.keep_heavy.push_back(ht);
      }
    }
    P.max_tc = std::max(P.max_tc, dg.ntc);
    P.groups.push_back(dg);
  }
}

static void run_groups(Context* ctx, const qf_table& binding, int64_t num_positive, const qf_table* relations,
                       std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o, uint64_t* h_counts,
                       uint64_t* d_bits) {
  int64_t max_rows = 0;
  for (const GroupSpec& g : specs) max_rows = std::max(max_rows, relations[g.relation].num_rows);
  const int R = o.radix_bits > 0 ? std::min(16, o.radix_bits) : auto_radix_bits(max_rows);
  const int tile = o.tile_rows > 0 ? std::min(4096, std::max(64, o.tile_rows)) : 1024;

  DevBuf counts(ctx, std::max(1, num_slots) * 4 * sizeof(uint64_t));
  fill_u64(ctx, counts.as<uint64_t>(), 0, (int64_t)num_slots * 4);
  const int64_t words = ceil_div(binding.num_rows, 64);
  if (d_bits && num_slots > 0) fill_u64(ctx, d_bits, 0, (int64_t)num_slots * words);

  if (!specs.empty() && binding.num_rows > 0) {
    Prepared P;
    prepare(ctx, binding, relations, specs, R, o.reuse_partitions != 0, P);
    DevBuf dgroups(ctx, P.groups.size() * sizeof(DevGroup)), dlits(ctx, std::max<size_t>(1, P.lits.size()) * sizeof(DevLit));
    QF_CHECK_HIP(hipMemcpyAsync(dgroups.get(), P.groups.data(), P.groups.size() * sizeof(DevGroup),
                                hipMemcpyHostToDevice, ctx->stream));
    QF_CHECK_HIP(hipMemcpyAsync(dlits.get(), P.lits.data(), P.lits.size() * sizeof(DevLit), hipMemcpyHostToDevice,
                                ctx->stream));
    const int64_t parts = (int64_t)1 << R;
    const size_t lds_bytes = (EV_TBL_WORDS + (size_t)(P.max_tc + 1) * tile) * 4;
    QF_REQUIRE(lds_bytes <= 160 * 1024, "work item does not fit in LDS");
    QF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_eval),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    const int max_blocks = ctx->num_cus * std::max(1, (int)((160 * 1024) / lds_bytes));
    const bool batched = o.mode == QF_EVAL_BATCHED;
    const int per_launch = batched ? (int)P.groups.size() : 1;
    DevBuf cell_counts(ctx, (size_t)per_launch * parts * 8), cell_starts(ctx, (size_t)per_launch * parts * 8);
    EvalArgs a{};
    a.lits = dlits.as<DevLit>();
    a.R = R;
    a.num_positive = num_positive;
    a.counts_out = reinterpret_cast<unsigned long long*>(counts.get());
    a.bits = reinterpret_cast<unsigned long long*>(d_bits);
    a.bit_words = words;
    a.tile_rows = tile;
    for (size_t g0 = 0; g0 < P.groups.size(); g0 += per_launch) {
      const int ng = (int)std::min<size_t>(per_launch, P.groups.size() - g0);
      const int64_t cells = (int64_t)ng * parts;
      a.groups = dgroups.as<DevGroup>() + g0;
      a.num_cells = cells;
      a.starts = cell_starts.as<int64_t>();
      a.counts = cell_counts.as<int64_t>();
      PhaseScope ps(ctx, "probe_count");
      hipLaunchKernelGGL(k_count_items, dim3((unsigned)ceil_div(cells, 256)), dim3(256), 0, ctx->stream, a.groups, ng,
                         R, tile, cell_counts.as<int64_t>());
      QF_CHECK_HIP(hipGetLastError());
      scan_exclusive_i64(ctx, cell_counts.as<int64_t>(), cell_starts.as<int64_t>(), cells);
      // Enough blocks to fill the GPU; each block walks items with a grid stride.
      const int blocks = (int)std::min<int64_t>(max_blocks, std::max<int64_t>(1, cells));
      hipLaunchKernelGGL(k_eval, dim3(blocks), dim3(EV_THREADS), lds_bytes, ctx->stream, a);
      QF_CHECK_HIP(hipGetLastError());
    }
    sync(ctx);  // the prepared buffers are released when this scope ends
  }
  copy_d2h(ctx, h_counts, counts.get(), (size_t)num_slots * 4 * sizeof(uint64_t));
}

}  // namespace qf

using namespace qf;

extern "C" {

void qf_eval_options_default(qf_eval_options* o) {
  if (!o) return;
  o->mode = QF_EVAL_BATCHED;
  o->radix_bits = 0;
  o->tile_rows = 0;
  o->use_graph = 0;
  o->reuse_partitions = 0;
}

static void check_table(const qf_table* t, const char* what) {
  QF_REQUIRE(t != nullptr, std::string(what) + " is null");
  QF_REQUIRE(t->num_rows >= 0 && t->num_rows < ((int64_t)1 << 31), std::string(what) + " row count out of range");
  QF_REQUIRE(t->num_columns >= 1 && t->columns != nullptr, std::string(what) + " has no columns");
  for (int c = 0; c < t->num_columns; ++c)
    QF_REQUIRE(t->num_rows == 0 || t->columns[c] != nullptr, std::string(what) + " has a null column");
}

int qf_evaluate_literals(qf_context* c, const qf_table* binding, int64_t num_positive, const qf_table* relations,
                         const int* arg_types, int num_relations, const qf_literal* literals, int num_literals,
                         const qf_eval_options* opts, uint64_t* h_counts, uint64_t* d_semi_bits) {
  return guarded(c, [&](Context* ctx) {
    check_table(binding, "binding table");
    QF_REQUIRE(num_positive >= 0 && num_positive <= binding->num_rows, "num_positive out of range");
    QF_REQUIRE(num_literals >= 0 && (num_literals == 0 || (literals && h_counts)), "null literals or counts");
    QF_REQUIRE(num_relations >= 0 && (num_relations == 0 || relations), "null relations");
    qf_eval_options o;
    qf_eval_options_default(&o);
    if (opts) o = *opts;
    std::vector<int> type_start(num_relations + 1, 0);
    for (int r = 0; r < num_relations; ++r) {
      check_table(&relations[r], "relation");
      QF_REQUIRE(relations[r].num_columns <= QF_MAX_ARITY, "relation arity above QF_MAX_ARITY");
      type_start[r + 1] = type_start[r] + relations[r].num_columns;
    }
    // Group literals by (relation, join position, clause join variable)
    // (learner/QuickFoil.cpp:142-161, CandidateLiteralEvaluator.cpp:281-289); PLAIN_LITERAL
    // keeps one group per literal.
    std::vector<GroupSpec> specs;
    std::map<std::tuple<int, int, int>, int> index;
    for (int l = 0; l < num_literals; ++l) {
      const qf_literal& lit = literals[l];
      QF_REQUIRE(lit.relation >= 0 && lit.relation < num_relations, "literal relation out of range");
      const int arity = relations[lit.relation].num_columns;
      const int* types = arg_types ? arg_types + type_start[lit.relation] : nullptr;
      for (int i = 0; i < arity; ++i)
        QF_REQUIRE(lit.vars[i] >= -1 && lit.vars[i] < binding->num_columns, "literal variable out of range");
      const int jk = literal_join_key(lit, arity, types);
      QF_REQUIRE(jk >= 0, "literal has no bound variable");
      LitSpec ls;
      ls.slot = l;
      for (int i = 0; i < arity; ++i)
        if (lit.vars[i] >= 0) {
          ls.pcols.push_back(i);
          ls.bcols.push_back(lit.vars[i]);
        }
      normalise_key(ls);
      const auto key = std::make_tuple(lit.relation, jk, lit.vars[jk]);
      int gi;
      if (o.mode == QF_EVAL_PLAIN_LITERAL || !index.count(key)) {
        gi = (int)specs.size();
        GroupSpec g;
        g.relation = lit.relation;
        g.join_col = jk;
        g.join_var = lit.vars[jk];
        specs.push_back(g);
        if (o.mode != QF_EVAL_PLAIN_LITERAL) index[key] = gi;
      } else {
        gi = index[key];
      }
      specs[gi].lits.push_back(ls);
    }
    run_groups(ctx, *binding, num_positive, relations, specs, num_literals, o, h_counts, d_semi_bits);
    for (int l = 0; l < num_literals; ++l)
      if (h_counts[4 * l + 2] > 0x7fffffffull || h_counts[4 * l + 3] > 0x7fffffffull)
        throw Error(QF_ERR_COUNT_RANGE, "a binding count exceeds int32 (values are exact in uint64)");
  });
}

int qf_count_aggregate(qf_context* c, const qf_table* binding, int64_t num_positive, const qf_table* relations,
                       int num_relations, const qf_join_group_plan* plans, int num_plans, int num_slots,
                       const qf_eval_options* opts, uint64_t* h_counts, uint64_t* d_semi_bits) {
  return guarded(c, [&](Context* ctx) {
    check_table(binding, "binding table");
    QF_REQUIRE(num_positive >= 0 && num_positive <= binding->num_rows, "num_positive out of range");
    QF_REQUIRE(num_plans >= 0 && num_slots >= 0 && (num_slots == 0 || h_counts), "bad plan arguments");
    qf_eval_options o;
    qf_eval_options_default(&o);
    if (opts) o = *opts;
    if (o.mode == QF_EVAL_PLAIN_LITERAL) o.mode = QF_EVAL_PLAIN_GROUP;
    std::vector<GroupSpec> specs;
    std::vector<int> seen(num_slots, 0);
    for (int p = 0; p < num_plans; ++p) {
      const qf_join_group_plan& pl = plans[p];
      QF_REQUIRE(pl.relation >= 0 && pl.relation < num_relations, "plan relation out of range");
      const qf_table& rel = relations[pl.relation];
      check_table(&rel, "relation");
      QF_REQUIRE(rel.num_columns <= QF_MAX_ARITY, "relation arity above QF_MAX_ARITY");
      QF_REQUIRE(pl.join_column >= 0 && pl.join_column < rel.num_columns, "plan join column out of range");
      QF_REQUIRE(pl.binding_column >= 0 && pl.binding_column < binding->num_columns, "plan binding column out of range");
      QF_REQUIRE(pl.num_atoms >= 0 && pl.num_nodes >= pl.num_atoms, "plan node counts");
      GroupSpec g;
      g.relation = pl.relation;
      g.join_col = pl.join_column;
      g.join_var = pl.binding_column;
      // Atom set of every node: atoms are leaves, conjunctions take the union of children
      // (learner/PredicateEvaluationPlan.hpp:56-80; CountAggregator.cpp:205-237).
      std::vector<std::vector<int>> atoms_of(pl.num_nodes);
      for (int x = 0; x < pl.num_nodes; ++x) {
        if (x < pl.num_atoms) {
          QF_REQUIRE(pl.atom_bg_col[x] >= 0 && pl.atom_bg_col[x] < rel.num_columns, "atom column out of range");
          QF_REQUIRE(pl.atom_binding_col[x] >= 0 && pl.atom_binding_col[x] < binding->num_columns,
                     "atom binding column out of range");
          atoms_of[x] = {x};
        } else {
          const int lft = pl.node_left[x], rgt = pl.node_right[x];
          QF_REQUIRE(lft >= 0 && lft < x && rgt >= 0 && rgt < x, "conjunction children must precede the node");
          std::set<int> u(atoms_of[lft].begin(), atoms_of[lft].end());
          u.insert(atoms_of[rgt].begin(), atoms_of[rgt].end());
          atoms_of[x].assign(u.begin(), u.end());
        }
      }
      auto add = [&](int slot, const std::vector<int>& atoms) {
        QF_REQUIRE(slot >= 0 && slot < num_slots, "literal slot out of range");
        QF_REQUIRE(!seen[slot], "literal slot used twice");
        seen[slot] = 1;
        LitSpec ls;
        ls.slot = slot;
        ls.pcols.push_back(pl.join_column);
        ls.bcols.push_back(pl.binding_column);
        for (int a : atoms) {
          ls.pcols.push_back(pl.atom_bg_col[a]);
          ls.bcols.push_back(pl.atom_binding_col[a]);
        }
        normalise_key(ls);
        g.lits.push_back(ls);
      };
      if (pl.root_literal >= 0) add(pl.root_literal, {});
      for (int x = 0; x < pl.num_nodes; ++x)
        if (pl.node_literal[x] >= 0) add(pl.node_literal[x], atoms_of[x]);
      if (!g.lits.empty()) specs.push_back(g);
    }
    run_groups(ctx, *binding, num_positive, relations, specs, num_slots, o, h_counts, d_semi_bits);
  });
}

}  // extern "C"
