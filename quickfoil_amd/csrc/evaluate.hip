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
// Layout.  Relations and the binding table are co-partitioned by the top bits of mix32(join
// value): the binding table once per clause join variable with R_T bits, carrying the referenced
// columns and the row id; each relation on each join column with R_u <= R_T bits (chosen from
// its size; the batched form gives every relation R_T), carrying all its columns.  Top bits
// nest, so relation partition q covers the binding partitions [q << (R_T - R_u), (q + 1) <<
// (R_T - R_u)).
//
// Work.  A unit is (relation, join column) with the literal groups of every join variable that
// uses it (one literal for the plain form).  Units with the same radix bits form a batch: they
// see the same binding cells, so a work item (batch, cell q, chunk of <= tile binding rows of q
// across the batch's join variables) stages its binding rows in LDS once for all its units.
// Item counts per cell are computed and scanned on the device and expanded into an item list
// that a persistent grid of 256-thread workgroups walks.  Per unit, an item first probes a table
// of the partition's join column for every staged row and lists the rows that match (every
// literal of the unit binds the join column, so the other rows match nothing), counting the
// literals keyed on the join column alone; then, per wider key subset, it builds an
// open-addressing count table of the partition's facts in LDS (or uses the HBM table of a heavy
// partition, built once per (relation, join column, subset) and cached) and the four waves split
// the (literal, 64-row block of listed rows) entries of that subset literal-major.  A literal of
// three or more key columns skips the blocks where its parent (its key minus one column) matched
// nothing, and a subset none of whose literals can match (no orphan, no parent hit in the item)
// is skipped whole.  Covered counts come from ballots; binding counts from per-lane sums reduced
// once per literal run and added by four lanes into the workgroup's copy of the counters, which
// are summed after the launches.
#include "qf_internal.h"

#include <algorithm>
#include <list>
#include <map>
#include <type_traits>
#include <set>

namespace qf {

#ifndef QF_EV_THREADS
#define QF_EV_THREADS 256
#endif
constexpr int EV_THREADS = QF_EV_THREADS;  // threads of one evaluation block (a multiple of 64)
constexpr int EV_WAVES = EV_THREADS / 64;
constexpr int EV_MAXK = QF_MAX_ARITY;  // key columns of one subset tuple (distinct relation columns)
constexpr int EV_MAXEQ = 4;            // binding-side equalities of one literal (repeated columns)
constexpr int EV_MAXSUB = 32;          // key subsets per unit (bits of a uint32)
constexpr int EV_MAXTC = 12;           // staged binding columns (plus the row id)
constexpr int EV_MAXG = EV_MAXTC;      // join variables per unit
constexpr int EV_MAXLU = 512;          // literals per unit
constexpr int EV_TBL_WORDS = 2048;     // LDS words for one count table (8 KB)
constexpr int EV_PST_WORDS = 1536;     // LDS words for the partition's facts (6 KB)
constexpr int EV_MIN_CAP = 16;
constexpr int EV_MAXMASK = 64;         // literals per unit whose hit masks prune other literals
constexpr int EV_MAXBU = 32;           // units per batch (units of equal radix bits share staged rows)
#ifndef QF_EV_STAGE
#define QF_EV_STAGE 4
#endif
constexpr int EV_STAGE = QF_EV_STAGE;  // binding words a thread stages per round (loads in flight)

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
  int tcol[EV_MAXK];    // staged binding column per key position
  int eq[EV_MAXEQ][2];  // staged binding columns that must be equal in the row
  int pmask;            // hit-mask slot of a literal whose key is this key minus one column, or -1
  int mslot;            // this literal's hit-mask slot (it prunes others), or -1
};

struct DevUGroup {
  int jtcol;                // staged column of the join variable
  int sbeg[EV_MAXSUB + 1];  // literals of subset s: lits[sbeg[s], sbeg[s + 1]); subset 0 = {join column}
};

struct DevUnit {
  const int32_t* pcols[QF_MAX_ARITY];  // relation, partitioned on the join column
  const int64_t* poff;
  const int32_t* hq;  // per partition: heavy ordinal or -1
  int R;              // relation radix bits
  int shift;          // binding digits are `shift` bits finer
  int arity;
  int ngroups;
  int nsub;
  int jcol;               // the join column (subset 0's key)
  uint32_t orphan;        // bit s: subset s has a literal without a parent mask
  uint32_t child_subsets[EV_MAXMASK];  // per mask slot: the subsets of the literals it prunes
  int ug_of_bg[EV_MAXG];  // the unit's group for each group of its batch, or -1
  DevUGroup g[EV_MAXG];
  DevSubset sub[EV_MAXSUB];
};

// HBM table layout (cap slots): state[cap] | cnt[cap] | keys[nk][cap].  state = 1 + partition-
// relative row that claimed the slot; empty slots have cnt == 0.
__device__ inline uint32_t tuple_hash(const int32_t* v, int nk) {
  uint32_t h = 0x2545f491u;
  for (int k = 0; k < nk; ++k) h = (h ^ (uint32_t)v[k]) * 0x9e3779b1u + (uint32_t)k;
  return mix32(h);
}

// Read-only kernel inputs (work list, units, literals) are read through the constant address
// space, so uniform reads become scalar loads.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ inline cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}
// Marks a block-uniform value (read from LDS or memory) as wave-uniform for the compiler.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uni64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// LDS tables: every table takes the whole table region, EV_LDS_CAP = 2048 slots, so its load
// factor is at most 1/4 (512 facts) and 1/8 at the usual 256.  A wave's probe loop runs as long
// as its longest probe: over ~50 lanes that is ~3 slots at load 1/8 and ~4.6 at load 1/4 (a miss
// averages 1.15 and 1.4 slots).  One word per slot: fingerprint (10 bits) << 22 | (1 + staged
// fact row) << 11 | count (11 bits); 0 = empty.  Facts and counts stay below 2048 because an LDS
// partition holds at most 512 facts.
constexpr int EV_LDS_MAX_FACTS = 512;
constexpr uint32_t EV_FP_SHIFT = 22;
constexpr uint32_t EV_REP_SHIFT = 11;
constexpr uint32_t EV_FIELD_MASK = 0x7ffu;
static_assert(EV_LDS_MAX_FACTS <= (int)EV_FIELD_MASK, "fact rows and counts must fit 11 bits");
constexpr uint32_t EV_FP_MASK = ~((1u << EV_FP_SHIFT) - 1u);

// Hash of the LDS tables: one 32-bit multiply per key column.  Only the table's own build and
// probes see it, and every fingerprint hit is checked against the staged key, so it needs spread
// rather than the HBM tables' full mixing.  The fingerprint is the product's top 10 bits, already
// in their place in a slot word; the slot is the log2(cap) bits below them (cap <= 2048: bits
// 11-21 at most), which every key bit below them feeds.
__device__ inline uint32_t lds_hash(const int32_t* v, int nk) {
  uint32_t h = (uint32_t)v[0] * 0x9e3779b1u;
  for (int k = 1; k < nk; ++k) h = (h ^ (uint32_t)v[k]) * 0x85ebca6bu;
  return h;
}
__device__ inline uint32_t lds_slot(uint32_t h, uint32_t cap) {
  return (h >> (EV_FP_SHIFT - __builtin_ctz(cap))) & (cap - 1);  // cap is a power of two
}
__device__ inline uint32_t lds_fp(uint32_t h) { return h & EV_FP_MASK; }
static_assert(EV_TBL_WORDS <= (1 << EV_FP_SHIFT), "an LDS table's slot bits sit below the fingerprint");
constexpr uint32_t EV_LDS_CAP = EV_TBL_WORDS;
static_assert((EV_LDS_CAP & (EV_LDS_CAP - 1)) == 0 && EV_LDS_CAP >= 4 * EV_LDS_MAX_FACTS,
              "LDS tables are a power of two, at load <= 1/4");
static_assert(EV_LDS_CAP % (4 * QF_EV_THREADS) == 0, "a block clears a table in 16-byte words");
// Clears an LDS table (16 bytes per thread and step; the table starts the block's LDS).
__device__ inline void clear_lds_table(uint32_t* tbl) {
#pragma unroll
  for (int i = threadIdx.x; i < (int)EV_LDS_CAP / 4; i += QF_EV_THREADS)
    reinterpret_cast<uint4*>(tbl)[i] = make_uint4(0u, 0u, 0u, 0u);
}
// HBM tables: cap = P2(4 * facts) slots (load <= 1/4), P2(2 * facts) above 2^29 facts.
__host__ __device__ inline uint32_t table_cap_hbm(int64_t np) {
  uint32_t c = next_pow2_u32((uint32_t)(np < (1ll << 29) ? 4 * np : 2 * np));
  return c < EV_MIN_CAP ? EV_MIN_CAP : c;
}

__device__ inline uint32_t table_claim(uint32_t* state, uint32_t cap, const int32_t* vals, int nk,
                                       const int32_t* const* pcols, const int* pcol, int64_t p0, int64_t r) {
  uint32_t slot = tuple_hash(vals, nk) & (cap - 1);
  while (true) {
    // A claimed slot never changes, so a plain read settles most rows of a skewed partition
    // (which repeat one key) without queueing an atomic on that slot; an empty read is claimed.
    uint32_t prev = __hip_atomic_load(&state[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == 0) prev = atomicCAS(&state[slot], 0u, (uint32_t)(r + 1));
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
  uint32_t out = 0;
  bool live = true;
  // A slot's count and keys are loaded together (one memory round trip per slot).
  while (live) {
    const uint32_t c = cnt[slot];
    bool eq = c != 0;
    for (int k = 0; k < nk; ++k) eq &= keys[(size_t)k * cap + slot] == vals[k];
    if (eq) out = c;
    live = c != 0 && !eq;
    slot = (slot + 1) & (cap - 1);
  }
  return out;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the 64 lanes of a wave whose lanes are all active: DPP adds within each row of 16
// lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror 4, row_ror 8 leave every lane with
// its row's sum), then the four row sums read out of lanes 0, 16, 32 and 48.
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
         (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}
__device__ inline uint64_t wave_sum(uint64_t v) { return wave_sum_u64(v); }
__device__ inline uint64_t wave_sum(uint32_t v) { return wave_sum_u32(v); }

// Lane i takes lane i - kShr of its row of 16 lanes (lanes shifted in from outside the row read
// 0).
template <int kShr>
__device__ inline int32_t row_shr(int32_t x) {
  return __builtin_amdgcn_update_dpp(0, x, 0x110 + kShr, 0xf, 0xf, false);  // lanes shifted in read 0
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
// Work items.

// Units with the same radix bits see the same binding cells; a batch of them stages each chunk
// of binding rows once.  Its groups are the join variables of its units.
struct DevBatch {
  int R;
  int shift;  // binding digits are `shift` bits finer
  int ngroups;
  int nunits;
  int heavy_only;  // streamed units: only the cells the streamed form leaves (light[q] == 0)
  const int32_t* tcols[EV_MAXG][EV_MAXTC + 1];  // binding table partitioned on the group's variable; [ntc] = row id
  const int64_t* toff[EV_MAXG];
  int unit[EV_MAXBU];
};

struct DevItem {
  int batch;
  int chunk;
  int64_t q;
};

__device__ inline int find_unit(const int64_t* unit_cells, int num_units, int64_t cell) {
  int lo = 0, hi = num_units - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (unit_cells[mid] <= cell) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_count_items(const DevBatch* batches, const DevUnit* units, const int64_t* unit_cells,
                              int num_units, int tile, const uint8_t* light, int64_t* counts) {
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= unit_cells[num_units]) return;
  const int b = find_unit(unit_cells, num_units, cell);
  const DevBatch& B = batches[b];
  const int64_t q = cell - unit_cells[b];
  if (B.heavy_only && light[q]) {  // the streamed form counts this cell (its batches have shift 0)
    counts[cell] = 0;
    return;
  }
  bool facts = false;
  for (int k = 0; k < B.nunits; ++k) {
    const DevUnit& U = units[B.unit[k]];
    facts = facts || U.poff[q + 1] > U.poff[q];
  }
  int64_t rows = 0;
  if (facts)
    for (int g = 0; g < B.ngroups; ++g) rows += B.toff[g][(q + 1) << B.shift] - B.toff[g][q << B.shift];
  counts[cell] = (rows + tile - 1) / tile;
}

// ------------------------------------------------------------------------------------------
// Streamed units: counting from the relation side.  Within one binding cell the binding side is
// small (the rows of every join variable whose join value falls in the cell) and shared by every
// unit, while a relation's facts there can be many.  A work item (cell q, chunk of binding rows,
// part of the cell's relation entries) stages its rows once and builds two LDS tables of them:
//  * the value table: (join value x, join variable v) -> positive and negative rows;
//  * the pair table: one entry (row r, partner column w) per staged row of join variable v and
//    each binding column w that some literal binds beside v, keyed (T_v(r), T_w(r)).
// Each relation side is reduced once (cold, cached) to its distinct key tuples with their
// multiplicities, per key subset of the unit and per cell (a subset of one column gives runs of
// equal join values).  Every entry of the cell then probes once, whatever the clause variables:
//  * a run (x, c) of unit u adds, for every join variable v with literals keyed on the join
//    column alone, binding += c * rows(x, v) and covered += rows(x, v) (split by label);
//  * a tuple t = (P_j, P_k1, ...) with count c probes (P_j, P_k1); every matching entry (r, w)
//    with r in the rows of join variable v is a match of the literals of (u, subset, group of v)
//    whose column k1 is bound to w and whose further columns equal row r: binding += c and
//    covered += 1 (distinct tuples, so each (literal, row) pair matches at most one tuple).
// A cell is streamed when its binding rows fit TS_CMAX chunks (its relation entries are streamed
// once per chunk); the others (heavily repeated binding values) are left to k_eval, whose cost
// follows the binding rows.  Whether a workload is better off without the streamed form (many
// binding rows per cell against few relation entries, or hot values on both sides) is the
// autotuner's choice.

constexpr int TS_THREADS = 512;
constexpr int TS_ROWS = 1024;          // staged binding rows per work item (all join variables)
constexpr int TS_VCAP = 2048;          // value-table slots (at most TS_ROWS entries: load <= 1/2)
constexpr int TS_VLOG = 11;
constexpr int TS_PCAP = 8192;          // pair-table slots
constexpr int TS_PLOG = 13;
constexpr int TS_PMAX = TS_PCAP / 2;   // pair entries per item (load <= 1/2)
constexpr int TS_MAXL = 128;           // entry lists per launch
constexpr int TS_MAXT = 256;           // (unit, join variable) counters of join-column literals per launch
constexpr int TS_PART = 512;           // wave chunks of relation entries per work item
constexpr int TS_CMAX = 4;             // binding chunks of a cell that the streamed form takes
constexpr int TS_RUNROW = 16;          // runs per row of lanes (a row's runs belong to one list)
constexpr int TS_RUNROWS = 4;          // rows per wave chunk
constexpr int TS_TUPCH = 256;          // tuples per wave chunk (four per lane)
static_assert(TS_VCAP == (1 << TS_VLOG) && TS_PCAP == (1 << TS_PLOG), "table sizes are powers of two");
static_assert(TS_ROWS <= 2047 && TS_ROWS <= 0xffff && EV_MAXTC <= 16 && EV_MAXG <= 16 && 2 * EV_MAXG < 256,
              "slot words hold rows, columns and variables; row flags hold the variable");
static_assert(TS_VCAP >= 2 * TS_ROWS, "value table at load <= 1/2");
static_assert(TS_ROWS % TS_THREADS == 0, "staged rows are split evenly over the threads");

// Distinct key tuples of a (relation, join column) partition for one key subset, in cell order.
struct TsList {
  const int32_t* col[EV_MAXK];  // col[0]: join column value; col[1]: the pair partner; then the rest
  const int32_t* cnt;           // facts with the tuple; bit 31 (m >= 3): the first two columns differ
                                // from the previous tuple's
  const int64_t* off;           // per cell: first entry (cells + 1)
  int m;                        // key columns (1: runs of the join column)
  int unit;                     // DevUnit of the literal ranges (m >= 2)
  int subset;                   // the unit's subset (m >= 2)
  int kidx[EV_MAXK];            // list column i is key position kidx[i] of the subset's literals
  int psubset;                  // m >= 3: the unit's subset of the first two columns, counted here too, or -1
  int pkidx1;                   // its key position of list column 1
  int target[EV_MAXG];          // m == 1: per join variable, its counter or -1
  int vgroup[EV_MAXG];          // m >= 2: per join variable, the unit's group or -1
};

struct TsTarget {
  int l0, l1;  // literals [l0, l1) receive the counter
};

struct TsLaunch {
  const TsList* lists;  // run lists (m == 1) first, then tuple lists
  int nlists;
  int nrun;
  const TsTarget* targets;
  int ntargets;
  int nvars;   // join variables (staged segments)
  int ntc;     // staged binding columns
  int tr;      // rows of each join variable per work item
  int jt[EV_MAXG];          // staged column of each join variable
  uint32_t wmask[EV_MAXG];  // per join variable: the staged columns paired with it
  const int32_t* tcols[EV_MAXG][EV_MAXTC + 1];  // binding partitioned on each join variable; [ntc] = row id
  const int64_t* toff[EV_MAXG];
  int64_t cells;
};


// Which cells the streamed form takes (light[q] = 1; k_eval skips them): those whose binding rows
// fit TS_CMAX chunks (every streamed launch stages the same binding rows).
__global__ void k_ts_light(const TsLaunch* L_, uint8_t* light) {
  cptr<TsLaunch> L = as_const(L_);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L->cells) return;
  int64_t chunks = 0;
  for (int v = 0; v < L->nvars; ++v) chunks = max(chunks, (L->toff[v][q + 1] - L->toff[v][q] + L->tr - 1) / L->tr);
  light[q] = chunks <= TS_CMAX ? 1 : 0;
}

// The relation side of a streamed launch's cells, fixed by its key-tuple lists (made once per
// plan): cellmat[q * nlists + l] = (first entry, entries) of list l in cell q, so that a work item
// reads its cell's lists in one pass, and parts[q] = the parts of the cell's wave chunks (its
// runs, all run lists one after another in rows of TS_RUNROW, TS_RUNROWS rows per chunk, then its
// tuples in chunks of TS_TUPCH) in TS_PART chunks each.
__global__ void k_ts_cells(const TsLaunch* L_, int32_t* parts, int2* cellmat) {
  cptr<TsLaunch> L = as_const(L_);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L->cells) return;
  cptr<TsList> lists = as_const(L->lists);
  int64_t rows = 0, nt = 0;
  for (int l = 0; l < L->nlists; ++l) {
    const int64_t b = lists[l].off[q], n = lists[l].off[q + 1] - b;
    cellmat[q * L->nlists + l] = make_int2((int)b, (int)n);
    if (l < L->nrun) rows += (n + TS_RUNROW - 1) / TS_RUNROW;
    else nt += n;
  }
  const int64_t wc = (rows + TS_RUNROWS - 1) / TS_RUNROWS + (nt + TS_TUPCH - 1) / TS_TUPCH;
  parts[q] = (int32_t)max<int64_t>(1, (wc + TS_PART - 1) / TS_PART);
}

// Work items per cell that the streamed form takes: (binding chunks) x (parts of the cell).
__global__ void k_ts_count(const TsLaunch* L_, const uint8_t* light, const int32_t* parts, int64_t* counts) {
  cptr<TsLaunch> L = as_const(L_);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L->cells) return;
  int64_t chunks = 0;
  for (int v = 0; v < L->nvars; ++v) {
    const int64_t rows = L->toff[v][q + 1] - L->toff[v][q];
    chunks = max(chunks, (rows + L->tr - 1) / L->tr);
  }
  counts[q] = light[q] ? chunks * parts[q] : 0;
}

// One thread per work item: item i of cell q (the last cell whose scanned start is <= i) is
// binding chunk (i - start) / parts[q], part (i - start) % parts[q].
__global__ void k_ts_expand(const int64_t* counts, const int64_t* starts, const int32_t* parts, int64_t cells,
                            DevItem* items, int64_t* total) {
  const int64_t all = starts[cells - 1] + counts[cells - 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    total[0] = all;
    total[1] = 0;  // the next item to hand out (k_stream's workgroups take items as they finish)
  }
  if (i >= all) return;
  int64_t lo = 0, hi = cells - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int idx = (int)(i - starts[lo]), p = parts[lo];
  items[i] = DevItem{idx % p, idx / p, lo};
}

__device__ inline uint32_t ts_vhash(int32_t x, int v) { return ((uint32_t)x * 0x9e3779b1u + (uint32_t)v) * 0x85ebca6bu; }
__device__ inline uint32_t ts_phash(int32_t x, int32_t y) {
  return (((uint32_t)x * 0x9e3779b1u) ^ (uint32_t)y) * 0x85ebca6bu;
}
// Value slot (64 bits): the join value in the low word; above it the value's negative rows (bits
// 32-44), positive rows (45-57) and join variable (58-61) of the item, and bit 63 set once the
// slot is taken (0 = empty).  A probe compares value and variable in one read.
// Pair slot word: fingerprint (13 bits) << 19 | join variable << 15 | partner column << 11 |
// (1 + row).  0 = empty.
constexpr unsigned long long TS_VTAKEN = 1ull << 63;
constexpr unsigned long long TS_VKEYBITS = 0xffffffffull | (0xfull << 58) | TS_VTAKEN;
constexpr int TS_VNEG = 32, TS_VPOS = 45, TS_VVAR = 58;
static_assert(TS_ROWS < (1 << 13), "a value's rows fit 13 bits");
__device__ inline unsigned long long ts_vkey(int32_t x, int vj) {
  return TS_VTAKEN | ((unsigned long long)vj << TS_VVAR) | (uint32_t)x;
}
constexpr uint32_t TS_PFP_MASK = ~((1u << 19) - 1u);

struct TsArgs {
  const TsLaunch* launch;
  const int2* cellmat;  // per cell and list: first entry, entries
  const DevUnit* units;
  const DevLit* lits;
  const DevItem* items;
  int64_t* num_items;  // [0]: items; [1]: the next item to hand out
  int64_t max_items;   // the item list's capacity
  int num_positive;
  unsigned long long* counts_out;  // ncopies x num_slots x 4: block b adds into copy b % ncopies
  int64_t copy_words;
  int ncopies;
};

// A tuple matched staged row r of join variable vj through partner column w: the literals of
// the list's subset in the unit's group of vj whose column k1 is bound to w and whose further
// key columns equal the row's.
__device__ __attribute__((noinline)) void ts_tuple_match(const TsArgs& a, cptr<TsList> D, int64_t pos, int r, int vj,
                                                         int w, const int32_t* tstage, const uint8_t* rflag,
                                                         unsigned long long* copy) {
  const int g = D->vgroup[vj];
  if (g < 0) return;
  cptr<DevUnit> U = as_const(a.units) + D->unit;
  cptr<DevLit> lits = as_const(a.lits);
  const int m = D->m;
  const uint32_t cw = (uint32_t)D->cnt[pos];
  const unsigned long long facts = cw & 0x7fffffffu;
  const bool p = (rflag[r] & 1) != 0;
  if (D->psubset >= 0) {
    // The first two columns' subset: every tuple of a (column 0, column 1) group adds its facts,
    // the group's first tuple the covered row.
    const int l0 = U->g[g].sbeg[D->psubset], l1 = U->g[g].sbeg[D->psubset + 1];
    for (int li = l0; li < l1; ++li) {
      if (lits[li].tcol[D->pkidx1] != w) continue;
      unsigned long long* c = copy + (int64_t)lits[li].slot * 4;
      if (cw >> 31) atomicAdd(c + (p ? 0 : 1), 1ull);
      atomicAdd(c + (p ? 2 : 3), facts);
    }
  }
  const int l0 = U->g[g].sbeg[D->subset], l1 = U->g[g].sbeg[D->subset + 1];
  for (int li = l0; li < l1; ++li) {
    if (lits[li].tcol[D->kidx[1]] != w) continue;
    bool ok = true;
    for (int i = 2; i < m && ok; ++i) ok = D->col[i][pos] == tstage[lits[li].tcol[D->kidx[i]] * TS_ROWS + r];
    if (!ok) continue;
    unsigned long long* c = copy + (int64_t)lits[li].slot * 4;
    atomicAdd(c + (p ? 0 : 1), 1ull);
    atomicAdd(c + (p ? 2 : 3), facts);
  }
}

// Inclusive sums within each row of 16 lanes (DPP row shifts); lane 15 of a row ends with the
// row's total.  Every lane must be active.
__device__ inline uint32_t row_sum16(uint32_t v) {
  v += (uint32_t)row_shr<1>((int32_t)v);
  v += (uint32_t)row_shr<2>((int32_t)v);
  v += (uint32_t)row_shr<4>((int32_t)v);
  v += (uint32_t)row_shr<8>((int32_t)v);
  return v;
}

// A tuple whose first two columns are equal (x == y) matches rows of join variable vj whose
// literal binds column k1 to vj itself: the pair table leaves those pairs out (a cell holds few
// distinct join values, so they would pile up on few slots), and the rows of vj with value x are
// found by scanning vj's staged rows.
__device__ __attribute__((noinline)) void ts_equal_pair(const TsArgs& a, cptr<TsList> D, int64_t pos, int32_t x,
                                                        int nvars, const int* seg_off, const int32_t* tstage,
                                                        const uint8_t* rflag, const int* jt, unsigned long long* copy) {
  for (int vj = 0; vj < nvars; ++vj) {
    if (D->vgroup[vj] < 0) continue;
    const int32_t* col = tstage + jt[vj] * TS_ROWS;
    for (int r = seg_off[vj]; r < seg_off[vj + 1]; ++r)
      if (col[r] == x) ts_tuple_match(a, D, pos, r, vj, jt[vj], tstage, rflag, copy);
  }
}

// Probes of the pair table for up to TS_TJ tuples of a lane, interleaved (their slot reads in
// flight together); every matching entry is handed to ts_tuple_match.
constexpr int TS_TJ = TS_TUPCH / 64;
__device__ inline void ts_pair_probes(const TsArgs& a, cptr<TsList> lists, const int (&lj)[TS_TJ],
                                      const int64_t (&pos)[TS_TJ], const int32_t (&x)[TS_TJ],
                                      const int32_t (&y)[TS_TJ], const uint32_t* pword, const int32_t* tstage,
                                      const uint8_t* rflag, const int* jt, unsigned long long* copy) {
  uint32_t slot[TS_TJ], fp[TS_TJ];
  bool live[TS_TJ];
#pragma unroll
  for (int j = 0; j < TS_TJ; ++j) {
    const uint32_t h = ts_phash(x[j], y[j]);
    fp[j] = ((h >> 4) & 0x1fffu) << 19;
    slot[j] = h >> (32 - TS_PLOG);
    live[j] = pos[j] >= 0;
  }
  // The table is read four slots (one aligned 16-byte group) at a time: a probe checks the slots
  // from its current one to the group's end and stops at the first empty slot.  The common case
  // (no fingerprint hit) is branch-free: bit masks of the group's empty slots and fingerprint
  // hits; only a hit goes to the key check.
  // Every lane reads all four groups (a slot is always inside the table), so that the reads need
  // no exec-mask changes; a finished probe's group is read and ignored.
  const uint4* groups = reinterpret_cast<const uint4*>(pword);
  while (live[0] || live[1] || live[2] || live[3]) {
    uint4 g[TS_TJ];
#pragma unroll
    for (int j = 0; j < TS_TJ; ++j) g[j] = groups[slot[j] >> 2];
#pragma unroll
    for (int j = 0; j < TS_TJ; ++j) {
      const uint32_t wds[4] = {g[j].x, g[j].y, g[j].z, g[j].w};
      uint32_t empty = 0, hit = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        empty |= (wds[s] == 0u ? 1u : 0u) << s;
        hit |= ((wds[s] & TS_PFP_MASK) == fp[j] ? 1u : 0u) << s;
      }
      const uint32_t from = 0xfu << (slot[j] & 3u);       // slots at or after the current one
      const uint32_t stop = empty & from;                  // the probe ends at the first of these
      const uint32_t before = stop ? (stop & (0u - stop)) - 1u : 0xfu;  // slots before that end
      uint32_t cand = live[j] ? hit & from & before & ~empty : 0u;
      while (cand) {
        const int s = __ffs((int)cand) - 1;
        cand &= cand - 1;
        const uint32_t wd = wds[s];
        const int r = (int)(wd & 0x7ffu) - 1, w = (int)((wd >> 11) & 15u), vj = (int)((wd >> 15) & 15u);
        if (tstage[jt[vj] * TS_ROWS + r] == x[j] && tstage[w * TS_ROWS + r] == y[j])
          ts_tuple_match(a, lists + lj[j], pos[j], r, vj, w, tstage, rflag, copy);
      }
      live[j] = live[j] && stop == 0u;
      slot[j] = ((slot[j] | 3u) + 1u) & (TS_PCAP - 1);
    }
  }
}
static_assert(TS_TJ == 4, "the probe loop tests four tuples");

// Four waves per SIMD: two 512-thread workgroups per CU, as LDS allows.
__global__ __launch_bounds__(TS_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_stream(TsArgs a) {
  extern __shared__ uint32_t lds[];
  cptr<TsLaunch> L = as_const(a.launch);
  const int ntc = L->ntc;
  int32_t* tstage = reinterpret_cast<int32_t*>(lds);   // ntc x TS_ROWS
  unsigned long long* vtab = reinterpret_cast<unsigned long long*>(lds + ntc * TS_ROWS);  // TS_VCAP
  uint32_t* pword = lds + ntc * TS_ROWS + 2 * TS_VCAP;                                       // TS_PCAP
  uint8_t* rflag = reinterpret_cast<uint8_t*>(pword + TS_PCAP);  // TS_ROWS: join variable << 1 | positive
  __shared__ unsigned long long acc[TS_MAXT][4];
  __shared__ const int32_t* lcol0[TS_MAXL];  // per list: join values (runs) or first key column (tuples)
  __shared__ const int32_t* lcol1[TS_MAXL];  // counts (runs) or second key column (tuples)
  __shared__ int lbase[TS_MAXL];             // per list, in the item's cell: first entry (a list holds
                                             // fewer than 2^31 entries),
  __shared__ int lcnt[TS_MAXL];              // entries,
  __shared__ int lstart[TS_MAXL + 1];        // first row (run lists) or first position (tuple lists)
  __shared__ int16_t ltgt[TS_MAXL][4];       // per run list: the counters of join variables 0-3
  __shared__ uint8_t cmap[TS_PART + 1];      // per wave chunk of the item (and the next): its first list
  __shared__ const int32_t* tcol[EV_MAXG][EV_MAXTC + 1];
  __shared__ const int64_t* toff[EV_MAXG];
  __shared__ int seg_off[EV_MAXG + 1];
  __shared__ int jt[EV_MAXG];
  __shared__ int item_rows, item_tuples, next_chunk;
  static_assert(TS_MAXT < 32768 && TS_MAXL <= 256, "target entries and list indices fit their types");

  const int lane = threadIdx.x & 63;
  const int wave = uni(threadIdx.x >> 6);
  const int nvars = L->nvars, nlists = L->nlists, nrun = L->nrun;
  const int64_t total = min(uni64(a.num_items[0]), a.max_items);
  cptr<DevItem> items = as_const(a.items);
  cptr<TsList> lists = as_const(L->lists);
  unsigned long long* copy = a.counts_out + (int64_t)(blockIdx.x % a.ncopies) * a.copy_words;
  // Items are handed out one at a time from a device counter (their costs vary with the cells);
  // thread 0 asks for the next one while the current one streams.
  __shared__ int64_t item_next;
  int64_t next = 0;
  if (threadIdx.x == 0) next = (int64_t)atomicAdd(reinterpret_cast<unsigned long long*>(a.num_items + 1), 1ull);
  for (int i = threadIdx.x; i < TS_MAXT * 4; i += TS_THREADS) (&acc[0][0])[i] = 0;
  for (int i = threadIdx.x; i < nlists; i += TS_THREADS) {
    lcol0[i] = lists[i].col[0];
    lcol1[i] = lists[i].m == 1 ? lists[i].cnt : lists[i].col[1];
#pragma unroll
    for (int v = 0; v < 4; ++v) ltgt[i][v] = (int16_t)(v < nvars ? lists[i].target[v] : -1);
  }
  for (int i = threadIdx.x; i < nvars * (ntc + 1); i += TS_THREADS) tcol[i / (ntc + 1)][i % (ntc + 1)] = L->tcols[i / (ntc + 1)][i % (ntc + 1)];
  if ((int)threadIdx.x < nvars) {
    jt[threadIdx.x] = L->jt[threadIdx.x];
    toff[threadIdx.x] = L->toff[threadIdx.x];
  }

  while (true) {
    if (threadIdx.x == 0) item_next = next;
    __syncthreads();  // the previous item is done with the tables, rows and lists
    const int64_t it = item_next;
    if (it >= total) break;
    const int64_t q = items[it].q;
    const int chunk = items[it].chunk, part = items[it].batch;  // binding chunk, part of the cell's entries
    // Binding rows of the chunk: the join variables' segments one after another (wave-uniform
    // loop); each thread notes the segment and source row of its two staged rows.
    constexpr int RPT = TS_ROWS / TS_THREADS;  // staged rows per thread
    int rv[RPT];
    int64_t rsrc[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      rv[j] = -1;
      rsrc[j] = 0;
    }
    int rows_all = 0;
    for (int v = 0; v < nvars; ++v) {
      const int64_t t0 = L->toff[v][q], t1 = L->toff[v][q + 1];
      const int64_t s = t0 + (int64_t)chunk * L->tr;
      const int n = (int)max<int64_t>(0, min<int64_t>(L->tr, t1 - s));
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int r = (int)threadIdx.x + j * TS_THREADS;
        if (r >= rows_all && r < rows_all + n) {
          rv[j] = v;
          rsrc[j] = s + (r - rows_all);
        }
      }
      if (threadIdx.x == 0) seg_off[v] = rows_all;
      rows_all += n;
    }
    if (threadIdx.x == 0) seg_off[nvars] = rows_all;
    if (wave == 0) {
      // The cell's entries of every list (two lists per lane): runs in rows of TS_RUNROW (each
      // list's runs start a row), tuples one after another.
      int2 bn[2] = {make_int2(0, 0), make_int2(0, 0)};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int l = lane + 64 * j;
        if (l < nlists) bn[j] = a.cellmat[q * nlists + l];
      }
      int rows = 0, tuples = 0;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int l = lane + 64 * j;
        const int n = l < nlists ? bn[j].y : 0;
        const int nrw = l < nrun ? (n + TS_RUNROW - 1) / TS_RUNROW : 0;
        const int ntp = l >= nrun ? n : 0;
        int xr = nrw, xt = ntp;
        for (int off = 1; off < 64; off <<= 1) {
          const int yr = __shfl_up(xr, off, 64), yt = __shfl_up(xt, off, 64);
          if (lane >= off) {
            xr += yr;
            xt += yt;
          }
        }
        if (l < nlists) {
          lbase[l] = bn[j].x;
          lcnt[l] = n;
          lstart[l] = l < nrun ? rows + xr - nrw : tuples + xt - ntp;
        }
        rows += __shfl(xr, 63, 64);
        tuples += __shfl(xt, 63, 64);
      }
      if (lane == 0) {
        lstart[nlists] = tuples;
        item_rows = rows;
        item_tuples = tuples;
        next_chunk = part * TS_PART;
      }
    }
    // Stage the rows, two per thread (all columns of a row in flight together): the ntc binding
    // columns and a flag (join variable << 1 | positive, from the row id); then empty the tables.
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int r = (int)threadIdx.x + j * TS_THREADS;
      if (rv[j] < 0) continue;
      const int v = rv[j];
      const int64_t sr = rsrc[j];
      // Columns in groups of five (loads in flight together); column ntc is the row id.
      for (int cb = 0; cb <= ntc; cb += 5) {
        int32_t val[5];
#pragma unroll
        for (int k = 0; k < 5; ++k)
          if (cb + k <= ntc) val[k] = tcol[v][cb + k][sr];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const int c = cb + k;
          if (c < ntc) tstage[c * TS_ROWS + r] = val[k];
          else if (c == ntc) rflag[r] = (uint8_t)((v << 1) | (val[k] < a.num_positive ? 1 : 0));
        }
      }
    }
    {
      const int n16 = (2 * TS_VCAP + TS_PCAP) / 4;
      uint4* t = reinterpret_cast<uint4*>(vtab);
      for (int i = threadIdx.x; i < n16; i += TS_THREADS) t[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    const int nrows = item_rows, nt = item_tuples;
    const int rc = (nrows + TS_RUNROWS - 1) / TS_RUNROWS;
    const int c_all = rc + (nt + TS_TUPCH - 1) / TS_TUPCH;
    const int c0 = part * TS_PART, c1 = min(c0 + TS_PART, c_all);
    // The first list of each of the item's wave chunks and of the chunk after them.
    for (int t = threadIdx.x; t <= c1 - c0; t += TS_THREADS) {
      const int k = c0 + t;
      int lo, hi, key;
      if (k < rc) {
        lo = 0, hi = nrun - 1, key = k * TS_RUNROWS;
      } else {
        lo = nrun, hi = nlists - 1, key = (k - rc) * TS_TUPCH;
      }
      if (k >= c_all) lo = hi = nlists - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (lstart[mid] <= key) lo = mid;
        else hi = mid - 1;
      }
      cmap[t] = (uint8_t)lo;
    }
    // Build both tables.  A row's first value-table CAS and the CASes of up to four pair entries
    // are in flight together; collisions then walk on one slot at a time.
    for (int r = threadIdx.x; r < rows_all; r += TS_THREADS) {
      const int fl = rflag[r];
      const int vj = fl >> 1;
      const int jc = jt[vj];
      const int32_t x = tstage[jc * TS_ROWS + r];
      const unsigned long long vkey = ts_vkey(x, vj);
      const unsigned long long vcount = 1ull << ((fl & 1) ? TS_VPOS : TS_VNEG);
      uint32_t vslot = ts_vhash(x, vj) >> (32 - TS_VLOG);
      uint32_t wm = L->wmask[vj] & ~(1u << jc);  // pairs of the join variable with itself: see ts_equal_pair
      uint32_t ps[4], pw[4], pp[4];
      int np = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ps[k] = pw[k] = 0;
        if (wm) {
          const int w = __ffs((int)wm) - 1;
          wm &= wm - 1;
          const uint32_t h = ts_phash(x, tstage[w * TS_ROWS + r]);
          pw[k] = (((h >> 4) & 0x1fffu) << 19) | ((uint32_t)vj << 15) | ((uint32_t)w << 11) | (uint32_t)(r + 1);
          ps[k] = h >> (32 - TS_PLOG);
          np = k + 1;
        }
      }
      // A new value claims its slot with its first row counted; a repeated one adds its row.
      unsigned long long vprev = atomicCAS(&vtab[vslot], 0ull, vkey | vcount);
#pragma unroll
      for (int k = 0; k < 4; ++k) pp[k] = k < np ? atomicCAS(&pword[ps[k]], 0u, pw[k]) : 0u;
      while (vprev != 0 && (vprev & TS_VKEYBITS) != vkey) {
        vslot = (vslot + 1) & (TS_VCAP - 1);
        vprev = atomicCAS(&vtab[vslot], 0ull, vkey | vcount);
      }
      if (vprev != 0) atomicAdd(&vtab[vslot], vcount);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        while (pp[k] != 0) {
          ps[k] = (ps[k] + 1) & (TS_PCAP - 1);
          pp[k] = atomicCAS(&pword[ps[k]], 0u, pw[k]);
        }
      while (wm) {  // partners beyond the first four
        const int w = __ffs((int)wm) - 1;
        wm &= wm - 1;
        const uint32_t h = ts_phash(x, tstage[w * TS_ROWS + r]);
        const uint32_t word = (((h >> 4) & 0x1fffu) << 19) | ((uint32_t)vj << 15) | ((uint32_t)w << 11) | (uint32_t)(r + 1);
        uint32_t slot = h >> (32 - TS_PLOG);
        while (atomicCAS(&pword[slot], 0u, word) != 0) slot = (slot + 1) & (TS_PCAP - 1);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) next = (int64_t)atomicAdd(reinterpret_cast<unsigned long long*>(a.num_items + 1), 1ull);
    // Stream the item's wave chunks, each wave taking the next one as it finishes: run chunks of
    // TS_RUNROWS rows of TS_RUNROW lanes (a row holds runs of one list), then tuple chunks of
    // TS_TUPCH tuples (TS_TJ consecutive ones per lane).  A lane finds its list by binary search
    // between the chunk's first list and the next chunk's.
    while (true) {
      int k = 0;
      if (lane == 0) k = atomicAdd(&next_chunk, 1);
      k = __shfl(k, 0, 64);
      if (k >= c1) break;
      const bool run_chunk = k < rc;
      int lo = cmap[k - c0];
      int hi = (k + 1 < c_all && (k + 1 < rc) == run_chunk) ? (int)cmap[k + 1 - c0] : run_chunk ? nrun - 1 : nlists - 1;
      if (run_chunk) {
        const int row = k * TS_RUNROWS + (lane >> 4);
        const bool rin = row < nrows;
        if (rin)
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (lstart[mid] <= row) lo = mid;
            else hi = mid - 1;
          }
        const int l = lo;
        const int e = (row - lstart[l]) * TS_RUNROW + (lane & 15);
        const bool in = rin && e < lcnt[l];
        int32_t x = 0;
        uint32_t c = 0;
        if (in) {
          const int64_t pos = lbase[l] + e;
          x = lcol0[l][pos];
          c = (uint32_t)lcol1[l][pos];
        }
        // Join variables four at a time: their value-table reads in flight together.
        // Binding sums of 16 lanes fit 32 bits when every run count is below 2^28 / tr (a value's
        // rows of one variable are at most tr).
        const bool small = !__ballot(in && c >= (1u << 28) / (uint32_t)L->tr);
        for (int vg = 0; vg < nvars; vg += 4) {
          int t[4];
          unsigned long long key[4];
          uint32_t slot[4], pn[4];
          bool look[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int vj = vg + u;
            t[u] = !rin || vj >= nvars ? -1 : vj < 4 ? ltgt[l][vj] : lists[l].target[vj];
            key[u] = ts_vkey(x, vj);
            slot[u] = ts_vhash(x, vj) >> (32 - TS_VLOG);
            look[u] = in && t[u] >= 0;
            pn[u] = 0;
          }
          while (look[0] || look[1] || look[2] || look[3]) {
            unsigned long long wd[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) wd[u] = vtab[slot[u]];  // always inside the table; read unmasked
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              if (!look[u]) continue;
              if (wd[u] == 0) {
                look[u] = false;
              } else if ((wd[u] & TS_VKEYBITS) == key[u]) {
                // positive rows << 16 | negative rows
                pn[u] = (uint32_t)((wd[u] >> TS_VPOS) & 0x1fffu) << 16 | (uint32_t)((wd[u] >> TS_VNEG) & 0x1fffu);
                look[u] = false;
              } else {
                slot[u] = (slot[u] + 1) & (TS_VCAP - 1);
              }
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (!__ballot(t[u] >= 0 && pn[u] != 0)) continue;
            // Row sums: covered (packed; a row's runs are distinct values of one list, so its
            // positive and negative rows stay below the chunk's rows of the variable) and
            // binding, in one 32-bit sum when every run count is small, else in 21-bit pieces (a
            // run count is below 2^31 and rows below 2^11).
            const uint32_t cov = row_sum16(pn[u]);
            uint64_t sbp, sbn;
            if (small) {
              sbp = row_sum16(c * (pn[u] >> 16));
              sbn = row_sum16(c * (pn[u] & 0xffffu));
            } else {
              const uint64_t bp = (uint64_t)c * (pn[u] >> 16), bn = (uint64_t)c * (pn[u] & 0xffffu);
              sbp = ((uint64_t)row_sum16((uint32_t)(bp >> 21)) << 21) + row_sum16((uint32_t)(bp & 0x1fffffu));
              sbn = ((uint64_t)row_sum16((uint32_t)(bn >> 21)) << 21) + row_sum16((uint32_t)(bn & 0x1fffffu));
            }
            if ((lane & 15) == 15 && t[u] >= 0 && cov != 0) {
              atomicAdd(&acc[t[u]][0], (unsigned long long)(cov >> 16));
              atomicAdd(&acc[t[u]][1], (unsigned long long)(cov & 0xffffu));
              atomicAdd(&acc[t[u]][2], (unsigned long long)sbp);
              atomicAdd(&acc[t[u]][3], (unsigned long long)sbn);
            }
          }
        }
      } else {
        const int i0 = (k - rc) * TS_TUPCH + TS_TJ * lane;  // the lane's first tuple
        if (i0 < nt)
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (lstart[mid] <= i0) lo = mid;
            else hi = mid - 1;
          }
        int l = lo;
        int32_t x[TS_TJ], y[TS_TJ];
        int64_t pos[TS_TJ];
        int lj[TS_TJ];
#pragma unroll
        for (int j = 0; j < TS_TJ; ++j) {
          const int i = i0 + j;
          x[j] = y[j] = 0;
          pos[j] = -1;
          lj[j] = l;
          if (i < nt) {
            while (lstart[l + 1] <= i) ++l;  // lists with no tuple in this cell are passed over
            lj[j] = l;
            pos[j] = lbase[l] + (i - lstart[l]);
            x[j] = lcol0[l][pos[j]];
            y[j] = lcol1[l][pos[j]];
          }
        }
        ts_pair_probes(a, lists, lj, pos, x, y, pword, tstage, rflag, jt, copy);
#pragma unroll
        for (int j = 0; j < TS_TJ; ++j)
          if (pos[j] >= 0 && x[j] == y[j])
            ts_equal_pair(a, lists + lj[j], pos[j], x[j], nvars, seg_off, tstage, rflag, jt, copy);
      }
    }
  }
  __syncthreads();  // every item's counters are in acc
  cptr<TsTarget> targets = as_const(L->targets);
  cptr<DevLit> lits = as_const(a.lits);
  for (int i = threadIdx.x; i < L->ntargets * 4; i += TS_THREADS) {
    const unsigned long long v = acc[i >> 2][i & 3];
    if (v == 0) continue;
    for (int li = targets[i >> 2].l0; li < targets[i >> 2].l1; ++li)
      atomicAdd(copy + (int64_t)lits[li].slot * 4 + (i & 3), v);
  }
}

// Distinct tuples of sorted rows: heads[i] = 1 where a new tuple starts.
struct TsCols {
  const int32_t* c[EV_MAXK];
  int m;
};
__global__ void k_tuple_heads(TsCols s, int64_t n, int64_t* flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool head = i == 0;
  for (int k = 0; k < s.m && !head; ++k) head = s.c[k][i] != s.c[k][i - 1];
  flags[i] = head ? 1 : 0;
}

// Tuples from the scanned heads (starts[i] = heads before row i): each tuple's columns and first
// row, and per cell its first tuple (cell p starts at row poff[p]; equal tuples share their
// cell).  Threads cover max(n, cells + 1).
struct TsOut {
  int32_t* c[EV_MAXK];
};
__global__ void k_tuple_compact(TsCols s, int64_t n, const int64_t* flags, const int64_t* starts, TsOut out,
                                int32_t* first_row, const int64_t* poff, int64_t cells, int64_t* tuple_off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) {
    const int64_t t = starts[i];
    for (int k = 0; k < s.m; ++k) out.c[k][t] = s.c[k][i];
    first_row[t] = (int32_t)i;
  }
  if (i <= cells) {
    const int64_t row = poff[i];
    tuple_off[i] = row < n ? starts[row] : (n > 0 ? starts[n - 1] + flags[n - 1] : 0);
  }
}

// Tuple multiplicities from consecutive first rows (the last tuple ends at row n); with three or
// more key columns, bit 31 marks a tuple whose first two columns differ from the previous
// tuple's (the first tuple of a (column 0, column 1) group).
__global__ void k_tuple_counts(TsOut tup, int m, const int32_t* first_row, int64_t tuples, int64_t n, int32_t* cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tuples) return;
  uint32_t c = (uint32_t)((t + 1 < tuples ? (int64_t)first_row[t + 1] : n) - first_row[t]);
  if (m >= 3 && (t == 0 || tup.c[0][t] != tup.c[0][t - 1] || tup.c[1][t] != tup.c[1][t - 1])) c |= 0x80000000u;
  cnt[t] = (int32_t)c;
}

// One thread per work item (the grid covers an upper bound of the item count): item i belongs to
// the last cell whose scanned start is <= i (cells without items share their start with the
// next cell).  Thread 0 writes the total.  A cell with many items (a frequent join value) is
// expanded by many threads.
__global__ void k_expand_items(const int64_t* unit_cells, int num_units, const int64_t* counts,
                               const int64_t* starts, DevItem* items, int64_t* total) {
  const int64_t cells = unit_cells[num_units];
  const int64_t all = starts[cells - 1] + counts[cells - 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *total = all;
  if (i >= all) return;
  int64_t lo = 0, hi = cells - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int u = find_unit(unit_cells, num_units, lo);
  items[i] = DevItem{u, (int)(i - starts[lo]), lo - unit_cells[u]};
}

struct EvalArgs {
  const DevBatch* batches;
  const DevUnit* units;
  const DevLit* lits;
  const DevItem* items;
  const int64_t* num_items;
  int ntc;                // staged binding columns (the row id follows)
  int num_positive;  // binding rows are below 2^31
  unsigned long long* counts_out;  // ncopies x num_slots x 4: block b adds into copy b % ncopies
  int64_t copy_words;              // num_slots x 4
  int ncopies;
  unsigned long long* bits;        // num_slots x bit_words, or null
  int64_t bit_words;
  int tile_rows;
  int cidx_off;  // LDS word offsets: compacted row indices (uint16 x tile), hit masks (uint64)
  int mask_off;
  int mstride;   // 64-row blocks per mask slot (tile / 64)
};

// Per-item shared state, filled by the block's first threads.
struct ItemShared {
  int seg_group[EV_MAXG];      // batch group of the segment
  int seg_ug[EV_MAXG];         // the current unit's group for the segment, or -1
  int seg_off[EV_MAXG + 1];
  int seg_rows[EV_MAXG];
  int64_t seg_src[EV_MAXG];
  const int32_t* seg_tid[EV_MAXG];  // the segment's first binding row id (HBM), for covered bitvectors
  int seg_hits[EV_MAXG];       // rows of the segment whose join value occurs in the partition
  int blk_start[EV_MAXG + 1];  // join phase: 64-row blocks of all rows, per segment
  int ent_start[EV_MAXG + 1];  // subset phase: (literal, block of hit rows) entries, per segment
  uint32_t need;               // bit s: a parent of a literal of subset s matched a row of this item
  int64_t u_p0[EV_MAXBU];      // per unit of the batch: first fact of cell q, fact count, heavy ordinal
  int u_np[EV_MAXBU];
  int u_hq[EV_MAXBU];
  int8_t u_ug[EV_MAXBU][EV_MAXG];  // per unit and batch group: the unit's group, or -1
  int nseg;
};

// Probe of an LDS table (packed slot words); the representative fact's key columns are compared
// in the staged facts (poff[k] = word offset of key column k there).
template <int NK>
__device__ inline uint32_t lds_probe(const uint32_t* tbl, uint32_t cap, const int32_t* pstage, const int* poff,
                                     const int32_t* vals, int nk) {
  const uint32_t h = lds_hash(vals, NK ? NK : nk);
  const uint32_t fp = lds_fp(h);
  uint32_t slot = lds_slot(h, cap);
  uint32_t out = 0;
  bool live = true;
  // One exit test per slot; the key check runs only on a fingerprint match.
  while (live) {
    const uint32_t w = tbl[slot];
    bool eq = w != 0 && (w & EV_FP_MASK) == fp;
    if (eq) {
      const int r = (int)((w >> EV_REP_SHIFT) & EV_FIELD_MASK) - 1;
#pragma unroll
      for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
        if (NK || k < nk) eq &= (pstage[poff[k] + r] == vals[k]);
      if (eq) out = w & EV_FIELD_MASK;
    }
    live = w != 0 && !eq;
    slot = (slot + 1) & (cap - 1);
  }
  return out;
}

// Wave-uniform registers of the literal being probed.  Its row-side equalities (rare: only a
// plan that binds one relation column to two binding columns has them) are read when needed.
struct LitRegs {
  int slot;
  int neq;
  int pmask, mslot;
  int tb[EV_MAXK];  // word offsets of the key's binding columns in the staged tile
};

template <int NK>
__device__ inline void load_literal(cptr<DevLit> lits, int li, int nk, int tile, LitRegs& L) {
  cptr<DevLit> p = lits + li;
  L.slot = p->slot;
  L.neq = p->neq;
  L.pmask = p->pmask;
  L.mslot = p->mslot;
#pragma unroll
  for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
    if (NK || k < nk) L.tb[k] = p->tcol[k] * tile;
}

// Whether staged row `row` meets literal li's row-side equalities.
__device__ inline bool row_equalities(cptr<DevLit> lits, int li, int neq, int tile, const int32_t* tstage, int row) {
  cptr<DevLit> p = lits + li;
  bool ok = true;
  for (int q = 0; q < neq; ++q) ok &= tstage[p->eq[q][0] * tile + row] == tstage[p->eq[q][1] * tile + row];
  return ok;
}

// Per-lane binding sums: 32 bits for LDS tables (a count is below 2^11 and a run covers at most
// one tile of rows, so a wave's sum stays below 2^32), 64 bits for HBM tables.
template <bool kLds>
using LaneSum = typename std::conditional<kLds, uint32_t, uint64_t>::type;

// Adds a run's counts to literals [l0, l1) (several literals of one group can share a key):
// covered counts come from ballots (wave-uniform), binding counts from per-lane sums of cnt
// (positives, negatives) reduced across the wave here.  Lanes 0-3 add the four counters.
template <class Acc>
__device__ inline void flush_counts(const EvalArgs& a, cptr<DevLit> lits, int l0, int l1, uint64_t& cp,
                                    uint64_t& cn, Acc& lbp, Acc& lbn, int lane) {
  if (cp | cn) {  // no covered row means no binding either
    const uint64_t bp = wave_sum(lbp), bn = wave_sum(lbn);
    const uint64_t v = lane == 0 ? cp : lane == 1 ? cn : lane == 2 ? bp : bn;
    if (lane < 4 && v != 0)
      for (int li = l0; li < l1; ++li)
        atomicAdd(a.counts_out + (int64_t)lits[li].slot * 4 + lane, (unsigned long long)v);
  }
  cp = cn = 0;
  lbp = lbn = 0;
}

__device__ inline uint32_t lane_prefix(uint64_t m, int lane) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Join phase of an item: every staged row probes the table of the join column alone.  Rows
// whose join value occurs in the partition are listed per segment in `cidx` (segment-relative
// row numbers); every literal of the unit binds the join column, so the other rows match no
// literal.  The literals keyed on the join column alone (subset 0) are counted here.
// Two single-column probes of an LDS table interleaved, so that their slot and key reads are in
// flight together.  Finished lanes keep reading in-range words.
__device__ inline void lds_probe1_x2(const uint32_t* tbl, uint32_t cap, const int32_t* pkey, int32_t va, int32_t vb,
                                     bool live_a, bool live_b, uint32_t& ca, uint32_t& cb) {
  const uint32_t ha = lds_hash(&va, 1), hb = lds_hash(&vb, 1);
  const uint32_t fa = lds_fp(ha), fb = lds_fp(hb);
  uint32_t sa = lds_slot(ha, cap), sb = lds_slot(hb, cap);
  ca = cb = 0;
  while (live_a || live_b) {
    const uint32_t wa = tbl[sa], wb = tbl[sb];
    const int ra = (int)((wa >> EV_REP_SHIFT) & EV_FIELD_MASK) - 1, rb = (int)((wb >> EV_REP_SHIFT) & EV_FIELD_MASK) - 1;
    const int32_t ka = pkey[ra < 0 ? 0 : ra], kb = pkey[rb < 0 ? 0 : rb];
    const bool hit_a = wa != 0 && (wa & EV_FP_MASK) == fa && ka == va;
    const bool hit_b = wb != 0 && (wb & EV_FP_MASK) == fb && kb == vb;
    if (live_a && hit_a) ca = wa & EV_FIELD_MASK;
    if (live_b && hit_b) cb = wb & EV_FIELD_MASK;
    live_a = live_a && wa != 0 && !hit_a;
    live_b = live_b && wb != 0 && !hit_b;
    sa = (sa + 1) & (cap - 1);
    sb = (sb + 1) & (cap - 1);
  }
}

// Join phase of an item: every staged row probes the table of the join column alone.  Rows
// whose join value occurs in the partition are listed per segment in `cidx` (segment-relative
// row numbers); every literal of the unit binds the join column, so the other rows match no
// literal.  The literals keyed on the join column alone (subset 0) are counted here.  A wave
// takes two 64-row blocks of a segment per step.
template <bool kLds>
__device__ void join_phase(const EvalArgs& a, cptr<DevUnit> U, ItemShared& sh, const int32_t* tstage,
                           const uint8_t* posf, uint16_t* cidx,
                           const uint32_t* tbl, const uint32_t* cnt, const int32_t* keys, uint32_t cap,
                           const int32_t* pstage, int np, int lane, int wave) {
  const int nseg = uni(sh.nseg);
  const int E = uni(sh.blk_start[nseg]);
  const int e0 = (int)((int64_t)E * wave / EV_WAVES), e1 = (int)((int64_t)E * (wave + 1) / EV_WAVES);
  if (e0 >= e1) return;
  const int tile = a.tile_rows;
  cptr<DevLit> lits = as_const(a.lits);
  const int32_t* pkey = pstage + U->jcol * np;
  int sg = 0;
  while (e0 >= uni(sh.blk_start[sg + 1])) ++sg;
  int blk = e0 - uni(sh.blk_start[sg]);
  int seg_rows = uni(sh.seg_rows[sg]), seg_off = uni(sh.seg_off[sg]), nblk = (seg_rows + 63) >> 6;
  int g = uni(sh.seg_ug[sg]);
  int jt = U->g[g].jtcol * tile, l0 = U->g[g].sbeg[0], l1 = U->g[g].sbeg[1];
  uint64_t cp = 0, cn = 0;
  LaneSum<kLds> lbp = 0, lbn = 0;
  for (int e = e0; e < e1;) {
    const bool two = e + 1 < e1 && blk + 1 < nblk;
    const int ra = blk * 64 + lane, rb = ra + 64;
    const bool va_ok = ra < seg_rows, vb_ok = two && rb < seg_rows;
    const int row_a = seg_off + (va_ok ? ra : 0), row_b = seg_off + (vb_ok ? rb : 0);
    const int32_t va = tstage[jt + row_a], vb = tstage[jt + row_b];
    uint32_t ca = 0, cb = 0;
    if (kLds) {
      lds_probe1_x2(tbl, cap, pkey, va, vb, va_ok, vb_ok, ca, cb);
    } else {
      if (va_ok) ca = table_probe(cnt, keys, cap, &va, 1);
      if (vb_ok) cb = table_probe(cnt, keys, cap, &vb, 1);
    }
    const uint64_t ma = __ballot(ca != 0), mb = __ballot(cb != 0);
    if (ma | mb) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&sh.seg_hits[sg], __popcll(ma) + __popcll(mb));
      base = __builtin_amdgcn_readlane(base, 0);
      if (ca != 0) cidx[seg_off + base + (int)lane_prefix(ma, lane)] = (uint16_t)ra;
      if (cb != 0) cidx[seg_off + base + __popcll(ma) + (int)lane_prefix(mb, lane)] = (uint16_t)rb;
      if (l1 > l0) {
        const bool pa = posf[row_a] != 0, pb = posf[row_b] != 0;
        cp += __popcll(__ballot(ca != 0 && pa)) + __popcll(__ballot(cb != 0 && pb));
        cn += __popcll(__ballot(ca != 0 && !pa)) + __popcll(__ballot(cb != 0 && !pb));
        lbp += (pa ? ca : 0u) + (pb ? cb : 0u);
        lbn += (pa ? 0u : ca) + (pb ? 0u : cb);
        if (a.bits != nullptr) {
          const int32_t* tids = sh.seg_tid[sg];
          const int32_t ta = ca != 0 ? tids[row_a - seg_off] : 0, tb = cb != 0 ? tids[row_b - seg_off] : 0;
          for (int li = l0; li < l1; ++li) {
            unsigned long long* bits = a.bits + (int64_t)lits[li].slot * a.bit_words;
            if (ca != 0) atomicOr(&bits[ta >> 6], 1ull << (ta & 63));
            if (cb != 0) atomicOr(&bits[tb >> 6], 1ull << (tb & 63));
          }
        }
      }
    }
    const int step = two ? 2 : 1;
    e += step;
    blk += step;
    if (blk == nblk) {  // the segment's blocks end
      flush_counts(a, lits, l0, l1, cp, cn, lbp, lbn, lane);
      blk = 0;
      if (e < e1) {
        do ++sg;
        while (e >= uni(sh.blk_start[sg + 1]));  // segments of groups the unit lacks have no blocks
        seg_rows = uni(sh.seg_rows[sg]);
        seg_off = uni(sh.seg_off[sg]);
        nblk = (seg_rows + 63) >> 6;
        g = uni(sh.seg_ug[sg]);
        jt = U->g[g].jtcol * tile;
        l0 = U->g[g].sbeg[0];
        l1 = U->g[g].sbeg[1];
      }
    }
  }
  flush_counts(a, lits, l0, l1, cp, cn, lbp, lbn, lane);
}

// Probes subset s's table for this wave's share of the (literal, 64-row block of join-hit
// rows) entries, which run literal-major within each segment.  A literal with a parent mask
// skips blocks where its parent matched no row and probes only rows where it did; a literal
// with its own mask slot records its hit rows for its children.  kLds: packed LDS table `tbl`;
// else the HBM table (cnt, keys).  NK > 0 fixes the key width; NK == 0 reads it from `nk`.
template <int NK, bool kLds>
__device__ void probe_entries(const EvalArgs& a, cptr<DevUnit> U, ItemShared& sh, int s, int nk,
                              const int32_t* tstage, const uint8_t* posf, const uint16_t* cidx, uint64_t* masks, const uint32_t* tbl,
                              const uint32_t* cnt, const int32_t* keys, uint32_t cap, const int32_t* pstage, int np,
                              int lane, int wave) {
  const int nseg = uni(sh.nseg);
  const int E = uni(sh.ent_start[nseg]);
  const int e0 = (int)((int64_t)E * wave / EV_WAVES), e1 = (int)((int64_t)E * (wave + 1) / EV_WAVES);
  if (e0 >= e1) return;
  const int tile = a.tile_rows;
  const int mstride = a.mstride;
  cptr<DevLit> lits = as_const(a.lits);
  int poff[EV_MAXK];
#pragma unroll
  for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
    if (NK || k < nk) poff[k] = U->sub[s].pcol[k] * np;

  // Position of entry e0: segment, literal within the segment, block of hit rows.
  int sg = 0;
  while (e0 >= uni(sh.ent_start[sg + 1])) ++sg;
  int seg_hits = uni(sh.seg_hits[sg]), seg_off = uni(sh.seg_off[sg]), seg_end = uni(sh.ent_start[sg + 1]);
  int nblk = (seg_hits + 63) >> 6;
  int lit = (e0 - uni(sh.ent_start[sg])) / nblk;
  int blk = (e0 - uni(sh.ent_start[sg])) - lit * nblk;
  int lit0 = U->g[uni(sh.seg_ug[sg])].sbeg[s];
  LitRegs L;
  load_literal<NK>(lits, lit0 + lit, nk, tile, L);
  uint64_t cp = 0, cn = 0;
  LaneSum<kLds> lbp = 0, lbn = 0;
  for (int e = e0; e < e1; ++e) {
    const int rloc = blk * 64 + lane;
    uint64_t pm = ~0ull;
    if (L.pmask >= 0) pm = uni64((int64_t)masks[L.pmask * mstride + blk]);
    uint32_t c = 0;
    if (pm != 0) {
      const bool valid = rloc < seg_hits && ((pm >> lane) & 1);
      const int row = seg_off + (rloc < seg_hits ? (int)cidx[seg_off + rloc] : 0);
      if (valid) {
        const bool ok = L.neq == 0 || row_equalities(lits, lit0 + lit, L.neq, tile, tstage, row);
        if (ok) {
          int32_t vals[EV_MAXK];
#pragma unroll
          for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
            if (NK || k < nk) vals[k] = tstage[L.tb[k] + row];
          if (kLds) c = lds_probe<NK>(tbl, cap, pstage, poff, vals, nk);
          else c = table_probe(cnt, keys, cap, vals, NK ? NK : nk);
        }
      }
      if (__ballot(c != 0)) {
        const bool pos = posf[row] != 0;
        cp += __popcll(__ballot(c != 0 && pos));
        cn += __popcll(__ballot(c != 0 && !pos));
        lbp += pos ? c : 0u;
        lbn += pos ? 0u : c;
        if (a.bits != nullptr && c != 0) {
          const int32_t tid = sh.seg_tid[sg][row - seg_off];
          atomicOr(&a.bits[(int64_t)L.slot * a.bit_words + (tid >> 6)], 1ull << (tid & 63));
        }
      }
    }
    if (L.mslot >= 0) {
      const uint64_t hm = __ballot(c != 0);
      if (lane == 0) {
        masks[L.mslot * mstride + blk] = hm;
        if (hm != 0) atomicOr(&sh.need, U->child_subsets[L.mslot]);
      }
    }
    if (++blk == nblk) {  // the literal's run in this segment ends
      blk = 0;
      flush_counts(a, lits, lit0 + lit, lit0 + lit + 1, cp, cn, lbp, lbn, lane);
      ++lit;
      if (e + 1 < e1) {
        if (e + 1 >= seg_end) {
          do ++sg;
          while (e + 1 >= uni(sh.ent_start[sg + 1]));
          lit = 0;
          seg_hits = uni(sh.seg_hits[sg]);
          seg_off = uni(sh.seg_off[sg]);
          seg_end = uni(sh.ent_start[sg + 1]);
          nblk = (seg_hits + 63) >> 6;
          lit0 = U->g[uni(sh.seg_ug[sg])].sbeg[s];
        }
        load_literal<NK>(lits, lit0 + lit, nk, tile, L);
      }
    }
  }
  flush_counts(a, lits, lit0 + lit, lit0 + lit + 1, cp, cn, lbp, lbn, lane);  // no-op after a completed run
}

template <int NK>
__device__ void build_lds_table(uint32_t* tbl, uint32_t cap, const int32_t* pstage, const int* poff, int np, int nk) {
  for (int r = threadIdx.x; r < np; r += EV_THREADS) {
    int32_t vals[EV_MAXK];
#pragma unroll
    for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
      if (NK || k < nk) vals[k] = pstage[poff[k] + r];
    const uint32_t h = lds_hash(vals, NK ? NK : nk);
    const uint32_t fp = lds_fp(h);
    const uint32_t mine = fp | ((uint32_t)(r + 1) << EV_REP_SHIFT);
    uint32_t slot = lds_slot(h, cap);
    while (true) {
      const uint32_t prev = atomicCAS(&tbl[slot], 0u, mine);
      if (prev == 0) break;
      if ((prev & EV_FP_MASK) == fp) {
        const int rep = (int)((prev >> EV_REP_SHIFT) & EV_FIELD_MASK) - 1;
        bool eq = true;
#pragma unroll
        for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
          if (NK || k < nk) eq = eq && (pstage[poff[k] + rep] == vals[k]);
        if (eq) break;
      }
      slot = (slot + 1) & (cap - 1);
    }
    atomicAdd(&tbl[slot], 1u);
  }
}

// One key subset s >= 1 of an item: build its count table (LDS partitions; the table words were
// cleared before the barrier that precedes this call), then probe.
template <int NK>
__device__ inline void eval_subset(const EvalArgs& a, cptr<DevUnit> U, ItemShared& sh, int s, int nk, int hq,
                                   uint32_t* tbl, const int32_t* tstage, const uint8_t* posf, const uint16_t* cidx,
                                   uint64_t* masks,
                                   const int32_t* pstage, int np, int lane, int wave) {
  if (hq < 0) {
    const uint32_t cap = EV_LDS_CAP;
    int poff[EV_MAXK];
#pragma unroll
    for (int k = 0; k < (NK ? NK : EV_MAXK); ++k)
      if (NK || k < nk) poff[k] = U->sub[s].pcol[k] * np;
    build_lds_table<NK>(tbl, cap, pstage, poff, np, nk);
    __syncthreads();  // table ready
    probe_entries<NK, true>(a, U, sh, s, nk, tstage, posf, cidx, masks, tbl, nullptr, nullptr, cap, pstage, np, lane,
                            wave);
  } else {
    const uint32_t cap = table_cap_hbm(np);
    const uint32_t* base = U->sub[s].hbuf + U->sub[s].hbase[hq];
    probe_entries<NK, false>(a, U, sh, s, nk, tstage, posf, cidx, masks, nullptr, base + cap,
                             reinterpret_cast<const int32_t*>(base + 2 * (int64_t)cap), cap, pstage, np, lane, wave);
  }
}

// Exclusive prefix sums over lanes 0-15, the segments of an item (at most EV_MAXG of them), by
// DPP shifts within the wave's first row of 16 lanes (row_shr 1, 2, 4, 8: no LDS round trips);
// lane 15 ends with its exclusive sum, so lane 15's value plus its own is the total.  Lanes 16-63
// hold sums over their own rows, which callers ignore.  Every lane of the wave must be active.
static_assert(EV_MAXG <= 16, "segments fit one DPP row");
__device__ inline int row_excl_prefix(int v) {
  int x = v;
  x += row_shr<1>(x);
  x += row_shr<2>(x);
  x += row_shr<4>(x);
  x += row_shr<8>(x);
  return x - v;
}
template <int kShr>
__device__ inline int64_t row_shr64(int64_t x) {
  const uint32_t lo = (uint32_t)row_shr<kShr>((int32_t)(uint32_t)x);
  const uint32_t hi = (uint32_t)row_shr<kShr>((int32_t)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ inline int64_t row_excl_prefix64(int64_t v) {
  int64_t x = v;
  x += row_shr64<1>(x);
  x += row_shr64<2>(x);
  x += row_shr64<4>(x);
  x += row_shr64<8>(x);
  return x - v;
}

// Facts of an LDS partition, prefetched into registers while the previous unit is evaluated:
// thread t holds flattened words t + EV_THREADS * j of the staged layout [column][row].
#ifndef QF_EV_PF
#define QF_EV_PF 3
#endif
constexpr int EV_PF = QF_EV_PF;  // fact words per thread carried in registers to the next unit
struct FactPrefetch {
  int32_t v[EV_PF];
  int unit = -1;
  int64_t q = -1;
};

// Register j of thread t carries fact row t of column j (columns below EV_PF, rows below
// EV_THREADS); the other words are loaded when the unit starts.
__device__ inline void prefetch_facts(cptr<DevUnit> U, int unit, int64_t q, int64_t p0, int np, FactPrefetch& pf) {
  const int arity = U->arity;
  if ((int)threadIdx.x < np) {
#pragma unroll
    for (int j = 0; j < EV_PF; ++j)
      if (j < arity) pf.v[j] = U->pcols[j][p0 + threadIdx.x];
  }
  pf.unit = unit;
  pf.q = q;
}

// 256-thread blocks: a block's barriers wait for four waves only, which measured faster than
// eight-wave blocks at the same LDS per CU.  At the tiles the autotuner picks for wide steps
// (768 rows) LDS holds five blocks per CU, so the attribute asks for five waves per SIMD: that
// leaves up to 96 VGPRs per lane, which the kernel fits without spilling to scratch (at six, 80
// VGPRs, it spilled and measured 1 % slower).
#ifndef QF_EV_WAVES_PER_EU
#define QF_EV_WAVES_PER_EU 5
#endif
// kBits: covered bitvectors were asked for (a.bits); the instance without them compiles their
// code out of the probe loops.
template <bool kBits>
__global__ __launch_bounds__(EV_THREADS) __attribute__((amdgpu_waves_per_eu(QF_EV_WAVES_PER_EU))) void k_eval(
    EvalArgs a) {
  if (!kBits) a.bits = nullptr;
  extern __shared__ uint32_t lds[];
  uint32_t* tbl = lds;                                                                 // EV_TBL_WORDS
  int32_t* pstage = reinterpret_cast<int32_t*>(lds + EV_TBL_WORDS);                   // EV_PST_WORDS
  int32_t* tstage = reinterpret_cast<int32_t*>(lds + EV_TBL_WORDS + EV_PST_WORDS);   // ntc x tile
  uint8_t* posf = reinterpret_cast<uint8_t*>(tstage + a.ntc * a.tile_rows);          // tile bytes
  uint16_t* cidx = reinterpret_cast<uint16_t*>(lds + a.cidx_off);                    // tile
  uint64_t* masks = reinterpret_cast<uint64_t*>(lds + a.mask_off);                   // slots x mstride
  __shared__ ItemShared sh;

  const int lane = threadIdx.x & 63;
  const int wave = uni(threadIdx.x >> 6);
  const int64_t total = uni64(*a.num_items);
  const int ntc = a.ntc;
  cptr<DevItem> items = as_const(a.items);
  FactPrefetch pf;
  // This block's copy of the counters: blocks spread their atomics over separate copies, so that
  // few of them queue on one address.
  a.counts_out += (int64_t)(blockIdx.x % a.ncopies) * a.copy_words;

  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int item_batch = items[it].batch, item_chunk = items[it].chunk;
    const int64_t q = items[it].q;
    cptr<DevBatch> B = as_const(a.batches) + item_batch;

    __syncthreads();  // the previous item is done with LDS
    // Segments: the chunk's rows over the batch's join variables, in group order (wave 0, one
    // lane per group).
    if (wave == 0) {
      const int64_t c0 = (int64_t)item_chunk * a.tile_rows, c1 = c0 + a.tile_rows;
      const int g = lane;
      const bool ok = g < B->ngroups;
      int64_t t0 = 0, t1 = 0;
      if (ok) {
        t0 = B->toff[g][q << B->shift];
        t1 = B->toff[g][(q + 1) << B->shift];
      }
      const int64_t run = row_excl_prefix64(t1 - t0);
      const int64_t lo_ = max(run, c0), hi_ = min(run + (t1 - t0), c1);
      const int n = (ok && hi_ > lo_) ? (int)(hi_ - lo_) : 0;
      const uint64_t has = __ballot(n > 0);
      const int idx = __popcll(has & ((1ull << lane) - 1ull));
      const int off = row_excl_prefix(n);
      if (n > 0) {
        sh.seg_group[idx] = g;
        sh.seg_off[idx] = off;
        sh.seg_rows[idx] = n;
        sh.seg_src[idx] = t0 + (lo_ - run);
        sh.seg_tid[idx] = B->tcols[g][ntc] + (t0 + (lo_ - run));
      }
      if (lane == 15) {
        sh.seg_off[__popcll(has)] = off + n;
        sh.nseg = __popcll(has);
      }
    } else if (wave == 1) {  // the units' partitions at q, one lane per unit
      if (lane < B->nunits) {
        cptr<DevUnit> V = as_const(a.units) + B->unit[lane];
        const int64_t v0 = V->poff[q];
        sh.u_p0[lane] = v0;
        sh.u_np[lane] = (int)(V->poff[q + 1] - v0);
        sh.u_hq[lane] = V->hq[q];
      }
    } else if (wave == 2) {  // the units' group maps, one lane per (unit, batch group)
      for (int l = lane; l < B->nunits * B->ngroups; l += 64) {
        const int k = l / B->ngroups, g = l - k * B->ngroups;
        sh.u_ug[k][g] = (int8_t)(as_const(a.units) + B->unit[k])->ug_of_bg[g];
      }
    }
    __syncthreads();
    const int nseg = uni(sh.nseg);
    // Stage the binding rows of every segment (same staged columns in every group).  Word w of a
    // segment's ntc + 1 columns of n rows (column-major; the last column is the row id, staged as
    // a positive flag: a row is positive below num_positive) goes to thread w % EV_THREADS, which
    // takes EV_STAGE words at a time so that their loads are in flight together.  Lane c of every
    // wave holds column c's source address; the rounds are block-uniform, so every lane takes
    // part in the shuffles that hand the addresses out.
    for (int sgi = 0; sgi < nseg; ++sgi) {
      const int g = uni(sh.seg_group[sgi]);
      const int64_t src0 = uni64(sh.seg_src[sgi]);
      const int n = uni(sh.seg_rows[sgi]), dst0 = uni(sh.seg_off[sgi]);
      const uint64_t colp = lane <= ntc ? (uint64_t)(uintptr_t)(B->tcols[g][lane] + src0) : 0;
      const int words = (ntc + 1) * n;
      for (int w0 = 0; w0 < words; w0 += EV_STAGE * EV_THREADS) {
        int32_t v[EV_STAGE];
        int at[EV_STAGE];
#pragma unroll
        for (int j = 0; j < EV_STAGE; ++j) {
          const int w = w0 + (int)threadIdx.x + j * EV_THREADS;
          const bool ok = w < words;
          const int c = ok ? w / n : 0;
          const int i = ok ? w - c * n : 0;  // words past the end read row 0 of column 0
          const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)colp, c, 64);
          const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(colp >> 32), c, 64);
          v[j] = reinterpret_cast<const int32_t*>((uintptr_t)(((uint64_t)hi << 32) | lo))[i];
          at[j] = !ok ? -1 : c < ntc ? c * a.tile_rows + dst0 + i : -2 - (dst0 + i);
        }
#pragma unroll
        for (int j = 0; j < EV_STAGE; ++j) {
          if (at[j] >= 0) tstage[at[j]] = v[j];
          else if (at[j] <= -2) posf[-2 - at[j]] = v[j] < a.num_positive ? 1 : 0;
        }
      }
    }

   for (int k = 0; k < B->nunits; ++k) {
    cptr<DevUnit> U = as_const(a.units) + B->unit[k];
    const int64_t p0 = uni64(sh.u_p0[k]);
    const int np = uni(sh.u_np[k]);
    if (np == 0) continue;
    const int hq = uni(sh.u_hq[k]);
    __syncthreads();  // staged rows ready; the previous unit is done with the table, facts, lists and masks
    if (wave == 0) {  // segment map of this unit, one lane per segment
      const bool ok = lane < nseg;
      int ug = -1, blocks = 0;
      if (ok) {
        ug = sh.u_ug[k][sh.seg_group[lane]];
        blocks = ug >= 0 ? (sh.seg_rows[lane] + 63) >> 6 : 0;
        sh.seg_ug[lane] = ug;
        sh.seg_hits[lane] = 0;
      }
      const int bs = row_excl_prefix(blocks);
      if (ok) sh.blk_start[lane] = bs;
      if (lane == 15) sh.blk_start[nseg] = bs + blocks;
      if (lane == 0) sh.need = 0;
    }
    // The partition's facts (all relation columns) for an LDS table, from the registers the
    // previous unit prefetched them into when it could.
    if (hq < 0) {
      const int unit = B->unit[k];
      if (pf.unit == unit && pf.q == q) {
        const int arity = U->arity;
        if ((int)threadIdx.x < np) {
#pragma unroll
          for (int j = 0; j < EV_PF; ++j)
            if (j < arity) pstage[j * np + threadIdx.x] = pf.v[j];
        }
        for (int c = 0; c < arity; ++c) {
          const int32_t* src = U->pcols[c] + p0;
          for (int i = threadIdx.x + (c < EV_PF ? EV_THREADS : 0); i < np; i += EV_THREADS) pstage[c * np + i] = src[i];
        }
      } else {
        for (int c = 0; c < U->arity; ++c) {
          const int32_t* src = U->pcols[c] + p0;
          for (int i = threadIdx.x; i < np; i += EV_THREADS) pstage[c * np + i] = src[i];
        }
      }
      clear_lds_table(tbl);
    }
    // Start loading the next unit's facts; they arrive while this unit is evaluated.
    {
      int nk2 = -1;
      for (int j = k + 1; j < B->nunits && nk2 < 0; ++j)
        if (uni(sh.u_np[j]) > 0 && uni(sh.u_hq[j]) < 0) nk2 = j;
      if (nk2 >= 0) {
        cptr<DevUnit> V = as_const(a.units) + B->unit[nk2];
        prefetch_facts(V, B->unit[nk2], q, uni64(sh.u_p0[nk2]), uni(sh.u_np[nk2]), pf);
      }
    }
    __syncthreads();  // segment map, facts and cleared table ready

    // Join phase (subset 0: the join column alone).
    if (hq < 0) {
      const uint32_t cap = EV_LDS_CAP;
      const int poff = U->jcol * np;
      build_lds_table<1>(tbl, cap, pstage, &poff, np, 1);
      __syncthreads();  // table ready
      join_phase<true>(a, U, sh, tstage, posf, cidx, tbl, nullptr, nullptr, cap, pstage, np, lane, wave);
    } else {
      const uint32_t cap = table_cap_hbm(np);
      const uint32_t* base = U->sub[0].hbuf + U->sub[0].hbase[hq];
      join_phase<false>(a, U, sh, tstage, posf, cidx, nullptr, base + cap,
                        reinterpret_cast<const int32_t*>(base + 2 * (int64_t)cap), cap, pstage, np, lane, wave);
    }

    for (int s = 1; s < U->nsub; ++s) {
      __syncthreads();  // hit lists / previous subset's probes, masks and need bits done
      // Whether any literal of the subset can match: one without a parent, or one whose parent
      // matched a row of this item (a parent's subset has fewer columns, so it ran earlier).
      if (!(((U->orphan | uni((int)sh.need)) >> s) & 1u)) continue;
      if (wave == 0) {  // entry list, one lane per segment
        const bool ok = lane < nseg;
        const int g = ok ? sh.seg_ug[lane] : -1;
        const int e = g >= 0 ? (U->g[g].sbeg[s + 1] - U->g[g].sbeg[s]) * ((sh.seg_hits[lane] + 63) >> 6) : 0;
        const int es = row_excl_prefix(e);
        if (ok) sh.ent_start[lane] = es;
        if (lane == 15) sh.ent_start[nseg] = es + e;
      }
      if (hq < 0)
        clear_lds_table(tbl);
      __syncthreads();  // entry list and cleared table ready
      if (uni(sh.ent_start[nseg]) == 0) continue;
      const int nk = U->sub[s].nk;
      switch (nk) {
        case 1: eval_subset<1>(a, U, sh, s, nk, hq, tbl, tstage, posf, cidx, masks, pstage, np, lane, wave); break;
        case 2: eval_subset<2>(a, U, sh, s, nk, hq, tbl, tstage, posf, cidx, masks, pstage, np, lane, wave); break;
        case 3: eval_subset<3>(a, U, sh, s, nk, hq, tbl, tstage, posf, cidx, masks, pstage, np, lane, wave); break;
        default: eval_subset<0>(a, U, sh, s, nk, hq, tbl, tstage, posf, cidx, masks, pstage, np, lane, wave); break;
      }
    }
   }
  }
}

// Sums the blocks' counter copies: out[w] = sum over c of copies[c * words + w].  A block takes
// 64 words; its 16 waves split the copies (loads coalesced over the words), then add in LDS.
constexpr int SUM_WORDS = 64, SUM_GROUPS = 16;
__global__ __launch_bounds__(SUM_WORDS * SUM_GROUPS) void k_sum_copies(const uint64_t* copies, int ncopies,
                                                                     int64_t words, uint64_t* out) {
  __shared__ uint64_t part[SUM_GROUPS][SUM_WORDS];
  const int x = threadIdx.x & (SUM_WORDS - 1), g = threadIdx.x / SUM_WORDS;
  const int64_t w = (int64_t)blockIdx.x * SUM_WORDS + x;
  uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (w < words) {
    int c = g;
    for (; c + 3 * SUM_GROUPS < ncopies; c += 4 * SUM_GROUPS) {
      s0 += copies[(int64_t)c * words + w];
      s1 += copies[(int64_t)(c + SUM_GROUPS) * words + w];
      s2 += copies[(int64_t)(c + 2 * SUM_GROUPS) * words + w];
      s3 += copies[(int64_t)(c + 3 * SUM_GROUPS) * words + w];
    }
    for (; c < ncopies; c += SUM_GROUPS) s0 += copies[(int64_t)c * words + w];
  }
  part[g][x] = s0 + s1 + s2 + s3;
  __syncthreads();
  if (g == 0 && w < words) {
    uint64_t sum = 0;
    for (int k = 0; k < SUM_GROUPS; ++k) sum += part[k][x];
    out[w] = sum;
  }
}

// ------------------------------------------------------------------------------------------
// Host side.

struct LitSpec {
  int slot;
  std::vector<int> pcols;                // distinct relation columns of the key (sorted)
  std::vector<int> bcols;                // binding column per key position
  std::vector<std::pair<int, int>> eqs;  // binding columns that must be equal (repeated columns)
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
  const size_t n = ls.pcols.size();
  std::pair<int, int> small[16];
  std::vector<std::pair<int, int>> large;
  std::pair<int, int>* kv = small;
  if (n > 16) {
    large.resize(n);
    kv = large.data();
  }
  for (size_t i = 0; i < n; ++i) kv[i] = std::make_pair(ls.pcols[i], ls.bcols[i]);
  std::sort(kv, kv + n);
  ls.pcols.clear();
  ls.bcols.clear();
  for (size_t i = 0; i < n; ++i) {
    const std::pair<int, int>& p = kv[i];
    if (!ls.pcols.empty() && ls.pcols.back() == p.first) {
      if (ls.bcols.back() != p.second) ls.eqs.emplace_back(ls.bcols.back(), p.second);
      continue;
    }
    ls.pcols.push_back(p.first);
    ls.bcols.push_back(p.second);
  }
  QF_REQUIRE((int)ls.eqs.size() <= EV_MAXEQ, "too many repeated columns in one literal");
}

static int auto_radix_bits(int64_t rows, int target) {
  // About `target` facts per partition (default 256): most partitions fit the LDS tables.
  int r = 0;
  while (r < 16 && (rows >> r) > target) ++r;
  return r;
}

// Autotuner state: best (partition_rows, tile_rows, streamed form on) per workload shape.
struct TuneChoice {
  int partition_rows = 0, tile_rows = 0;
  bool stream = true;
  double ms = 0.0;
};
struct TuneCache {
  std::map<std::string, TuneChoice> best;
  TuneChoice last;
};

// HBM count tables of the heavy partitions of one (relation, join column) for one subset.
struct HeavyTables {
  DevBuf buf;
  DevBuf base;  // int64 per heavy ordinal, then the build and finalize job lists (kept with the
                // table, so that nothing waits for the build kernels)
};

// A relation partition's distinct key tuples for one key subset (join column first), cell by
// cell, with their multiplicities (streamed units).
struct TupleList {
  std::vector<DevBuf> cols;  // int32 per tuple and key column
  DevBuf cnt;                // int32 per tuple: facts with the tuple
  DevBuf off;                // int64 per cell + 1: first tuple
  int64_t tuples = 0;
};

// A relation partitioned on one join column, with its heavy partitions and their tables.
struct RelPart {
  std::shared_ptr<PartitionedTable> pt;
  int join_col = 0;
  DevBuf hq;                     // int32 per partition
  std::vector<int64_t> heavy_q;  // heavy partitions
  std::map<std::vector<int>, std::shared_ptr<HeavyTables>> tables;
  std::map<std::vector<int>, std::shared_ptr<TupleList>> tuples;
};

// A streamed launch's cell-by-list matrix and per-cell parts (k_ts_cells): they depend on the
// launch's key-tuple lists only.
struct CellMat {
  DevBuf parts, cellmat;
};

struct RelCache {
  // Key: rows, join column, radix bits, then every column address.
  std::map<std::vector<int64_t>, std::shared_ptr<RelPart>> parts;
  // Cell matrices of recent streamed launches, most recent first, keyed by cells, run lists and
  // every list's per-cell offsets array (arrays of the cached key-tuple lists, which live as long
  // as this cache).
  std::list<std::pair<std::vector<int64_t>, std::shared_ptr<CellMat>>> cellmats;
};
constexpr size_t EV_CELLMAT_CACHE = 4;

static RelCache& rel_cache(Context* ctx) {
  if (!ctx->eval_cache) ctx->eval_cache = std::make_shared<RelCache>();
  return *std::static_pointer_cast<RelCache>(ctx->eval_cache);
}

static std::shared_ptr<RelPart> partition_relation(Context* ctx, const qf_table& rel, int join_col, int R,
                                                   bool cache) {
  std::vector<int64_t> key{rel.num_rows, join_col, R};
  for (int c = 0; c < rel.num_columns; ++c) key.push_back((int64_t)(uintptr_t)rel.columns[c]);
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
  rp->join_col = join_col;
  // Heavy: the partition's facts or its LDS table do not fit.
  const int64_t parts = (int64_t)1 << R;
  std::vector<int32_t> hq(parts, -1);
  for (int64_t q = 0; q < parts; ++q) {
    const int64_t np = pt->h_offsets[q + 1] - pt->h_offsets[q];
    if (np > 0 && (np > EV_LDS_MAX_FACTS || np * rel.num_columns > EV_PST_WORDS)) {
      hq[q] = (int32_t)rp->heavy_q.size();
      rp->heavy_q.push_back(q);
    }
  }
  rp->hq = DevBuf(ctx, parts * 4);
  copy_h2d(ctx, rp->hq.get(), hq.data(), parts * 4);
  if (cache) rc.parts[key] = rp;
  return rp;
}

// The partition's rows sorted by the key columns (each by all 32 bits of its mix32; the join
// column, key[0], last, so that its top radix bits keep every cell's rows in place) and cut into
// distinct tuples; made once per (relation partition, key) and cached with the partition.
static std::shared_ptr<TupleList> relation_tuples(Context* ctx, RelPart& rp, const std::vector<int>& key) {
  auto it = rp.tuples.find(key);
  if (it != rp.tuples.end()) return it->second;
  PhaseScope ps(ctx, "relation_tuples");
  const int m = (int)key.size();
  QF_REQUIRE(m >= 1 && m <= EV_MAXK && key[0] == rp.join_col, "bad tuple key");
  const int64_t n = rp.pt->num_rows;
  const int64_t parts = (int64_t)1 << rp.pt->radix_bits;
  auto tl = std::make_shared<TupleList>();
  QF_REQUIRE(n < ((int64_t)1 << 31), "relation rows must stay below 2^31");

  std::vector<DevBuf> buf[2];
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < m; ++k) buf[s].emplace_back(ctx, std::max<int64_t>(1, n) * 4);
  std::vector<const int32_t*> cur(m);
  for (int k = 0; k < m; ++k) cur[k] = rp.pt->columns[key[k]].as<int32_t>();
  if (n > 0) {
    int side = 0;
    for (int k = m - 1; k >= 0; --k) {  // stable LSD passes, the join column last
      std::vector<int32_t*> out(m);
      for (int c = 0; c < m; ++c) out[c] = buf[side][c].as<int32_t>();
      stable_partition(ctx, cur.data(), m, k, false, n, 1, 32, out.data(), 1, nullptr);
      for (int c = 0; c < m; ++c) cur[c] = out[c];
      side ^= 1;
    }
  }
  TsCols sc{};
  sc.m = m;
  for (int k = 0; k < m; ++k) sc.c[k] = cur[k];
  DevBuf flags(ctx, std::max<int64_t>(1, n) * 8), starts(ctx, std::max<int64_t>(1, n) * 8);
  int64_t tuples = 0;
  if (n > 0) {
    hipLaunchKernelGGL(k_tuple_heads, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, sc, n,
                       flags.as<int64_t>());
    QF_CHECK_HIP(hipGetLastError());
    scan_exclusive_i64(ctx, flags.as<int64_t>(), starts.as<int64_t>(), n);
    int64_t last[2];
    copy_d2h(ctx, &last[0], starts.as<int64_t>() + n - 1, 8);
    copy_d2h(ctx, &last[1], flags.as<int64_t>() + n - 1, 8);
    tuples = last[0] + last[1];
  }
  TsOut out{};
  for (int k = 0; k < m; ++k) {
    tl->cols.emplace_back(ctx, std::max<int64_t>(1, tuples) * 4);
    out.c[k] = tl->cols.back().as<int32_t>();
  }
  tl->cnt = DevBuf(ctx, std::max<int64_t>(1, tuples) * 4);
  tl->off = DevBuf(ctx, (parts + 1) * 8);
  tl->tuples = tuples;
  DevBuf first(ctx, std::max<int64_t>(1, tuples) * 4);
  hipLaunchKernelGGL(k_tuple_compact, dim3((unsigned)ceil_div(std::max(n, parts + 1), 256)), dim3(256), 0,
                     ctx->stream, sc, n, flags.as<int64_t>(), starts.as<int64_t>(), out, first.as<int32_t>(),
                     rp.pt->offsets.as<int64_t>(), parts, tl->off.as<int64_t>());
  QF_CHECK_HIP(hipGetLastError());
  if (tuples > 0) {
    hipLaunchKernelGGL(k_tuple_counts, dim3((unsigned)ceil_div(tuples, 256)), dim3(256), 0, ctx->stream, out, m,
                       first.as<int32_t>(), tuples, n, tl->cnt.as<int32_t>());
    QF_CHECK_HIP(hipGetLastError());
  }
  rp.tuples[key] = tl;
  return tl;
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
  static_assert(sizeof(HeavyJob) % 8 == 0, "job lists follow the int64 offsets");
  const size_t base_bytes = std::max<size_t>(1, base.size()) * 8;
  ht->base = DevBuf(ctx, base_bytes + (build.size() + fin.size()) * sizeof(HeavyJob));
  if (!base.empty()) {
    PhaseScope ps(ctx, "heavy_tables");
    // Offsets and both job lists in one upload.
    std::vector<uint8_t> host(base_bytes + (build.size() + fin.size()) * sizeof(HeavyJob));
    std::memcpy(host.data(), base.data(), base.size() * 8);
    std::memcpy(host.data() + base_bytes, build.data(), build.size() * sizeof(HeavyJob));
    std::memcpy(host.data() + base_bytes + build.size() * sizeof(HeavyJob), fin.data(), fin.size() * sizeof(HeavyJob));
    copy_h2d(ctx, ht->base.get(), host.data(), host.size());
    const HeavyJob* dbuild = reinterpret_cast<const HeavyJob*>(ht->base.as<uint8_t>() + base_bytes);
    const HeavyJob* dfin = dbuild + build.size();
    fill_i32(ctx, ht->buf.as<int32_t>(), 0, words);
    DevRel rel{};
    for (int c = 0; c < rp.pt->num_columns; ++c) rel.pcols[c] = rp.pt->columns[c].as<int32_t>();
    rel.poff = rp.pt->offsets.as<int64_t>();
    DevSubset sub{};
    sub.nk = nk;
    for (int k = 0; k < nk; ++k) sub.pcol[k] = pcols[k];
    hipLaunchKernelGGL(k_heavy_build, dim3((unsigned)build.size()), dim3(EV_THREADS), 0, ctx->stream, rel, sub,
                       dbuild, ht->buf.as<uint32_t>());
    QF_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_heavy_finalize, dim3((unsigned)fin.size()), dim3(EV_THREADS), 0, ctx->stream, rel, sub,
                       dfin, ht->buf.as<uint32_t>());
    QF_CHECK_HIP(hipGetLastError());
  }
  rp.tables[pcols] = ht;
  return ht;
}

// Binding table partitioned on column `var`, carrying `cols` (must include var) and row ids, into
// the caller's buffers (ncols + 1 columns of rows int32, offsets 2^R + 1 int64).
static void partition_binding(Context* ctx, const qf_table& b, int var, const std::vector<int>& cols, int R,
                              const std::vector<int32_t*>& out_cols, int64_t* d_offsets, Scratch* scratch) {
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
    std::vector<int32_t*> out{out_cols[key_index]};
    std::vector<int> batch;
    while (pos < others.size() && batch.size() < 6) batch.push_back(others[pos++]);
    for (int c : batch) {
      in.push_back(b.columns[c]);
      out.push_back(out_cols[std::find(cols.begin(), cols.end(), c) - cols.begin()]);
    }
    if (first) out.push_back(out_cols[cols.size()]);
    if (scratch) scratch->used = 0;  // the previous batch's kernels precede on this stream
    stable_partition(ctx, in.data(), (int)in.size(), 0, first, b.num_rows, 1, R, out.data(), 1,
                     first ? d_offsets : nullptr, scratch);
    first = false;
  } while (pos < others.size());
}

// Pinned host staging of a context.  A plan's small uploads (descriptors, lists, targets) are
// copied into it and sent with asynchronous copies, without a stream synchronisation each; the
// next plan waits for them (an event) before reusing the space.  An upload that does not fit goes
// synchronously, and the region grows before the next plan.  `counts` receives the counts of
// steps run without a graph (a captured step keeps its own).
struct HostStage {
  char* base = nullptr;
  size_t cap = 0, used = 0, want = 0;
  hipEvent_t done = nullptr, before = nullptr;
  hipStream_t stream = nullptr;  // the uploads' stream (default priority)
  uint64_t releases = 0;         // pool releases the upload stream is ordered after
  bool pending = false;
  uint64_t* counts = nullptr;
  size_t counts_cap = 0;
  ~HostStage() {
    if (pending && done) (void)hipEventSynchronize(done);
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    if (done) (void)hipEventDestroy(done);
    if (before) (void)hipEventDestroy(before);
    if (base) (void)hipHostFree(base);
    if (counts) (void)hipHostFree(counts);
  }
};

static HostStage& host_stage(Context* ctx) {
  if (!ctx->host_stage) ctx->host_stage = std::make_shared<HostStage>();
  return *std::static_pointer_cast<HostStage>(ctx->host_stage);
}

// The uploads run on a stream of their own, after the work queued on the context stream before
// the plan began (begin_uploads) and before the work queued after it ended (end_uploads) or
// after uploads_before, so that they overlap a binding partition queued in between.

// Starts a plan's uploads: the previous plan's copies are done with the region.
static void begin_uploads(Context* ctx) {
  HostStage& H = host_stage(ctx);
  if (!H.stream) QF_CHECK_HIP(hipStreamCreateWithFlags(&H.stream, hipStreamNonBlocking));
  if (!H.done) QF_CHECK_HIP(hipEventCreateWithFlags(&H.done, hipEventDisableTiming));
  if (!H.before) QF_CHECK_HIP(hipEventCreateWithFlags(&H.before, hipEventDisableTiming));
  if (H.pending) {
    QF_CHECK_HIP(hipEventSynchronize(H.done));
    H.pending = false;
  } else if (H.used > 0) {
    QF_CHECK_HIP(hipStreamSynchronize(H.stream));  // a plan that stopped early left copies without the event
  }
  QF_CHECK_HIP(hipEventRecord(H.before, ctx->stream));
  QF_CHECK_HIP(hipStreamWaitEvent(H.stream, H.before, 0));
  H.releases = ctx->pool->releases();
  if (H.want > H.cap) {
    if (H.base) QF_CHECK_HIP(hipHostFree(H.base));
    H.base = nullptr;
    H.cap = 0;
    const size_t cap = std::max<size_t>(H.want * 2, (size_t)1 << 20);
    QF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&H.base), cap, hipHostMallocDefault));
    H.cap = cap;
  }
  H.used = 0;
  H.want = 0;
}

// Host-to-device copy of a plan upload on the upload stream (see HostStage).  The destination
// may be a pool block released since the upload stream last waited for the context stream (the
// pool hands blocks to later work of the same stream without a wait): then it waits again.
static void upload(Context* ctx, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  HostStage& H = host_stage(ctx);
  if (ctx->pool->releases() != H.releases) {
    QF_CHECK_HIP(hipEventRecord(H.before, ctx->stream));
    QF_CHECK_HIP(hipStreamWaitEvent(H.stream, H.before, 0));
    H.releases = ctx->pool->releases();
  }
  const size_t at = (H.used + 255) & ~(size_t)255;
  if (H.base && at + bytes <= H.cap) {
    std::memcpy(H.base + at, src, bytes);
    QF_CHECK_HIP(hipMemcpyAsync(dst, H.base + at, bytes, hipMemcpyHostToDevice, H.stream));
    H.used = at + bytes;
  } else {
    H.want = std::max(H.want, at + bytes);
    QF_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, H.stream));
    QF_CHECK_HIP(hipStreamSynchronize(H.stream));
  }
}

// The context stream's later work waits for the uploads so far.
static void uploads_before(Context* ctx) {
  HostStage& H = host_stage(ctx);
  QF_CHECK_HIP(hipEventRecord(H.done, H.stream));
  QF_CHECK_HIP(hipStreamWaitEvent(ctx->stream, H.done, 0));
}

// Ends a plan's uploads: the context stream's later work waits for them; the event also marks
// when the region may be reused.
static void end_uploads(Context* ctx) {
  HostStage& H = host_stage(ctx);
  QF_CHECK_HIP(hipEventRecord(H.done, H.stream));
  QF_CHECK_HIP(hipStreamWaitEvent(ctx->stream, H.done, 0));
  H.pending = true;
}

// Pinned counts for a step run without a graph (the call synchronises before it returns).
static uint64_t* stage_counts(Context* ctx, size_t bytes) {
  HostStage& H = host_stage(ctx);
  if (bytes > H.counts_cap) {
    if (H.counts) QF_CHECK_HIP(hipHostFree(H.counts));
    H.counts = nullptr;
    H.counts_cap = 0;
    QF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&H.counts), bytes, hipHostMallocDefault));
    H.counts_cap = bytes;
  }
  return H.counts;
}

// Everything one evaluation call needs on the device; kept alive by a cached graph.
struct StepPlan {
  int R_T = 0;
  int tile = 1024;
  std::vector<int> bcols;                  // staged binding columns
  std::vector<int> join_vars;
  std::vector<std::vector<DevBuf>> tparts;  // per join variable: staged columns + row id
  std::vector<DevBuf> toffs;
  std::vector<std::vector<int>> pgroups;    // join variables partitioned together (indices)
  std::map<std::pair<int, int>, int> unit_R;  // radix bits per (relation, join column)
  bool partitioned = false;                 // the binding partition is already enqueued
  std::vector<DevBuf> tscratch;             // per group: its partition's temporaries
  std::vector<std::shared_ptr<RelPart>> keep_rel;
  std::vector<std::shared_ptr<HeavyTables>> keep_heavy;
  std::vector<DevUnit> units;
  std::vector<char> unit_streamed;          // per unit: counted by k_stream in light cells, k_eval in the rest
  std::vector<std::shared_ptr<TupleList>> keep_tuples;
  struct StreamLaunch {
    DevBuf d_launch, d_lists, d_targets;
    std::shared_ptr<CellMat> cm;  // per cell: parts; per cell and list: (first entry, entries)
    int64_t items_bound = 0;
  };
  std::vector<StreamLaunch> slaunches;      // units of the streamed form, in launches of bounded lists
  DevBuf d_ts_counts, d_ts_starts, d_ts_items, d_ts_num, d_light;
  int64_t ts_items = 0;                     // upper bound of any streamed launch's work items
  size_t ts_lds = 0;
  int ts_blocks = 0;
  std::vector<DevBatch> batches;
  std::vector<DevLit> lits;
  std::vector<std::pair<int, int>> launches;  // unit ranges launched together
  std::vector<int64_t> unit_cells;            // per launch-local prefix, concatenated per launch
  std::vector<int64_t> launch_items;          // per launch: upper bound of its work items
  DevBuf d_units, d_batches, d_lits, d_unit_cells, d_cell_counts, d_cell_starts, d_items, d_num_items, d_counts;
  DevBuf d_copies;  // ncopies x num_slots x 4 partial counters, summed into d_counts after the launches
  int ncopies = 1;
  int64_t max_cells = 0, max_items = 0;
  size_t lds_bytes = 0;
  int max_blocks = 0;
  int max_mask = 0;
  int cidx_off = 0, mask_off = 0, mstride = 0;
  int num_slots = 0;
  int64_t bit_words = 0;
  uint64_t* pinned = nullptr;
  bool own_pinned = false;   // a captured step's own pinned counts (else the context's)
  hipGraphExec_t graph = nullptr;
  std::vector<DevBuf> held;  // pool blocks a captured graph uses
  ~StepPlan() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (pinned && own_pinned) (void)hipHostFree(pinned);
  }
};

// Units: (relation, join column) with a group per join variable (batched), one group per
// literal-group (plain group), or one literal each (plain literal).
// stream: the batched form without covered bitvectors, whose units k_stream counts from their
// relations' distinct key tuples in the cells whose binding rows fit its work items.
// The plan in two parts: plan_binding fixes what the binding partition needs (radix bits,
// staged columns, join variables, the partitions' buffers), so that an uncaptured step can start
// the partition while plan_units builds the rest.
static void plan_binding(Context* ctx, const qf_table& binding, const qf_table* relations,
                         const std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o, StepPlan& P) {
  begin_uploads(ctx);
  P.num_slots = num_slots;
  P.tile = o.tile_rows > 0 ? std::min(4096, std::max(64, o.tile_rows)) : 1024;
  P.tile = (P.tile + 63) & ~63;  // whole 64-row blocks
  std::map<std::pair<int, int>, int>& unit_R = P.unit_R;
  for (const GroupSpec& g : specs) {
    const int R = o.radix_bits > 0 ? std::min(16, o.radix_bits)
                                   : auto_radix_bits(relations[g.relation].num_rows,
                                                     o.partition_rows > 0 ? o.partition_rows : 256);
    unit_R[std::make_pair(g.relation, g.join_col)] = R;
    P.R_T = std::max(P.R_T, R);
  }
  // Batched: every relation takes the binding's R_T bits.  Units of all sizes then share batches
  // (a batch stages the binding rows once, so fewer distinct R means fewer passes over them), and
  // binding rows skip a small relation's empty partitions instead of probing its table.
  if (o.mode == QF_EVAL_BATCHED && o.radix_bits <= 0)
    for (auto& kv : unit_R) kv.second = P.R_T;
  std::set<int> bset, vset;
  for (const GroupSpec& g : specs) {
    vset.insert(g.join_var);
    bset.insert(g.join_var);
    for (const LitSpec& l : g.lits) {
      bset.insert(l.bcols.begin(), l.bcols.end());
      for (auto& e : l.eqs) {
        bset.insert(e.first);
        bset.insert(e.second);
      }
    }
  }
  P.bcols.assign(bset.begin(), bset.end());
  P.join_vars.assign(vset.begin(), vset.end());
  QF_REQUIRE((int)P.bcols.size() <= EV_MAXTC, "too many binding columns referenced by one step");
  const int64_t rows = std::max<int64_t>(1, binding.num_rows);
  // The binding partitions of the join variables: up to PT_MAXKEYS of them in one group whose
  // first pass reads the binding columns once, when the staged columns and the row id fit the
  // shared pass (six columns); otherwise one variable per group.  Groups are independent and run
  // side by side, each on its own stream with its temporaries in a region of its own (a batch of
  // at most 6 data columns plus the key and the row id is the widest single-variable pass).
  for (size_t v = 0; v < P.join_vars.size(); ++v) {
    P.tparts.emplace_back();
    for (size_t c = 0; c <= P.bcols.size(); ++c) P.tparts.back().emplace_back(ctx, rows * 4);
    P.toffs.emplace_back(ctx, (((size_t)1 << P.R_T) + 1) * 8);
  }
  const bool shared_pass = P.bcols.size() + 1 <= 6;
  const int per_group = shared_pass ? PT_MAXKEYS : 1;
  for (int v = 0; v < (int)P.join_vars.size(); v += per_group) {
    P.pgroups.emplace_back();
    for (int k = v; k < std::min<int>(v + per_group, (int)P.join_vars.size()); ++k) P.pgroups.back().push_back(k);
    const int nk = (int)P.pgroups.back().size();
    const int part_cols = shared_pass ? (int)P.bcols.size() + 1 : (int)std::min<size_t>(P.bcols.size(), 7) + 1;
    P.tscratch.emplace_back(ctx, std::max<size_t>(256, partition_scratch_bytes(binding.num_rows, part_cols, P.R_T, nk)));
  }
}

static void plan_units(Context* ctx, const qf_table& binding, const qf_table* relations,
                       const std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o, StepPlan& P,
                       bool stream) {
  const bool cache = o.reuse_partitions != 0;
  const std::map<std::pair<int, int>, int>& unit_R = P.unit_R;
  auto var_index = [&](int v) { return (int)(std::find(P.join_vars.begin(), P.join_vars.end(), v) - P.join_vars.begin()); };
  auto tidx = [&](int c) { return (int)(std::find(P.bcols.begin(), P.bcols.end(), c) - P.bcols.begin()); };

  std::map<std::pair<int, int>, std::shared_ptr<RelPart>> rel_parts;
  for (auto& kv : unit_R)
    rel_parts[kv.first] = partition_relation(ctx, relations[kv.first.first], kv.first.second, kv.second, cache);
  for (auto& kv : rel_parts) P.keep_rel.push_back(kv.second);

  // Unit membership.
  struct UnitSpec {
    std::pair<int, int> key;
    std::vector<const GroupSpec*> groups;
    std::vector<std::vector<const LitSpec*>> lits;
  };
  std::vector<UnitSpec> us;
  // Streamed units: their entry lists (targets numbered within the unit) and pair columns.
  struct UnitStream {
    std::vector<TsList> lists;
    std::vector<int64_t> sizes;  // entries of each list
    std::vector<TsTarget> targets;
    uint32_t wmask[EV_MAXG] = {};
  };
  std::vector<UnitStream> ustreams;
  std::vector<std::vector<int>> unit_vars;  // per unit: the join variable of each of its groups
  std::map<std::pair<int, int>, int> uindex;
  for (const GroupSpec& g : specs) {
    const auto key = std::make_pair(g.relation, g.join_col);
    if (o.mode == QF_EVAL_BATCHED) {
      auto it = uindex.find(key);
      int ui;
      if (it == uindex.end()) {
        ui = (int)us.size();
        us.push_back(UnitSpec{key, {}, {}});
        uindex[key] = ui;
      } else {
        ui = it->second;
      }
      UnitSpec& U = us[ui];
      auto gi = std::find_if(U.groups.begin(), U.groups.end(), [&](const GroupSpec* x) { return x->join_var == g.join_var; });
      if (gi == U.groups.end()) {
        U.groups.push_back(&g);
        U.lits.emplace_back();
        gi = U.groups.end() - 1;
      }
      for (const LitSpec& l : g.lits) U.lits[gi - U.groups.begin()].push_back(&l);
    } else if (o.mode == QF_EVAL_PLAIN_GROUP) {
      us.push_back(UnitSpec{key, {&g}, {{}}});
      for (const LitSpec& l : g.lits) us.back().lits[0].push_back(&l);
    } else {
      for (const LitSpec& l : g.lits) us.push_back(UnitSpec{key, {&g}, {{&l}}});
    }
  }

  for (const UnitSpec& S : us) {
    RelPart& rp = *rel_parts[S.key];
    DevUnit du{};
    for (int c = 0; c < rp.pt->num_columns; ++c) du.pcols[c] = rp.pt->columns[c].as<int32_t>();
    du.poff = rp.pt->offsets.as<int64_t>();
    du.hq = rp.hq.as<int32_t>();
    du.R = rp.pt->radix_bits;
    du.shift = P.R_T - du.R;
    du.ngroups = (int)S.groups.size();
    QF_REQUIRE(du.ngroups <= EV_MAXG, "too many join variables in one unit");
    du.arity = rp.pt->num_columns;
    // Key subsets of the unit: subset 0 is the join column alone (every literal of the unit binds
    // it); the others follow by key width, then first use.  Literals with row-side equalities
    // get subsets of their own.
    using SubKey = std::pair<std::vector<int>, bool>;
    auto sub_of = [](const std::vector<SubKey>& subs, const LitSpec* l) {
      for (size_t i = 0; i < subs.size(); ++i)
        if (subs[i].second == !l->eqs.empty() && subs[i].first == l->pcols) return (int)i;
      return -1;
    };
    const int jc = S.key.second;
    std::vector<SubKey> subsets{SubKey(std::vector<int>{jc}, false)};
    for (int gi = 0; gi < du.ngroups; ++gi)
      for (const LitSpec* l : S.lits[gi])
        if (sub_of(subsets, l) < 0) subsets.emplace_back(l->pcols, !l->eqs.empty());
    std::stable_sort(subsets.begin() + 1, subsets.end(),
                     [](const SubKey& x, const SubKey& y) { return x.first.size() < y.first.size(); });
    QF_REQUIRE((int)subsets.size() <= EV_MAXSUB, "too many key subsets in one unit");
    std::vector<std::vector<int>> lsub(du.ngroups);  // each literal's subset
    for (int gi = 0; gi < du.ngroups; ++gi)
      for (const LitSpec* l : S.lits[gi]) lsub[gi].push_back(sub_of(subsets, l));
    du.jcol = jc;
    int unit_lits = 0, nmask = 0;
    const size_t unit_first = P.lits.size();
    for (int gi = 0; gi < du.ngroups; ++gi) {
      const int vi = var_index(S.groups[gi]->join_var);
      DevUGroup& G = du.g[gi];
      (void)vi;
      G.jtcol = tidx(S.groups[gi]->join_var);
      // The group's literals ordered by subset; sbeg gives each subset's range.
      std::vector<const LitSpec*> order;
      const int first = (int)P.lits.size();
      for (int s = 0; s < (int)subsets.size(); ++s) {
        G.sbeg[s] = (int)P.lits.size();
        for (size_t li = 0; li < S.lits[gi].size(); ++li) {
          if (lsub[gi][li] != s) continue;
          const LitSpec* l = S.lits[gi][li];
          DevLit dl{};
          dl.slot = l->slot;
          dl.subset = s;
          dl.nk = (int)l->pcols.size();
          for (int k = 0; k < dl.nk; ++k) dl.tcol[k] = tidx(l->bcols[k]);
          dl.neq = (int)l->eqs.size();
          for (int e = 0; e < dl.neq; ++e) {
            dl.eq[e][0] = tidx(l->eqs[e].first);
            dl.eq[e][1] = tidx(l->eqs[e].second);
          }
          dl.pmask = -1;
          dl.mslot = -1;
          P.lits.push_back(dl);
          order.push_back(l);
          ++unit_lits;
        }
      }
      for (int s = (int)subsets.size(); s <= EV_MAXSUB; ++s) G.sbeg[s] = (int)P.lits.size();
      // Parents: a literal of three or more key columns is pruned by a literal of the same group
      // whose key is its key minus one column, bound to the same variables, with no row-side
      // equalities (its matches contain the child's).  Two-column keys are already pruned by the
      // join phase.
      for (size_t i = 0; i < order.size(); ++i) {
        const LitSpec* c = order[i];
        if (c->pcols.size() < 3) continue;
        for (size_t j = 0; j < order.size() && P.lits[first + i].pmask < 0; ++j) {
          const LitSpec* par = order[j];
          if (par->pcols.size() + 1 != c->pcols.size() || !par->eqs.empty()) continue;
          bool contained = true;
          for (size_t k = 0; k < par->pcols.size() && contained; ++k) {
            const auto pos = std::find(c->pcols.begin(), c->pcols.end(), par->pcols[k]) - c->pcols.begin();
            contained = pos < (ptrdiff_t)c->pcols.size() && c->bcols[pos] == par->bcols[k];
          }
          if (!contained) continue;
          DevLit& pd = P.lits[first + j];
          if (pd.mslot < 0) {
            if (nmask >= EV_MAXMASK) break;
            pd.mslot = nmask++;
          }
          P.lits[first + i].pmask = pd.mslot;
        }
      }
    }
    QF_REQUIRE(unit_lits <= EV_MAXLU, "too many literals in one unit");
    P.max_mask = std::max(P.max_mask, nmask);
    du.orphan = 0;
    for (int m = 0; m < EV_MAXMASK; ++m) du.child_subsets[m] = 0;
    for (size_t i = unit_first; i < P.lits.size(); ++i) {
      const DevLit& dl = P.lits[i];
      if (dl.pmask < 0) du.orphan |= 1u << dl.subset;
      else du.child_subsets[dl.pmask] |= 1u << dl.subset;
    }
    du.nsub = (int)subsets.size();
    // The streamed form takes the unit when its cells are the binding's and no literal carries a
    // row-side equality (the tables pair a join variable with one more column).
    bool eqs = false;
    for (const auto& gl : S.lits)
      for (const LitSpec* l : gl) eqs = eqs || !l->eqs.empty();
    const bool streamed = stream && !eqs && du.R == P.R_T;
    P.unit_streamed.push_back(streamed ? 1 : 0);
    if (streamed) {
      UnitStream ust;
      const int unit_index = (int)P.units.size();
      TsList run{};
      for (int v = 0; v < EV_MAXG; ++v) run.target[v] = run.vgroup[v] = -1;
      for (int gi = 0; gi < du.ngroups; ++gi)
        if (du.g[gi].sbeg[1] > du.g[gi].sbeg[0]) {
          run.target[var_index(S.groups[gi]->join_var)] = (int)ust.targets.size();
          ust.targets.push_back(TsTarget{du.g[gi].sbeg[0], du.g[gi].sbeg[1]});
        }
      if (!ust.targets.empty()) {
        auto tl = relation_tuples(ctx, rp, std::vector<int>{jc});
        P.keep_tuples.push_back(tl);
        run.m = 1;
        run.col[0] = tl->cols[0].as<int32_t>();
        run.cnt = tl->cnt.as<int32_t>();
        run.off = tl->off.as<int64_t>();
        run.unit = unit_index;
        run.psubset = -1;
        ust.lists.push_back(run);
        ust.sizes.push_back(tl->tuples);
      }
      // List keys: the join column, then the subset's other columns ascending.  A two-column
      // subset whose columns begin a wider subset's key is counted from that subset's list.
      std::vector<std::vector<int>> keys(subsets.size());
      for (int s = 1; s < (int)subsets.size(); ++s) {
        keys[s].push_back(jc);
        for (int c : subsets[s].first)
          if (c != jc) keys[s].push_back(c);
      }
      std::vector<int> merged_into(subsets.size(), -1), prefix_of(subsets.size(), -1);
      for (int s = 1; s < (int)subsets.size(); ++s) {
        if (keys[s].size() != 2) continue;
        for (int t = 1; t < (int)subsets.size() && merged_into[s] < 0; ++t)
          if (keys[t].size() >= 3 && keys[t][1] == keys[s][1] && prefix_of[t] < 0) {
            merged_into[s] = t;
            prefix_of[t] = s;
          }
      }
      for (int s = 1; s < (int)subsets.size(); ++s) {
        if (merged_into[s] >= 0) continue;
        const std::vector<int>& pc = subsets[s].first;  // ascending relation columns, the join column among them
        const std::vector<int>& key = keys[s];
        auto tl = relation_tuples(ctx, rp, key);
        P.keep_tuples.push_back(tl);
        TsList L{};
        L.m = (int)key.size();
        for (int i = 0; i < L.m; ++i) {
          L.col[i] = tl->cols[i].as<int32_t>();
          L.kidx[i] = (int)(std::find(pc.begin(), pc.end(), key[i]) - pc.begin());
        }
        L.cnt = tl->cnt.as<int32_t>();
        L.off = tl->off.as<int64_t>();
        L.unit = unit_index;
        L.subset = s;
        L.psubset = prefix_of[s];
        if (L.psubset >= 0) {
          const std::vector<int>& ppc = subsets[L.psubset].first;
          L.pkidx1 = (int)(std::find(ppc.begin(), ppc.end(), key[1]) - ppc.begin());
        }
        for (int v = 0; v < EV_MAXG; ++v) L.target[v] = L.vgroup[v] = -1;
        for (int gi = 0; gi < du.ngroups; ++gi) {
          const int vi = var_index(S.groups[gi]->join_var);
          L.vgroup[vi] = gi;
          for (int li = du.g[gi].sbeg[s]; li < du.g[gi].sbeg[s + 1]; ++li) ust.wmask[vi] |= 1u << P.lits[li].tcol[L.kidx[1]];
          if (L.psubset >= 0)
            for (int li = du.g[gi].sbeg[L.psubset]; li < du.g[gi].sbeg[L.psubset + 1]; ++li)
              ust.wmask[vi] |= 1u << P.lits[li].tcol[L.pkidx1];
        }
        ust.lists.push_back(L);
        ust.sizes.push_back(tl->tuples);
      }
      ustreams.push_back(ust);
    }
    for (int s = 0; s < du.nsub; ++s) {
      du.sub[s].nk = (int)subsets[s].first.size();
      for (int k = 0; k < du.sub[s].nk; ++k) du.sub[s].pcol[k] = subsets[s].first[k];
      if (!rp.heavy_q.empty()) {
        auto ht = heavy_tables(ctx, rp, subsets[s].first);
        du.sub[s].hbuf = ht->buf.as<uint32_t>();
        du.sub[s].hbase = ht->base.as<int64_t>();
        P.keep_heavy.push_back(ht);
      }
    }
    P.units.push_back(du);
    std::vector<int> vars;
    for (const GroupSpec* g : S.groups) vars.push_back(g->join_var);
    unit_vars.push_back(vars);
  }

  // Batches: in the batched form, units with equal radix bits (they see the same binding cells,
  // so each chunk of binding rows is staged once for all of them); otherwise one unit each.
  std::vector<std::vector<int>> members;
  if (o.mode == QF_EVAL_BATCHED) {
    // Streamed units form batches of their own, which k_eval runs in the cells the streamed form
    // leaves.
    std::map<std::pair<int, int>, std::vector<int>> by_r;
    for (int u = 0; u < (int)P.units.size(); ++u) by_r[std::make_pair(P.units[u].R, (int)P.unit_streamed[u])].push_back(u);
    for (auto& kv : by_r) {
      std::vector<int> cur;
      std::set<int> cur_vars;
      for (int u : kv.second) {
        std::set<int> vs = cur_vars;
        vs.insert(unit_vars[u].begin(), unit_vars[u].end());
        if (!cur.empty() && ((int)cur.size() == EV_MAXBU || (int)vs.size() > EV_MAXG)) {
          members.push_back(cur);
          cur.clear();
          vs = std::set<int>(unit_vars[u].begin(), unit_vars[u].end());
        }
        cur.push_back(u);
        cur_vars = vs;
      }
      if (!cur.empty()) members.push_back(cur);
    }
  } else {
    for (int u = 0; u < (int)P.units.size(); ++u) members.push_back({u});
  }
  for (const auto& m : members) {
    DevBatch b{};
    b.R = P.units[m[0]].R;
    b.heavy_only = P.unit_streamed[m[0]];
    b.shift = P.R_T - b.R;
    std::vector<int> bvars;
    for (int u : m)
      for (int v : unit_vars[u])
        if (std::find(bvars.begin(), bvars.end(), v) == bvars.end()) bvars.push_back(v);
    b.ngroups = (int)bvars.size();
    for (int g = 0; g < b.ngroups; ++g) {
      const int vi = var_index(bvars[g]);
      for (size_t c = 0; c <= P.bcols.size(); ++c) b.tcols[g][c] = P.tparts[vi][c].as<int32_t>();
      b.toff[g] = P.toffs[vi].as<int64_t>();
    }
    b.nunits = (int)m.size();
    for (int k = 0; k < b.nunits; ++k) {
      const int u = m[k];
      b.unit[k] = u;
      for (int g = 0; g < EV_MAXG; ++g) P.units[u].ug_of_bg[g] = -1;
      for (int g = 0; g < b.ngroups; ++g) {
        auto it = std::find(unit_vars[u].begin(), unit_vars[u].end(), bvars[g]);
        if (it != unit_vars[u].end()) P.units[u].ug_of_bg[g] = (int)(it - unit_vars[u].begin());
      }
    }
    P.batches.push_back(b);
  }

  // Launch ranges: every batch at once (batched) or one per launch (plain forms).
  const int NB = (int)P.batches.size();
  if (o.mode == QF_EVAL_BATCHED) {
    if (NB > 0) P.launches.emplace_back(0, NB);
  } else {
    for (int b = 0; b < NB; ++b) P.launches.emplace_back(b, b + 1);
  }
  for (auto& lr : P.launches) {
    int64_t c = 0, items = 0;
    for (int b = lr.first; b < lr.second; ++b) {
      P.unit_cells.push_back(c);
      c += (int64_t)1 << P.batches[b].R;
      // Items of a batch: at most one per cell plus one per tile of its binding rows.
      items += ((int64_t)1 << P.batches[b].R) + (int64_t)P.batches[b].ngroups * ceil_div(binding.num_rows, P.tile);
    }
    P.unit_cells.push_back(c);
    P.max_cells = std::max(P.max_cells, c);
    P.max_items = std::max(P.max_items, items);
    P.launch_items.push_back(items);
  }
  if (!ustreams.empty()) {
    // Launches of at most TS_MAXL lists and TS_MAXT targets; every launch stages all join
    // variables with the same rows per item, so that all agree on which cells are light.
    const int64_t cells = (int64_t)1 << P.R_T;
    const int nvars = (int)P.join_vars.size();
    struct Pack {
      std::vector<TsList> lists;
      std::vector<int64_t> sizes;
      int nrun = 0;
      int64_t entries = 0;
      std::vector<TsTarget> targets;
      uint32_t wmask[EV_MAXG] = {};
    };
    std::vector<Pack> packs(1);
    for (const UnitStream& u : ustreams) {
      int64_t entries = 0;
      for (int64_t s : u.sizes) entries += s;
      QF_REQUIRE(entries < ((int64_t)1 << 31), "a unit's key tuples must stay below 2^31");
      if (packs.back().lists.size() + u.lists.size() > (size_t)TS_MAXL ||
          packs.back().targets.size() + u.targets.size() > (size_t)TS_MAXT ||
          packs.back().entries + entries >= ((int64_t)1 << 31))
        packs.emplace_back();
      Pack& pk = packs.back();
      const int t0 = (int)pk.targets.size();
      for (size_t i = 0; i < u.lists.size(); ++i) {
        TsList L = u.lists[i];
        for (int v = 0; v < EV_MAXG; ++v)
          if (L.target[v] >= 0) L.target[v] += t0;
        // Run lists first: the kernel concatenates a cell's runs, then its tuples.
        const size_t at = L.m == 1 ? (size_t)pk.nrun++ : pk.lists.size();
        pk.lists.insert(pk.lists.begin() + at, L);
        pk.sizes.insert(pk.sizes.begin() + at, u.sizes[i]);
        pk.entries += u.sizes[i];
      }
      pk.targets.insert(pk.targets.end(), u.targets.begin(), u.targets.end());
      for (int v = 0; v < EV_MAXG; ++v) pk.wmask[v] |= u.wmask[v];
    }
    int tr = TS_ROWS / std::max(1, nvars);
    for (const Pack& pk : packs) {
      int pairs = 0;
      for (int v = 0; v < nvars; ++v) pairs += __builtin_popcount(pk.wmask[v] & ~(1u << tidx(P.join_vars[v])));
      if (pairs > 0) tr = std::min(tr, TS_PMAX / pairs);
    }
    QF_REQUIRE(tr >= 1, "too many paired binding columns for the streamed form");
    int64_t bound = 0;
    for (const Pack& pk : packs) {
      StepPlan::StreamLaunch SL;
      SL.d_lists = DevBuf(ctx, std::max<size_t>(1, pk.lists.size()) * sizeof(TsList));
      if (!pk.lists.empty()) upload(ctx, SL.d_lists.get(), pk.lists.data(), pk.lists.size() * sizeof(TsList));
      SL.d_targets = DevBuf(ctx, std::max<size_t>(1, pk.targets.size()) * sizeof(TsTarget));
      if (!pk.targets.empty())
        upload(ctx, SL.d_targets.get(), pk.targets.data(), pk.targets.size() * sizeof(TsTarget));
      TsLaunch tl{};
      tl.lists = SL.d_lists.as<TsList>();
      tl.nlists = (int)pk.lists.size();
      tl.nrun = pk.nrun;
      tl.targets = SL.d_targets.as<TsTarget>();
      tl.ntargets = (int)pk.targets.size();
      tl.nvars = nvars;
      tl.ntc = (int)P.bcols.size();
      tl.tr = tr;
      for (int v = 0; v < nvars; ++v) {
        tl.jt[v] = tidx(P.join_vars[v]);
        tl.wmask[v] = pk.wmask[v];
        for (size_t c = 0; c <= P.bcols.size(); ++c) tl.tcols[v][c] = P.tparts[v][c].as<int32_t>();
        tl.toff[v] = P.toffs[v].as<int64_t>();
      }
      tl.cells = cells;
      SL.d_launch = DevBuf(ctx, sizeof(TsLaunch));
      upload(ctx, SL.d_launch.get(), &tl, sizeof(TsLaunch));
      // The cells' list ranges depend on the relations' key-tuple lists only: with cached
      // relations a matrix made for the same lists before is taken again.
      std::vector<int64_t> cm_key{cells, pk.nrun};
      for (const TsList& l : pk.lists) cm_key.push_back((int64_t)(uintptr_t)l.off);
      RelCache& rc = rel_cache(ctx);
      if (cache)
        for (auto it = rc.cellmats.begin(); it != rc.cellmats.end(); ++it)
          if (it->first == cm_key) {
            SL.cm = it->second;
            rc.cellmats.splice(rc.cellmats.begin(), rc.cellmats, it);
            break;
          }
      if (!SL.cm) {
        SL.cm = std::make_shared<CellMat>();
        SL.cm->parts = DevBuf(ctx, cells * 4);
        SL.cm->cellmat = DevBuf(ctx, (size_t)cells * std::max<size_t>(1, pk.lists.size()) * sizeof(int2));
        uploads_before(ctx);
        hipLaunchKernelGGL(k_ts_cells, dim3((unsigned)ceil_div(cells, 256)), dim3(256), 0, ctx->stream,
                           SL.d_launch.as<TsLaunch>(), SL.cm->parts.as<int32_t>(), SL.cm->cellmat.as<int2>());
        QF_CHECK_HIP(hipGetLastError());
        if (cache) {
          rc.cellmats.emplace_front(std::move(cm_key), SL.cm);
          if (rc.cellmats.size() > EV_CELLMAT_CACHE) rc.cellmats.pop_back();
        }
      }
      // Items: at most TS_CMAX binding chunks per cell times its parts, which number at most
      // one per cell with entries plus the cell's wave chunks over TS_PART (a cell's chunks are
      // at most its entries / TS_RUNROW, plus one row per run list, over TS_RUNROWS, plus two).
      const int64_t chunks = (pk.entries / TS_RUNROW + cells * (int64_t)pk.nrun) / TS_RUNROWS + 2 * cells;
      SL.items_bound = (int64_t)TS_CMAX * (cells + chunks / TS_PART + 1);
      bound = std::max(bound, SL.items_bound);
      P.slaunches.push_back(std::move(SL));
    }
    P.d_ts_counts = DevBuf(ctx, cells * 8);
    P.d_ts_starts = DevBuf(ctx, cells * 8);
    P.d_light = DevBuf(ctx, cells);

    P.ts_items = bound;
    P.d_ts_items = DevBuf(ctx, bound * sizeof(DevItem));
    P.d_ts_num = DevBuf(ctx, 16);
    P.ts_lds = (size_t)P.bcols.size() * TS_ROWS * 4 + (size_t)(2 * TS_VCAP + TS_PCAP) * 4 + TS_ROWS;
    QF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_stream),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.ts_lds));
    int per_cu = 0;
    QF_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stream, TS_THREADS, P.ts_lds));
    P.ts_blocks = ctx->num_cus * std::max(1, per_cu);
  }
  P.d_units = DevBuf(ctx, std::max<size_t>(1, P.units.size()) * sizeof(DevUnit));
  P.d_batches = DevBuf(ctx, std::max<size_t>(1, P.batches.size()) * sizeof(DevBatch));
  if (!P.batches.empty()) upload(ctx, P.d_batches.get(), P.batches.data(), P.batches.size() * sizeof(DevBatch));
  P.d_lits = DevBuf(ctx, std::max<size_t>(1, P.lits.size()) * sizeof(DevLit));
  P.d_unit_cells = DevBuf(ctx, std::max<size_t>(1, P.unit_cells.size()) * 8);
  if (!P.units.empty()) upload(ctx, P.d_units.get(), P.units.data(), P.units.size() * sizeof(DevUnit));
  if (!P.lits.empty()) upload(ctx, P.d_lits.get(), P.lits.data(), P.lits.size() * sizeof(DevLit));
  if (!P.unit_cells.empty()) upload(ctx, P.d_unit_cells.get(), P.unit_cells.data(), P.unit_cells.size() * 8);
  P.d_cell_counts = DevBuf(ctx, std::max<int64_t>(1, P.max_cells) * 8);
  P.d_cell_starts = DevBuf(ctx, std::max<int64_t>(1, P.max_cells) * 8);
  P.d_items = DevBuf(ctx, std::max<int64_t>(1, P.max_items) * sizeof(DevItem));
  P.d_num_items = DevBuf(ctx, 8);
  P.d_counts = DevBuf(ctx, std::max(1, num_slots) * 4 * sizeof(uint64_t));
  // LDS words: table | facts | staged rows (ntc x tile) | positive flags (uint8 x tile) |
  // hit-row list (uint16 x tile) | hit masks (uint64 x max_mask x tile / 64).
  P.cidx_off = EV_TBL_WORDS + EV_PST_WORDS + (int)P.bcols.size() * P.tile + P.tile / 4;
  P.mask_off = P.cidx_off + P.tile / 2;
  P.mstride = P.tile / 64;
  P.lds_bytes = ((size_t)P.mask_off + (size_t)P.max_mask * P.mstride * 2) * 4;
  QF_REQUIRE(P.lds_bytes <= 160 * 1024 - 4096, "work item does not fit in LDS");
  for (const void* fn : {reinterpret_cast<const void*>(k_eval<false>), reinterpret_cast<const void*>(k_eval<true>)})
    QF_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds_bytes));
  int per_cu = 0, per_cu_bits = 0;
  QF_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eval<false>, EV_THREADS, P.lds_bytes));
  QF_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_bits, k_eval<true>, EV_THREADS, P.lds_bytes));
  per_cu = std::min(per_cu, per_cu_bits);
  P.max_blocks = ctx->num_cus * std::max(1, per_cu);
  // Counter copies: up to one per CU, within 32 MiB.
  const int64_t copy_bytes = (int64_t)std::max(1, num_slots) * 4 * (int64_t)sizeof(uint64_t);
  P.ncopies = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(ctx->num_cus, P.max_blocks), (32ll << 20) / copy_bytes));
  P.d_copies = DevBuf(ctx, (size_t)(P.ncopies * copy_bytes));
  P.bit_words = ceil_div(binding.num_rows, 64);
  end_uploads(ctx);
}

static void plan_step(Context* ctx, const qf_table& binding, const qf_table* relations,
                      const std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o, StepPlan& P,
                      bool stream) {
  plan_binding(ctx, binding, relations, specs, num_slots, o, P);
  plan_units(ctx, binding, relations, specs, num_slots, o, P, stream);
}

// Auxiliary streams for forked work: n - 1 streams and n events (event 0 forks, event v joins
// stream v - 1 back).  Default priority, non-blocking.
static void ensure_aux_streams(Context* ctx, int n) {
  while ((int)ctx->aux_streams.size() < n - 1) {
    hipStream_t st;
    QF_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    ctx->aux_streams.push_back(st);
  }
  while ((int)ctx->aux_events.size() < n) {
    hipEvent_t ev;
    QF_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    ctx->aux_events.push_back(ev);
  }
}

// The binding partitions of the plan's join variables (group by group, forked streams joined
// back to the context stream).
static void enqueue_partition(Context* ctx, const qf_table& binding, StepPlan& P) {
  PhaseScope ps(ctx, "partition_binding");
  const int ng = (int)P.pgroups.size();
  ensure_aux_streams(ctx, ng);
  hipStream_t main = ctx->stream;
  if (ng > 1) QF_CHECK_HIP(hipEventRecord(ctx->aux_events[0], main));
  for (int g = 0; g < ng; ++g) {
    // Group g on aux stream g - 1 (the first on the context stream), after all work already
    // queued on the context stream.
    hipStream_t st = g == 0 ? main : ctx->aux_streams[g - 1];
    if (g > 0) QF_CHECK_HIP(hipStreamWaitEvent(st, ctx->aux_events[0], 0));
    Scratch scratch;
    scratch.base = P.tscratch[g].as<char>();
    scratch.capacity = P.tscratch[g].bytes();
    ctx->stream = st;
    try {
      const std::vector<int>& grp = P.pgroups[g];
      if (grp.size() > 1) {
        std::vector<const int32_t*> in;
        for (int c : P.bcols) in.push_back(binding.columns[c]);
        std::vector<int> keys;
        std::vector<int32_t*> outs;
        std::vector<int64_t*> offs;
        for (int v : grp) {
          keys.push_back((int)(std::find(P.bcols.begin(), P.bcols.end(), P.join_vars[v]) - P.bcols.begin()));
          for (auto& b : P.tparts[v]) outs.push_back(b.as<int32_t>());
          offs.push_back(P.toffs[v].as<int64_t>());
        }
        stable_partition_multi(ctx, in.data(), (int)in.size(), keys.data(), (int)keys.size(), true,
                               binding.num_rows, 1, P.R_T, outs.data(), offs.data(), &scratch);
      } else {
        const int v = grp[0];
        std::vector<int32_t*> outs;
        for (auto& b : P.tparts[v]) outs.push_back(b.as<int32_t>());
        partition_binding(ctx, binding, P.join_vars[v], P.bcols, P.R_T, outs, P.toffs[v].as<int64_t>(), &scratch);
      }
    } catch (...) {
      ctx->stream = main;
      throw;
    }
    ctx->stream = main;
    if (g > 0) {
      QF_CHECK_HIP(hipEventRecord(ctx->aux_events[g], st));
      QF_CHECK_HIP(hipStreamWaitEvent(main, ctx->aux_events[g], 0));
    }
  }
}

// Device work of one step on the context stream, the binding partitions forked to auxiliary
// streams and joined back (capturable: kernels, device copies and event waits).
static void enqueue_step(Context* ctx, const qf_table& binding, int64_t num_positive, StepPlan& P, uint64_t* d_bits,
                         uint64_t* h_out) {
  fill_u64(ctx, P.d_copies.as<uint64_t>(), 0, (int64_t)P.ncopies * P.num_slots * 4);
  if (d_bits && P.num_slots > 0) fill_u64(ctx, d_bits, 0, (int64_t)P.num_slots * P.bit_words);
  if (!P.units.empty() && binding.num_rows > 0) {
    if (!P.partitioned) enqueue_partition(ctx, binding, P);
    // Streamed units: the cells they take, then launch by launch.
    const int64_t cells_T = (int64_t)1 << P.R_T;
    if (!P.slaunches.empty()) {
      PhaseScope ps(ctx, "work_items");
      hipLaunchKernelGGL(k_ts_light, dim3((unsigned)ceil_div(cells_T, 256)), dim3(256), 0, ctx->stream,
                         P.slaunches[0].d_launch.as<TsLaunch>(), P.d_light.as<uint8_t>());
      QF_CHECK_HIP(hipGetLastError());
    }
    for (size_t li = 0; li < P.slaunches.size(); ++li) {
      const StepPlan::StreamLaunch& SL = P.slaunches[li];
      {
        PhaseScope ps(ctx, "work_items");
        hipLaunchKernelGGL(k_ts_count, dim3((unsigned)ceil_div(cells_T, 256)), dim3(256), 0, ctx->stream,
                           SL.d_launch.as<TsLaunch>(), P.d_light.as<uint8_t>(), SL.cm->parts.as<int32_t>(),
                           P.d_ts_counts.as<int64_t>());
        QF_CHECK_HIP(hipGetLastError());
        scan_exclusive_i64(ctx, P.d_ts_counts.as<int64_t>(), P.d_ts_starts.as<int64_t>(), cells_T);
        hipLaunchKernelGGL(k_ts_expand, dim3((unsigned)ceil_div(SL.items_bound, 256)), dim3(256), 0, ctx->stream,
                           P.d_ts_counts.as<int64_t>(), P.d_ts_starts.as<int64_t>(), SL.cm->parts.as<int32_t>(), cells_T,
                           P.d_ts_items.as<DevItem>(), P.d_ts_num.as<int64_t>());
        QF_CHECK_HIP(hipGetLastError());
      }
      TsArgs ta{};
      ta.launch = SL.d_launch.as<TsLaunch>();
      ta.cellmat = SL.cm->cellmat.as<int2>();
      ta.units = P.d_units.as<DevUnit>();
      ta.lits = P.d_lits.as<DevLit>();
      ta.items = P.d_ts_items.as<DevItem>();
      ta.num_items = P.d_ts_num.as<int64_t>();
      ta.max_items = P.ts_items;
      ta.num_positive = (int)num_positive;
      ta.counts_out = reinterpret_cast<unsigned long long*>(P.d_copies.get());
      ta.copy_words = (int64_t)P.num_slots * 4;
      ta.ncopies = P.ncopies;
      PhaseScope ps(ctx, "probe_count_stream");
      const int blocks = (int)std::min<int64_t>(P.ts_blocks, std::max<int64_t>(1, SL.items_bound));
      hipLaunchKernelGGL(k_stream, dim3((unsigned)blocks), dim3(TS_THREADS), P.ts_lds, ctx->stream, ta);
      QF_CHECK_HIP(hipGetLastError());
    }
    EvalArgs a{};
    a.lits = P.d_lits.as<DevLit>();
    a.ntc = (int)P.bcols.size();
    a.num_positive = (int)num_positive;
    a.counts_out = reinterpret_cast<unsigned long long*>(P.d_copies.get());
    a.copy_words = (int64_t)P.num_slots * 4;
    a.ncopies = P.ncopies;
    a.bits = reinterpret_cast<unsigned long long*>(d_bits);
    a.bit_words = P.bit_words;
    a.tile_rows = P.tile;
    a.cidx_off = P.cidx_off;
    a.mask_off = P.mask_off;
    a.mstride = P.mstride;
    a.items = P.d_items.as<DevItem>();
    a.num_items = P.d_num_items.as<int64_t>();
    size_t uc = 0, li = 0;
    for (auto& lr : P.launches) {
      const int64_t items_bound = P.launch_items[li++];
      const int nu = lr.second - lr.first;
      const int64_t cells = P.unit_cells[uc + nu];
      a.units = P.d_units.as<DevUnit>();  // batches name units by their index in the step
      a.batches = P.d_batches.as<DevBatch>() + lr.first;
      const int64_t* unit_cells = P.d_unit_cells.as<int64_t>() + uc;
      uc += nu + 1;
      if (cells == 0) continue;
      {
        PhaseScope ps(ctx, "work_items");
        const dim3 grid((unsigned)ceil_div(cells, 256));
        hipLaunchKernelGGL(k_count_items, grid, dim3(256), 0, ctx->stream, a.batches, a.units, unit_cells, nu, P.tile,
                           P.d_light.as<uint8_t>(), P.d_cell_counts.as<int64_t>());
        QF_CHECK_HIP(hipGetLastError());
        scan_exclusive_i64(ctx, P.d_cell_counts.as<int64_t>(), P.d_cell_starts.as<int64_t>(), cells);
        hipLaunchKernelGGL(k_expand_items, dim3((unsigned)std::max<int64_t>(1, ceil_div(items_bound, 256))), dim3(256),
                           0, ctx->stream, unit_cells, nu, P.d_cell_counts.as<int64_t>(),
                           P.d_cell_starts.as<int64_t>(), P.d_items.as<DevItem>(), P.d_num_items.as<int64_t>());
        QF_CHECK_HIP(hipGetLastError());
      }
      // Blocks loop over the items the device listed; more blocks than the bound would idle.
      const int blocks = (int)std::min<int64_t>(P.max_blocks, std::max<int64_t>(1, items_bound));
      PhaseScope ps(ctx, "probe_count");
      if (a.bits != nullptr)
        hipLaunchKernelGGL(k_eval<true>, dim3(blocks), dim3(EV_THREADS), P.lds_bytes, ctx->stream, a);
      else
        hipLaunchKernelGGL(k_eval<false>, dim3(blocks), dim3(EV_THREADS), P.lds_bytes, ctx->stream, a);
      QF_CHECK_HIP(hipGetLastError());
    }
  }
  if (P.num_slots > 0) {
    const int64_t words = (int64_t)P.num_slots * 4;
    hipLaunchKernelGGL(k_sum_copies, dim3((unsigned)ceil_div(words, SUM_WORDS)), dim3(SUM_WORDS * SUM_GROUPS), 0,
                       ctx->stream,
                       P.d_copies.as<uint64_t>(), P.ncopies, words, P.d_counts.as<uint64_t>());
    QF_CHECK_HIP(hipGetLastError());
  }
  QF_CHECK_HIP(hipMemcpyAsync(h_out, P.d_counts.get(), (size_t)P.num_slots * 4 * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
}

struct GraphCache {
  std::vector<int64_t> key;
  std::unique_ptr<StepPlan> plan;
  // The raw arguments of the qf_evaluate_literals call the graph last served (empty when it
  // was made by another entry point): an identical call replays it before grouping literals.
  std::string call;
};

// Replays a captured step on the private stream after the caller's queued work and copies its
// counts out.
static void replay_graph(Context* ctx, const GraphCache& gc, uint64_t* h_counts, int num_slots) {
  QF_CHECK_HIP(hipEventRecord(ctx->graph_order_event, ctx->stream));
  QF_CHECK_HIP(hipStreamWaitEvent(ctx->graph_stream, ctx->graph_order_event, 0));
  QF_CHECK_HIP(hipGraphLaunch(gc.plan->graph, ctx->graph_stream));
  QF_CHECK_HIP(hipStreamSynchronize(ctx->graph_stream));
  std::memcpy(h_counts, gc.plan->pinned, (size_t)num_slots * 4 * sizeof(uint64_t));
}

// Appends the raw arguments of an evaluation call (sizes, options, column addresses, argument
// types and literal bytes) to `out`; two calls with equal strings compute the same step.
static void call_signature(std::string& out, const qf_table& binding, int64_t num_positive, const qf_table* relations,
                           const int* arg_types, int num_relations, const qf_literal* literals, int num_literals,
                           const qf_eval_options& o, const uint64_t* d_bits) {
  auto put = [&](const void* p, size_t n) { out.append(static_cast<const char*>(p), n); };
  const int64_t head[] = {binding.num_rows, binding.num_columns, num_positive, num_relations, num_literals,
                          (int64_t)(uintptr_t)d_bits, arg_types ? 1 : 0};
  out.clear();
  put(head, sizeof(head));
  put(&o, sizeof(o));
  put(binding.columns, (size_t)binding.num_columns * sizeof(void*));
  int64_t ntypes = 0;
  for (int r = 0; r < num_relations; ++r) {
    QF_REQUIRE(relations[r].num_columns >= 1 && relations[r].num_columns <= QF_MAX_ARITY && relations[r].columns,
               "relation has no columns or more than QF_MAX_ARITY");
    const int64_t shape[] = {relations[r].num_rows, relations[r].num_columns};
    put(shape, sizeof(shape));
    put(relations[r].columns, (size_t)relations[r].num_columns * sizeof(void*));
    ntypes += relations[r].num_columns;
  }
  if (arg_types) put(arg_types, (size_t)ntypes * sizeof(int));
  put(literals, (size_t)num_literals * sizeof(qf_literal));
}

static void run_groups_impl(Context* ctx, const qf_table& binding, int64_t num_positive, const qf_table* relations,
                            std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o,
                            uint64_t* h_counts, uint64_t* d_bits, bool allow_stream = true);

static TuneCache& tune_cache(Context* ctx) {
  if (!ctx->tune_cache) ctx->tune_cache = std::make_shared<TuneCache>();
  return *std::static_pointer_cast<TuneCache>(ctx->tune_cache);
}

// Times one configuration of the step (device work only, HIP events on the context stream).
static double time_config(Context* ctx, const qf_table& binding, int64_t num_positive, const qf_table* relations,
                          std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o, uint64_t* h_counts,
                          bool stream) {
  qf_eval_options t = o;
  t.autotune = 0;
  t.use_graph = 0;
  run_groups_impl(ctx, binding, num_positive, relations, specs, num_slots, t, h_counts, nullptr, stream);  // warm caches
  hipEvent_t e0, e1;
  QF_CHECK_HIP(hipEventCreate(&e0));
  QF_CHECK_HIP(hipEventCreate(&e1));
  QF_CHECK_HIP(hipEventRecord(e0, ctx->stream));
  run_groups_impl(ctx, binding, num_positive, relations, specs, num_slots, t, h_counts, nullptr, stream);
  QF_CHECK_HIP(hipEventRecord(e1, ctx->stream));
  QF_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QF_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

static void run_groups(Context* ctx, const qf_table& binding, int64_t num_positive, const qf_table* relations,
                       std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o, uint64_t* h_counts,
                       uint64_t* d_bits) {
  if (num_slots == 0) return;
  if (!o.autotune || o.radix_bits > 0 || o.tile_rows > 0 || o.partition_rows > 0 || specs.empty()) {
    run_groups_impl(ctx, binding, num_positive, relations, specs, num_slots, o, h_counts, d_bits);
    return;
  }
  // Workload shape: binding rows and columns, mode, and every group's relation size and literals.
  std::string shape = std::to_string(binding.num_rows) + "/" + std::to_string(binding.num_columns) + "/" +
                      std::to_string(o.mode) + "|";
  for (const GroupSpec& g : specs)
    shape += std::to_string(relations[g.relation].num_rows) + ":" + std::to_string(relations[g.relation].num_columns) +
             ":" + std::to_string(g.join_col) + ":" + std::to_string(g.lits.size()) + ";";
  TuneCache& tc = tune_cache(ctx);
  auto it = tc.best.find(shape);
  if (it == tc.best.end()) {
    // Grid: facts per relation partition (radix bits) x binding rows per work item, with the
    // streamed form on or off (the batched form without covered bitvectors uses it).
    TuneChoice best;
    best.ms = 1e30;
    const bool streamable = o.mode == QF_EVAL_BATCHED && d_bits == nullptr;
    for (int s = streamable ? 1 : 0; s >= 0; --s)
      for (int pr : {128, 256, 512})
        for (int tile : {512, 768, 1024, 1536, 2048}) {
          if (s && tile != 768 && tile != 1536) continue;  // with the streamed form the tile sizes only k_eval's rest
          qf_eval_options t = o;
          t.partition_rows = pr;
          t.tile_rows = tile;
          const double ms = time_config(ctx, binding, num_positive, relations, specs, num_slots, t, h_counts, s != 0);
          if (ms < best.ms) best = TuneChoice{pr, tile, s != 0, ms};
        }
    it = tc.best.emplace(shape, best).first;
  }
  tc.last = it->second;
  qf_eval_options t = o;
  t.autotune = 0;
  t.partition_rows = it->second.partition_rows;
  t.tile_rows = it->second.tile_rows;
  run_groups_impl(ctx, binding, num_positive, relations, specs, num_slots, t, h_counts, d_bits, it->second.stream);
}

static void run_groups_impl(Context* ctx, const qf_table& binding, int64_t num_positive, const qf_table* relations,
                            std::vector<GroupSpec>& specs, int num_slots, const qf_eval_options& o,
                            uint64_t* h_counts, uint64_t* d_bits, bool allow_stream) {
  if (num_slots == 0) return;
  // A captured step is replayed when the same call repeats: same binding and relation
  // columns, literal set, options and output pointer.
  // The key lists every input of the step (sizes, options, column addresses, literal keys),
  // with a length before each variable-length part so that different steps never share a key.
  std::vector<int64_t> key;
  QF_REQUIRE(!o.use_graph || o.reuse_partitions, "use_graph needs reuse_partitions (relations must stay unchanged)");
  if (o.use_graph && o.mode == QF_EVAL_BATCHED) {
    key.reserve(64 + 16 * (size_t)num_slots);
    key.insert(key.end(), {binding.num_rows, num_positive, o.radix_bits, o.tile_rows, o.partition_rows,
                           o.reuse_partitions, (int64_t)(uintptr_t)d_bits, binding.num_columns, allow_stream ? 1 : 0});
    for (int c = 0; c < binding.num_columns; ++c) key.push_back((int64_t)(uintptr_t)binding.columns[c]);
    key.push_back((int64_t)specs.size());
    for (const GroupSpec& g : specs) {
      const qf_table& rel = relations[g.relation];
      key.insert(key.end(), {(int64_t)rel.num_rows, (int64_t)rel.num_columns, g.join_col, g.join_var,
                             (int64_t)g.lits.size()});
      for (int c = 0; c < rel.num_columns; ++c) key.push_back((int64_t)(uintptr_t)rel.columns[c]);
      for (const LitSpec& l : g.lits) {
        key.insert(key.end(), {l.slot, (int64_t)l.pcols.size(), (int64_t)l.eqs.size()});
        for (size_t k = 0; k < l.pcols.size(); ++k) key.insert(key.end(), {l.pcols[k], l.bcols[k]});
        for (auto& e : l.eqs) key.insert(key.end(), {e.first, e.second});
      }
    }
    if (!ctx->graph_stream) {
      QF_CHECK_HIP(hipStreamCreateWithFlags(&ctx->graph_stream, hipStreamNonBlocking));
      QF_CHECK_HIP(hipEventCreateWithFlags(&ctx->graph_order_event, hipEventDisableTiming));
    }
    auto* gc = static_cast<GraphCache*>(ctx->graph_cache.get());
    if (gc && gc->key == key && gc->plan && gc->plan->graph) {
      replay_graph(ctx, *gc, h_counts, num_slots);
      return;
    }
  }
  std::unique_ptr<StepPlan> P(new StepPlan);
  const bool streamable = allow_stream && o.mode == QF_EVAL_BATCHED && d_bits == nullptr;
  if (key.empty()) {
    // Uncaptured: the binding partition runs while the rest of the plan is built.
    plan_binding(ctx, binding, relations, specs, num_slots, o, *P);
    if (binding.num_rows > 0 && !specs.empty()) {
      enqueue_partition(ctx, binding, *P);
      P->partitioned = true;
    }
    plan_units(ctx, binding, relations, specs, num_slots, o, *P, streamable);
  } else {
    plan_step(ctx, binding, relations, specs, num_slots, o, *P, streamable);
  }
  const size_t count_bytes = (size_t)num_slots * 4 * sizeof(uint64_t);
  if (!key.empty()) {
    QF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&P->pinned), count_bytes, hipHostMallocDefault));
    P->own_pinned = true;
  } else {
    P->pinned = stage_counts(ctx, count_bytes);
  }
  if (!key.empty()) {
    // Capture the step's device work once; replays skip all host planning.
    // Temporaries released during capture stay reserved for the graph's lifetime.  Phase
    // timing is off inside the graph (HIP cannot time event nodes of a graph); replays are
    // timed as a whole by the caller.
    const bool timing = ctx->timing;
    ctx->timing = false;
    ctx->capture_hold = &P->held;
    hipStream_t user_stream = ctx->stream;
    // Planning may have queued work on the caller's stream: order the graph after it too.
    QF_CHECK_HIP(hipEventRecord(ctx->graph_order_event, user_stream));
    QF_CHECK_HIP(hipStreamWaitEvent(ctx->graph_stream, ctx->graph_order_event, 0));
    ctx->stream = ctx->graph_stream;  // the step's work is captured from the private stream
    QF_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    try {
      enqueue_step(ctx, binding, num_positive, *P, d_bits, P->pinned);
    } catch (...) {
      hipGraph_t g;
      (void)hipStreamEndCapture(ctx->stream, &g);
      ctx->stream = user_stream;
      ctx->timing = timing;
      ctx->capture_hold = nullptr;
      throw;
    }
    hipGraph_t g;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    ctx->stream = user_stream;
    ctx->timing = timing;
    ctx->capture_hold = nullptr;
    QF_CHECK_HIP(ec);
    hipError_t e = hipGraphInstantiate(&P->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    QF_CHECK_HIP(e);
    QF_CHECK_HIP(hipGraphLaunch(P->graph, ctx->graph_stream));
    QF_CHECK_HIP(hipStreamSynchronize(ctx->graph_stream));
    std::memcpy(h_counts, P->pinned, (size_t)num_slots * 4 * sizeof(uint64_t));
    auto gc = std::make_shared<GraphCache>();
    gc->key = key;
    gc->plan = std::move(P);
    ctx->graph_cache = gc;
    return;
  }
  enqueue_step(ctx, binding, num_positive, *P, d_bits, P->pinned);
  sync(ctx);
  std::memcpy(h_counts, P->pinned, (size_t)num_slots * 4 * sizeof(uint64_t));
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
  o->partition_rows = 0;
  o->autotune = 0;
}

int qf_autotune_choice(qf_context* c, int* partition_rows, int* tile_rows, double* best_ms) {
  return guarded(c, [&](Context* ctx) {
    const TuneCache& tc = tune_cache(ctx);
    if (partition_rows) *partition_rows = tc.last.partition_rows;
    if (tile_rows) *tile_rows = tc.last.tile_rows;
    if (best_ms) *best_ms = tc.last.ms;
  });
}

static void check_table(const qf_table* t, const char* what) {
  QF_REQUIRE(t != nullptr, std::string(what) + " is null");
  QF_REQUIRE(t->num_rows >= 0 && t->num_rows < ((int64_t)1 << 31), std::string(what) + " row count out of range");
  QF_REQUIRE(t->num_columns >= 1 && t->columns != nullptr, std::string(what) + " has no columns");
  for (int c = 0; c < t->num_columns; ++c)
    QF_REQUIRE(t->num_rows == 0 || t->columns[c] != nullptr, std::string(what) + " has a null column");
}

}  // extern "C"

namespace qf {

// Literal groups of a step: literals grouped by (relation, join position, clause join variable)
// (learner/QuickFoil.cpp:142-161, CandidateLiteralEvaluator.cpp:281-289); PLAIN_LITERAL keeps one
// group per literal.
static std::vector<GroupSpec> literal_groups(int binding_columns, const qf_table* relations, const int* arg_types,
                                             int num_relations, const qf_literal* literals, int num_literals,
                                             const qf_eval_options& o) {
  std::vector<int> type_start(num_relations + 1, 0);
  for (int r = 0; r < num_relations; ++r) {
    check_table(&relations[r], "relation");
    QF_REQUIRE(relations[r].num_columns <= QF_MAX_ARITY, "relation arity above QF_MAX_ARITY");
    type_start[r + 1] = type_start[r] + relations[r].num_columns;
  }
  std::vector<GroupSpec> specs;
  // Group of each (relation, join position, clause join variable).
  std::vector<int> index((size_t)std::max(1, num_relations) * QF_MAX_ARITY * std::max(1, binding_columns), -1);
  for (int l = 0; l < num_literals; ++l) {
    const qf_literal& lit = literals[l];
    QF_REQUIRE(lit.relation >= 0 && lit.relation < num_relations, "literal relation out of range");
    const int arity = relations[lit.relation].num_columns;
    const int* types = arg_types ? arg_types + type_start[lit.relation] : nullptr;
    for (int i = 0; i < arity; ++i)
      QF_REQUIRE(lit.vars[i] >= -1 && lit.vars[i] < binding_columns, "literal variable out of range");
    const int jk = literal_join_key(lit, arity, types);
    QF_REQUIRE(jk >= 0, "literal has no bound variable");
    LitSpec ls;
    ls.slot = l;
    ls.pcols.reserve(arity);
    ls.bcols.reserve(arity);
    for (int i = 0; i < arity; ++i)
      if (lit.vars[i] >= 0) {
        ls.pcols.push_back(i);
        ls.bcols.push_back(lit.vars[i]);
      }
    normalise_key(ls);
    int& slot_of_key = index[((size_t)lit.relation * QF_MAX_ARITY + jk) * binding_columns + lit.vars[jk]];
    int gi;
    if (o.mode == QF_EVAL_PLAIN_LITERAL || slot_of_key < 0) {
      gi = (int)specs.size();
      GroupSpec g;
      g.relation = lit.relation;
      g.join_col = jk;
      g.join_var = lit.vars[jk];
      specs.push_back(g);
      if (o.mode != QF_EVAL_PLAIN_LITERAL) slot_of_key = gi;
    } else {
      gi = slot_of_key;
    }
    specs[gi].lits.push_back(std::move(ls));
  }
  return specs;
}

static void check_count_range(const uint64_t* h_counts, int num_literals) {
  for (int l = 0; l < num_literals; ++l)
    if (h_counts[4 * l + 2] > 0x7fffffffull || h_counts[4 * l + 3] > 0x7fffffffull)
      throw Error(QF_ERR_COUNT_RANGE, "a binding count exceeds int32 (values are exact in uint64)");
}

}  // namespace qf

extern "C" {

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
    // A call identical to the one the captured step last served replays it without regrouping
    // the literals or rebuilding the step's key.
    std::string call;
    if (o.use_graph && o.reuse_partitions && o.mode == QF_EVAL_BATCHED && num_literals > 0) {
      call_signature(call, *binding, num_positive, relations, arg_types, num_relations, literals, num_literals, o,
                     d_semi_bits);
      const auto* gc = static_cast<const GraphCache*>(ctx->graph_cache.get());
      if (gc && gc->plan && gc->plan->graph && gc->call == call) {
        replay_graph(ctx, *gc, h_counts, num_literals);
        check_count_range(h_counts, num_literals);
        return;
      }
    }
    std::vector<GroupSpec> specs =
        literal_groups(binding->num_columns, relations, arg_types, num_relations, literals, num_literals, o);
    run_groups(ctx, *binding, num_positive, relations, specs, num_literals, o, h_counts, d_semi_bits);
    // The cached graph now computes this call's step (replayed or just captured).
    if (!call.empty() && ctx->graph_cache) static_cast<GraphCache*>(ctx->graph_cache.get())->call = std::move(call);
    check_count_range(h_counts, num_literals);
  });
}

int qf_evaluate_literals_split(qf_context* c, const qf_table* positives, const qf_table* negatives,
                               const qf_table* relations, const int* arg_types, int num_relations,
                               const qf_literal* literals, int num_literals, const qf_eval_options* opts,
                               uint64_t* h_counts, uint64_t* d_semi_bits_pos, uint64_t* d_semi_bits_neg) {
  return guarded(c, [&](Context* ctx) {
    check_table(positives, "positive binding table");
    check_table(negatives, "negative binding table");
    QF_REQUIRE(positives->num_columns == negatives->num_columns, "the two binding tables differ in columns");
    QF_REQUIRE(num_literals >= 0 && (num_literals == 0 || (literals && h_counts)), "null literals or counts");
    QF_REQUIRE(num_relations >= 0 && (num_relations == 0 || relations), "null relations");
    qf_eval_options o;
    qf_eval_options_default(&o);
    if (opts) o = *opts;
    o.use_graph = 0;  // the two passes would replace each other's captured step
    std::vector<GroupSpec> specs =
        literal_groups(positives->num_columns, relations, arg_types, num_relations, literals, num_literals, o);
    // ExecuteOnPositives, then ExecuteOnNegatives (CandidateLiteralEvaluator.cpp:345-405): every
    // row of the first table is positive, none of the second.
    std::vector<uint64_t> pos((size_t)std::max(1, num_literals) * 4), neg((size_t)std::max(1, num_literals) * 4);
    run_groups(ctx, *positives, positives->num_rows, relations, specs, num_literals, o, pos.data(), d_semi_bits_pos);
    run_groups(ctx, *negatives, 0, relations, specs, num_literals, o, neg.data(), d_semi_bits_neg);
    for (int l = 0; l < num_literals; ++l) {
      h_counts[4 * l + 0] = pos[4 * l + 0];
      h_counts[4 * l + 1] = neg[4 * l + 1];
      h_counts[4 * l + 2] = pos[4 * l + 2];
      h_counts[4 * l + 3] = neg[4 * l + 3];
    }
    check_count_range(h_counts, num_literals);
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
