// Materialising joins and semi-joins: CreateLabelAwareBindingTables / CreateBindingTable
// (operations/MultiColumnHashJoin.cpp), LeftSemiJoin / RightSemiJoin, and the dedup table of
// BuildHashTableAfterSemiJoin.
//
// Row order.  The reference walks the outer (probe) side in row order and, for each outer row,
// the chain of its bucket in the inner (build) table, which lists equal-key inner rows in
// descending row order (head insertion, operations/BuildHashTable.cpp:118-122 with
// DoBlockJoin, MultiColumnHashJoin.cpp:496-528).  Here the inner rows are stably sorted by the
// HBM hash-table slot of their key, so each distinct key owns one segment in ascending row
// order; an exclusive scan of per-outer-row match counts gives output offsets, and one thread
// per output row finds its outer row by binary search and reads its inner row from the end of
// the segment.  The output is the reference's, row for row.
#include "qf_internal.h"

#include <algorithm>

namespace qf {

struct Keys {
  const int32_t* c[QF_MAX_ARITY];
  int n;
};

__device__ inline uint32_t key_hash(const Keys& k, int64_t row) {
  uint32_t h = 0x7f4a7c15u;
  for (int i = 0; i < k.n; ++i) h = (h ^ (uint32_t)k.c[i][row]) * 0x9e3779b1u + (uint32_t)i;
  return mix32(h);
}

__device__ inline bool key_eq(const Keys& a, int64_t ra, const Keys& b, int64_t rb) {
  for (int i = 0; i < a.n; ++i)
    if (a.c[i][ra] != b.c[i][rb]) return false;
  return true;
}

// Inserts every row of `k` (rows with bit clear in `only` are skipped when only != null) into
// an open-addressing set; state[slot] = 1 + representative row.  slot_of[row] = its slot or -1.
__global__ void k_keyset_build(Keys k, int64_t n, const uint64_t* only, uint32_t* state, uint32_t cap,
                               int32_t* slot_of) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (only && !((only[i >> 6] >> (i & 63)) & 1ull)) {
    if (slot_of) slot_of[i] = -1;
    return;
  }
  uint32_t slot = key_hash(k, i) & (cap - 1);
  while (true) {
    const uint32_t prev = atomicCAS(&state[slot], 0u, (uint32_t)(i + 1));
    if (prev == 0 || key_eq(k, i, k, prev - 1)) break;
    slot = (slot + 1) & (cap - 1);
  }
  if (slot_of) slot_of[i] = (int32_t)slot;
}

__device__ inline int32_t keyset_find(const Keys& probe, int64_t row, const Keys& inner, const uint32_t* state,
                                      uint32_t cap) {
  uint32_t slot = key_hash(probe, row) & (cap - 1);
  while (true) {
    const uint32_t st = state[slot];
    if (st == 0) return -1;
    if (key_eq(probe, row, inner, st - 1)) return (int32_t)slot;
    slot = (slot + 1) & (cap - 1);
  }
}

// Semi-join bitvector over `probe` rows: bit = key present in the set; 64 rows per ballot.
__global__ void k_semi_bits(Keys probe, int64_t n, Keys inner, const uint32_t* state, uint32_t cap, uint64_t* bits,
                            unsigned long long* ones) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool found = i < n && keyset_find(probe, i, inner, state, cap) >= 0;
  const uint64_t word = __ballot(found);
  if ((threadIdx.x & 63) == 0 && i < n) {
    bits[i >> 6] = word;
    if (word) atomicAdd(ones, (unsigned long long)__popcll(word));
  }
}

// Per outer row: matching inner segment [seg_begin, seg_end) and its size; counts[n] = 0, so
// that an exclusive scan over n + 1 entries ends with the total.
__global__ void k_match_counts(Keys outer, int64_t n, Keys inner, const uint32_t* state, uint32_t cap,
                               const int64_t* seg_offsets, int64_t* counts, int64_t* seg_end) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == n) counts[n] = 0;
  if (i >= n) return;
  const int32_t slot = keyset_find(outer, i, inner, state, cap);
  if (slot < 0) {
    counts[i] = 0;
    seg_end[i] = 0;
  } else {
    counts[i] = seg_offsets[slot + 1] - seg_offsets[slot];
    seg_end[i] = seg_offsets[slot + 1];
  }
}

struct ProjectArgs {
  const int32_t* outer_cols[16];
  const int32_t* inner_cols[16];
  int32_t* out[16];
  int from_outer[16];  // 1: column of the outer row, 0: of the inner row
  int ncols;
};

__global__ void k_project(ProjectArgs p, const int64_t* out_off, const int64_t* seg_end, const int32_t* inner_sorted,
                          int64_t n_outer, int64_t total) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  // Outer row o: last o with out_off[o] <= e.
  int64_t lo = 0, hi = n_outer - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (out_off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const int64_t o = lo;
  const int64_t k = e - out_off[o];
  const int64_t inner = inner_sorted[seg_end[o] - 1 - k];  // descending inner rows
  for (int c = 0; c < p.ncols; ++c) p.out[c][e] = p.from_outer[c] ? p.outer_cols[c][o] : p.inner_cols[c][inner];
}

// k_project for a caller-presized output: the row count is read on the device (out_off[n_outer],
// the scan's last entry) and recorded in *total_out; the rows are written only when they fit
// `capacity`, otherwise nothing is (the caller retries with the exact count).
__global__ void k_project_presized(ProjectArgs p, const int64_t* out_off, const int64_t* seg_end,
                                   const int32_t* inner_sorted, int64_t n_outer, int64_t capacity,
                                   int64_t* total_out) {
  const int64_t total = out_off[n_outer];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e == 0) *total_out = total;
  if (total > capacity || e >= total) return;
  int64_t lo = 0, hi = n_outer - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (out_off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const int64_t k = e - out_off[lo];
  const int64_t inner = inner_sorted[seg_end[lo] - 1 - k];
  for (int c = 0; c < p.ncols; ++c) p.out[c][e] = p.from_outer[c] ? p.outer_cols[c][lo] : p.inner_cols[c][inner];
}

__global__ void k_popcount_total(const uint64_t* bits, int64_t words, unsigned long long* ones) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = i < words ? (uint32_t)__popcll(bits[i]) : 0u;
  if (c) atomicAdd(ones, (unsigned long long)c);
}

// ------------------------------------------------------------------------------------------

static Keys make_keys(const std::vector<const int32_t*>& cols) {
  QF_REQUIRE(!cols.empty() && cols.size() <= QF_MAX_ARITY, "1..6 key columns");
  Keys k{};
  k.n = (int)cols.size();
  for (int i = 0; i < k.n; ++i) k.c[i] = cols[i];
  return k;
}

static int log2_u32(uint32_t p) {
  int b = 0;
  while ((1u << b) < p) ++b;
  return b;
}

struct KeySet {
  DevBuf state;
  uint32_t cap = 0;
};

static KeySet build_keyset(Context* ctx, const Keys& k, int64_t n, const uint64_t* only, int32_t* slot_of) {
  KeySet s;
  s.cap = next_pow2_u32((uint32_t)std::max<int64_t>(16, 2 * n));
  s.state = DevBuf(ctx, (size_t)s.cap * 4);
  fill_i32(ctx, s.state.as<int32_t>(), 0, s.cap);
  if (n > 0)
    hipLaunchKernelGGL(k_keyset_build, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, k, n, only,
                       s.state.as<uint32_t>(), s.cap, slot_of);
  QF_CHECK_HIP(hipGetLastError());
  return s;
}

// The inner side of an ordered join: its key set, and its rows stably sorted by key slot (each
// distinct key owns one segment of ascending rows).
struct InnerSide {
  Keys k{};
  int64_t n = 0;
  KeySet set;
  DevBuf sorted_row, seg;
};

static void prepare_inner(Context* ctx, const std::vector<const int32_t*>& inner_keys, int64_t n_inner, InnerSide& in) {
  in.k = make_keys(inner_keys);
  in.n = n_inner;
  if (n_inner == 0) return;
  DevBuf slot_of(ctx, n_inner * 4);
  in.set = build_keyset(ctx, in.k, n_inner, nullptr, slot_of.as<int32_t>());
  DevBuf sorted_slot(ctx, n_inner * 4);
  in.sorted_row = DevBuf(ctx, n_inner * 4);
  in.seg = DevBuf(ctx, ((size_t)in.set.cap + 1) * 8);
  const int32_t* src[1] = {slot_of.as<int32_t>()};
  int32_t* out[2] = {sorted_slot.as<int32_t>(), in.sorted_row.as<int32_t>()};
  stable_partition(ctx, src, 1, 0, true, n_inner, 0, log2_u32(in.set.cap), out, 1, in.seg.as<int64_t>());
}

// The outer side: per outer row the size and end of its inner segment, and the output offsets.
struct OuterSide {
  int64_t n = 0;
  int64_t total = 0;
  DevBuf counts, seg_end, out_off;
};

// Counts the output rows of outer x inner: the scan's last entry, copied to the host unless
// `to_host` is false (the count then stays on the device at out_off[n_outer]).
static int64_t count_outer(Context* ctx, const std::vector<const int32_t*>& outer_keys, int64_t n_outer,
                           const InnerSide& in, OuterSide& o, bool to_host = true) {
  o.n = n_outer;
  o.total = 0;
  if (n_outer == 0 || in.n == 0) return 0;
  const Keys ko = make_keys(outer_keys);
  o.counts = DevBuf(ctx, (n_outer + 1) * 8);
  o.seg_end = DevBuf(ctx, n_outer * 8);
  o.out_off = DevBuf(ctx, (n_outer + 1) * 8);
  hipLaunchKernelGGL(k_match_counts, dim3((unsigned)ceil_div(n_outer + 1, 256)), dim3(256), 0, ctx->stream, ko,
                     n_outer, in.k, in.set.state.as<uint32_t>(), in.set.cap, in.seg.as<int64_t>(),
                     o.counts.as<int64_t>(), o.seg_end.as<int64_t>());
  QF_CHECK_HIP(hipGetLastError());
  scan_exclusive_i64(ctx, o.counts.as<int64_t>(), o.out_off.as<int64_t>(), n_outer + 1);
  if (to_host) copy_d2h(ctx, &o.total, o.out_off.as<int64_t>() + n_outer, 8);
  return o.total;
}

// Writes the joined rows (outer rows in order, equal-key inner rows in descending order).
static void project(Context* ctx, const OuterSide& o, const InnerSide& in, const ProjectArgs& proj) {
  if (o.total == 0) return;
  hipLaunchKernelGGL(k_project, dim3((unsigned)ceil_div(o.total, 256)), dim3(256), 0, ctx->stream, proj,
                     o.out_off.as<int64_t>(), o.seg_end.as<int64_t>(), in.sorted_row.as<int32_t>(), o.n, o.total);
  QF_CHECK_HIP(hipGetLastError());
}

// Outer rows in order, inner equal-key rows in descending order; projections per output column.
// Returns the row count; writes only when out != null and the count equals expect (expect < 0:
// any count).
static int64_t ordered_join(Context* ctx, const std::vector<const int32_t*>& outer_keys, int64_t n_outer,
                            const std::vector<const int32_t*>& inner_keys, int64_t n_inner,
                            const ProjectArgs& proj, bool write, int64_t expect, int64_t capacity = -1) {
  if (n_outer == 0 || n_inner == 0) return 0;
  InnerSide in;
  prepare_inner(ctx, inner_keys, n_inner, in);
  OuterSide o;
  const int64_t total = count_outer(ctx, outer_keys, n_outer, in, o);
  if (!write || total == 0 || (expect >= 0 && total != expect) || (capacity >= 0 && total > capacity)) return total;
  project(ctx, o, in, proj);
  sync(ctx);
  return total;
}

static void check_tab(const qf_table* t, const char* what) {
  QF_REQUIRE(t && t->columns && t->num_columns >= 1, std::string(what) + ": bad table");
  QF_REQUIRE(t->num_rows >= 0 && t->num_rows < ((int64_t)1 << 31), std::string(what) + ": row count out of range");
}

}  // namespace qf

using namespace qf;

extern "C" {

int qf_create_label_aware_binding_table(qf_context* c, const qf_table* binding, int64_t num_positive,
                                        const qf_table* relation, const qf_literal* literal, int64_t num_binding_pos,
                                        int64_t num_binding_neg, int32_t* const* d_out_columns) {
  return guarded(c, [&](Context* ctx) {
    check_tab(binding, "binding table");
    check_tab(relation, "relation");
    QF_REQUIRE(literal && d_out_columns, "null literal or outputs");
    QF_REQUIRE(num_positive >= 0 && num_positive <= binding->num_rows, "num_positive out of range");
    const int C = binding->num_columns, arity = relation->num_columns;
    QF_REQUIRE(arity <= QF_MAX_ARITY, "arity above QF_MAX_ARITY");
    const int64_t npos = num_positive, nneg = binding->num_rows - num_positive, nbg = relation->num_rows;
    std::vector<const int32_t*> bg_keys, pos_keys, neg_keys;
    std::vector<int> unbound;
    for (int i = 0; i < arity; ++i) {
      const int v = literal->vars[i];
      QF_REQUIRE(v >= -1 && v < C, "literal variable out of range");
      if (v >= 0) {
        bg_keys.push_back(relation->columns[i]);
        pos_keys.push_back(binding->columns[v]);
        neg_keys.push_back(binding->columns[v] + npos);
      } else {
        unbound.push_back(i);
      }
    }
    QF_REQUIRE(!bg_keys.empty(), "literal has no bound variable");
    const int ncols = C + (int)unbound.size();
    QF_REQUIRE(ncols <= 16, "too many output columns");
    QF_REQUIRE(num_binding_pos >= 0 && num_binding_neg >= 0, "negative sizes");
    // Which side is probed follows MultiColumnHashJoin.cpp:160-249.  Each section is joined
    // once: both sizes are checked against the evaluator's counts before anything is written (the
    // reference writes into its exact presize unchecked, :144-159).
    const bool bg_outer = npos < nbg && nneg < nbg;
    InnerSide inner[2];
    OuterSide outer[2];
    ProjectArgs proj[2];
    if (!bg_outer) prepare_inner(ctx, bg_keys, nbg, inner[0]);  // the relation is the inner side of both
    for (int label = 0; label < 2; ++label) {
      const int64_t nb = label == 0 ? npos : nneg;
      const int64_t base = label == 0 ? 0 : npos;
      const int64_t expect = label == 0 ? num_binding_pos : num_binding_neg;
      ProjectArgs& p = proj[label];
      p = ProjectArgs{};
      p.ncols = ncols;
      for (int col = 0; col < ncols; ++col) {
        p.out[col] = d_out_columns[col] + (label == 0 ? 0 : num_binding_pos);
        const bool clause_col = col < C;
        const int32_t* src = clause_col ? binding->columns[col] + base : relation->columns[unbound[col - C]];
        p.from_outer[col] = bg_outer ? !clause_col : clause_col;
        if (p.from_outer[col]) p.outer_cols[col] = src;
        else p.inner_cols[col] = src;
      }
      const std::vector<const int32_t*>& bkeys = label == 0 ? pos_keys : neg_keys;
      int64_t total = 0;
      if (nb > 0 && nbg > 0) {
        if (bg_outer) {
          prepare_inner(ctx, bkeys, nb, inner[label]);
          total = count_outer(ctx, bg_keys, nbg, inner[label], outer[label]);
        } else {
          total = count_outer(ctx, bkeys, nb, inner[0], outer[label]);
        }
      }
      if (total != expect)
        throw Error(QF_ERR_OVERFLOW, "join produced " + std::to_string(total) + " rows for a " +
                                         (label == 0 ? "positive" : "negative") + " section sized " +
                                         std::to_string(expect));
    }
    for (int label = 0; label < 2; ++label)
      if (outer[label].total > 0) project(ctx, outer[label], inner[bg_outer ? label : 0], proj[label]);
    sync(ctx);
  });
}

static int create_binding_table(qf_context* c, const qf_table* binding, const qf_table* relation,
                                const qf_literal* literal, int32_t* const* d_out_columns, int64_t* h_num_rows,
                                int64_t capacity) {
  // A presized call whose rows exceed the capacity is an expected outcome (the caller retries
  // with the exact size): it returns QF_ERR_OVERFLOW without going through an exception.
  bool over_capacity = false;
  const int st = guarded(c, [&](Context* ctx) {
    check_tab(binding, "binding table");
    check_tab(relation, "relation");
    QF_REQUIRE(literal && h_num_rows, "null literal or row count");
    const int C = binding->num_columns, arity = relation->num_columns;
    QF_REQUIRE(arity <= QF_MAX_ARITY, "arity above QF_MAX_ARITY");
    std::vector<const int32_t*> lit_keys, bind_keys;
    std::vector<int> unbound;
    for (int i = 0; i < arity; ++i) {
      const int v = literal->vars[i];
      if (v >= 0 && v < C) {  // CreateBindingTable: ids below the column count are join keys
        lit_keys.push_back(relation->columns[i]);
        bind_keys.push_back(binding->columns[v]);
      } else {
        unbound.push_back(i);
      }
    }
    QF_REQUIRE(!lit_keys.empty(), "literal has no bound variable");
    const int ncols = C + (int)unbound.size();
    QF_REQUIRE(ncols <= 16, "too many output columns");
    const int64_t N = binding->num_rows, nbg = relation->num_rows;
    const bool binding_outer = nbg < N;  // MultiColumnHashJoin.cpp:55-98
    ProjectArgs p{};
    p.ncols = ncols;
    for (int col = 0; col < ncols; ++col) {
      p.out[col] = d_out_columns ? d_out_columns[col] : nullptr;
      const bool clause_col = col < C;
      const int32_t* src = clause_col ? binding->columns[col] : relation->columns[unbound[col - C]];
      p.from_outer[col] = binding_outer ? clause_col : !clause_col;
      if (p.from_outer[col]) p.outer_cols[col] = src;
      else p.inner_cols[col] = src;
    }
    const bool write = d_out_columns != nullptr;
    const int64_t expect = write && capacity < 0 ? *h_num_rows : -1;
    int64_t total;
    if (binding_outer) total = ordered_join(ctx, bind_keys, N, lit_keys, nbg, p, write, expect, capacity);
    else total = ordered_join(ctx, lit_keys, nbg, bind_keys, N, p, write, expect, capacity);
    *h_num_rows = total;
    if (write && capacity < 0 && total != expect) throw Error(QF_ERR_OVERFLOW, "output size differs from the row count");
    over_capacity = write && capacity >= 0 && total > capacity;
  });
  if (st == QF_OK && over_capacity) {
    reinterpret_cast<Context*>(c)->last_error = "output rows exceed the capacity";
    return QF_ERR_OVERFLOW;
  }
  return st;
}

int qf_create_binding_table(qf_context* c, const qf_table* binding, const qf_table* relation,
                            const qf_literal* literal, int32_t* const* d_out_columns, int64_t* h_num_rows) {
  return create_binding_table(c, binding, relation, literal, d_out_columns, h_num_rows, -1);
}

int qf_create_binding_table_presized(qf_context* c, const qf_table* binding, const qf_table* relation,
                                     const qf_literal* literal, int64_t capacity, int32_t* const* d_out_columns,
                                     int64_t* h_num_rows) {
  if (capacity < 0 || !d_out_columns) return QF_ERR_INVALID;
  return create_binding_table(c, binding, relation, literal, d_out_columns, h_num_rows, capacity);
}

int qf_create_binding_tables_presized(qf_context* c, const qf_table* binding, const qf_table* relations,
                                      int num_relations, const qf_literal* literals, int num_literals,
                                      const int64_t* capacities, int32_t* const* const* d_out_columns,
                                      int64_t* h_num_rows, uint8_t* h_fits) {
  bool any_over = false;
  const int st = guarded(c, [&](Context* ctx) {
    check_tab(binding, "binding table");
    QF_REQUIRE(num_literals >= 0 && num_relations >= 0, "negative counts");
    QF_REQUIRE(num_literals == 0 || (literals && capacities && d_out_columns && h_num_rows && h_fits),
               "null literal arguments");
    const int C = binding->num_columns;
    const int64_t N = binding->num_rows;
    // Every join is enqueued before any result is read.  A join's temporaries go back to the
    // pool after its launches: later work on the stream reuses them only after those finish.
    DevBuf totals(ctx, (size_t)std::max(1, num_literals) * 8);
    fill_i64(ctx, totals.as<int64_t>(), 0, std::max(1, num_literals));
    for (int i = 0; i < num_literals; ++i) {
      const qf_literal& lit = literals[i];
      QF_REQUIRE(lit.relation >= 0 && lit.relation < num_relations, "literal relation out of range");
      const qf_table& rel = relations[lit.relation];
      check_tab(&rel, "relation");
      QF_REQUIRE(rel.num_columns <= QF_MAX_ARITY, "arity above QF_MAX_ARITY");
      QF_REQUIRE(capacities[i] >= 0 && d_out_columns[i], "bad capacity or outputs");
      std::vector<const int32_t*> lit_keys, bind_keys;
      std::vector<int> unbound;
      for (int k = 0; k < rel.num_columns; ++k) {
        const int v = lit.vars[k];
        if (v >= 0 && v < C) {  // CreateBindingTable: ids below the column count are join keys
          lit_keys.push_back(rel.columns[k]);
          bind_keys.push_back(binding->columns[v]);
        } else {
          unbound.push_back(k);
        }
      }
      QF_REQUIRE(!lit_keys.empty(), "literal has no bound variable");
      const int ncols = C + (int)unbound.size();
      QF_REQUIRE(ncols <= 16, "too many output columns");
      const int64_t nbg = rel.num_rows;
      const bool binding_outer = nbg < N;  // MultiColumnHashJoin.cpp:55-98
      ProjectArgs p{};
      p.ncols = ncols;
      for (int col = 0; col < ncols; ++col) {
        p.out[col] = d_out_columns[i][col];
        const bool clause_col = col < C;
        const int32_t* src = clause_col ? binding->columns[col] : rel.columns[unbound[col - C]];
        p.from_outer[col] = binding_outer ? clause_col : !clause_col;
        if (p.from_outer[col]) p.outer_cols[col] = src;
        else p.inner_cols[col] = src;
      }
      if (N == 0 || nbg == 0) continue;  // no rows; its total stays 0
      InnerSide in;
      OuterSide out;
      if (binding_outer) {
        prepare_inner(ctx, lit_keys, nbg, in);
        count_outer(ctx, bind_keys, N, in, out, false);
      } else {
        prepare_inner(ctx, bind_keys, N, in);
        count_outer(ctx, lit_keys, nbg, in, out, false);
      }
      hipLaunchKernelGGL(k_project_presized, dim3((unsigned)std::max<int64_t>(1, ceil_div(capacities[i], 256))),
                         dim3(256), 0, ctx->stream, p, out.out_off.as<int64_t>(), out.seg_end.as<int64_t>(),
                         in.sorted_row.as<int32_t>(), out.n, capacities[i], totals.as<int64_t>() + i);
      QF_CHECK_HIP(hipGetLastError());
    }
    if (num_literals > 0) copy_d2h(ctx, h_num_rows, totals.get(), (size_t)num_literals * 8);
    for (int i = 0; i < num_literals; ++i) {
      h_fits[i] = h_num_rows[i] <= capacities[i] ? 1 : 0;
      any_over = any_over || !h_fits[i];
    }
  });
  if (st == QF_OK && any_over) {
    reinterpret_cast<Context*>(c)->last_error = "some joins exceed their capacity (see h_fits)";
    return QF_ERR_OVERFLOW;
  }
  return st;
}

static int semi_join(qf_context* c, const qf_table* probe_keys, const qf_table* build_keys, uint64_t* d_out_bits,
                     int64_t* h_num_ones, bool left) {
  return guarded(c, [&](Context* ctx) {
    check_tab(probe_keys, "probe keys");
    check_tab(build_keys, "build keys");
    QF_REQUIRE(probe_keys->num_columns == build_keys->num_columns, "key column counts differ");
    QF_REQUIRE(d_out_bits, "null output");
    std::vector<const int32_t*> pk(probe_keys->columns, probe_keys->columns + probe_keys->num_columns);
    std::vector<const int32_t*> bk(build_keys->columns, build_keys->columns + build_keys->num_columns);
    // Left: bits over probe rows, set = build keys.  Right: bits over build rows, set = probe keys.
    const std::vector<const int32_t*>& out_side = left ? pk : bk;
    const std::vector<const int32_t*>& set_side = left ? bk : pk;
    const int64_t n_out = left ? probe_keys->num_rows : build_keys->num_rows;
    const int64_t n_set = left ? build_keys->num_rows : probe_keys->num_rows;
    const Keys ko = make_keys(out_side), ks = make_keys(set_side);
    DevBuf ones(ctx, 8);
    fill_u64(ctx, ones.as<uint64_t>(), 0, 1);
    if (n_out > 0) {
      KeySet set = build_keyset(ctx, ks, n_set, nullptr, nullptr);
      hipLaunchKernelGGL(k_semi_bits, dim3((unsigned)ceil_div(n_out, 256)), dim3(256), 0, ctx->stream, ko, n_out, ks,
                         set.state.as<uint32_t>(), set.cap, d_out_bits, ones.as<unsigned long long>());
      QF_CHECK_HIP(hipGetLastError());
      sync(ctx);
    }
    int64_t h = 0;
    copy_d2h(ctx, &h, ones.get(), 8);
    if (h_num_ones) *h_num_ones = h;
  });
}

int qf_left_semi_join(qf_context* c, const qf_table* probe_keys, const qf_table* build_keys, uint64_t* d_out_bits,
                      int64_t* h_num_ones) {
  return semi_join(c, probe_keys, build_keys, d_out_bits, h_num_ones, true);
}

int qf_right_semi_join(qf_context* c, const qf_table* probe_keys, const qf_table* build_keys, uint64_t* d_out_bits,
                       int64_t* h_num_ones) {
  return semi_join(c, probe_keys, build_keys, d_out_bits, h_num_ones, false);
}

}  // extern "C"

namespace qf {

// pass 0: minimum set row per key slot; pass 1: first-occurrence flag and sort key
// (1 + reference bucket for inserted rows, 0 otherwise).
__global__ void k_first_of_key(Keys k, int64_t n, const uint64_t* bits, const int32_t* slot_of, uint32_t* minrow,
                               int pass, uint8_t* first, int32_t* sort_key, uint32_t mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool set = (bits[i >> 6] >> (i & 63)) & 1ull;
  if (pass == 0) {
    if (set) atomicMin(&minrow[slot_of[i]], (uint32_t)i);
    return;
  }
  const bool f = set && minrow[slot_of[i]] == (uint32_t)i;
  first[i] = f ? 1 : 0;
  if (f) {
    // Reference bucket: Hash<num_keys> with HashCombine, & (P2(n) - 1) (BuildHashTable.cpp:139).
    uint32_t h = ref_hash(k.c[0][i]);
    for (int c = 1; c < k.n; ++c) h = ref_hash_combine(h, k.c[c][i]);
    sort_key[i] = (int32_t)(h & mask) + 1;
  } else {
    sort_key[i] = 0;
  }
}

// Rows sorted by (1 + bucket), ascending row within a bucket; key 0 = not inserted.
__global__ void k_link_inserted(const int32_t* skey, const int32_t* srow, int64_t n, int32_t* buckets,
                                int32_t* next) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t key = skey[k];
  if (key == 0) {
    next[srow[k]] = 0;
    return;
  }
  const bool has_prev = k > 0 && skey[k - 1] == key;
  next[srow[k]] = has_prev ? srow[k - 1] + 1 : 0;
  if (k == n - 1 || skey[k + 1] != key) buckets[key - 1] = srow[k] + 1;
}

}  // namespace qf

extern "C" int qf_build_hash_table_after_semi_join(qf_context* c, const qf_table* keys, const uint64_t* d_bits,
                                                   int32_t* d_buckets, int32_t* d_next, uint8_t* d_first) {
  return guarded(c, [&](Context* ctx) {
    check_tab(keys, "keys");
    QF_REQUIRE(keys->num_columns <= QF_MAX_ARITY, "1..6 key columns");
    QF_REQUIRE(d_bits && d_buckets && d_next && d_first, "null argument");
    const int64_t n = keys->num_rows;
    if (n == 0) return;
    std::vector<const int32_t*> kc(keys->columns, keys->columns + keys->num_columns);
    const Keys k = make_keys(kc);
    DevBuf slot_of(ctx, n * 4);
    KeySet set = build_keyset(ctx, k, n, d_bits, slot_of.as<int32_t>());
    DevBuf minrow(ctx, (size_t)set.cap * 4);
    fill_i32(ctx, minrow.as<int32_t>(), -1, set.cap);  // 0xffffffff as uint32
    const uint32_t nb = next_pow2_u32((uint32_t)n);
    DevBuf skey(ctx, n * 4);
    const dim3 grid((unsigned)ceil_div(n, 256));
    hipLaunchKernelGGL(k_first_of_key, grid, dim3(256), 0, ctx->stream, k, n, d_bits, slot_of.as<int32_t>(),
                       minrow.as<uint32_t>(), 0, d_first, skey.as<int32_t>(), nb - 1u);
    hipLaunchKernelGGL(k_first_of_key, grid, dim3(256), 0, ctx->stream, k, n, d_bits, slot_of.as<int32_t>(),
                       minrow.as<uint32_t>(), 1, d_first, skey.as<int32_t>(), nb - 1u);
    QF_CHECK_HIP(hipGetLastError());
    fill_i32(ctx, d_buckets, 0, nb);
    DevBuf sk(ctx, n * 4), sr(ctx, n * 4);
    const int32_t* in[1] = {skey.as<int32_t>()};
    int32_t* out[2] = {sk.as<int32_t>(), sr.as<int32_t>()};
    stable_sort_by_key(ctx, in, 1, 0, true, n, log2_u32(nb) + 1, out);
    hipLaunchKernelGGL(k_link_inserted, grid, dim3(256), 0, ctx->stream, sk.as<int32_t>(), sr.as<int32_t>(), n,
                       d_buckets, d_next);
    QF_CHECK_HIP(hipGetLastError());
    sync(ctx);
  });
}

// Example coverage of one literal (the semi-join chain of learner/QuickFoil.cpp:471-573 and
// learner/LiteralSelector.cpp:150-210): binding rows [row_begin, row_end) that match the literal
// on its bound arguments, projected to their first examples->num_columns columns (the head), then
// bit e of d_out_bits = example row e equals one of those projections.
extern "C" int qf_example_coverage(qf_context* c, const qf_table* binding, int64_t row_begin, int64_t row_end,
                                   const qf_table* relation, const qf_literal* literal, const qf_table* examples,
                        uint64_t* d_out_bits, int64_t* h_num_covered) {
  return guarded(c, [&](Context* ctx) {
    check_tab(binding, "binding");
    check_tab(relation, "relation");
    check_tab(examples, "examples");
    QF_REQUIRE(literal && d_out_bits, "null argument");
    QF_REQUIRE(row_begin >= 0 && row_begin <= row_end && row_end <= binding->num_rows, "row range out of bounds");
    QF_REQUIRE(examples->num_columns <= binding->num_columns, "examples wider than the binding table");
    std::vector<const int32_t*> bkeys, rkeys, head, ex;
    for (int pos = 0; pos < relation->num_columns; ++pos) {
      const int v = literal->vars[pos];
      if (v < 0) continue;
      QF_REQUIRE(v < binding->num_columns, "literal variable outside the binding table");
      bkeys.push_back(binding->columns[v] + row_begin);
      rkeys.push_back(relation->columns[pos]);
    }
    QF_REQUIRE(!bkeys.empty(), "literal has no bound variable");
    for (int col = 0; col < examples->num_columns; ++col) {
      head.push_back(binding->columns[col] + row_begin);
      ex.push_back(examples->columns[col]);
    }
    const int64_t n = row_end - row_begin, ne = examples->num_rows;
    DevBuf ones(ctx, 8);
    fill_u64(ctx, ones.as<uint64_t>(), 0, 1);
    if (ne > 0) fill_u64(ctx, d_out_bits, 0, ceil_div(ne, 64));
    if (n > 0 && ne > 0 && relation->num_rows > 0) {
      // Binding rows that match the literal.
      DevBuf matched(ctx, (size_t)ceil_div(n, 64) * 8);
      DevBuf scratch(ctx, 8);
      const Keys kb = make_keys(bkeys), kr = make_keys(rkeys);
      KeySet facts = build_keyset(ctx, kr, relation->num_rows, nullptr, nullptr);
      hipLaunchKernelGGL(k_semi_bits, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, kb, n, kr,
                         facts.state.as<uint32_t>(), facts.cap, matched.as<uint64_t>(),
                         scratch.as<unsigned long long>());
      QF_CHECK_HIP(hipGetLastError());
      // Their head projections, probed by the examples.
      const Keys kh = make_keys(head), ke = make_keys(ex);
      KeySet heads = build_keyset(ctx, kh, n, matched.as<uint64_t>(), nullptr);
      hipLaunchKernelGGL(k_semi_bits, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0, ctx->stream, ke, ne, kh,
                         heads.state.as<uint32_t>(), heads.cap, d_out_bits, ones.as<unsigned long long>());
      QF_CHECK_HIP(hipGetLastError());
      sync(ctx);
    }
    int64_t h = 0;
    copy_d2h(ctx, &h, ones.get(), 8);
    if (h_num_covered) *h_num_covered = h;
  });
}
