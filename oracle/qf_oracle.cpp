/*
 * CPU oracle for the candidate-literal evaluation path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch restatement of the QuickFOIL operator pipeline (jbalint/quickfoil) used to
 * check the HIP product.  It is loaded only by tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg.  Nothing in the product links or calls it.
 *
 * What pins each function (checks that do not retype the function's own formula):
 *  - orc_hash_*: operations/RadixPartition_test.cpp:106-168 presumes the identity hash for
 *    int32 (value i lands in partition i % 2^r); HashCombine is checked in the tests against
 *    boost's published hash_combine on hand-computed values.
 *  - orc_radix_partition: the three known-answer cases of RadixPartition_test.cpp:106-168
 *    (r in {1,3,7,12}, sizes 10..100000), plus multiset and stability invariants.
 *  - orc_build_hash_table_*: every key is reachable from its bucket and each chain holds
 *    exactly the rows of that bucket in descending position (brute force per bucket).
 *  - orc_evaluate_pipeline: equal to orc_evaluate_bruteforce (nested loops over the definition
 *    of the counts) on small random inputs and on the uncle sample data, and to
 *    orc_evaluate_grouped at any size; invariants covered <= binding, covered_pos <= N_pos.
 *  - orc_generate_plan: every literal slot's conjunction equals the literal's atom set.
 *  - orc_create_*binding_table: output multiset equals the brute-force extended join and the
 *    row count equals binding_pos + binding_neg from the brute-force counts.
 *  - semi-joins and dedup: brute force (exists / first occurrence) on small inputs.
 */
#include "qf_oracle.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <unordered_map>
#include <vector>

namespace {

typedef std::vector<uint64_t> Bits;  // dynamic_bitset<unsigned long> block layout

inline void bit_set(Bits& b, int64_t i) { b[i >> 6] |= (1ull << (i & 63)); }
inline bool bit_test(const Bits& b, int64_t i) { return (b[i >> 6] >> (i & 63)) & 1ull; }
// BitVector::test_set: returns the previous value and sets the bit.
inline bool bit_test_set(Bits& b, int64_t i) {
  const bool old = bit_test(b, i);
  bit_set(b, i);
  return old;
}
inline int64_t nwords(int64_t n) { return (n + 63) >> 6; }

// ---------------------------------------------------------------------------------------
// Chained hash table: buckets hold 1-based heads, next holds 1-based links, 0 ends a chain
// (storage/FoilHashTable.cpp:34-49).
struct ChainTable {
  uint32_t mask = 0;
  std::vector<int32_t> buckets;
  std::vector<int32_t> next;
};

// FoilHashTable(size, radix_bits): P2(size) buckets, mask = (P2(size)-1) << radix_bits.  The
// reference requires size > 0 (storage/FoilHashTable.cpp:36); an empty table (a label side with no
// rows) gets one empty bucket here, so that probing it finds nothing.
ChainTable make_table(int64_t size, int radix_bits) {
  ChainTable t;
  const uint32_t nb = size > 0 ? orc_next_pow2((uint32_t)size) : 1u;
  t.mask = (nb - 1u) << radix_bits;
  t.buckets.assign(nb, 0);
  t.next.assign(size, 0);
  return t;
}

uint32_t hash_keys(const std::vector<const int32_t*>& cols, int64_t row) {
  // Hash<num_values>(values, tid) (utility/Hash.hpp:35-55)
  uint32_t h = orc_hash_int(cols[0][row]);
  for (size_t i = 1; i < cols.size(); ++i) h = orc_hash_combine(h, cols[i][row]);
  return h;
}

bool keys_equal(const std::vector<const int32_t*>& a, int64_t ra,
                const std::vector<const int32_t*>& b, int64_t rb) {
  // VectorEqualAt (utility/Hash.hpp:57-83)
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i][ra] != b[i][rb]) return false;
  return true;
}

// BuildHashTableOnTableImpl (operations/BuildHashTable.cpp:97-125).
ChainTable build_on_table(const std::vector<const int32_t*>& keys, int64_t n) {
  ChainTable t = make_table(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t b = hash_keys(keys, i) & t.mask;
    t.next[i] = t.buckets[b];
    t.buckets[b] = (int32_t)(i + 1);
  }
  return t;
}

// ---------------------------------------------------------------------------------------
// Radix partitions of one column: (value, tid) pairs in partition order.
struct Partitions {
  std::vector<int32_t> pairs;     // 2 per tuple
  std::vector<int64_t> offsets;   // 2^r + 1
  int64_t size(int p) const { return offsets[p + 1] - offsets[p]; }
  int32_t value(int64_t i) const { return pairs[2 * i]; }
  int32_t tid(int64_t i) const { return pairs[2 * i + 1]; }
};

Partitions partition_column(const int32_t* values, int64_t n, int radix_bits) {
  Partitions p;
  p.pairs.resize(2 * n);
  p.offsets.resize(((size_t)1 << radix_bits) + 1);
  orc_radix_partition(values, n, radix_bits, p.pairs.data(), p.offsets.data());
  return p;
}

// ---------------------------------------------------------------------------------------
// Plan (learner/PredicateEvaluationPlan.hpp:31-112) in index form.
struct Plan {
  int root_literal = -1;                 // literal with no atom
  std::vector<std::pair<int, int>> atoms;  // (background column, binding column)
  std::vector<int> left, right, literal;   // per tree node; left/right -1 for atoms
  int num_atoms() const { return (int)atoms.size(); }
};

struct NodeInfo {
  std::set<int> atoms, literal_ids, mergeable;
  int literal = -1, left = -1, right = -1, refs = 0;
};

// GeneratePredicateEvaluationPlan (learner/CandidateLiteralEvaluator.cpp:105-262), with two
// deliberate differences that do not change any count: ordered sets replace unordered ones
// (ties between equally good pairs are broken by ascending node id), and a pair whose union
// repeats an existing node's atom set is skipped, which makes the greedy loop terminate for
// literals with three or more atoms (the reference loop re-merges the same pair forever there).
Plan generate_plan(int arity, int join_key, const std::vector<std::vector<int>>& lit_vars) {
  Plan plan;
  std::vector<NodeInfo> nodes;
  std::vector<std::set<int>> lit_atoms(lit_vars.size());
  std::map<std::pair<int, int>, int> atom_ids;
  std::set<int> remaining;
  for (int l = 0; l < (int)lit_vars.size(); ++l) {
    for (int vid = 0; vid < arity; ++vid) {
      const int var = lit_vars[l][vid];
      if (var >= 0 && vid != join_key) {
        auto key = std::make_pair(vid, var);
        auto it = atom_ids.find(key);
        if (it == atom_ids.end()) {
          const int id = (int)atom_ids.size();
          nodes.emplace_back();
          nodes.back().atoms.insert(id);
          plan.atoms.push_back(key);
          it = atom_ids.emplace(key, id).first;
        }
        nodes[it->second].literal_ids.insert(l);
        lit_atoms[l].insert(it->second);
      }
    }
    if (lit_atoms[l].empty()) {
      plan.root_literal = l;
    } else if (lit_atoms[l].size() == 1) {
      const int a = *lit_atoms[l].begin();
      nodes[a].literal = l;
      nodes[a].literal_ids.erase(l);
    } else {
      remaining.insert(l);
    }
  }
  const int num_atoms = (int)nodes.size();
  for (int i = 0; i < num_atoms; ++i)
    for (int k = 0; k < num_atoms; ++k)
      if (k != i) nodes[i].mergeable.insert(k);

  while (!remaining.empty()) {
    int best = 0, b1 = -1, b2 = -1;
    for (int f = 0; f + 1 < (int)nodes.size(); ++f) {
      for (int s : nodes[f].mergeable) {
        std::set<int> uni = nodes[f].atoms;
        uni.insert(nodes[s].atoms.begin(), nodes[s].atoms.end());
        bool repeat = false;
        for (const NodeInfo& n : nodes)
          if (n.atoms == uni) { repeat = true; break; }
        if (repeat) continue;
        int inter = 0;
        for (int l : nodes[s].literal_ids) inter += nodes[f].literal_ids.count(l);
        if (inter > best) { best = inter; b1 = f; b2 = s; }
      }
    }
    if (b1 < 0) break;  // cannot happen for well-formed groups
    ++nodes[b1].refs;
    ++nodes[b2].refs;
    const int nid = (int)nodes.size();
    NodeInfo ni;
    ni.left = b1;
    ni.right = b2;
    ni.atoms = nodes[b1].atoms;
    ni.atoms.insert(nodes[b2].atoms.begin(), nodes[b2].atoms.end());
    for (int l : nodes[b1].literal_ids)
      if (nodes[b2].literal_ids.count(l)) ni.literal_ids.insert(l);
    for (auto it = ni.literal_ids.begin(); it != ni.literal_ids.end();) {
      if (ni.literal < 0 && lit_atoms[*it] == ni.atoms) {
        ni.literal = *it;
        remaining.erase(*it);
        for (int i = 0; i < nid; ++i) nodes[i].literal_ids.erase(*it);
        it = ni.literal_ids.erase(it);
      } else {
        ++it;
      }
    }
    nodes.push_back(ni);
    for (int i = 0; i < nid; ++i) {
      if (i == b1 || i == b2) continue;
      bool disjoint = true;
      for (int a : nodes[nid].atoms)
        if (nodes[i].atoms.count(a)) { disjoint = false; break; }
      if (disjoint) nodes[i].mergeable.insert(nid);
    }
  }
  // Prune conjunctions that own no literal and feed nothing (:235-242).
  for (int id = (int)nodes.size() - 1; id >= num_atoms; --id) {
    if (nodes[id].literal < 0 && nodes[id].refs == 0) {
      --nodes[nodes[id].left].refs;
      --nodes[nodes[id].right].refs;
    }
  }
  std::vector<int> remap(nodes.size(), -1);
  for (int i = 0; i < num_atoms; ++i) {
    remap[i] = i;
    plan.left.push_back(-1);
    plan.right.push_back(-1);
    plan.literal.push_back(nodes[i].literal);
  }
  for (int i = num_atoms; i < (int)nodes.size(); ++i) {
    if (nodes[i].literal >= 0 || nodes[i].refs != 0) {
      remap[i] = (int)plan.left.size();
      plan.left.push_back(remap[nodes[i].left]);
      plan.right.push_back(remap[nodes[i].right]);
      plan.literal.push_back(nodes[i].literal);
    }
  }
  return plan;
}

// ---------------------------------------------------------------------------------------
// Per-literal outputs and per-plan aggregation state (operations/CountAggregator.cpp).
struct LiteralCounts {
  int64_t covered_pos = 0, covered_neg = 0, binding_pos = 0, binding_neg = 0;
};

struct PlanState {
  Plan plan;
  std::vector<int> slot_to_literal;   // group literal slot -> global literal index
  int rel = -1, join_key = -1;
  int saved_partition = -1;
  // Semi-bitvectors per tree node (index num_nodes = root), per label.
  std::vector<Bits> pos_semi, neg_semi;
};

struct Chunk {
  std::vector<int32_t> probe_tids, build_tids, build_rel_tids;
};

// ---------------------------------------------------------------------------------------
int literal_join_key_impl(int arity, const int* vars, const int* types) {
  // FoilLiteral::UpdateJoinKey (schema/FoilLiteral.hpp:133-144)
  int jk = -1;
  for (int pos = 0; pos < arity; ++pos) {
    if (vars[pos] < 0) continue;
    const int t = types ? types[pos] : 0;
    if (jk == -1 || (types ? types[jk] : 0) < t) jk = pos;
  }
  return jk;
}

const int* rel_types(const orc_eval_input* in, int rel) {
  return in->arg_types ? in->arg_types + in->rel_col_start[rel] : nullptr;
}
const int32_t* rel_col(const orc_eval_input* in, int rel, int c) {
  return in->rel_columns[in->rel_col_start[rel] + c];
}
const int* lit_vars_of(const orc_eval_input* in, int l) { return in->lit_vars + (size_t)l * ORC_MAX_ARITY; }

// Runs CountAggregator::Execute (or ExecuteOnOneLabel) for one clause join variable.
// label_mode: 0 = consecutive table (Execute), 1 = positives only, 2 = negatives only.
void run_aggregator(const orc_eval_input* in, int join_var, std::vector<PlanState*>& plans,
                    const std::vector<int32_t*>& binding_cols, int64_t num_rows,
                    int64_t num_positive, int64_t tid_base, int label_mode, int radix_bits,
                    int chunk_size, std::map<std::pair<int, int>, Partitions>& bg_cache,
                    std::vector<LiteralCounts>& out, std::vector<Bits>* semi_global) {
  // RadixPartition + BuildHashTableOnPartitions on the binding join column
  // (learner/CandidateLiteralEvaluator.cpp:316-323).
  Partitions bp = partition_column(binding_cols[join_var], num_rows, radix_bits);
  const int nparts = 1 << radix_bits;
  std::vector<ChainTable> tables(nparts);
  for (int p = 0; p < nparts; ++p) {
    const int64_t np = bp.size(p);
    if (np == 0) continue;
    tables[p] = make_table(np, radix_bits);
    for (int64_t i = 0; i < np; ++i) {
      const uint32_t h = orc_hash_int(bp.value(bp.offsets[p] + i));
      const uint32_t b = (h & tables[p].mask) >> radix_bits;  // HASH_BIT_MODULO
      tables[p].next[i] = tables[p].buckets[b];
      tables[p].buckets[b] = (int32_t)(i + 1);
    }
  }
  for (PlanState* ps : plans) ps->saved_partition = -1;

  // Tables in group order, join groups per table (PartitionAssigner.hpp:77-138).
  std::vector<int> table_order;
  std::map<int, std::vector<PlanState*>> per_table;
  for (PlanState* ps : plans) {
    if (!per_table.count(ps->rel)) table_order.push_back(ps->rel);
    per_table[ps->rel].push_back(ps);
  }

  for (int p = 0; p < nparts; ++p) {
    for (int rel : table_order) {
      for (PlanState* ps : per_table[rel]) {
        const Partitions& pp = bg_cache.at(std::make_pair(rel, ps->join_key));
        const int64_t psize = pp.size(p);
        for (int64_t off = 0; off < psize; off += chunk_size) {
          const int64_t len = std::min<int64_t>(chunk_size, psize - off);
          // HashJoin::Next (operations/HashJoin.cpp:23-86): skip empty build partitions.
          if (bp.size(p) == 0) continue;
          Chunk ch;
          const ChainTable& t = tables[p];
          for (int64_t k = 0; k < len; ++k) {
            const int64_t pi = pp.offsets[p] + off + k;
            const int32_t pv = pp.value(pi);
            const uint32_t b = (orc_hash_int(pv) & t.mask) >> radix_bits;
            for (int pos = t.buckets[b] - 1; pos >= 0; pos = t.next[pos] - 1) {
              if (bp.value(bp.offsets[p] + pos) == pv) {
                ch.build_tids.push_back(bp.tid(bp.offsets[p] + pos));
                ch.probe_tids.push_back(pp.tid(pi));
                ch.build_rel_tids.push_back(pos);
              }
            }
          }
          if (ch.build_tids.empty()) continue;
          const int64_t m = (int64_t)ch.build_tids.size();
          const int64_t bpsize = bp.size(p);

          if (ps->saved_partition != p) {  // ResetSemiVectors (CountAggregator.cpp:28-54)
            const int nn = (int)ps->plan.left.size();
            ps->pos_semi.assign(nn + 1, Bits(nwords(bpsize), 0));
            ps->neg_semi.assign(nn + 1, Bits(nwords(bpsize), 0));
            ps->saved_partition = p;
          }
          // Filter::Next: one bitvector per atom (operations/Filter.cpp:17-38,
          // expressions/ComparisonPredicate.hpp:47-87).
          const int na = ps->plan.num_atoms();
          const int nn = (int)ps->plan.left.size();
          std::vector<Bits> nb(nn, Bits(nwords(m), 0));
          for (int a = 0; a < na; ++a) {
            const int32_t* pc = rel_col(in, rel, ps->plan.atoms[a].first);
            const int32_t* bc = binding_cols[ps->plan.atoms[a].second];
            for (int64_t i = 0; i < m; ++i)
              if (pc[ch.probe_tids[i]] == bc[ch.build_tids[i] - tid_base]) bit_set(nb[a], i);
          }
          for (int x = na; x < nn; ++x)
            for (int64_t w = 0; w < nwords(m); ++w)
              nb[x][w] = nb[ps->plan.left[x]][w] & nb[ps->plan.right[x]][w];

          // Label of each match (LabelBitVector, CountAggregator.cpp:58-79).
          Bits label(nwords(m), 0);
          for (int64_t i = 0; i < m; ++i) {
            bool positive;
            if (label_mode == 0) positive = ch.build_tids[i] < num_positive;
            else positive = (label_mode == 1);
            if (positive) bit_set(label, i);
          }

          auto update = [&](int slot, int semi_index, const Bits* filter) {
            const int lit = ps->slot_to_literal[slot];
            LiteralCounts& lc = out[lit];
            for (int64_t i = 0; i < m; ++i) {
              if (filter && !bit_test(*filter, i)) continue;
              const bool pos = bit_test(label, i);
              const int32_t rt = ch.build_rel_tids[i];
              if (pos) {
                ++lc.binding_pos;
                if (!bit_test_set(ps->pos_semi[semi_index], rt)) {
                  ++lc.covered_pos;
                  if (semi_global) bit_set((*semi_global)[lit], bp.tid(bp.offsets[p] + rt) + 0);
                }
              } else {
                ++lc.binding_neg;
                if (!bit_test_set(ps->neg_semi[semi_index], rt)) {
                  ++lc.covered_neg;
                  if (semi_global) bit_set((*semi_global)[lit], bp.tid(bp.offsets[p] + rt) + 0);
                }
              }
            }
          };
          // Fast path / root literal (CountAggregator.cpp:127-150, 158-174).
          if (ps->plan.root_literal >= 0) update(ps->plan.root_literal, nn, nullptr);
          // Atom and conjunction nodes (CountAggregator.cpp:176-237).
          for (int x = 0; x < nn; ++x)
            if (ps->plan.literal[x] >= 0) update(ps->plan.literal[x], x, &nb[x]);
        }
      }
    }
  }
}

}  // namespace

// =======================================================================================
extern "C" {

uint32_t orc_hash_int(int32_t v) {
  // std::hash<int> is the identity in libstdc++; Hash<T> casts it to uint32 (Hash.hpp:20-23).
  return (uint32_t)v;
}

uint32_t orc_hash_combine(uint32_t seed, int32_t v) {
  // HashCombine (utility/Hash.hpp:25-28), computed in uint32.
  return seed ^ (orc_hash_int(v) + 0x9e3779b9u + (seed << 6) + (seed >> 2));
}

uint32_t orc_next_pow2(uint32_t v) {
  // NEXT_POW_2 (storage/FoilHashTable.cpp:12-26); v must be > 0.
  v--;
  v |= v >> 1;
  v |= v >> 2;
  v |= v >> 4;
  v |= v >> 8;
  v |= v >> 16;
  return v + 1;
}

void orc_radix_partition(const int32_t* values, int64_t n, int radix_bits, int32_t* out_pairs,
                         int64_t* out_offsets) {
  // Pass 1: histogram of Hash(v) & mask (RadixPartition.cpp:65-74); pass 2 writes each tuple
  // after all earlier tuples of its partition, which is what the write-combining buffers of
  // RadixPartition.cpp:98-154 produce.
  const uint32_t np = 1u << radix_bits, mask = np - 1u;
  std::vector<int64_t> hist(np, 0);
  for (int64_t i = 0; i < n; ++i) ++hist[orc_hash_int(values[i]) & mask];
  int64_t off = 0;
  for (uint32_t p = 0; p < np; ++p) {
    out_offsets[p] = off;
    off += hist[p];
  }
  out_offsets[np] = off;
  std::vector<int64_t> cur(out_offsets, out_offsets + np);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t p = orc_hash_int(values[i]) & mask;
    const int64_t d = cur[p]++;
    out_pairs[2 * d] = values[i];
    out_pairs[2 * d + 1] = (int32_t)i;
  }
}

int64_t orc_partition_bucket_offsets(const int64_t* offsets, int num_partitions,
                                     int64_t* bucket_offsets) {
  int64_t total = 0;
  for (int p = 0; p < num_partitions; ++p) {
    if (bucket_offsets) bucket_offsets[p] = total;
    const int64_t n = offsets[p + 1] - offsets[p];
    if (n > 0) total += orc_next_pow2((uint32_t)n);
  }
  if (bucket_offsets) bucket_offsets[num_partitions] = total;
  return total;
}

void orc_build_hash_table_on_partitions(const int32_t* pairs, const int64_t* offsets,
                                        int radix_bits, const int64_t* bucket_offsets,
                                        int32_t* buckets, int32_t* next) {
  // BuildHashTableOnPartitions (operations/BuildHashTable.cpp:54-93): per non-empty partition,
  // P2(n) zeroed buckets and head insertion next[i] = buckets[b]; buckets[b] = i + 1.
  const int np = 1 << radix_bits;
  for (int p = 0; p < np; ++p) {
    const int64_t n = offsets[p + 1] - offsets[p];
    if (n == 0) continue;
    const uint32_t nb = orc_next_pow2((uint32_t)n);
    const uint32_t mask = (nb - 1u) << radix_bits;
    int32_t* bk = buckets + bucket_offsets[p];
    int32_t* nx = next + offsets[p];
    std::fill(bk, bk + nb, 0);
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t b = (orc_hash_int(pairs[2 * (offsets[p] + i)]) & mask) >> radix_bits;
      nx[i] = bk[b];
      bk[b] = (int32_t)(i + 1);
    }
  }
}

void orc_build_hash_table_on_table(const int32_t* const* key_columns, int num_keys, int64_t n,
                                   int32_t* buckets, int32_t* next) {
  std::vector<const int32_t*> keys(key_columns, key_columns + num_keys);
  ChainTable t = build_on_table(keys, n);
  std::copy(t.buckets.begin(), t.buckets.end(), buckets);
  std::copy(t.next.begin(), t.next.end(), next);
}

int orc_literal_join_key(int arity, const int* vars, const int* arg_types) {
  return literal_join_key_impl(arity, vars, arg_types);
}

int orc_generate_plan(int arity, int join_key, const int* lit_vars, int n, int max_atoms,
                      int max_nodes, int* num_atoms, int* atom_bg_col, int* atom_binding_col,
                      int* root_literal, int* node_left, int* node_right, int* node_literal) {
  std::vector<std::vector<int>> lv(n);
  for (int l = 0; l < n; ++l) lv[l].assign(lit_vars + (size_t)l * ORC_MAX_ARITY, lit_vars + (size_t)l * ORC_MAX_ARITY + arity);
  Plan p = generate_plan(arity, join_key, lv);
  if (p.num_atoms() > max_atoms || (int)p.left.size() > max_nodes) return -1;
  *num_atoms = p.num_atoms();
  for (int a = 0; a < p.num_atoms(); ++a) {
    atom_bg_col[a] = p.atoms[a].first;
    atom_binding_col[a] = p.atoms[a].second;
  }
  *root_literal = p.root_literal;
  for (size_t x = 0; x < p.left.size(); ++x) {
    node_left[x] = p.left[x];
    node_right[x] = p.right[x];
    node_literal[x] = p.literal[x];
  }
  return (int)p.left.size();
}

typedef std::map<std::pair<int, int>, Partitions> RelationPartitions;  // TableView::partitions_at cache

struct orc_relation_cache {
  RelationPartitions parts;
};

static int evaluate_pipeline(const orc_eval_input* in, int radix_bits, int chunk_size, RelationPartitions& bg_cache,
                             uint64_t* counts, uint64_t* semi_bits);

int orc_evaluate_pipeline(const orc_eval_input* in, int radix_bits, int chunk_size,
                          uint64_t* counts, uint64_t* semi_bits) {
  RelationPartitions bg_cache;
  return evaluate_pipeline(in, radix_bits, chunk_size, bg_cache, counts, semi_bits);
}

orc_relation_cache* orc_relation_cache_create(void) { return new orc_relation_cache(); }
void orc_relation_cache_destroy(orc_relation_cache* c) { delete c; }

// The background partitions of every (relation, join key) the literals use, made before any
// evaluation, as the evaluator keeps them on the fact tables across refinement steps
// (learner/CandidateLiteralEvaluator.cpp:291-299).
void orc_relation_cache_fill(orc_relation_cache* c, const orc_eval_input* in, int radix_bits) {
  for (int l = 0; l < in->num_literals; ++l) {
    const int rel = in->lit_rel[l];
    const int jk = literal_join_key_impl(in->rel_arity[rel], lit_vars_of(in, l), rel_types(in, rel));
    if (jk < 0) continue;
    const auto key = std::make_pair(rel, jk);
    if (!c->parts.count(key)) c->parts.emplace(key, partition_column(rel_col(in, rel, jk), in->rel_rows[rel], radix_bits));
  }
}

int orc_evaluate_pipeline_cached(const orc_eval_input* in, orc_relation_cache* c, int radix_bits, int chunk_size,
                                 uint64_t* counts, uint64_t* semi_bits) {
  return evaluate_pipeline(in, radix_bits, chunk_size, c->parts, counts, semi_bits);
}

static int evaluate_pipeline(const orc_eval_input* in, int radix_bits, int chunk_size, RelationPartitions& bg_cache,
                             uint64_t* counts, uint64_t* semi_bits) {
  const int L = in->num_literals;
  const int64_t N = in->num_bindings;
  std::vector<LiteralCounts> out(L);
  std::vector<Bits> semi;
  if (semi_bits) semi.assign(L, Bits(nwords(N), 0));

  // Group literals by the clause variable at their join key (learner/QuickFoil.cpp:142-161),
  // then by predicate and join key (CandidateLiteralEvaluator.cpp:281-289).  Groups keep the
  // order of first appearance; the reference's order (pointer-keyed hash maps) does not
  // change any count.
  std::map<int, std::vector<int>> by_var;
  std::vector<int> jk(L);
  for (int l = 0; l < L; ++l) {
    const int rel = in->lit_rel[l];
    jk[l] = literal_join_key_impl(in->rel_arity[rel], lit_vars_of(in, l), rel_types(in, rel));
    if (jk[l] < 0) continue;  // all-unbound literals are never enumerated
    by_var[lit_vars_of(in, l)[jk[l]]].push_back(l);
  }
  std::vector<int32_t*> bcols(in->num_binding_columns);
  for (int c = 0; c < in->num_binding_columns; ++c) bcols[c] = const_cast<int32_t*>(in->binding_columns[c]);

  for (auto& kv : by_var) {
    const int v = kv.first;
    std::vector<std::unique_ptr<PlanState>> states;
    std::map<std::pair<int, int>, PlanState*> index;
    std::vector<std::pair<int, int>> order;
    std::map<std::pair<int, int>, std::vector<int>> members;
    for (int l : kv.second) {
      auto key = std::make_pair(in->lit_rel[l], jk[l]);
      if (!members.count(key)) order.push_back(key);
      members[key].push_back(l);
    }
    // Keep relations contiguous (one literal group per predicate).
    std::stable_sort(order.begin(), order.end(),
                     [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                       auto fa = std::find_if(order.begin(), order.end(), [&](const std::pair<int, int>& x) { return x.first == a.first; });
                       auto fb = std::find_if(order.begin(), order.end(), [&](const std::pair<int, int>& x) { return x.first == b.first; });
                       return fa < fb;
                     });
    std::vector<PlanState*> plans;
    for (auto& key : order) {
      const int rel = key.first, j = key.second;
      if (!bg_cache.count(key))
        bg_cache.emplace(key, partition_column(rel_col(in, rel, j), in->rel_rows[rel], radix_bits));
      std::unique_ptr<PlanState> ps(new PlanState);
      ps->rel = rel;
      ps->join_key = j;
      ps->slot_to_literal = members[key];
      std::vector<std::vector<int>> lv;
      for (int l : members[key]) lv.emplace_back(lit_vars_of(in, l), lit_vars_of(in, l) + in->rel_arity[rel]);
      ps->plan = generate_plan(in->rel_arity[rel], j, lv);
      plans.push_back(ps.get());
      states.push_back(std::move(ps));
    }
    run_aggregator(in, v, plans, bcols, N, in->num_positive, 0, 0, radix_bits, chunk_size, bg_cache,
                   out, semi_bits ? &semi : nullptr);
  }
  int status = 0;
  for (int l = 0; l < L; ++l) {
    const LiteralCounts& c = out[l];
    counts[4 * l + 0] = (uint64_t)c.covered_pos;
    counts[4 * l + 1] = (uint64_t)c.covered_neg;
    counts[4 * l + 2] = (uint64_t)c.binding_pos;
    counts[4 * l + 3] = (uint64_t)c.binding_neg;
    if (c.binding_pos > INT_MAX || c.binding_neg > INT_MAX) status = -1;
  }
  if (semi_bits)
    for (int l = 0; l < L; ++l) std::copy(semi[l].begin(), semi[l].end(), semi_bits + (size_t)l * nwords(N));
  return status;
}

void orc_evaluate_bruteforce(const orc_eval_input* in, uint64_t* counts, uint64_t* semi_bits) {
  const int64_t N = in->num_bindings;
  for (int l = 0; l < in->num_literals; ++l) {
    const int rel = in->lit_rel[l];
    const int* vars = lit_vars_of(in, l);
    uint64_t c[4] = {0, 0, 0, 0};
    uint64_t* bits = semi_bits ? semi_bits + (size_t)l * nwords(N) : nullptr;
    if (bits) std::fill(bits, bits + nwords(N), 0ull);
    for (int64_t r = 0; r < N; ++r) {
      uint64_t cnt = 0;
      for (int64_t s = 0; s < in->rel_rows[rel]; ++s) {
        bool match = true;
        for (int i = 0; i < in->rel_arity[rel] && match; ++i)
          if (vars[i] >= 0 && rel_col(in, rel, i)[s] != in->binding_columns[vars[i]][r]) match = false;
        cnt += match;
      }
      const bool pos = r < in->num_positive;
      c[pos ? 2 : 3] += cnt;
      c[pos ? 0 : 1] += (cnt > 0);
      if (bits && cnt > 0) bits[r >> 6] |= 1ull << (r & 63);
    }
    std::copy(c, c + 4, counts + 4 * l);
  }
}

void orc_evaluate_grouped(const orc_eval_input* in, uint64_t* counts, uint64_t* semi_bits) {
  const int64_t N = in->num_bindings;
  struct VecHash {
    size_t operator()(const std::vector<int32_t>& v) const {
      uint64_t h = 1469598103934665603ull;
      for (int32_t x : v) h = (h ^ (uint32_t)x) * 1099511628211ull;
      return (size_t)h;
    }
  };
  for (int l = 0; l < in->num_literals; ++l) {
    const int rel = in->lit_rel[l];
    const int* vars = lit_vars_of(in, l);
    std::vector<int> pos_list;
    for (int i = 0; i < in->rel_arity[rel]; ++i)
      if (vars[i] >= 0) pos_list.push_back(i);
    std::unordered_map<std::vector<int32_t>, uint64_t, VecHash> freq;
    std::vector<int32_t> key(pos_list.size());
    for (int64_t s = 0; s < in->rel_rows[rel]; ++s) {
      for (size_t k = 0; k < pos_list.size(); ++k) key[k] = rel_col(in, rel, pos_list[k])[s];
      ++freq[key];
    }
    uint64_t c[4] = {0, 0, 0, 0};
    uint64_t* bits = semi_bits ? semi_bits + (size_t)l * nwords(N) : nullptr;
    if (bits) std::fill(bits, bits + nwords(N), 0ull);
    for (int64_t r = 0; r < N; ++r) {
      for (size_t k = 0; k < pos_list.size(); ++k) key[k] = in->binding_columns[vars[pos_list[k]]][r];
      auto it = freq.find(key);
      const uint64_t cnt = it == freq.end() ? 0 : it->second;
      const bool pos = r < in->num_positive;
      c[pos ? 2 : 3] += cnt;
      c[pos ? 0 : 1] += (cnt > 0);
      if (bits && cnt > 0) bits[r >> 6] |= 1ull << (r & 63);
    }
    std::copy(c, c + 4, counts + 4 * l);
  }
}

// ---- MultiColumnHashJoin ------------------------------------------------------------------
namespace {

// DoBlockJoin (MultiColumnHashJoin.cpp:496-528): probe rows in order, build positions in chain
// order, emitting (probe, build) pairs for equal keys.
void block_join(const std::vector<const int32_t*>& probe_keys, int64_t num_probe,
                const std::vector<const int32_t*>& build_keys, const ChainTable& t,
                std::vector<int64_t>& probe_tids, std::vector<int64_t>& build_tids) {
  for (int64_t pr = 0; pr < num_probe; ++pr) {
    const uint32_t b = hash_keys(probe_keys, pr) & t.mask;
    for (int pos = t.buckets[b] - 1; pos >= 0; pos = t.next[pos] - 1)
      if (keys_equal(probe_keys, pr, build_keys, pos)) {
        probe_tids.push_back(pr);
        build_tids.push_back(pos);
      }
  }
}

}  // namespace

int64_t orc_create_label_aware_binding_table(const orc_eval_input* in, int literal,
                                             int64_t num_binding_pos, int64_t num_binding_neg,
                                             int32_t* const* out_columns) {
  const int rel = in->lit_rel[literal];
  const int arity = in->rel_arity[rel];
  const int* vars = lit_vars_of(in, literal);
  const int C = in->num_binding_columns;
  const int64_t npos = in->num_positive, nneg = in->num_bindings - in->num_positive;
  const int64_t nbg = in->rel_rows[rel];
  std::vector<const int32_t*> bg_keys, pos_keys, neg_keys, bg_cols, pos_cols, neg_cols;
  std::vector<int> unbound;
  for (int i = 0; i < arity; ++i) {
    if (vars[i] >= 0) {
      bg_keys.push_back(rel_col(in, rel, i));
      pos_keys.push_back(in->binding_columns[vars[i]]);
      neg_keys.push_back(in->binding_columns[vars[i]] + npos);
    } else {
      unbound.push_back(i);
    }
  }
  for (int i = 0; i < arity; ++i) bg_cols.push_back(rel_col(in, rel, i));
  for (int c = 0; c < C; ++c) {
    pos_cols.push_back(in->binding_columns[c]);
    neg_cols.push_back(in->binding_columns[c] + npos);
  }
  if (bg_keys.empty()) return -1;  // a literal with no bound variable is never extended
  const int ncols = C + (int)unbound.size();
  // Outputs: positives at [0, num_binding_pos), negatives after (:144-159).
  std::vector<std::vector<int32_t>> pos_out(ncols), neg_out(ncols);
  auto emit = [&](std::vector<std::vector<int32_t>>& dst, const std::vector<const int32_t*>& clause_cols,
                  int64_t clause_row, int64_t bg_row) {
    for (int c = 0; c < C; ++c) dst[c].push_back(clause_cols[c][clause_row]);
    for (size_t u = 0; u < unbound.size(); ++u) dst[C + u].push_back(bg_cols[unbound[u]][bg_row]);
  };
  if (npos < nbg && nneg < nbg) {
    // Background probes hash tables on both label tables: CollaborateJoin (:160-193, 557-627).
    ChainTable tp = build_on_table(pos_keys, npos), tn = build_on_table(neg_keys, nneg);
    std::vector<int64_t> pp, pb, np_, nb;
    block_join(bg_keys, nbg, pos_keys, tp, pp, pb);
    block_join(bg_keys, nbg, neg_keys, tn, np_, nb);
    for (size_t i = 0; i < pb.size(); ++i) emit(pos_out, pos_cols, pb[i], pp[i]);
    for (size_t i = 0; i < nb.size(); ++i) emit(neg_out, neg_cols, nb[i], np_[i]);
  } else {
    // Hash on the background table, positives then negatives probe it (:194-249).
    ChainTable tb = build_on_table(bg_keys, nbg);
    std::vector<int64_t> pp, pb, np_, nb;
    block_join(pos_keys, npos, bg_keys, tb, pp, pb);
    block_join(neg_keys, nneg, bg_keys, tb, np_, nb);
    for (size_t i = 0; i < pp.size(); ++i) emit(pos_out, pos_cols, pp[i], pb[i]);
    for (size_t i = 0; i < np_.size(); ++i) emit(neg_out, neg_cols, np_[i], nb[i]);
  }
  const int64_t got_pos = ncols ? (int64_t)pos_out[0].size() : 0;
  const int64_t got_neg = ncols ? (int64_t)neg_out[0].size() : 0;
  if (got_pos != num_binding_pos || got_neg != num_binding_neg) return -1;
  if (out_columns) {
    for (int c = 0; c < ncols; ++c) {
      std::copy(pos_out[c].begin(), pos_out[c].end(), out_columns[c]);
      std::copy(neg_out[c].begin(), neg_out[c].end(), out_columns[c] + got_pos);
    }
  }
  return got_pos + got_neg;
}

int64_t orc_create_binding_table(const orc_eval_input* in, int literal, int32_t* const* out_columns) {
  const int rel = in->lit_rel[literal];
  const int arity = in->rel_arity[rel];
  const int* vars = lit_vars_of(in, literal);
  const int C = in->num_binding_columns;
  const int64_t N = in->num_bindings, nbg = in->rel_rows[rel];
  std::vector<const int32_t*> lit_keys, bind_keys, bg_cols, bind_cols;
  std::vector<int> unbound;
  for (int i = 0; i < arity; ++i) {
    // CreateBindingTable (:26-115): variables with id < #columns are join keys.
    if (vars[i] < C && vars[i] >= 0) {
      lit_keys.push_back(rel_col(in, rel, i));
      bind_keys.push_back(in->binding_columns[vars[i]]);
    } else {
      unbound.push_back(i);
    }
  }
  if (lit_keys.empty()) return -1;  // a literal with no bound variable is never extended
  for (int i = 0; i < arity; ++i) bg_cols.push_back(rel_col(in, rel, i));
  for (int c = 0; c < C; ++c) bind_cols.push_back(in->binding_columns[c]);
  std::vector<int64_t> pt, bt;
  int64_t rows;
  const int ncols = C + (int)unbound.size();
  if (nbg < N) {
    ChainTable t = build_on_table(lit_keys, nbg);
    block_join(bind_keys, N, lit_keys, t, pt, bt);  // binding probes, background build
    rows = (int64_t)pt.size();
    if (out_columns)
      for (int64_t i = 0; i < rows; ++i) {
        for (int c = 0; c < C; ++c) out_columns[c][i] = bind_cols[c][pt[i]];
        for (size_t u = 0; u < unbound.size(); ++u) out_columns[C + u][i] = bg_cols[unbound[u]][bt[i]];
      }
  } else {
    ChainTable t = build_on_table(bind_keys, N);
    block_join(lit_keys, nbg, bind_keys, t, pt, bt);  // background probes, binding build
    rows = (int64_t)pt.size();
    if (out_columns)
      for (int64_t i = 0; i < rows; ++i) {
        for (int c = 0; c < C; ++c) out_columns[c][i] = bind_cols[c][bt[i]];
        for (size_t u = 0; u < unbound.size(); ++u) out_columns[C + u][i] = bg_cols[unbound[u]][pt[i]];
      }
  }
  (void)ncols;
  return rows;
}

void orc_left_semi_join(const int32_t* const* probe_keys, int64_t num_probe,
                        const int32_t* const* build_keys, int64_t num_build, int num_keys,
                        uint64_t* out_bits) {
  std::vector<const int32_t*> pk(probe_keys, probe_keys + num_keys), bk(build_keys, build_keys + num_keys);
  std::fill(out_bits, out_bits + nwords(num_probe), 0ull);
  if (num_build == 0) return;
  ChainTable t = build_on_table(bk, num_build);
  for (int64_t pr = 0; pr < num_probe; ++pr) {
    const uint32_t b = hash_keys(pk, pr) & t.mask;
    for (int pos = t.buckets[b] - 1; pos >= 0; pos = t.next[pos] - 1)
      if (keys_equal(pk, pr, bk, pos)) {
        out_bits[pr >> 6] |= 1ull << (pr & 63);
        break;
      }
  }
}

void orc_right_semi_join(const int32_t* const* probe_keys, int64_t num_probe,
                         const int32_t* const* build_keys, int64_t num_build, int num_keys,
                         uint64_t* out_bits) {
  std::vector<const int32_t*> pk(probe_keys, probe_keys + num_keys), bk(build_keys, build_keys + num_keys);
  std::fill(out_bits, out_bits + nwords(num_build), 0ull);
  if (num_build == 0) return;
  ChainTable t = build_on_table(bk, num_build);
  for (int64_t pr = 0; pr < num_probe; ++pr) {
    const uint32_t b = hash_keys(pk, pr) & t.mask;
    for (int pos = t.buckets[b] - 1; pos >= 0; pos = t.next[pos] - 1)
      if (keys_equal(pk, pr, bk, pos)) out_bits[pos >> 6] |= 1ull << (pos & 63);
  }
}

void orc_dedup_after_semi_join(const int32_t* const* keys, int num_keys, int64_t n,
                               const uint64_t* bits, uint8_t* out_first) {
  // InsertIfNotPresent over the set rows in ascending order (BuildHashTable.cpp:127-150).
  std::vector<const int32_t*> k(keys, keys + num_keys);
  ChainTable t = make_table(n > 0 ? n : 1, 0);
  std::fill(out_first, out_first + n, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (!((bits[i >> 6] >> (i & 63)) & 1ull)) continue;
    const uint32_t b = hash_keys(k, i) & t.mask;
    bool exist = false;
    for (int pos = t.buckets[b] - 1; pos >= 0; pos = t.next[pos] - 1)
      if (keys_equal(k, i, k, pos)) { exist = true; break; }
    if (!exist) {
      t.next[i] = t.buckets[b];
      t.buckets[b] = (int32_t)(i + 1);
      out_first[i] = 1;
    }
  }
}

}  // extern "C"
