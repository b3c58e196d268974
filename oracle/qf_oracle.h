/*
 * CPU oracle for the candidate-literal evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and from scratch, the
 * algorithms of the QuickFOIL operators (jbalint/quickfoil) so that the HIP product can be
 * checked against them.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it, and only as the checker.  The product never calls it.
 *
 * Every function cites the reference file:line it follows.  Pinning (what checks each
 * function against something other than itself) is listed in the header comment of
 * qf_oracle.cpp.
 */
#ifndef QF_ORACLE_H_
#define QF_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_ARITY 6

/* ---- hashing (utility/Hash.hpp:16-69) ---- */
uint32_t orc_hash_int(int32_t v);
uint32_t orc_hash_combine(uint32_t seed, int32_t v);
uint32_t orc_next_pow2(uint32_t v);

/* ---- RadixPartition (operations/RadixPartition.cpp:48-186) ----
 * Stable partition of one int32 column into (value, tuple_id) pairs by
 * Hash(v) & (2^radix_bits - 1).  out_pairs has 2*n int32 (value, tid interleaved, the
 * PartitionTuple layout of storage/PartitionTuple.hpp:20-42); out_offsets has 2^r + 1 entries. */
void orc_radix_partition(const int32_t* values, int64_t n, int radix_bits,
                         int32_t* out_pairs, int64_t* out_offsets);

/* ---- BuildHashTableOnPartitions (operations/BuildHashTable.cpp:54-93) ----
 * bucket_offsets[p] is the start of partition p's buckets (P2(n_p) of them, 0 if empty);
 * returns the total bucket count when bucket_offsets is non-null or null. */
int64_t orc_partition_bucket_offsets(const int64_t* offsets, int num_partitions,
                                     int64_t* bucket_offsets);
void orc_build_hash_table_on_partitions(const int32_t* pairs, const int64_t* offsets,
                                        int radix_bits, const int64_t* bucket_offsets,
                                        int32_t* buckets, int32_t* next);

/* ---- BuildHashTableOnTable (operations/BuildHashTable.cpp:97-125,193-210) ---- */
void orc_build_hash_table_on_table(const int32_t* const* key_columns, int num_keys,
                                   int64_t n, int32_t* buckets, int32_t* next);

/* ---- candidate-literal evaluation ----
 * Binding table: num_binding_columns int32 columns of n rows, rows < num_positive positive.
 * Relations: rel_columns[rel_col_start[i] + c] is column c of relation i.
 * Literal l: relation lit_rel[l], arity = relation arity, lit_vars[l*ORC_MAX_ARITY + i]
 * is the clause variable at position i (-1 = new variable).
 * arg_types[rel_col_start[i] + c] is the variable type of argument c (may be NULL: all 0).
 * counts[l*4 + k]: k=0 covered_pos, 1 covered_neg, 2 binding_pos, 3 binding_neg
 * (field order of learner/CandidateLiteralInfo.hpp:21-25).
 * semi_bits (nullable): literal l's covered bitvector over binding rows, ceil(n/64) words. */
typedef struct {
  const int32_t* const* binding_columns;
  int num_binding_columns;
  int64_t num_bindings;
  int64_t num_positive;
  const int32_t* const* rel_columns;
  const int* rel_col_start;
  const int* rel_arity;
  const int64_t* rel_rows;
  const int* arg_types;
  int num_relations;
  const int* lit_rel;
  const int* lit_vars;
  int num_literals;
} orc_eval_input;

/* Reference pipeline: RadixPartition -> BuildHashTableOnPartitions -> PartitionAssigner ->
 * HashJoin -> Filter -> CountAggregator under GeneratePredicateEvaluationPlan, driven as
 * CandidateLiteralEvaluator::Evaluate drives it (learner/CandidateLiteralEvaluator.cpp:264-406).
 * Returns 0, or -1 when a count does not fit the reference's int32 size_type. */
int orc_evaluate_pipeline(const orc_eval_input* in, int radix_bits, int chunk_size,
                          uint64_t* counts, uint64_t* semi_bits);
/* The same pipeline over background partitions made once (RadixPartition of every (relation,
 * join key) the literals use) and kept across calls, as CandidateLiteralEvaluator keeps them on
 * the fact tables (learner/CandidateLiteralEvaluator.cpp:291-299).  A cache belongs to one set of
 * relations (the same rel_columns and rel_rows in every call). */
typedef struct orc_relation_cache orc_relation_cache;
orc_relation_cache* orc_relation_cache_create(void);
void orc_relation_cache_destroy(orc_relation_cache* cache);
void orc_relation_cache_fill(orc_relation_cache* cache, const orc_eval_input* in, int radix_bits);
int orc_evaluate_pipeline_cached(const orc_eval_input* in, orc_relation_cache* cache, int radix_bits,
                                 int chunk_size, uint64_t* counts, uint64_t* semi_bits);
/* Nested-loop definition of the four counts (brute force; tiny inputs only). */
void orc_evaluate_bruteforce(const orc_eval_input* in, uint64_t* counts, uint64_t* semi_bits);
/* Grouped-count closed form: cnt(r) = #{s : P[B](s) = T[v_B](r)}; any skew. */
void orc_evaluate_grouped(const orc_eval_input* in, uint64_t* counts, uint64_t* semi_bits);

/* Literal join key (schema/FoilLiteral.hpp:133-144): the bound position with the largest
 * variable type id, first on ties; -1 when nothing is bound. */
int orc_literal_join_key(int arity, const int* vars, const int* arg_types);

/* ---- GeneratePredicateEvaluationPlan (learner/CandidateLiteralEvaluator.cpp:105-262) ----
 * For one join group of n literals (all with the same relation, join key and clause join
 * variable), writes the plan: atoms (bg column, binding column), tree nodes, node literal
 * slots.  node_left/node_right are -1 for atom nodes.  Returns the number of tree nodes, or
 * -1 when a buffer is too small.  literal slot ids index the group's literal list. */
int orc_generate_plan(int arity, int join_key, const int* lit_vars /* n x ORC_MAX_ARITY */,
                      int n, int max_atoms, int max_nodes,
                      int* num_atoms, int* atom_bg_col, int* atom_binding_col,
                      int* root_literal, int* node_left, int* node_right, int* node_literal);

/* ---- MultiColumnHashJoin::CreateLabelAwareBindingTables
 *      (operations/MultiColumnHashJoin.cpp:117-256, 403-627) ----
 * New binding table for clause + literal: clause columns then one column per new variable
 * (in position order), positives then negatives, in the reference's row order.
 * out_columns has num_binding_columns + #new columns pointers of out_rows int32 each.
 * Returns the number of rows written, or -1 when out_rows is not the exact size. */
int64_t orc_create_label_aware_binding_table(const orc_eval_input* in, int literal,
                                             int64_t num_binding_pos, int64_t num_binding_neg,
                                             int32_t* const* out_columns);
/* CreateBindingTable (operations/MultiColumnHashJoin.cpp:26-115): join a binding table with a
 * literal whose variables with id >= num_binding_columns are new columns; returns row count
 * (out_columns sized by the caller from a first call with out_columns == NULL). */
int64_t orc_create_binding_table(const orc_eval_input* in, int literal,
                                 int32_t* const* out_columns);

/* ---- semi-joins (operations/LeftSemiJoin.hpp:56-141, RightSemiJoin.hpp:45-82) ----
 * Left: bit i of out_bits set iff probe row i has a build row with equal keys.
 * Right: bit j set iff build row j has a probe row with equal keys. */
void orc_left_semi_join(const int32_t* const* probe_keys, int64_t num_probe,
                        const int32_t* const* build_keys, int64_t num_build, int num_keys,
                        uint64_t* out_bits);
void orc_right_semi_join(const int32_t* const* probe_keys, int64_t num_probe,
                         const int32_t* const* build_keys, int64_t num_build, int num_keys,
                         uint64_t* out_bits);
/* BuildHashTableAfterSemiJoin dedup insert (operations/BuildHashTable.cpp:127-189,212-236):
 * out_first[i] = 1 iff row i is set in bits and no earlier set row has equal keys. */
void orc_dedup_after_semi_join(const int32_t* const* keys, int num_keys, int64_t n,
                               const uint64_t* bits, uint8_t* out_first);

#ifdef __cplusplus
}
#endif

#endif  /* QF_ORACLE_H_ */
