/*
 * quickfoil_amd: MI355X (gfx950) implementation of QuickFOIL's candidate-literal evaluation
 * path behind a C ABI.
 *
 * Conventions
 *  - Every function returns an int status (qf_status).  On failure qf_last_error(ctx) holds a
 *    message.  No C++ exception crosses this boundary.
 *  - Pointers named d_* are device (HBM) pointers owned by the caller; h_* are host pointers.
 *    Table columns are int32 arrays, column-major (SoA), as the reference's TableView holds them
 *    (storage/TableView.hpp:22-88).
 *  - All work is ordered on the context's stream.  A context is not thread-safe; calls that
 *    return host results (counts, sizes) synchronise the stream before returning.
 *  - Counts are uint64.  The reference keeps int32 size_type counts
 *    (learner/CandidateLiteralInfo.hpp:21-25); qf_evaluate_literals reports QF_ERR_COUNT_RANGE
 *    beside the (correct) 64-bit values when a count does not fit int32.
 */
#ifndef QUICKFOIL_AMD_H_
#define QUICKFOIL_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QF_MAX_ARITY 6

typedef enum {
  QF_OK = 0,
  QF_ERR_INVALID = 1,        /* bad argument */
  QF_ERR_DEVICE = 2,         /* HIP runtime error */
  QF_ERR_OOM = 3,            /* device allocation failed */
  QF_ERR_OVERFLOW = 4,       /* caller-given output size does not match the result */
  QF_ERR_COUNT_RANGE = 5,    /* a count exceeds int32 (values are still exact in uint64) */
  QF_ERR_UNSUPPORTED = 6
} qf_status;

typedef struct qf_context qf_context;

/* Opaque context: device, stream, device memory pool, cached partitions. */
int qf_context_create(int device, qf_context** out);
void qf_context_destroy(qf_context* ctx);
/* Use a caller-owned hipStream_t (default priority) instead of the context's own stream. */
int qf_context_set_stream(qf_context* ctx, void* hip_stream);
int qf_synchronize(qf_context* ctx);
const char* qf_last_error(const qf_context* ctx);
/* Releases cached partitions and the pooled device memory. */
int qf_context_trim(qf_context* ctx);

/* Phase timing with HIP events recorded on the context's stream around each phase of the
 * evaluation (relation partitioning, binding partitioning, HBM table build, probe/count).
 * qf_context_phase_times copies up to max_phases (name, accumulated milliseconds, launches)
 * entries and returns the number of phases recorded since the last reset. */
int qf_context_enable_timing(qf_context* ctx, int enable);
int qf_context_reset_timing(qf_context* ctx);
int qf_context_phase_times(qf_context* ctx, const char** names, double* ms, int64_t* launches, int max_phases);

/* A table of int32 columns in device memory (TableView, storage/TableView.hpp:22-88). */
typedef struct {
  const int32_t* const* columns; /* host array of num_columns device pointers */
  int num_columns;
  int64_t num_rows;
} qf_table;

/* ---------------------------------------------------------------------------------------
 * Operators with the reference's layouts.
 * ------------------------------------------------------------------------------------- */

/* RadixPartition(int column_id, TableView*) (operations/RadixPartition.hpp:13-14,
 * operations/RadixPartition.cpp:48-186).  Stable partition of d_values by
 * Hash(v) & (2^radix_bits - 1) with the identity int32 hash (utility/Hash.hpp:20-23).
 * d_out_pairs: 2*n int32, PartitionTuple {value, tuple_id} interleaved
 * (storage/PartitionTuple.hpp:20-42).  d_offsets: 2^radix_bits + 1 int64 partition starts.
 * radix_bits in [1, 16]. */
int qf_radix_partition(qf_context* ctx, const int32_t* d_values, int64_t n, int radix_bits,
                       int32_t* d_out_pairs, int64_t* d_offsets);

/* qf_radix_partition for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops or harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same arguments and results as qf_radix_partition for n <= QF_NOALLOC_MAX_ROWS; larger inputs
 * return QF_ERR_INVALID.  It calls no device allocator: its temporaries live in a fixed 36 MiB
 * device array that the library holds on every GPU it runs on until the process ends.  All calls
 * on one GPU share that array, so calls that overlap on one GPU (other contexts or streams)
 * race; calls ordered on one stream do not. */
#define QF_NOALLOC_MAX_ROWS 4194304
int qf_radix_partition_noalloc(qf_context* ctx, const int32_t* d_values, int64_t n, int radix_bits,
                               int32_t* d_out_pairs, int64_t* d_offsets);

/* Bucket layout of BuildHashTableOnPartitions: partition p gets P2(n_p) buckets (none when
 * empty; storage/FoilHashTable.cpp:34-49).  Host-only: h_offsets are the partition offsets,
 * h_bucket_offsets receives num_partitions + 1 entries.  Returns the total bucket count
 * (or a negative qf_status). */
int64_t qf_partition_bucket_layout(const int64_t* h_offsets, int num_partitions,
                                   int64_t* h_bucket_offsets);

/* BuildHashTableOnPartitions(int column_id, TableView*) (operations/BuildHashTable.cpp:54-93).
 * Chained tables with the reference's exact layout: d_buckets (1-based heads, 0 = empty, at
 * d_bucket_offsets[p] for partition p) and d_next (1-based links, partition-relative, at the
 * partition's tuple offset).  Bit-identical to sequential head insertion in partition order. */
int qf_build_hash_table_on_partitions(qf_context* ctx, const int32_t* d_pairs,
                                      const int64_t* d_offsets, int radix_bits,
                                      const int64_t* d_bucket_offsets, int64_t total_buckets,
                                      int64_t n, int32_t* d_buckets, int32_t* d_next);

/* BuildHashTableOnTable(keys, table) (operations/BuildHashTable.cpp:97-125,193-210):
 * P2(n) buckets, bucket = Hash<num_keys>(keys, row) & (P2(n) - 1) with HashCombine
 * (utility/Hash.hpp:25-55).  num_keys in [1, QF_MAX_ARITY]. */
int qf_build_hash_table_on_table(qf_context* ctx, const qf_table* keys, int32_t* d_buckets,
                                 int32_t* d_next);

/* ---------------------------------------------------------------------------------------
 * Candidate-literal evaluation.
 * ------------------------------------------------------------------------------------- */

/* One candidate literal P(v_0..v_{a-1}): a relation index and, per argument position, the
 * clause variable (= binding-table column) or -1 for a new variable
 * (schema/FoilLiteral.hpp:21-164, schema/FoilVariable.hpp). */
typedef struct {
  int relation;
  int vars[QF_MAX_ARITY];
} qf_literal;

typedef enum {
  QF_EVAL_BATCHED = 0,       /* whole step in one persistent launch (graph-captured) */
  QF_EVAL_PLAIN_GROUP = 1,   /* one launch per join group (relation, join column, variable) */
  QF_EVAL_PLAIN_LITERAL = 2  /* one launch per literal */
} qf_eval_mode;

typedef struct {
  int mode;              /* qf_eval_mode */
  int radix_bits;        /* 0 = from partition_rows; else partition bits (at most 16) for both
                            sides */
  int tile_rows;         /* binding rows per work item; 0 = 1024, or the autotuner's choice when
                            autotune = 1 */
  int use_graph;         /* batched mode: capture the step's device work as a hipGraph and
                            replay it when the same call repeats (same columns, literals,
                            options); needs reuse_partitions = 1 */
  int reuse_partitions;  /* 1: cache relation partitions and their tables in the context, keyed
                            by (column addresses, rows, join column, radix bits), as the reference
                            caches them on the fact table (learner/CandidateLiteralEvaluator.cpp:
                            291-299).  Only for relations that stay unchanged at that address
                            until qf_context_trim; default 0. */
  int partition_rows;    /* 0 = default (256) or autotuned; target facts per relation partition,
                            from which each relation's radix bits follow (at most 16); the
                            batched mode then gives every relation the largest of them */
  int autotune;          /* 1: on the first call for a workload shape, time a grid of
                            (partition_rows, tile_rows) and keep the fastest in the context */
} qf_eval_options;

void qf_eval_options_default(qf_eval_options* opts);

/* The configuration the autotuner chose for the most recent autotuned call (0s if none). */
int qf_autotune_choice(qf_context* ctx, int* partition_rows, int* tile_rows, double* best_ms);

/* CandidateLiteralEvaluator::Evaluate (learner/CandidateLiteralEvaluator.cpp:264-406) for all
 * literals of a refinement step, over every clause join variable at once
 * (learner/QuickFoil.cpp:418-440).  binding: the clause's consecutive binding table, rows
 * < num_positive positive (schema/FoilClause.hpp:77-99).  relations[i]: background fact tables;
 * arg_types (nullable) holds each relation's argument type ids, used for the join-key rule
 * (schema/FoilLiteral.hpp:133-144), flattened relation-major.
 * h_counts: num_literals x 4 uint64 in CandidateLiteralInfo field order
 * {covered_pos, covered_neg, binding_pos, binding_neg}, literal order as given.
 * d_semi_bits (nullable): per literal, ceil(num_rows/64) uint64 words: bit r = 1 iff binding
 * row r has at least one matching fact (the dynamic_bitset block layout,
 * utility/BitVectorBuilder.hpp:46-54).  The positive part is rows [0, num_positive). */
int qf_evaluate_literals(qf_context* ctx, const qf_table* binding, int64_t num_positive,
                         const qf_table* relations, const int* arg_types, int num_relations,
                         const qf_literal* literals, int num_literals,
                         const qf_eval_options* opts, uint64_t* h_counts,
                         uint64_t* d_semi_bits);

/* CandidateLiteralEvaluator::Evaluate over a clause whose positive and negative bindings are two
 * tables (the restart clause, learner/QuickFoil.cpp:127-140, schema/FoilClause.hpp:152-161): the
 * passes of learner/CandidateLiteralEvaluator.cpp:345-405 (ExecuteOnPositives, then
 * ExecuteOnNegatives).  Counts as qf_evaluate_literals over the two tables concatenated
 * (positives first), without the copy.  d_semi_bits_pos / d_semi_bits_neg (nullable): per
 * literal, ceil(rows/64) words over the rows of each table.  opts->use_graph is ignored. */
int qf_evaluate_literals_split(qf_context* ctx, const qf_table* positives, const qf_table* negatives,
                               const qf_table* relations, const int* arg_types, int num_relations,
                               const qf_literal* literals, int num_literals, const qf_eval_options* opts,
                               uint64_t* h_counts, uint64_t* d_semi_bits_pos, uint64_t* d_semi_bits_neg);

/* CountAggregator::Execute over an explicit PredicateEvaluationPlan for one join group
 * (operations/CountAggregator.cpp:112-237, learner/PredicateEvaluationPlan.hpp:31-112).
 * Atoms are equalities relation[atom_bg_col[k]] == binding[atom_binding_col[k]] beside the join
 * equality relation[join_column] == binding[binding_column].  Nodes 0..num_atoms-1 are the
 * atoms; node x >= num_atoms is the conjunction of node_left[x] and node_right[x].
 * node_literal[x] / root_literal name an output slot (or -1).  h_counts: num_slots x 4. */
typedef struct {
  int relation;
  int join_column;
  int binding_column;
  int num_atoms;
  const int* atom_bg_col;
  const int* atom_binding_col;
  int num_nodes;
  const int* node_left;
  const int* node_right;
  const int* node_literal;
  int root_literal;
} qf_join_group_plan;

int qf_count_aggregate(qf_context* ctx, const qf_table* binding, int64_t num_positive,
                       const qf_table* relations, int num_relations,
                       const qf_join_group_plan* plans, int num_plans, int num_slots,
                       const qf_eval_options* opts, uint64_t* h_counts, uint64_t* d_semi_bits);

/* ---------------------------------------------------------------------------------------
 * Extending the binding set and coverage.
 * ------------------------------------------------------------------------------------- */

/* CreateLabelAwareBindingTables (operations/MultiColumnHashJoin.cpp:117-256).  Output
 * columns: the binding columns, then one per new variable of the literal in position order;
 * rows: positives then negatives, in the reference's row order.  num_binding_pos/neg are the
 * evaluator's counts and size the output exactly; a mismatch returns QF_ERR_OVERFLOW (the
 * reference writes past its buffer there).  d_out_columns: host array of device pointers,
 * each num_binding_pos + num_binding_neg int32. */
int qf_create_label_aware_binding_table(qf_context* ctx, const qf_table* binding,
                                        int64_t num_positive, const qf_table* relation,
                                        const qf_literal* literal, int64_t num_binding_pos,
                                        int64_t num_binding_neg, int32_t* const* d_out_columns);

/* CreateBindingTable (operations/MultiColumnHashJoin.cpp:26-115): variables with id >=
 * binding->num_columns are new columns.  Call with d_out_columns == NULL to get the row count
 * in *h_num_rows, then again with outputs of that size. */
int qf_create_binding_table(qf_context* ctx, const qf_table* binding, const qf_table* relation,
                            const qf_literal* literal, int32_t* const* d_out_columns,
                            int64_t* h_num_rows);

/* CreateBindingTable into caller buffers of `capacity` rows sized from an estimate (the
 * learned cardinality estimator presizes them): one pass when the output fits; otherwise
 * QF_ERR_OVERFLOW with the exact row count in *h_num_rows and nothing written, and the caller
 * retries with that size (exact fallback).  No counterpart in the reference, which has only the
 * growing-buffer form (operations/MultiColumnHashJoin.cpp:445-461). */
int qf_create_binding_table_presized(qf_context* ctx, const qf_table* binding, const qf_table* relation,
                                     const qf_literal* literal, int64_t capacity, int32_t* const* d_out_columns,
                                     int64_t* h_num_rows);

/* CreateBindingTable for several literals over one binding table, into caller-presized outputs
 * (capacities from an estimate: the learned estimator below), without a host round trip per
 * join: each join counts its rows on the device and writes them only when they fit.  No
 * counterpart in the reference, whose buffers grow by 1.5x (operations/MultiColumnHashJoin.cpp:
 * 445-461).  d_out_columns[i]: literal i's output column pointers (binding columns, then one per
 * new variable, as qf_create_binding_table), each of capacities[i] rows.  h_num_rows[i] receives
 * the exact row count; h_fits[i] = 1 when the rows were written, 0 when the capacity was too
 * small and nothing was written (the exact fallback: call again for those literals with their
 * exact counts).  Returns QF_ERR_OVERFLOW when any literal did not fit. */
int qf_create_binding_tables_presized(qf_context* ctx, const qf_table* binding, const qf_table* relations,
                                      int num_relations, const qf_literal* literals, int num_literals,
                                      const int64_t* capacities, int32_t* const* const* d_out_columns,
                                      int64_t* h_num_rows, uint8_t* h_fits);

/* LeftSemiJoin::Next over all chunks (operations/LeftSemiJoin.hpp:56-141): bit i of
 * d_out_bits (ceil(probe rows/64) words) = probe row i has a build row with equal keys.
 * probe_keys and build_keys hold the key columns (same count, 1..QF_MAX_ARITY). */
int qf_left_semi_join(qf_context* ctx, const qf_table* probe_keys, const qf_table* build_keys,
                      uint64_t* d_out_bits, int64_t* h_num_ones);

/* RightSemiJoin::Next (operations/RightSemiJoin.hpp:45-82): bit j = build row j has a probe
 * row with equal keys; d_out_bits has ceil(build rows/64) words. */
int qf_right_semi_join(qf_context* ctx, const qf_table* probe_keys, const qf_table* build_keys,
                       uint64_t* d_out_bits, int64_t* h_num_ones);

/* Example coverage of a literal: the semi-join chain of QuickFoil::ComputeCoverageOnUncoveredData
 * and LiteralSelector::ComputeCoveredPositives (learner/QuickFoil.cpp:471-573,
 * learner/LiteralSelector.cpp:150-210; SelectAndCreateSemiJoin + BuildHashTableAfterSemiJoin +
 * CreateSemiJoin).  Binding rows [row_begin, row_end) that match `literal` (vars index binding
 * columns; -1 = new) are projected to their first examples->num_columns columns; bit e of
 * d_out_bits (ceil(examples rows / 64) words) = example row e equals one of those projections.
 * *h_num_covered = number of set bits. */
int qf_example_coverage(qf_context* ctx, const qf_table* binding, int64_t row_begin, int64_t row_end,
                        const qf_table* relation, const qf_literal* literal, const qf_table* examples,
                        uint64_t* d_out_bits, int64_t* h_num_covered);

/* BuildHashTableAfterSemiJoin (operations/BuildHashTable.cpp:127-189,212-236): among rows set
 * in d_bits, keep the first row of each distinct key (InsertIfNotPresent order) and build the
 * chained table over them: d_buckets (P2(n)), d_next (n; entries of rows not inserted are 0),
 * d_first (n bytes, 1 for inserted rows). */
int qf_build_hash_table_after_semi_join(qf_context* ctx, const qf_table* keys,
                                        const uint64_t* d_bits, int32_t* d_buckets,
                                        int32_t* d_next, uint8_t* d_first);

/* ---------------------------------------------------------------------------------------
 * Data loading (host only).
 * ------------------------------------------------------------------------------------- */

/* LoadData (main/main.cpp:49-109): parses a relation file's text, one tuple per line with
 * num_fields '|'-separated int32 fields (std::stoi rules, types/FromString.hpp:23-30); blank lines
 * and lines starting with '#' are skipped; keep (nullable) drops skipped arguments.  Writes the
 * kept fields column-major into h_columns (each of `capacity` rows) and the row count into
 * *h_num_rows.  With h_columns == NULL it only counts rows.  A malformed line returns
 * QF_ERR_INVALID with its 1-based number in *h_error_line; too small a capacity returns
 * QF_ERR_OVERFLOW (the reference writes one row past its buffer there, main/main.cpp:73-80). */
int qf_parse_relation_text(const char* text, int64_t length, int num_fields, const int* keep,
                           int32_t* const* h_columns, int64_t capacity, int64_t* h_num_rows,
                           int64_t* h_error_line);

/* ---- Learned join-cardinality estimator (no counterpart in the reference) ----
 * A 40 -> 64 -> 64 -> 1 ReLU MLP, BF16 operands with FP32 accumulation on MFMA, predicting
 * log2(1 + output rows) of a join from 40 host-computed features (inputs row-major, float32).
 * Its only use is presizing (qf_create_binding_table_presized), so no result depends on it.
 * Parameters: qf_mlp_num_params() floats, W1[64][48] (columns 40..47 unused), b1[64],
 * W2[64][64], b2[64], w3[64], b3.  The model's device memory comes from the context, which must
 * outlive it. */
typedef struct qf_mlp qf_mlp;
int qf_mlp_num_params(void);
int qf_mlp_create(qf_context* ctx, const float* h_params, qf_mlp** out);
void qf_mlp_destroy(qf_mlp* mlp);
int qf_mlp_get_params(qf_context* ctx, qf_mlp* mlp, float* h_params);
/* Batched inference: d_out[i] = prediction for row i of d_x (n x 40). */
int qf_mlp_forward(qf_context* ctx, qf_mlp* mlp, const float* d_x, int64_t n, float* d_out);
/* Gradients of mean((pred - y)^2) over the n rows (h_grads: qf_mlp_num_params() floats, may be
 * NULL) and the loss. */
int qf_mlp_gradients(qf_context* ctx, qf_mlp* mlp, const float* d_x, const float* d_y, int64_t n, float* h_grads,
                     float* h_loss);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on the gradients of the n rows. */
int qf_mlp_train_step(qf_context* ctx, qf_mlp* mlp, const float* d_x, const float* d_y, int64_t n, float lr,
                      float* h_loss);

#ifdef __cplusplus
}
#endif

#endif /* QUICKFOIL_AMD_H_ */
