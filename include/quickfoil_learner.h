/* quickfoil_learner.h — C ABI of the host learner (QuickFoil search control) over a
 * pluggable evaluation backend.
 *
 * The learner restates learner/QuickFoil.cpp, learner/LiteralSelector.cpp and
 * learner/CandidateLiteralEnumerator.cpp; every count, binding table and coverage bitvector it
 * uses comes from the backend, so two backends with identical results learn identical clauses.
 * The device backend (qf_learner_device_backend) runs on libquickfoil_amd's operators;
 * any other backend fills the same table of functions. */
#ifndef QUICKFOIL_LEARNER_H_
#define QUICKFOIL_LEARNER_H_

#include <stdint.h>

#include "quickfoil_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Evaluation backend.  Tables are opaque handles owned by the backend (int32 SoA columns).
 * Relations are addressed by index (qf_literal.relation); every function returns a qf_status. */
typedef struct qf_learner_backend {
  void* self;
  /* Copies host columns into a new table. */
  int (*upload)(void* self, const int32_t* const* columns, int num_columns, int64_t num_rows, void** table);
  /* Copies a table's columns to host buffers of num_rows each. */
  int (*download)(void* self, void* table, int num_columns, int64_t num_rows, int32_t* const* columns);
  int (*release)(void* self, void* table);
  /* Rows of a followed by rows of b (same column count). */
  int (*concat)(void* self, void* a, int64_t rows_a, void* b, int64_t rows_b, int num_columns, void** out);
  /* The four counts of every literal over a binding table whose first num_positive rows are
   * positive (CandidateLiteralEvaluator::Evaluate, learner/CandidateLiteralEvaluator.cpp:264-406). */
  int (*evaluate)(void* self, void* binding, int num_columns, int64_t num_rows, int64_t num_positive,
                  const qf_literal* literals, int num_literals, uint64_t* counts);
  /* CreateLabelAwareBindingTables (operations/MultiColumnHashJoin.cpp:117-256): the extended
   * table (binding columns, then one column per new variable of the literal). */
  int (*extend)(void* self, void* binding, int num_columns, int64_t num_rows, int64_t num_positive,
                const qf_literal* literal, int64_t binding_pos, int64_t binding_neg, void** out);
  /* qf_example_coverage over binding rows [row_begin, row_end); bits over example rows. */
  int (*coverage)(void* self, void* binding, int num_columns, int64_t row_begin, int64_t row_end,
                  const qf_literal* literal, void* examples, int num_example_columns, int64_t num_examples,
                  uint64_t* h_bits, int64_t* num_covered);
  /* CreateBindingTable (operations/MultiColumnHashJoin.cpp:26-115): the binding table joined
   * with a literal whose variables >= num_columns are new columns; *out_rows = its rows. */
  int (*join)(void* self, void* binding, int num_columns, int64_t num_rows, const qf_literal* literal, void** out,
              int64_t* out_rows);
} qf_learner_backend;

/* Learner flags and their reference defaults (learner/QuickFoil.cpp:38-73,
 * learner/LiteralSelector.cpp:28-34). */
typedef struct qf_learner_options {
  double positive_threshold;                /* 0.8 */
  int max_iterations;                       /* 1000 */
  double minimum_inflated_precision;        /* 0.85 */
  double minimum_true_precision;            /* 0.8 */
  double minimum_f_score;                   /* 0.85 */
  int maximum_clause_length;                /* 25 */
  int maximum_random_literals;              /* 2 */
  int maximum_random_trials;                /* 5 */
  double minimum_coverage_for_tied_literal; /* 0.1 */
  int num_saved_literals;                   /* 5 */
  double random_margin;                     /* 0.03 */
} qf_learner_options;

/* A learning problem: background predicates 0..num_background-1 (relation index = predicate
 * index), then the target, whose first num_positive facts are positive (main/main.cpp:238-246). */
typedef struct qf_learner_problem {
  int num_background;
  const char* const* names;        /* num_background + 1 names, the target last */
  const int* arities;              /* num_background + 1 */
  const int* const* argument_types;/* num_background + 1 arrays */
  const int* keys;                 /* num_background + 1; -1 = no key */
  const int64_t* num_facts;        /* num_background + 1 */
  const int32_t* const* target_columns; /* host columns of the target facts */
  int64_t num_positive;
} qf_learner_problem;

void qf_learner_options_default(qf_learner_options* options);

/* QuickFoil::Learn (learner/QuickFoil.cpp:163-329).  The learnt clauses are written to `out` as
 * lines in FoilClause::ToString form (schema/FoilClause.cpp:31-41), NUL-terminated;
 * *num_clauses receives their number.  QF_ERR_OVERFLOW when `out` is too small. */
int qf_learn(const qf_learner_problem* problem, const qf_learner_options* options,
             const qf_learner_backend* backend, char* out, int64_t out_capacity, int* num_clauses);

/* QuickFoilTestRunner::RunTest (learner/QuickFoilTestRunner.cpp:135-160): test examples (head
 * arity columns) left uncovered by the clauses.  Clause c has body_lengths[c] body literals,
 * concatenated in `body` (variables index the clause's variables; the head binds 0..arity-1). */
int qf_learner_test(const qf_learner_backend* backend, const int32_t* const* test_columns, int head_arity,
                    int64_t num_tests, const qf_literal* body, const int* body_lengths, int num_clauses,
                    int64_t* num_uncovered);

/* Backend over libquickfoil_amd on a context: relations are the host columns given here (uploaded
 * once); tables live in device memory.  Release with qf_learner_device_backend_destroy. */
int qf_learner_device_backend_create(qf_context* ctx, int num_relations, const int* arities,
                                     const int32_t* const* const* relation_columns, const int64_t* relation_rows,
                                     const int* const* argument_types, qf_learner_backend* backend);
void qf_learner_device_backend_destroy(qf_learner_backend* backend);

#ifdef __cplusplus
}
#endif

#endif /* QUICKFOIL_LEARNER_H_ */
