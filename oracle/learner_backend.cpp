// TEST INFRASTRUCTURE ONLY: the learner's evaluation backend over the CPU oracle (host tables,
// the restated reference pipeline for counts, MultiColumnHashJoin for extension, the semi-join
// chain for example coverage).  Tests hand it to qf_learn beside the device backend; the two must
// learn identical clauses.
#include <cstring>
#include <vector>

#include "qf_oracle.h"
#include "quickfoil_learner.h"

namespace {

struct HostTable {
  std::vector<std::vector<int32_t>> cols;
  int64_t rows = 0;
};

struct CpuBackend {
  std::vector<HostTable> rels;
  std::vector<const int32_t*> rel_cols;
  std::vector<int> col_start, arity, types;
  std::vector<int64_t> rows;
};

std::vector<const int32_t*> ptrs(const HostTable& t, int ncols, int64_t offset = 0) {
  std::vector<const int32_t*> p;
  for (int c = 0; c < ncols; ++c) p.push_back(t.cols[c].data() + offset);
  return p;
}

orc_eval_input input(CpuBackend* b, const std::vector<const int32_t*>& bcols, int64_t rows, int64_t npos,
                     const std::vector<int>& lit_rel, const std::vector<int>& lit_vars) {
  orc_eval_input in;
  in.binding_columns = bcols.data();
  in.num_binding_columns = (int)bcols.size();
  in.num_bindings = rows;
  in.num_positive = npos;
  in.rel_columns = b->rel_cols.data();
  in.rel_col_start = b->col_start.data();
  in.rel_arity = b->arity.data();
  in.rel_rows = b->rows.data();
  in.arg_types = b->types.data();
  in.num_relations = (int)b->rels.size();
  in.lit_rel = lit_rel.data();
  in.lit_vars = lit_vars.data();
  in.num_literals = (int)lit_rel.size();
  return in;
}

int upload(void*, const int32_t* const* cols, int ncols, int64_t rows, void** out) {
  HostTable* t = new HostTable;
  t->rows = rows;
  for (int c = 0; c < ncols; ++c) t->cols.emplace_back(cols[c], cols[c] + rows);
  *out = t;
  return QF_OK;
}

int download(void*, void* table, int ncols, int64_t rows, int32_t* const* cols) {
  HostTable* t = static_cast<HostTable*>(table);
  for (int c = 0; c < ncols; ++c) std::memcpy(cols[c], t->cols[c].data(), (size_t)rows * 4);
  return QF_OK;
}

int release(void*, void* table) {
  delete static_cast<HostTable*>(table);
  return QF_OK;
}

int concat(void*, void* a, int64_t ra, void* b, int64_t rb, int ncols, void** out) {
  HostTable* t = new HostTable;
  t->rows = ra + rb;
  for (int c = 0; c < ncols; ++c) {
    std::vector<int32_t> col(static_cast<HostTable*>(a)->cols[c].begin(), static_cast<HostTable*>(a)->cols[c].begin() + ra);
    col.insert(col.end(), static_cast<HostTable*>(b)->cols[c].begin(), static_cast<HostTable*>(b)->cols[c].begin() + rb);
    t->cols.push_back(std::move(col));
  }
  *out = t;
  return QF_OK;
}

int evaluate(void* self, void* binding, int ncols, int64_t rows, int64_t npos, const qf_literal* lits, int n,
             uint64_t* counts) {
  CpuBackend* b = static_cast<CpuBackend*>(self);
  std::vector<const int32_t*> bc = ptrs(*static_cast<HostTable*>(binding), ncols);
  std::vector<int> lr, lv;
  for (int i = 0; i < n; ++i) {
    lr.push_back(lits[i].relation);
    for (int k = 0; k < ORC_MAX_ARITY; ++k) lv.push_back(lits[i].vars[k]);
  }
  orc_eval_input in = input(b, bc, rows, npos, lr, lv);
  // A binding count above int32 is still exact in the uint64 counts (the device backend does the same).
  (void)orc_evaluate_pipeline(&in, 5, 32768, counts, nullptr);
  return QF_OK;
}

int extend(void* self, void* binding, int ncols, int64_t rows, int64_t npos, const qf_literal* lit, int64_t bpos,
           int64_t bneg, void** out) {
  CpuBackend* b = static_cast<CpuBackend*>(self);
  std::vector<const int32_t*> bc = ptrs(*static_cast<HostTable*>(binding), ncols);
  std::vector<int> lr{lit->relation}, lv(lit->vars, lit->vars + ORC_MAX_ARITY);
  orc_eval_input in = input(b, bc, rows, npos, lr, lv);
  int extra = 0;
  for (int i = 0; i < b->arity[lit->relation]; ++i) extra += lit->vars[i] < 0 ? 1 : 0;
  HostTable* t = new HostTable;
  t->rows = bpos + bneg;
  std::vector<int32_t*> oc;
  for (int c = 0; c < ncols + extra; ++c) {
    t->cols.emplace_back((size_t)(bpos + bneg + 1), 0);
    oc.push_back(t->cols.back().data());
  }
  if (orc_create_label_aware_binding_table(&in, 0, bpos, bneg, oc.data()) != bpos + bneg) {
    delete t;
    return QF_ERR_OVERFLOW;
  }
  *out = t;
  return QF_OK;
}

int coverage(void* self, void* binding, int ncols, int64_t r0, int64_t r1, const qf_literal* lit, void* examples,
             int ex_cols, int64_t ex_rows, uint64_t* bits, int64_t* covered) {
  CpuBackend* b = static_cast<CpuBackend*>(self);
  const HostTable& bt = *static_cast<HostTable*>(binding);
  const HostTable& rel = b->rels[lit->relation];
  std::vector<const int32_t*> bk, rk;
  for (int i = 0; i < b->arity[lit->relation]; ++i)
    if (lit->vars[i] >= 0) {
      bk.push_back(bt.cols[lit->vars[i]].data() + r0);
      rk.push_back(rel.cols[i].data());
    }
  const int64_t n = r1 - r0;
  std::vector<uint64_t> matched((size_t)(n + 63) / 64 + 1, 0);
  if (n > 0) orc_left_semi_join(bk.data(), n, rk.data(), rel.rows, (int)bk.size(), matched.data());
  // Head projections of the matched rows.
  std::vector<std::vector<int32_t>> head(ex_cols);
  for (int64_t r = 0; r < n; ++r)
    if ((matched[r >> 6] >> (r & 63)) & 1ull)
      for (int c = 0; c < ex_cols; ++c) head[c].push_back(bt.cols[c][r0 + r]);
  std::vector<const int32_t*> hk, ek;
  for (int c = 0; c < ex_cols; ++c) {
    hk.push_back(head[c].data());
    ek.push_back(static_cast<HostTable*>(examples)->cols[c].data());
  }
  const int64_t words = (ex_rows + 63) / 64;
  std::memset(bits, 0, (size_t)words * 8);
  if (ex_rows > 0) orc_left_semi_join(ek.data(), ex_rows, hk.data(), (int64_t)head[0].size(), ex_cols, bits);
  int64_t ones = 0;
  for (int64_t w = 0; w < words; ++w) ones += __builtin_popcountll(bits[w]);
  *covered = ones;
  return QF_OK;
}

int join(void* self, void* binding, int ncols, int64_t rows, const qf_literal* lit, void** out, int64_t* out_rows) {
  CpuBackend* b = static_cast<CpuBackend*>(self);
  std::vector<const int32_t*> bc = ptrs(*static_cast<HostTable*>(binding), ncols);
  std::vector<int> lr{lit->relation}, lv(lit->vars, lit->vars + ORC_MAX_ARITY);
  orc_eval_input in = input(b, bc, rows, 0, lr, lv);
  const int64_t n = orc_create_binding_table(&in, 0, nullptr);
  if (n < 0) return QF_ERR_INVALID;
  int extra = 0;
  for (int i = 0; i < b->arity[lit->relation]; ++i) extra += lit->vars[i] >= ncols ? 1 : 0;
  HostTable* t = new HostTable;
  t->rows = n;
  std::vector<int32_t*> oc;
  for (int c = 0; c < ncols + extra; ++c) {
    t->cols.emplace_back((size_t)(n + 1), 0);
    oc.push_back(t->cols.back().data());
  }
  orc_create_binding_table(&in, 0, oc.data());
  *out = t;
  *out_rows = n;
  return QF_OK;
}

}  // namespace

extern "C" int orc_learner_backend_create(int num_relations, const int* arities,
                                          const int32_t* const* const* relation_columns, const int64_t* relation_rows,
                                          const int* const* argument_types, qf_learner_backend* backend) {
  CpuBackend* b = new CpuBackend;
  int start = 0;
  for (int r = 0; r < num_relations; ++r) {
    HostTable t;
    t.rows = relation_rows[r];
    for (int c = 0; c < arities[r]; ++c)
      t.cols.emplace_back(relation_columns[r][c], relation_columns[r][c] + relation_rows[r]);
    b->rels.push_back(std::move(t));
    b->col_start.push_back(start);
    b->arity.push_back(arities[r]);
    b->rows.push_back(relation_rows[r]);
    for (int c = 0; c < arities[r]; ++c) b->types.push_back(argument_types ? argument_types[r][c] : 0);
    start += arities[r];
  }
  for (int r = 0; r < num_relations; ++r)
    for (int c = 0; c < arities[r]; ++c) b->rel_cols.push_back(b->rels[r].cols[c].data());
  std::memset(backend, 0, sizeof(*backend));
  backend->self = b;
  backend->upload = upload;
  backend->download = download;
  backend->release = release;
  backend->concat = concat;
  backend->evaluate = evaluate;
  backend->extend = extend;
  backend->coverage = coverage;
  backend->join = join;
  return QF_OK;
}

extern "C" void orc_learner_backend_destroy(qf_learner_backend* backend) {
  delete static_cast<CpuBackend*>(backend->self);
  backend->self = nullptr;
}
