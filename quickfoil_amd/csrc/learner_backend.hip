// The learner's evaluation backend over this library's device operators: tables are device
// columns; evaluation, binding-table extension and example coverage run on the context's stream.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "qf_internal.h"
#include "quickfoil_learner.h"

namespace {

struct DevTable {
  std::vector<int32_t*> cols;
  int64_t rows = 0;
};

struct DeviceBackend {
  qf_context* ctx;
  std::vector<DevTable> relations;
  std::vector<qf_table> rel_tables;
  std::vector<std::vector<const int32_t*>> rel_ptrs;
  std::vector<int> types;  // argument types of every relation, concatenated
};

int alloc_table(int ncols, int64_t rows, DevTable** out) {
  DevTable* t = new DevTable;
  t->rows = rows;
  for (int c = 0; c < ncols; ++c) {
    int32_t* p = nullptr;
    if (hipMalloc(&p, (size_t)std::max<int64_t>(1, rows) * 4) != hipSuccess) {
      for (int32_t* q : t->cols) (void)hipFree(q);
      delete t;
      return QF_ERR_OOM;
    }
    t->cols.push_back(p);
  }
  *out = t;
  return QF_OK;
}

qf_table view(const DevTable* t, std::vector<const int32_t*>* ptrs, int ncols) {
  ptrs->assign(t->cols.begin(), t->cols.begin() + ncols);
  qf_table q;
  q.columns = ptrs->data();
  q.num_columns = ncols;
  q.num_rows = t->rows;
  return q;
}

int be_release(void* self, void* table) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  DevTable* t = static_cast<DevTable*>(table);
  if (!t) return QF_OK;
  (void)qf_synchronize(b->ctx);
  for (int32_t* p : t->cols) (void)hipFree(p);
  delete t;
  return QF_OK;
}

int be_upload(void* self, const int32_t* const* cols, int ncols, int64_t rows, void** out) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  DevTable* t = nullptr;
  int st = alloc_table(ncols, rows, &t);
  if (st != QF_OK) return st;
  if (qf_synchronize(b->ctx) != QF_OK) st = QF_ERR_DEVICE;
  for (int c = 0; c < ncols && st == QF_OK; ++c)
    if (rows > 0 && hipMemcpy(t->cols[c], cols[c], (size_t)rows * 4, hipMemcpyHostToDevice) != hipSuccess)
      st = QF_ERR_DEVICE;
  if (st != QF_OK) {
    be_release(self, t);
    return st;
  }
  *out = t;
  return QF_OK;
}

int be_download(void* self, void* table, int ncols, int64_t rows, int32_t* const* cols) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  DevTable* t = static_cast<DevTable*>(table);
  if (qf_synchronize(b->ctx) != QF_OK) return QF_ERR_DEVICE;
  for (int c = 0; c < ncols; ++c)
    if (rows > 0 && hipMemcpy(cols[c], t->cols[c], (size_t)rows * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return QF_ERR_DEVICE;
  return QF_OK;
}

int be_concat(void* self, void* a, int64_t ra, void* bb, int64_t rb, int ncols, void** out) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  DevTable* t = nullptr;
  int st = alloc_table(ncols, ra + rb, &t);
  if (st != QF_OK) return st;
  if (qf_synchronize(b->ctx) != QF_OK) st = QF_ERR_DEVICE;
  for (int c = 0; c < ncols && st == QF_OK; ++c) {
    if (ra > 0 && hipMemcpy(t->cols[c], static_cast<DevTable*>(a)->cols[c], (size_t)ra * 4, hipMemcpyDeviceToDevice))
      st = QF_ERR_DEVICE;
    else if (rb > 0 &&
             hipMemcpy(t->cols[c] + ra, static_cast<DevTable*>(bb)->cols[c], (size_t)rb * 4, hipMemcpyDeviceToDevice))
      st = QF_ERR_DEVICE;
  }
  if (st != QF_OK) {
    be_release(self, t);
    return st;
  }
  *out = t;
  return QF_OK;
}

int be_evaluate(void* self, void* binding, int ncols, int64_t rows, int64_t npos, const qf_literal* lits, int n,
                uint64_t* counts) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  std::vector<const int32_t*> p;
  DevTable* t = static_cast<DevTable*>(binding);
  qf_table bt = view(t, &p, ncols);
  bt.num_rows = rows;
  // The relations never change while the backend lives, so their partitions and tables are
  // cached in the context (the reference caches them on the fact table,
  // learner/CandidateLiteralEvaluator.cpp:291-299); the destroy call drops them.
  qf_eval_options o;
  qf_eval_options_default(&o);
  o.reuse_partitions = 1;
  const int st = qf_evaluate_literals(b->ctx, &bt, npos, b->rel_tables.data(), b->types.data(),
                                      (int)b->rel_tables.size(), lits, n, &o, counts, nullptr);
  // A binding count above int32 is still exact in the uint64 counts the learner reads.
  return st == QF_ERR_COUNT_RANGE ? QF_OK : st;
}

int be_extend(void* self, void* binding, int ncols, int64_t rows, int64_t npos, const qf_literal* lit, int64_t bpos,
              int64_t bneg, void** out) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  if (lit->relation < 0 || lit->relation >= (int)b->rel_tables.size()) return QF_ERR_INVALID;
  const qf_table& rel = b->rel_tables[lit->relation];
  int extra = 0;
  for (int i = 0; i < rel.num_columns; ++i) extra += lit->vars[i] < 0 ? 1 : 0;
  DevTable* o = nullptr;
  int st = alloc_table(ncols + extra, bpos + bneg, &o);
  if (st != QF_OK) return st;
  std::vector<const int32_t*> p;
  qf_table bt = view(static_cast<DevTable*>(binding), &p, ncols);
  bt.num_rows = rows;
  st = qf_create_label_aware_binding_table(b->ctx, &bt, npos, &rel, lit, bpos, bneg, o->cols.data());
  if (st == QF_OK) st = qf_synchronize(b->ctx);
  if (st != QF_OK) {
    be_release(self, o);
    return st;
  }
  *out = o;
  return QF_OK;
}

int be_coverage(void* self, void* binding, int ncols, int64_t r0, int64_t r1, const qf_literal* lit, void* examples,
                int ex_cols, int64_t ex_rows, uint64_t* h_bits, int64_t* covered) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  if (lit->relation < 0 || lit->relation >= (int)b->rel_tables.size()) return QF_ERR_INVALID;
  std::vector<const int32_t*> p, pe;
  DevTable* bt_ = static_cast<DevTable*>(binding);
  qf_table bt = view(bt_, &p, ncols);
  bt.num_rows = std::max(bt_->rows, r1);
  DevTable* et_ = static_cast<DevTable*>(examples);
  qf_table et = view(et_, &pe, ex_cols);
  et.num_rows = ex_rows;
  const int64_t words = std::max<int64_t>(1, (ex_rows + 63) / 64);
  uint64_t* d_bits = nullptr;
  if (hipMalloc(&d_bits, (size_t)words * 8) != hipSuccess) return QF_ERR_OOM;
  int st = qf_example_coverage(b->ctx, &bt, r0, r1, &b->rel_tables[lit->relation], lit, &et, d_bits, covered);
  if (st == QF_OK) st = qf_synchronize(b->ctx);
  if (st == QF_OK && hipMemcpy(h_bits, d_bits, (size_t)words * 8, hipMemcpyDeviceToHost) != hipSuccess)
    st = QF_ERR_DEVICE;
  (void)hipFree(d_bits);
  return st;
}

int be_join(void* self, void* binding, int ncols, int64_t rows, const qf_literal* lit, void** out, int64_t* out_rows) {
  DeviceBackend* b = static_cast<DeviceBackend*>(self);
  if (lit->relation < 0 || lit->relation >= (int)b->rel_tables.size()) return QF_ERR_INVALID;
  const qf_table& rel = b->rel_tables[lit->relation];
  int extra = 0;
  for (int i = 0; i < rel.num_columns; ++i) extra += lit->vars[i] >= ncols ? 1 : 0;
  std::vector<const int32_t*> p;
  qf_table bt = view(static_cast<DevTable*>(binding), &p, ncols);
  bt.num_rows = rows;
  int64_t n = 0;
  int st = qf_create_binding_table(b->ctx, &bt, &rel, lit, nullptr, &n);
  if (st != QF_OK) return st;
  DevTable* o = nullptr;
  st = alloc_table(ncols + extra, n, &o);
  if (st != QF_OK) return st;
  st = qf_create_binding_table(b->ctx, &bt, &rel, lit, o->cols.data(), &n);
  if (st == QF_OK) st = qf_synchronize(b->ctx);
  if (st != QF_OK) {
    be_release(self, o);
    return st;
  }
  *out = o;
  *out_rows = n;
  return QF_OK;
}

}  // namespace

extern "C" int qf_learner_device_backend_create(qf_context* ctx, int num_relations, const int* arities,
                                                const int32_t* const* const* relation_columns,
                                                const int64_t* relation_rows, const int* const* argument_types,
                                                qf_learner_backend* backend) {
  if (!ctx || !backend || num_relations < 0 || (num_relations > 0 && (!arities || !relation_columns || !relation_rows)))
    return QF_ERR_INVALID;
  DeviceBackend* b = new DeviceBackend;
  b->ctx = ctx;
  b->relations.resize(num_relations);
  b->rel_ptrs.resize(num_relations);
  auto fail = [&](int st) {
    for (DevTable& t : b->relations)
      for (int32_t* p : t.cols) (void)hipFree(p);
    delete b;
    return st;
  };
  for (int r = 0; r < num_relations; ++r) {
    if (arities[r] < 1 || arities[r] > QF_MAX_ARITY) return fail(QF_ERR_INVALID);
    DevTable* t = nullptr;
    void* h = nullptr;
    int st = be_upload(b, relation_columns[r], arities[r], relation_rows[r], &h);
    if (st != QF_OK) return fail(st);
    t = static_cast<DevTable*>(h);
    b->relations[r] = *t;
    delete t;
    for (int c = 0; c < arities[r]; ++c) b->types.push_back(argument_types ? argument_types[r][c] : 0);
  }
  for (int r = 0; r < num_relations; ++r) {
    b->rel_tables.push_back(view(&b->relations[r], &b->rel_ptrs[r], arities[r]));
  }
  std::memset(backend, 0, sizeof(*backend));
  backend->self = b;
  backend->upload = be_upload;
  backend->download = be_download;
  backend->release = be_release;
  backend->concat = be_concat;
  backend->evaluate = be_evaluate;
  backend->extend = be_extend;
  backend->coverage = be_coverage;
  backend->join = be_join;
  return QF_OK;
}

extern "C" void qf_learner_device_backend_destroy(qf_learner_backend* backend) {
  if (!backend || !backend->self) return;
  DeviceBackend* b = static_cast<DeviceBackend*>(backend->self);
  // Cached partitions are keyed by column address; a later allocation may reuse these.
  (void)qf_context_trim(b->ctx);
  for (DevTable& t : b->relations)
    for (int32_t* p : t.cols) (void)hipFree(p);
  delete b;
  backend->self = nullptr;
}
