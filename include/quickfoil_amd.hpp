/*
 * quickfoil_amd.hpp: a C++ adapter over the C ABI of quickfoil_amd.h that keeps the reference's
 * operator names and calling shapes, for a host C++ caller such as a restated
 * CandidateLiteralEvaluator:
 *
 *   RadixPartition(column_id, table)              operations/RadixPartition.hpp:13-14
 *   BuildHashTableOnPartitions(column_id, table)  operations/BuildHashTable.hpp:19-21
 *   BuildHashTableOnTable(build_keys, table)      operations/BuildHashTable.hpp:23-25
 *   CountAggregator::Execute / ExecuteOnPositives / ExecuteOnNegatives
 *                                                 operations/CountAggregator.hpp:18-53, writing
 *                                                 into CandidateLiteralInfo (learner/CandidateLiteralInfo.hpp:15-26)
 *   EvaluateCandidateLiterals                     CandidateLiteralEvaluator::Evaluate
 *                                                 (learner/CandidateLiteralEvaluator.cpp:264-406)
 *
 * Each operator takes a Context first (the reference's operators are free functions over global
 * state) and, like the reference's, has finished its device work when it returns (it synchronises
 * the context's stream, which host copies on the default stream do not wait for).  Tables are
 * device-resident int32 columns viewed by DeviceTableView, which carries the per-column partition
 * and hash-table caches that TableView carries (storage/TableView.hpp:22-88).
 * Counts are uint64; ToSizeType converts one to the reference's int32 size_type
 * (schema/TypeDefs.hpp:13) and refuses values that would wrap.  Where the reference aborts
 * through glog CHECK, the adapter throws quickfoil_amd::Error (status and qf_last_error text).
 *
 * Header-only: compile with the HIP runtime headers (-D__HIP_PLATFORM_AMD__ -I/opt/rocm/include)
 * and link libquickfoil_amd.so and libamdhip64.so.
 */
#ifndef QUICKFOIL_AMD_HPP_
#define QUICKFOIL_AMD_HPP_

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "quickfoil_amd.h"

namespace quickfoil_amd {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void CheckHip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(QF_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// A library context (device, stream, memory pool, cached relation partitions).
class Context {
 public:
  explicit Context(int device = 0) {
    const int st = qf_context_create(device, &ctx_);
    if (st != QF_OK) throw Error(st, "qf_context_create failed");
  }
  ~Context() { qf_context_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  qf_context* get() const { return ctx_; }
  void Check(int status, const char* what) const {
    if (status != QF_OK) throw Error(status, std::string(what) + ": " + qf_last_error(ctx_));
  }
  void Synchronize() const { Check(qf_synchronize(ctx_), "qf_synchronize"); }

  // A process-wide context on `device`, made on first use, for callers without one of their own.
  static Context& Default(int device = 0) {
    static Context ctx(device);
    return ctx;
  }

 private:
  qf_context* ctx_ = nullptr;
};

// Device memory owned by the adapter (hipMalloc / hipFree).
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(int64_t n) : n_(n) {
    CheckHip(hipMalloc(reinterpret_cast<void**>(&p_), (size_t)(n > 0 ? n : 1) * sizeof(T)), "hipMalloc");
  }
  ~DeviceArray() {
    if (p_) (void)hipFree(p_);
  }
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      if (p_) (void)hipFree(p_);
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
    }
    return *this;
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;

  T* data() const { return p_; }
  int64_t size() const { return n_; }
  std::vector<T> ToHost() const {
    std::vector<T> h((size_t)n_);
    if (n_ > 0) CheckHip(hipMemcpy(h.data(), p_, (size_t)n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
  }

 private:
  T* p_ = nullptr;
  int64_t n_ = 0;
};

// FoilHashTable (storage/FoilHashTable.hpp:19-59) in device memory: 1-based bucket heads
// (0 = empty) and 1-based chain links (0 = end).
struct DeviceFoilHashTable {
  DeviceArray<int32_t> buckets;
  DeviceArray<int32_t> next;
};

// TableView over int32 device columns (storage/TableView.hpp:22-88), with the caches of
// partitions_at / hash_tables_at per column.
class DeviceTableView {
 public:
  // PartitionTuple pairs {value, tuple id} in partition order and the 2^radix_bits + 1 offsets
  // (storage/PartitionTuple.hpp:20-42).
  struct Partitions {
    int radix_bits = 0;
    DeviceArray<int32_t> pairs;
    DeviceArray<int64_t> offsets;
    std::vector<int64_t> host_offsets;
    bool empty() const { return radix_bits == 0; }
  };
  // One FoilHashTable per partition, concatenated: partition p's buckets start at
  // bucket_offsets[p]; its links at the partition's tuple offset (BuildHashTable.cpp:54-93).
  struct PartitionHashTables {
    DeviceArray<int32_t> buckets;
    DeviceArray<int32_t> next;
    std::vector<int64_t> bucket_offsets;
    bool empty() const { return bucket_offsets.empty(); }
  };

  DeviceTableView(std::vector<const int32_t*> columns, int64_t num_tuples)
      : columns_(std::move(columns)), rows_(num_tuples), partitions_(columns_.size()), hash_tables_(columns_.size()) {
    if (columns_.empty()) throw Error(QF_ERR_INVALID, "a table needs at least one column");
  }
  DeviceTableView(const DeviceTableView&) = delete;
  DeviceTableView& operator=(const DeviceTableView&) = delete;

  int num_columns() const { return (int)columns_.size(); }
  int64_t num_tuples() const { return rows_; }
  bool empty() const { return rows_ == 0; }
  const int32_t* column_at(int i) const { return columns_.at(i); }
  // The C ABI's view of this table (valid while the object lives).
  qf_table table() const { return qf_table{columns_.data(), (int)columns_.size(), rows_}; }

  const Partitions& partitions_at(int column_id) const { return partitions_.at(column_id); }
  Partitions* mutable_partitions_at(int column_id) { return &partitions_.at(column_id); }
  const PartitionHashTables& hash_tables_at(int column_id) const { return hash_tables_.at(column_id); }
  PartitionHashTables* mutable_hash_tables_at(int column_id) { return &hash_tables_.at(column_id); }

 private:
  std::vector<const int32_t*> columns_;
  int64_t rows_;
  std::vector<Partitions> partitions_;
  std::vector<PartitionHashTables> hash_tables_;
};

// A device copy of host int32 columns that owns its memory.
class DeviceTable {
 public:
  static std::unique_ptr<DeviceTable> Upload(const std::vector<const int32_t*>& host_columns, int64_t rows) {
    std::unique_ptr<DeviceTable> t(new DeviceTable);
    std::vector<const int32_t*> ptrs;
    for (const int32_t* h : host_columns) {
      t->columns_.emplace_back(rows);
      if (rows > 0)
        CheckHip(hipMemcpy(t->columns_.back().data(), h, (size_t)rows * 4, hipMemcpyHostToDevice), "hipMemcpy");
      ptrs.push_back(t->columns_.back().data());
    }
    t->view_.reset(new DeviceTableView(ptrs, rows));
    return t;
  }
  const DeviceTableView& view() const { return *view_; }
  DeviceTableView* mutable_view() { return view_.get(); }

 private:
  DeviceTable() = default;
  std::vector<DeviceArray<int32_t>> columns_;
  std::unique_ptr<DeviceTableView> view_;
};

// Host tables that never change (the background facts), uploaded on first use and kept.
class DeviceTableCache {
 public:
  const DeviceTable& Get(const std::vector<const int32_t*>& host_columns, int64_t rows) {
    auto key = std::make_pair(host_columns, rows);
    auto it = tables_.find(key);
    if (it == tables_.end()) it = tables_.emplace(key, DeviceTable::Upload(host_columns, rows)).first;
    return *it->second;
  }

 private:
  std::map<std::pair<std::vector<const int32_t*>, int64_t>, std::unique_ptr<DeviceTable>> tables_;
};

// RadixPartition(int column_id, TableView*) (operations/RadixPartition.hpp:13-14): fills the
// table's partition cache for the column.  radix_bits defaults to the reference's
// FLAGS_num_radix_bits (operations/RadixPartition.cpp:26).
inline void RadixPartition(Context& ctx, int column_id, DeviceTableView* table, int radix_bits = 5) {
  DeviceTableView::Partitions* p = table->mutable_partitions_at(column_id);
  if (!p->empty()) throw Error(QF_ERR_INVALID, "column already partitioned");
  const int64_t n = table->num_tuples();
  p->pairs = DeviceArray<int32_t>(2 * n);
  p->offsets = DeviceArray<int64_t>(((int64_t)1 << radix_bits) + 1);
  ctx.Check(qf_radix_partition(ctx.get(), table->column_at(column_id), n, radix_bits, p->pairs.data(),
                               p->offsets.data()),
            "qf_radix_partition");
  ctx.Synchronize();
  p->host_offsets = p->offsets.ToHost();
  p->radix_bits = radix_bits;
}

// BuildHashTableOnPartitions(int column_id, TableView*) (operations/BuildHashTable.hpp:19-21):
// one chained table per partition of a partitioned column.
inline void BuildHashTableOnPartitions(Context& ctx, int column_id, DeviceTableView* table) {
  const DeviceTableView::Partitions& p = table->partitions_at(column_id);
  if (p.empty()) throw Error(QF_ERR_INVALID, "BuildHashTableOnPartitions needs RadixPartition first");
  DeviceTableView::PartitionHashTables* h = table->mutable_hash_tables_at(column_id);
  const int parts = 1 << p.radix_bits;
  h->bucket_offsets.assign(parts + 1, 0);
  const int64_t total = qf_partition_bucket_layout(p.host_offsets.data(), parts, h->bucket_offsets.data());
  if (total < 0) throw Error((int)-total, "qf_partition_bucket_layout failed");
  h->buckets = DeviceArray<int32_t>(total);
  h->next = DeviceArray<int32_t>(table->num_tuples());
  DeviceArray<int64_t> dev_offsets(parts + 1);
  CheckHip(hipMemcpy(dev_offsets.data(), h->bucket_offsets.data(), (size_t)(parts + 1) * 8, hipMemcpyHostToDevice),
           "hipMemcpy");
  ctx.Check(qf_build_hash_table_on_partitions(ctx.get(), p.pairs.data(), p.offsets.data(), p.radix_bits,
                                              dev_offsets.data(), total, table->num_tuples(), h->buckets.data(),
                                              h->next.data()),
            "qf_build_hash_table_on_partitions");
  ctx.Synchronize();  // dev_offsets is released on return
}

// FoilHashTable* BuildHashTableOnTable(build_keys, table) (operations/BuildHashTable.hpp:23-25):
// the caller owns the result.
inline std::unique_ptr<DeviceFoilHashTable> BuildHashTableOnTable(Context& ctx, const std::vector<int>& build_keys,
                                                                  const DeviceTableView& table) {
  std::vector<const int32_t*> keys;
  for (int k : build_keys) keys.push_back(table.column_at(k));
  const qf_table kt{keys.data(), (int)keys.size(), table.num_tuples()};
  int64_t buckets = 1;
  while (buckets < table.num_tuples()) buckets <<= 1;  // P2(n) buckets
  std::unique_ptr<DeviceFoilHashTable> t(new DeviceFoilHashTable);
  t->buckets = DeviceArray<int32_t>(buckets);
  t->next = DeviceArray<int32_t>(table.num_tuples());
  ctx.Check(qf_build_hash_table_on_table(ctx.get(), &kt, t->buckets.data(), t->next.data()),
            "qf_build_hash_table_on_table");
  ctx.Synchronize();
  return t;
}

// CandidateLiteralInfo (learner/CandidateLiteralInfo.hpp:15-26) with 64-bit counts.
struct CandidateLiteralInfo {
  qf_literal literal{};
  uint64_t num_covered_positive = 0;
  uint64_t num_covered_negative = 0;
  uint64_t num_binding_positive = 0;
  uint64_t num_binding_negative = 0;
};

// A count as the reference's size_type (int32_t, schema/TypeDefs.hpp:13); above INT32_MAX the
// reference's counters would wrap, so this throws QF_ERR_COUNT_RANGE instead.
inline int32_t ToSizeType(uint64_t count) {
  if (count > 0x7fffffffull)
    throw Error(QF_ERR_COUNT_RANGE, "count " + std::to_string(count) + " exceeds the reference's int32 size_type");
  return (int32_t)count;
}

// PredicateEvaluationPlan of one join group (learner/PredicateEvaluationPlan.hpp:88-112) in flat
// form: the join equality relation[join_column] == binding[binding_column]; atom k is the
// equality relation[atoms[k].first] == binding[atoms[k].second]; node ids are the atoms, then the
// conjunctions (children before parents); node_literals (one per node, may be null) and the root
// `literal` (the join equality alone) receive counts.
struct PredicateEvaluationPlan {
  int relation = 0;
  int join_column = 0;
  int binding_column = 0;
  std::vector<std::pair<int, int>> atoms;
  std::vector<std::pair<int, int>> conjunctions;
  std::vector<CandidateLiteralInfo*> node_literals;
  CandidateLiteralInfo* literal = nullptr;
};

// CountAggregator (operations/CountAggregator.hpp:18-53) over one binding table, the background
// relations and the join groups' plans: the Execute calls write the counts into the plans'
// CandidateLiteralInfo objects by side effect, as the reference's do.
class CountAggregator {
 public:
  CountAggregator(Context& ctx, const DeviceTableView& binding, std::vector<const DeviceTableView*> relations,
                  std::vector<PredicateEvaluationPlan> plans, const qf_eval_options* options = nullptr)
      : ctx_(ctx), binding_(binding), relations_(std::move(relations)), plans_(std::move(plans)) {
    qf_eval_options_default(&options_);
    if (options) options_ = *options;
  }

  // Execute(num_positive): rows below num_positive are positive (schema/FoilClause.hpp:77-99).
  void Execute(int64_t num_positive) { Run(num_positive, true, true); }
  // The two passes over separate positive and negative tables (CountAggregator.hpp:27-29).
  void ExecuteOnPositives() { Run(binding_.num_tuples(), true, false); }
  void ExecuteOnNegatives() { Run(0, false, true); }

 private:
  void Run(int64_t num_positive, bool positives, bool negatives) {
    std::vector<qf_table> rels;
    for (const DeviceTableView* r : relations_) rels.push_back(r->table());
    std::vector<CandidateLiteralInfo*> slots;
    std::map<CandidateLiteralInfo*, int> slot_of;
    auto slot = [&](CandidateLiteralInfo* info) {
      if (!info) return -1;
      auto it = slot_of.find(info);
      if (it != slot_of.end()) return it->second;
      slot_of[info] = (int)slots.size();
      slots.push_back(info);
      return (int)slots.size() - 1;
    };
    std::vector<std::vector<int>> bg(plans_.size()), bind(plans_.size()), left(plans_.size()),
        right(plans_.size()), lit(plans_.size());
    std::vector<qf_join_group_plan> flat(plans_.size());
    for (size_t p = 0; p < plans_.size(); ++p) {
      const PredicateEvaluationPlan& pl = plans_[p];
      const int na = (int)pl.atoms.size(), nn = na + (int)pl.conjunctions.size();
      if ((int)pl.node_literals.size() != nn) throw Error(QF_ERR_INVALID, "one literal entry per plan node");
      for (const auto& a : pl.atoms) {
        bg[p].push_back(a.first);
        bind[p].push_back(a.second);
      }
      for (int x = 0; x < nn; ++x) {
        left[p].push_back(x < na ? -1 : pl.conjunctions[x - na].first);
        right[p].push_back(x < na ? -1 : pl.conjunctions[x - na].second);
        lit[p].push_back(slot(pl.node_literals[x]));
      }
      flat[p] = qf_join_group_plan{pl.relation,   pl.join_column, pl.binding_column, na,
                                   bg[p].data(),  bind[p].data(), nn,                left[p].data(),
                                   right[p].data(), lit[p].data(), slot(pl.literal)};
    }
    std::vector<uint64_t> counts(4 * (slots.empty() ? 1 : slots.size()));
    const qf_table b = binding_.table();
    ctx_.Check(qf_count_aggregate(ctx_.get(), &b, num_positive, rels.data(), (int)rels.size(), flat.data(),
                                  (int)flat.size(), (int)slots.size(), &options_, counts.data(), nullptr),
               "qf_count_aggregate");
    for (size_t s = 0; s < slots.size(); ++s) {
      if (positives) {
        slots[s]->num_covered_positive = counts[4 * s + 0];
        slots[s]->num_binding_positive = counts[4 * s + 2];
      }
      if (negatives) {
        slots[s]->num_covered_negative = counts[4 * s + 1];
        slots[s]->num_binding_negative = counts[4 * s + 3];
      }
    }
  }

  Context& ctx_;
  const DeviceTableView& binding_;
  std::vector<const DeviceTableView*> relations_;
  std::vector<PredicateEvaluationPlan> plans_;
  qf_eval_options options_;
};

// CandidateLiteralEvaluator::Evaluate (learner/CandidateLiteralEvaluator.cpp:264-406) for all
// candidate literals of a refinement step in one batched call: results[i] receives literal i's
// counts.  arg_types: every relation's argument types, relation-major (empty: all equal), which
// choose each literal's join key (schema/FoilLiteral.hpp:133-144).
inline void EvaluateCandidateLiterals(Context& ctx, const DeviceTableView& binding, int64_t num_positive,
                                      const std::vector<const DeviceTableView*>& relations,
                                      const std::vector<int>& arg_types, const std::vector<qf_literal>& literals,
                                      std::vector<CandidateLiteralInfo>* results,
                                      const qf_eval_options* options = nullptr) {
  std::vector<qf_table> rels;
  for (const DeviceTableView* r : relations) rels.push_back(r->table());
  std::vector<uint64_t> counts(4 * (literals.empty() ? 1 : literals.size()));
  const qf_table b = binding.table();
  const int st = qf_evaluate_literals(ctx.get(), &b, num_positive, rels.data(),
                                      arg_types.empty() ? nullptr : arg_types.data(), (int)rels.size(),
                                      literals.data(), (int)literals.size(), options, counts.data(), nullptr);
  // QF_ERR_COUNT_RANGE still delivers exact counts; ToSizeType decides what an int32 caller does.
  if (st != QF_ERR_COUNT_RANGE) ctx.Check(st, "qf_evaluate_literals");
  results->assign(literals.size(), CandidateLiteralInfo());
  for (size_t i = 0; i < literals.size(); ++i) {
    CandidateLiteralInfo& r = (*results)[i];
    r.literal = literals[i];
    r.num_covered_positive = counts[4 * i + 0];
    r.num_covered_negative = counts[4 * i + 1];
    r.num_binding_positive = counts[4 * i + 2];
    r.num_binding_negative = counts[4 * i + 3];
  }
}

// The same over a clause whose positive and negative bindings are separate tables (the restart
// clause, learner/QuickFoil.cpp:127-140; CandidateLiteralEvaluator.cpp:345-405).
inline void EvaluateCandidateLiteralsSplit(Context& ctx, const DeviceTableView& positives,
                                           const DeviceTableView& negatives,
                                           const std::vector<const DeviceTableView*>& relations,
                                           const std::vector<int>& arg_types, const std::vector<qf_literal>& literals,
                                           std::vector<CandidateLiteralInfo>* results,
                                           const qf_eval_options* options = nullptr) {
  std::vector<qf_table> rels;
  for (const DeviceTableView* r : relations) rels.push_back(r->table());
  std::vector<uint64_t> counts(4 * (literals.empty() ? 1 : literals.size()));
  const qf_table p = positives.table(), n = negatives.table();
  const int st = qf_evaluate_literals_split(ctx.get(), &p, &n, rels.data(),
                                            arg_types.empty() ? nullptr : arg_types.data(), (int)rels.size(),
                                            literals.data(), (int)literals.size(), options, counts.data(), nullptr,
                                            nullptr);
  if (st != QF_ERR_COUNT_RANGE) ctx.Check(st, "qf_evaluate_literals_split");
  results->assign(literals.size(), CandidateLiteralInfo());
  for (size_t i = 0; i < literals.size(); ++i) {
    CandidateLiteralInfo& r = (*results)[i];
    r.literal = literals[i];
    r.num_covered_positive = counts[4 * i + 0];
    r.num_covered_negative = counts[4 * i + 1];
    r.num_binding_positive = counts[4 * i + 2];
    r.num_binding_negative = counts[4 * i + 3];
  }
}

}  // namespace quickfoil_amd

#endif  // QUICKFOIL_AMD_HPP_
