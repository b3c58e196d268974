// Test driver for the C++ adapter (include/quickfoil_amd.hpp) over the C ABI: a compiled caller
// of the library, as a reference-side C++ evaluator would be.  Reads a problem on stdin and
// prints what each operator returns; the GPU test compares the lines with the oracle.
//
// Input (whitespace separated):
//   BINDING <columns> <rows> <num_positive>   then the columns, each <rows> values
//   REL <arity> <rows>                        then the columns (repeatable; relation ids in order)
//   LIT <relation> <v0> .. <v5>               candidate literal (-1 = new variable; repeatable)
//   PLAN <relation> <join col> <binding col> <atoms> <conjunctions> <root literal or -1>
//     followed by <atoms> lines  "<bg col> <binding col> <literal or -1>"
//     and <conjunctions> lines   "<left node> <right node> <literal or -1>"
//   END
// Output lines:
//   EVAL <i> <covered+> <covered-> <binding+> <binding->     EvaluateCandidateLiterals
//   SPLIT <i> ...                                           EvaluateCandidateLiteralsSplit
//   AGG <i> ...                                             CountAggregator::Execute
//   AGG2 <i> ...                                            ExecuteOnPositives + ExecuteOnNegatives
//   OFFSETS <2^3 + 1 values>, PAIRS <2 x rows values>       RadixPartition(0, binding), 3 bits
//   PBUCKETS <...>, PNEXT <...>                             BuildHashTableOnPartitions(0, binding)
//   TBUCKETS <...>, TNEXT <...>                             BuildHashTableOnTable(all columns, relation 0)
//   SIZE <i> <int32 count or RANGE>                         ToSizeType(binding+ + binding-)
#include <iostream>
#include <string>
#include <vector>

#include "quickfoil_amd.hpp"

using namespace quickfoil_amd;

struct HostTable {
  int columns = 0;
  int64_t rows = 0;
  std::vector<std::vector<int32_t>> data;
  std::vector<const int32_t*> ptrs(int64_t begin, int64_t end) const {
    std::vector<const int32_t*> p;
    for (const auto& c : data) p.push_back(c.data() + begin);
    (void)end;
    return p;
  }
};

static HostTable read_table(std::istream& in, int columns, int64_t rows) {
  HostTable t;
  t.columns = columns;
  t.rows = rows;
  t.data.assign(columns, std::vector<int32_t>(rows > 0 ? rows : 1));
  for (int c = 0; c < columns; ++c)
    for (int64_t r = 0; r < rows; ++r) in >> t.data[c][r];
  return t;
}

static void print_counts(const char* tag, size_t i, const CandidateLiteralInfo& r) {
  std::cout << tag << " " << i << " " << r.num_covered_positive << " " << r.num_covered_negative << " "
            << r.num_binding_positive << " " << r.num_binding_negative << "\n";
}

template <class T>
static void print_array(const char* tag, const std::vector<T>& v) {
  std::cout << tag;
  for (const T& x : v) std::cout << " " << x;
  std::cout << "\n";
}

int main() {
  HostTable binding;
  int64_t num_positive = 0;
  std::vector<HostTable> relations;
  std::vector<qf_literal> literals;
  struct PlanIn {
    PredicateEvaluationPlan plan;
    int root = -1;
    std::vector<int> node_lits;
  };
  std::vector<PlanIn> plans;
  std::string tag;
  while (std::cin >> tag && tag != "END") {
    if (tag == "BINDING") {
      int c;
      int64_t r;
      std::cin >> c >> r >> num_positive;
      binding = read_table(std::cin, c, r);
    } else if (tag == "REL") {
      int a;
      int64_t r;
      std::cin >> a >> r;
      relations.push_back(read_table(std::cin, a, r));
    } else if (tag == "LIT") {
      qf_literal l{};
      std::cin >> l.relation;
      for (int k = 0; k < QF_MAX_ARITY; ++k) std::cin >> l.vars[k];
      literals.push_back(l);
    } else if (tag == "PLAN") {
      PlanIn p;
      int na, nc;
      std::cin >> p.plan.relation >> p.plan.join_column >> p.plan.binding_column >> na >> nc >> p.root;
      for (int k = 0; k < na; ++k) {
        int bg, bc, lit;
        std::cin >> bg >> bc >> lit;
        p.plan.atoms.emplace_back(bg, bc);
        p.node_lits.push_back(lit);
      }
      for (int k = 0; k < nc; ++k) {
        int l, r, lit;
        std::cin >> l >> r >> lit;
        p.plan.conjunctions.emplace_back(l, r);
        p.node_lits.push_back(lit);
      }
      plans.push_back(std::move(p));
    } else {
      std::cerr << "unknown input " << tag << "\n";
      return 2;
    }
  }
  try {
    Context ctx(0);
    auto dev_binding = DeviceTable::Upload(binding.ptrs(0, binding.rows), binding.rows);
    auto dev_pos = DeviceTable::Upload(binding.ptrs(0, num_positive), num_positive);
    auto dev_neg = DeviceTable::Upload(binding.ptrs(num_positive, binding.rows), binding.rows - num_positive);
    DeviceTableCache facts;  // the background tables: uploaded once
    std::vector<const DeviceTableView*> rels;
    for (const HostTable& r : relations) rels.push_back(&facts.Get(r.ptrs(0, r.rows), r.rows).view());

    std::vector<CandidateLiteralInfo> results;
    EvaluateCandidateLiterals(ctx, dev_binding->view(), num_positive, rels, {}, literals, &results);
    for (size_t i = 0; i < results.size(); ++i) print_counts("EVAL", i, results[i]);
    for (size_t i = 0; i < results.size(); ++i) {
      std::cout << "SIZE " << i << " ";
      try {
        std::cout << ToSizeType(results[i].num_binding_positive + results[i].num_binding_negative) << "\n";
      } catch (const Error& e) {
        std::cout << (e.status() == QF_ERR_COUNT_RANGE ? "RANGE" : "ERROR") << "\n";
      }
    }
    EvaluateCandidateLiteralsSplit(ctx, dev_pos->view(), dev_neg->view(), rels, {}, literals, &results);
    for (size_t i = 0; i < results.size(); ++i) print_counts("SPLIT", i, results[i]);

    // CountAggregator over the given plans; infos[i] receives literal i.
    if (!plans.empty()) {
      std::vector<CandidateLiteralInfo> infos(literals.size());
      auto with_infos = [&](std::vector<CandidateLiteralInfo>* target) {
        std::vector<PredicateEvaluationPlan> out;
        for (const PlanIn& p : plans) {
          PredicateEvaluationPlan q = p.plan;
          q.literal = p.root >= 0 ? &(*target)[p.root] : nullptr;
          for (int lit : p.node_lits) q.node_literals.push_back(lit >= 0 ? &(*target)[lit] : nullptr);
          out.push_back(q);
        }
        return out;
      };
      CountAggregator agg(ctx, dev_binding->view(), rels, with_infos(&infos));
      agg.Execute(num_positive);
      for (size_t i = 0; i < infos.size(); ++i) print_counts("AGG", i, infos[i]);
      std::vector<CandidateLiteralInfo> infos2(literals.size());
      CountAggregator on_pos(ctx, dev_pos->view(), rels, with_infos(&infos2));
      CountAggregator on_neg(ctx, dev_neg->view(), rels, with_infos(&infos2));
      on_pos.ExecuteOnPositives();
      on_neg.ExecuteOnNegatives();
      for (size_t i = 0; i < infos2.size(); ++i) print_counts("AGG2", i, infos2[i]);
    }

    // The reference's layouts.
    DeviceTableView* view = dev_binding->mutable_view();
    RadixPartition(ctx, 0, view, 3);
    print_array("OFFSETS", view->partitions_at(0).host_offsets);
    print_array("PAIRS", view->partitions_at(0).pairs.ToHost());
    BuildHashTableOnPartitions(ctx, 0, view);
    print_array("PBUCKETS", view->hash_tables_at(0).buckets.ToHost());
    print_array("PNEXT", view->hash_tables_at(0).next.ToHost());
    if (!rels.empty()) {
      std::vector<int> keys;
      for (int c = 0; c < rels[0]->num_columns(); ++c) keys.push_back(c);
      std::unique_ptr<DeviceFoilHashTable> t = BuildHashTableOnTable(ctx, keys, *rels[0]);
      print_array("TBUCKETS", t->buckets.ToHost());
      print_array("TNEXT", t->next.ToHost());
    }
  } catch (const Error& e) {
    std::cerr << "error " << e.status() << ": " << e.what() << "\n";
    return 1;
  }
  return 0;
}
