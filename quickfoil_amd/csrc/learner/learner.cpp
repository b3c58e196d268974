// QuickFoil search control over an evaluation backend.
//
// Restates learner/QuickFoil.{hpp,cpp} (constructor :88-121, Learn :163-329,
// AddBestCandidateLiteral :331-385, EvaluateAllCandidateLiterals :387-409,
// AddLiteralToBuildingClause :411-441, ComputeCoverageOnUncoveredData :471-573,
// AddBuildingClauseWithNewLiteral :575-721) and learner/QuickFoilState.hpp.  Binding tables,
// counts and coverage bits come from the backend; the binding data of a clause is always one
// table with its positives first (the reference's separate positive/negative blocks of the most
// general clause hold the same rows).
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "enumerator.hpp"
#include "quickfoil_learner.h"
#include "selector.hpp"

namespace qfl {
namespace {

struct LearnError : std::runtime_error {
  int status;
  LearnError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void check(int status, const char* what) {
  if (status != QF_OK) throw LearnError(status, std::string("backend call failed: ") + what);
}

struct Table {
  const qf_learner_backend* be;
  void* h = nullptr;
  int ncols;
  int64_t rows;
  Table(const qf_learner_backend* b, void* handle, int c, int64_t r) : be(b), h(handle), ncols(c), rows(r) {}
  ~Table() {
    if (h) be->release(be->self, h);
  }
};
using TablePtr = std::shared_ptr<Table>;

struct ClauseData {
  Clause clause;
  int64_t pos, neg;
  TablePtr binding;
};
using ClausePtr = std::shared_ptr<const ClauseData>;

// learner/QuickFoilState.hpp:17-60.
struct State {
  bool tied;
  ClausePtr clause;
  std::shared_ptr<State> previous;
  std::unique_ptr<Selector> selector;
  std::shared_ptr<LiteralSearchStats> stats;
  LiteralSet black;
  TablePtr uncovered;
};

// learner/QuickFoil.cpp:74-86.
struct Tied {
  std::unique_ptr<EvaluatedInfo> info;
  std::shared_ptr<State> state;
  std::shared_ptr<LiteralSearchStats> stats;
};

qf_literal to_qf(const Literal& l) {
  qf_literal q;
  q.relation = l.predicate()->index;
  for (int i = 0; i < QF_MAX_ARITY; ++i) q.vars[i] = i < l.num_vars() ? l.var(i).id : -1;
  return q;
}

class Learner {
 public:
  Learner(const qf_learner_problem& p, const qf_learner_options& o, const qf_learner_backend* be)
      : opt_(o), be_(be) {
    const int nb = p.num_background;
    preds_.resize(nb + 1);
    for (int i = 0; i <= nb; ++i) {
      Predicate& pr = preds_[i];
      pr.id = i;
      pr.index = i < nb ? i : -1;
      pr.name = p.names[i];
      pr.key = p.keys ? p.keys[i] : -1;
      pr.types.assign(p.argument_types[i], p.argument_types[i] + p.arities[i]);
      pr.num_facts = p.num_facts[i];
    }
    for (int i = 0; i < nb; ++i) background_.push_back(&preds_[i]);
    target_ = &preds_[nb];
    enumerator_.reset(new Enumerator(background_));
    const int64_t num_true = p.num_positive, num_false = target_->num_facts - p.num_positive;
    if (num_false <= 0) throw LearnError(QF_ERR_INVALID, "positive-only data is not supported");
    if (num_true <= 0) throw LearnError(QF_ERR_INVALID, "no positive facts");
    max_uncovered_ = (int64_t)(int32_t)((double)num_true * (1 - opt_.positive_threshold));
    const int h = target_->arity();
    void* all = nullptr;
    check(be_->upload(be_->self, p.target_columns, h, target_->num_facts, &all), "upload");
    TablePtr facts(new Table(be_, all, h, target_->num_facts));
    // Initial clause: the target facts, positives first (learner/QuickFoil.cpp:100-121).
    std::vector<const int32_t*> pos_cols, neg_cols;
    for (int c = 0; c < h; ++c) {
      pos_cols.push_back(p.target_columns[c]);
      neg_cols.push_back(p.target_columns[c] + num_true);
    }
    global_uncovered_ = upload(pos_cols, h, num_true);
    negatives_ = upload(neg_cols, h, num_false);
    first_stats_ = std::make_shared<LiteralSearchStats>();
    state_ = std::make_shared<State>();
    state_->tied = false;
    state_->clause = std::make_shared<ClauseData>(ClauseData{Clause(head()), num_true, num_false, facts});
    state_->stats = first_stats_;
    state_->uncovered = global_uncovered_;
  }

  void learn();
  const std::vector<std::string>& clauses() const { return clauses_; }

 private:
  Literal head() const {
    Literal l(target_);
    for (int i = 0; i < target_->arity(); ++i) l.add(Variable(i, target_->types[i]));
    return l;
  }
  TablePtr upload(const std::vector<const int32_t*>& cols, int ncols, int64_t rows) {
    void* t = nullptr;
    check(be_->upload(be_->self, cols.data(), ncols, rows, &t), "upload");
    return TablePtr(new Table(be_, t, ncols, rows));
  }
  bool continue_rule_search() const {
    return global_uncovered_->rows > max_uncovered_ && outer_ < opt_.max_iterations;
  }
  void most_general_clause();
  void evaluate_all(const std::vector<std::map<int, std::vector<const Literal*>>>& groups, Selector* sel,
                    std::unordered_set<const Literal*>* pruned, bool consider_random);
  bool add_best(bool is_tied, bool is_random, std::unique_ptr<EvaluatedInfo> best,
                const std::shared_ptr<LiteralSearchStats>& stats, std::unique_ptr<Selector>* selector);
  ClausePtr add_literal(const EvaluatedInfo& best, bool is_random);
  bool finish_clause_with(const EvaluatedInfo& best);
  int64_t coverage(const ClauseData& cd, bool positives, const Literal& l, const Table& examples,
                   std::vector<uint64_t>* bits);
  bool consider_as_last(const EvaluatedInfo& best) const {
    return !(best.precision() < opt_.minimum_inflated_precision &&
             state_->clause->clause.num_body() < opt_.maximum_clause_length);
  }

  qf_learner_options opt_;
  const qf_learner_backend* be_;
  std::vector<Predicate> preds_;
  std::vector<const Predicate*> background_;
  const Predicate* target_;
  std::unique_ptr<Enumerator> enumerator_;
  std::shared_ptr<State> state_;
  TablePtr global_uncovered_, negatives_;
  int64_t max_uncovered_ = 0;
  int outer_ = 0;
  std::shared_ptr<LiteralSearchStats> first_stats_;
  std::vector<std::unique_ptr<Tied>> tied_;
  std::vector<std::string> clauses_;
};

// learner/QuickFoil.cpp:128-143: the head alone over the still uncovered positives and every
// negative.
void Learner::most_general_clause() {
  void* t = nullptr;
  check(be_->concat(be_->self, global_uncovered_->h, global_uncovered_->rows, negatives_->h, negatives_->rows,
                    target_->arity(), &t),
        "concat");
  TablePtr both(new Table(be_, t, target_->arity(), global_uncovered_->rows + negatives_->rows));
  state_ = std::make_shared<State>();
  state_->tied = false;
  state_->clause = std::make_shared<ClauseData>(ClauseData{Clause(head()), global_uncovered_->rows, negatives_->rows, both});
  state_->stats = first_stats_;
  state_->uncovered = global_uncovered_;
}

// Coverage of examples by the clause's positive or negative binding rows that match `l`.
int64_t Learner::coverage(const ClauseData& cd, bool positives, const Literal& l, const Table& examples,
                          std::vector<uint64_t>* bits) {
  std::vector<uint64_t> local;
  std::vector<uint64_t>& b = bits ? *bits : local;
  b.assign((size_t)std::max<int64_t>(1, (examples.rows + 63) / 64), 0);
  const qf_literal q = to_qf(l);
  int64_t n = 0;
  const int64_t r0 = positives ? 0 : cd.pos, r1 = positives ? cd.pos : cd.pos + cd.neg;
  check(be_->coverage(be_->self, cd.binding->h, cd.clause.num_vars(), r0, r1, &q, examples.h, examples.ncols,
                      examples.rows, b.data(), &n),
        "coverage");
  return n;
}

// EvaluateAllCandidateLiterals (learner/QuickFoil.cpp:387-409), every group in one backend call
// (the batched step); results enter the selector in group, predicate, enumeration order.
void Learner::evaluate_all(const std::vector<std::map<int, std::vector<const Literal*>>>& groups, Selector* sel,
                           std::unordered_set<const Literal*>* pruned, bool consider_random) {
  std::vector<const Literal*> order;
  for (const auto& g : groups)
    for (const auto& kv : g)
      for (const Literal* l : kv.second) order.push_back(l);
  if (order.empty()) return;
  std::vector<qf_literal> q;
  for (const Literal* l : order) q.push_back(to_qf(*l));
  std::vector<uint64_t> counts(order.size() * 4);
  const ClauseData& cd = *state_->clause;
  check(be_->evaluate(be_->self, cd.binding->h, cd.clause.num_vars(), cd.pos + cd.neg, cd.pos, q.data(), (int)q.size(),
                      counts.data()),
        "evaluate");
  for (size_t i = 0; i < order.size(); ++i) {
    CandidateInfo c;
    c.literal = order[i];
    c.covered_pos = (int64_t)counts[4 * i + 0];
    c.covered_neg = (int64_t)counts[4 * i + 1];
    c.binding_pos = (int64_t)counts[4 * i + 2];
    c.binding_neg = (int64_t)counts[4 * i + 3];
    sel->insert(c, consider_random);
    if (c.covered_pos == 0) pruned->insert(order[i]);
  }
}

// AddLiteralToBuildingClause (learner/QuickFoil.cpp:411-441).
ClausePtr Learner::add_literal(const EvaluatedInfo& best, bool is_random) {
  const ClauseData& cd = *state_->clause;
  const qf_literal q = to_qf(best.literal);
  void* t = nullptr;
  check(be_->extend(be_->self, cd.binding->h, cd.clause.num_vars(), cd.pos + cd.neg, cd.pos, &q, best.binding_pos,
                    best.binding_neg, &t),
        "extend");
  Clause c = cd.clause;
  c.add_unbound_body_literal(best.literal, is_random);
  TablePtr tb(new Table(be_, t, c.num_vars(), best.binding_pos + best.binding_neg));
  return std::make_shared<ClauseData>(ClauseData{c, best.binding_pos, best.binding_neg, tb});
}

// AddBuildingClauseWithNewLiteral (learner/QuickFoil.cpp:575-721).
bool Learner::finish_clause_with(const EvaluatedInfo& best) {
  const ClauseData& cd = *state_->clause;
  std::vector<uint64_t> pos_bits;
  const int64_t cov_pos = coverage(cd, true, best.literal, *state_->uncovered, &pos_bits);
  const int64_t cov_neg = coverage(cd, false, best.literal, *negatives_, nullptr);
  const double precision = (double)cov_pos / (double)(cov_pos + cov_neg);
  const double recall = (double)cov_pos / (double)state_->uncovered->rows;
  const double f = 2 * precision * recall / (precision + recall);
  if (!(precision >= opt_.minimum_true_precision || f >= opt_.minimum_f_score)) return false;
  std::vector<uint64_t> bits;
  if (state_->uncovered != global_uncovered_) coverage(cd, true, best.literal, *global_uncovered_, &bits);
  else bits = pos_bits;
  // The global uncovered positives lose the covered ones, in row order.
  const int h = target_->arity();
  const int64_t n = global_uncovered_->rows;
  std::vector<std::vector<int32_t>> cols(h, std::vector<int32_t>((size_t)std::max<int64_t>(1, n)));
  std::vector<int32_t*> ptrs;
  for (auto& c : cols) ptrs.push_back(c.data());
  if (n > 0) check(be_->download(be_->self, global_uncovered_->h, h, n, ptrs.data()), "download");
  std::vector<std::vector<int32_t>> kept(h);
  for (int64_t r = 0; r < n; ++r)
    if (!((bits[r >> 6] >> (r & 63)) & 1ull))
      for (int c = 0; c < h; ++c) kept[c].push_back(cols[c][r]);
  std::vector<const int32_t*> kp;
  for (auto& c : kept) kp.push_back(c.data());
  global_uncovered_ = upload(kp, h, (int64_t)kept[0].size());
  Clause learnt = cd.clause;
  learnt.add_unbound_body_literal(best.literal, false);
  clauses_.push_back(learnt.to_string());
  return true;
}

// AddBestCandidateLiteral (learner/QuickFoil.cpp:331-385); true when the building clause ends.
bool Learner::add_best(bool is_tied, bool is_random, std::unique_ptr<EvaluatedInfo> best,
                       const std::shared_ptr<LiteralSearchStats>& stats, std::unique_ptr<Selector>* selector) {
  if (!is_random && consider_as_last(*best)) {
    if (finish_clause_with(*best)) {
      state_.reset();
      return true;
    } else if (is_tied || state_->tied || state_->clause->clause.num_body() >= opt_.maximum_clause_length) {
      state_.reset();
      return true;
    }
  }
  ClausePtr nc = add_literal(*best, is_random);
  std::shared_ptr<State> next = std::make_shared<State>();
  next->tied = is_tied || state_->tied;
  next->clause = nc;
  next->stats = stats;
  next->black = state_->black;
  next->uncovered = state_->uncovered;
  if (is_random) {
    state_->selector = std::move(*selector);
    next->previous = state_;
  }
  state_ = next;
  if (state_->clause->clause.num_body() == 1) first_stats_ = state_->stats;
  return false;
}

// QuickFoil::Learn (learner/QuickFoil.cpp:163-329).
void Learner::learn() {
  for (;;) {
    for (;;) {
      const ClauseData& cd = *state_->clause;
      auto entire = std::make_shared<LiteralMap>();
      LiteralPtrMap pruned;
      enumerator_->enumerate(cd.clause, *state_->stats, entire.get(), &pruned);
      const int64_t local_uncovered = state_->uncovered->rows;
      auto pruned_by_cov = std::make_shared<std::unordered_set<const Literal*>>();
      std::unique_ptr<Selector> selector(new Selector(
          local_uncovered, ClauseCounts{&cd.clause, cd.pos, cd.neg}, state_->black,
          SelectorOptions{opt_.num_saved_literals, opt_.random_margin}));
      // CreateLiteralEvaluationInfoGroups (learner/QuickFoil.cpp:145-161): by the clause
      // variable at each literal's join key, then by predicate.
      std::vector<std::map<int, std::vector<const Literal*>>> groups(cd.clause.num_vars());
      for (const auto& kv : pruned)
        for (const Literal* l : kv.second) groups[l->var(l->join_key()).id][kv.first].push_back(l);
      const bool consider_random = cd.clause.num_random() < opt_.maximum_random_literals &&
                                   (int)state_->black.size() < opt_.maximum_random_trials &&
                                   (int32_t)cd.pos / (int32_t)local_uncovered < 50;
      evaluate_all(groups, selector.get(), pruned_by_cov.get(), consider_random);
      auto stats = std::make_shared<LiteralSearchStats>();
      stats->generated = entire;
      stats->pruned_by_coverage = pruned_by_cov;

      std::vector<std::unique_ptr<EvaluatedInfo>> best_vec;
      std::shared_ptr<State> st = state_;
      bool is_random = selector->best(&best_vec, [&](const Literal& l) {
        return coverage(*st->clause, true, l, *st->uncovered, nullptr);
      });
      if (best_vec.empty()) {
        if (tied_.empty()) throw LearnError(QF_ERR_INVALID, "no valid candidate literal and no tied literal");
        break;
      }
      std::unique_ptr<EvaluatedInfo> best = std::move(best_vec.back());
      best_vec.pop_back();

      if (Selector::need_regrow(state_->clause->clause, *best)) {
        is_random = false;
        Selector* local = selector.get();
        best.reset();
        for (;;) {
          while (!best_vec.empty()) {
            best = std::move(best_vec.back());
            best_vec.pop_back();
            if (!Selector::need_regrow(state_->clause->clause, *best)) break;
            best.reset();
          }
          if (!best) {
            local->next_best(&best_vec);
            if (best_vec.empty()) {
              std::shared_ptr<State> prev = state_->previous;
              if (!prev) throw LearnError(QF_ERR_INVALID, "regrow without a previous state");
              prev->black.insert(state_->clause->clause.unbound_last_literal());
              stats = state_->stats;
              state_ = prev;
              local = state_->selector.get();
              if (!local) throw LearnError(QF_ERR_INVALID, "previous state without a selector");
              local->next_best(&best_vec);
            }
          } else {
            break;
          }
        }
      }
      if (!best) break;  // learner/QuickFoil.cpp:296-299

      for (auto& info : best_vec) {
        if ((double)info->covered_pos > opt_.minimum_coverage_for_tied_literal * (double)local_uncovered)
          tied_.emplace_back(new Tied{std::move(info), state_, stats});
      }
      best_vec.clear();
      if (add_best(false, is_random, std::move(best), stats, &selector)) break;
    }

    ++outer_;
    if (!continue_rule_search()) break;
    state_.reset();
    while (!tied_.empty()) {
      std::unique_ptr<Tied> t = std::move(tied_.back());
      tied_.pop_back();
      state_ = t->state;
      if (Selector::need_regrow(state_->clause->clause, *t->info)) {
        state_.reset();
        continue;
      }
      if (!add_best(true, false, std::move(t->info), t->stats, nullptr)) break;
    }
    if (!state_) most_general_clause();
  }
}

}  // namespace
}  // namespace qfl

extern "C" {

void qf_learner_options_default(qf_learner_options* o) {
  if (!o) return;
  o->positive_threshold = 0.8;
  o->max_iterations = 1000;
  o->minimum_inflated_precision = 0.85;
  o->minimum_true_precision = 0.8;
  o->minimum_f_score = 0.85;
  o->maximum_clause_length = 25;
  o->maximum_random_literals = 2;
  o->maximum_random_trials = 5;
  o->minimum_coverage_for_tied_literal = 0.1;
  o->num_saved_literals = 5;
  o->random_margin = 0.03;
}

int qf_learn(const qf_learner_problem* problem, const qf_learner_options* options, const qf_learner_backend* backend,
             char* out, int64_t out_capacity, int* num_clauses) {
  if (!problem || !backend || !out || out_capacity <= 0) return QF_ERR_INVALID;
  qf_learner_options o;
  qf_learner_options_default(&o);
  if (options) o = *options;
  try {
    qfl::Learner learner(*problem, o, backend);
    learner.learn();
    std::string text;
    for (const std::string& c : learner.clauses()) text += c + "\n";
    if (num_clauses) *num_clauses = (int)learner.clauses().size();
    if ((int64_t)text.size() + 1 > out_capacity) return QF_ERR_OVERFLOW;
    std::memcpy(out, text.c_str(), text.size() + 1);
    return QF_OK;
  } catch (const qfl::LearnError& e) {
    std::snprintf(out, (size_t)out_capacity, "%s", e.what());
    return e.status;
  } catch (const std::exception& e) {
    std::snprintf(out, (size_t)out_capacity, "%s", e.what());
    return QF_ERR_INVALID;
  }
}

}  // extern "C"

// QuickFoilTestRunner (learner/QuickFoilTestRunner.cpp:40-161): per clause, the test examples are
// joined with every body literal but the last (CreateBindingTable); the examples covered through
// the last literal (the coverage semi-join chain, ComputeUnCoveredData :44-133) leave the
// uncovered set, which keeps its row order.
extern "C" int qf_learner_test(const qf_learner_backend* be, const int32_t* const* test_columns, int h,
                               int64_t num_tests, const qf_literal* body, const int* body_lengths, int num_clauses,
                               int64_t* num_uncovered) {
  if (!be || !body_lengths || h < 1 || num_tests < 0 || (num_tests > 0 && !test_columns)) return QF_ERR_INVALID;
  using qfl::check;
  using qfl::Table;
  using TablePtr = std::shared_ptr<Table>;
  try {
    auto upload = [&](const std::vector<const int32_t*>& cols, int64_t rows) {
      void* t = nullptr;
      check(be->upload(be->self, cols.data(), h, rows, &t), "upload");
      return TablePtr(new Table(be, t, h, rows));
    };
    std::vector<const int32_t*> tc(test_columns, test_columns + h);
    TablePtr tests = upload(tc, num_tests);
    TablePtr uncovered = tests;
    int64_t off = 0;
    for (int c = 0; c < num_clauses; ++c) {
      const int nb = body_lengths[c];
      if (nb <= 0) return QF_ERR_INVALID;
      const qf_literal* lits = body + off;
      off += nb;
      TablePtr binding = tests;
      int ncols = h;
      bool empty = num_tests == 0;
      for (int i = 0; i + 1 < nb && !empty; ++i) {
        void* out = nullptr;
        int64_t rows = 0;
        check(be->join(be->self, binding->h, ncols, binding->rows, &lits[i], &out, &rows), "join");
        int extra = 0;
        for (int k = 0; k < QF_MAX_ARITY; ++k) extra += lits[i].vars[k] >= ncols ? 1 : 0;
        ncols += extra;
        binding = TablePtr(new Table(be, out, ncols, rows));
        empty = rows == 0;
      }
      if (empty) continue;  // the clause covers no test example
      qf_literal last = lits[nb - 1];
      for (int k = 0; k < QF_MAX_ARITY; ++k)
        if (last.vars[k] >= ncols) last.vars[k] = -1;
      std::vector<uint64_t> bits((size_t)std::max<int64_t>(1, (uncovered->rows + 63) / 64), 0);
      int64_t covered = 0;
      check(be->coverage(be->self, binding->h, ncols, 0, binding->rows, &last, uncovered->h, h, uncovered->rows,
                         bits.data(), &covered),
            "coverage");
      const int64_t n = uncovered->rows;
      std::vector<std::vector<int32_t>> cols(h, std::vector<int32_t>((size_t)std::max<int64_t>(1, n)));
      std::vector<int32_t*> ptrs;
      for (auto& col : cols) ptrs.push_back(col.data());
      if (n > 0) check(be->download(be->self, uncovered->h, h, n, ptrs.data()), "download");
      std::vector<std::vector<int32_t>> kept(h);
      for (int64_t r = 0; r < n; ++r)
        if (!((bits[r >> 6] >> (r & 63)) & 1ull))
          for (int k = 0; k < h; ++k) kept[k].push_back(cols[k][r]);
      std::vector<const int32_t*> kp;
      for (auto& k : kept) kp.push_back(k.data());
      uncovered = upload(kp, (int64_t)kept[0].size());
      if (uncovered->rows == 0) break;
    }
    if (num_uncovered) *num_uncovered = uncovered->rows;
    return QF_OK;
  } catch (const qfl::LearnError& e) {
    return e.status;
  } catch (const std::exception&) {
    return QF_ERR_INVALID;
  }
}
