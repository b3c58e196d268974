// Literal scoring and selection (learner/LiteralSelector.{hpp,cpp}).
#include "selector.hpp"

#include <algorithm>

namespace qfl {

namespace {
struct Greater {
  template <class E>
  bool operator()(const E& a, const E& b) const { return a->score > b->score; }
};
struct Less {
  template <class E>
  bool operator()(const E& a, const E& b) const { return a->score < b->score; }
};
}  // namespace

Selector::Selector(int64_t total_uncovered_positive, ClauseCounts clause, const LiteralSet& black, SelectorOptions o)
    : total_uncovered_positive_(total_uncovered_positive), clause_(clause), clause_area_(auec(clause.pos, clause.neg)),
      black_(black), opt_(o) {}

// learner/LiteralSelector.hpp:236-258 (MCC of the covered counts against the clause's bindings).
double Selector::mcc(int64_t cov_pos, int64_t cov_neg) const {
  if (cov_pos == 0) return -1;  // deviates from MCC, as the reference does
  if (cov_pos == clause_.pos && cov_neg == clause_.neg) return 0;
  const double tn = (double)(clause_.neg - cov_neg);
  const double fn = (double)(clause_.pos - cov_pos);
  const double total = (double)(cov_pos + cov_neg);
  return ((double)cov_pos * tn - (double)cov_neg * fn) /
         std::sqrt(total * ((double)(clause_.pos + clause_.neg) - total) * (double)clause_.neg * (double)clause_.pos);
}

// learner/LiteralSelector.hpp:268-281.
double Selector::auec(int64_t bind_pos, int64_t bind_neg) {
  if (bind_pos == 0) return 0;
  if (bind_neg == 0) return 1;
  const double p = (double)bind_pos / (double)(bind_pos + bind_neg);
  return ((1 - p) * (1 - p) * std::log2(1 - p) - p * p * std::log2(p)) * std::log(2) + p;
}

// learner/LiteralSelector.hpp:166-186.
bool Selector::need_regrow(const Clause& clause, int64_t covered_neg, const Literal& l) {
  if (clause.body().empty() || covered_neg == 0) return false;
  if (!clause.random_flags().back()) return false;
  for (int i = 0; i < l.num_vars(); ++i)
    if (l.var(i).id >= clause.num_vars_without_last() && i != l.predicate()->key) return false;
  return true;
}

// LiteralSelector::Insert (learner/LiteralSelector.cpp:46-146).
void Selector::insert(const CandidateInfo& info, bool consider_random) {
  const double raw = mcc(info.covered_pos, info.covered_neg);
  const double m = 1 + raw;
  const double a = 1 + entropy(info.binding_pos, info.binding_neg);
  if (m == 0 || a == 0) return;
  const double score = 5 * a * m / (m + 4 * a);
  auto make = [&]() {
    std::unique_ptr<Entry> e(new Entry{info, *info.literal, score});
    e->info.literal = &e->literal;
    return e;
  };
  if (consider_random) {
    const Literal& l = *info.literal;
    if (!l.is_bound() &&
        !(raw == 0 && info.binding_pos == info.covered_pos && info.binding_neg == info.covered_neg) &&
        raw < opt_.random_margin && raw > -opt_.random_margin && black_.count(l) == 0 &&
        !need_regrow(*clause_.clause, info.covered_neg, l)) {
      const double precision = (double)info.covered_pos / (double)(info.covered_pos + info.covered_neg);
      const double recall = (double)info.covered_pos / (double)clause_.pos;
      const double f = 2 * precision * recall / (precision + recall);
      if (!best_random_ ||
          (f > max_random_f_ ||
           (f == max_random_f_ && info.covered_pos < total_uncovered_positive_ &&
            info.binding_pos / info.covered_pos <= 2 && info.binding_pos > best_random_->info.binding_pos))) {
        max_random_f_ = f;
        best_random_ = make();
      }
    }
  }
  if ((int)heap_.size() < opt_.num_saved_literals) {
    heap_.push_back(make());
    std::push_heap(heap_.begin(), heap_.end(), Greater());
    return;
  }
  if (score < heap_.front()->score) return;
  const double min_score = heap_.front()->score;
  do {
    std::pop_heap(heap_.begin(), heap_.end(), Greater());
    heap_.pop_back();
  } while (!heap_.empty() && heap_.front()->score == min_score);
  heap_.push_back(make());
  std::push_heap(heap_.begin(), heap_.end(), Greater());
}

// learner/LiteralSelector.cpp:212-232.
bool Selector::choose_random(const Literal& random, const Literal& regular, int64_t uncovered,
                             const Coverage& covered_positives) const {
  const int64_t by_regular = covered_positives(regular);
  if ((double)by_regular >= 0.8 * (double)uncovered) return false;
  const int64_t by_random = covered_positives(random);
  if ((double)by_random < 1.2 * (double)by_regular) return false;
  return true;
}

bool Selector::best(std::vector<std::unique_ptr<EvaluatedInfo>>* out, const Coverage& covered_positives) {
  if (heap_.empty()) return false;
  std::sort(heap_.begin(), heap_.end(), Less());
  const double max_score = heap_.back()->score;
  double random_score = -1;
  bool use_random = false;
  if (best_random_) {
    const double m = 2 + mcc(best_random_->info.covered_pos, best_random_->info.covered_neg);
    const double a = 1 + entropy(best_random_->info.binding_pos, best_random_->info.binding_neg);
    random_score = 5 * a * m / (m + 4 * a);
  }
  const Entry& regular = *heap_.back();
  int idx;
  if (random_score > max_score && best_random_->info.covered_pos > regular.info.covered_pos &&
      (best_random_->info.binding_pos / best_random_->info.covered_pos < 50 ||
       (double)regular.info.covered_pos / (double)clause_.pos < 0.1) &&
      choose_random(best_random_->literal, regular.literal, total_uncovered_positive_, covered_positives)) {
    idx = (int)heap_.size() - 1;
    out->emplace_back(new EvaluatedInfo(best_random_->info, best_random_->score));
    use_random = true;
  } else {
    for (idx = (int)heap_.size() - 2; idx >= 0; --idx) {
      if (std::fabs(heap_[idx]->score - max_score) < 0.00001) out->emplace_back(new EvaluatedInfo(heap_[idx]->info, heap_[idx]->score));
      else break;
    }
    out->emplace_back(new EvaluatedInfo(regular.info, regular.score));
  }
  for (int i = 0; i <= idx; ++i) saved_.emplace_back(new EvaluatedInfo(heap_[i]->info, heap_[i]->score));
  return use_random;
}

void Selector::next_best(std::vector<std::unique_ptr<EvaluatedInfo>>* out) {
  if (empty()) return;
  const double max_score = saved_.back()->score;
  do {
    out->push_back(std::move(saved_.back()));
    saved_.pop_back();
  } while (!empty() && std::fabs(saved_.back()->score - max_score) < 0.00001);
}

}  // namespace qfl
