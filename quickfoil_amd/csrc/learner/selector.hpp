// Literal scoring and selection of the host learner.
//
// Restates learner/LiteralSelector.{hpp,cpp}: the MCC x area-under-entropy-curve score
// (learner/LiteralSelector.hpp:236-281, learner/LiteralSelector.cpp:68), the bounded min-heap of
// the best literals (learner/LiteralSelector.cpp:117-145), random literal candidates and the
// regrow rule.
#pragma once

#include <cmath>
#include <functional>
#include <memory>
#include <vector>

#include "foil.hpp"

namespace qfl {

// learner/CandidateLiteralInfo.hpp:15-26 (counts widened to 64 bits).
struct CandidateInfo {
  const Literal* literal;
  int64_t covered_pos = 0, covered_neg = 0, binding_pos = 0, binding_neg = 0;
};

// EvaluatedLiteralInfo (learner/LiteralSelector.hpp:38-60).
struct EvaluatedInfo {
  Literal literal;
  int64_t covered_pos, covered_neg, binding_pos, binding_neg;
  double score;
  EvaluatedInfo(const CandidateInfo& c, double s)
      : literal(*c.literal), covered_pos(c.covered_pos), covered_neg(c.covered_neg), binding_pos(c.binding_pos),
        binding_neg(c.binding_neg), score(s) {}
  double precision() const { return (double)binding_pos / (double)(binding_pos + binding_neg); }
};

struct SelectorOptions {
  int num_saved_literals = 5;  // learner/LiteralSelector.cpp:28-30
  double random_margin = 0.03;  // learner/LiteralSelector.cpp:32-34
};

// Clause binding counts the scores are relative to.
struct ClauseCounts {
  const Clause* clause;
  int64_t pos, neg;
};

class Selector {
 public:
  using Coverage = std::function<int64_t(const Literal&)>;  // covered uncovered positives

  Selector(int64_t total_uncovered_positive, ClauseCounts clause, const LiteralSet& black, SelectorOptions o);

  void insert(const CandidateInfo& info, bool consider_random);
  // GetBestLiteral (learner/LiteralSelector.hpp:76-135); returns whether the random literal won.
  bool best(std::vector<std::unique_ptr<EvaluatedInfo>>* out, const Coverage& covered_positives);
  // GetNextBestLiterals (learner/LiteralSelector.hpp:137-147).
  void next_best(std::vector<std::unique_ptr<EvaluatedInfo>>* out);
  bool empty() const { return saved_.empty(); }

  static bool need_regrow(const Clause& clause, int64_t covered_neg, const Literal& l);
  static bool need_regrow(const Clause& clause, const EvaluatedInfo& e) {
    return need_regrow(clause, e.covered_neg, e.literal);
  }

 private:
  struct Entry {
    CandidateInfo info;
    Literal literal;  // owned copy; info.literal points here
    double score;
  };
  double mcc(int64_t cov_pos, int64_t cov_neg) const;
  static double auec(int64_t bind_pos, int64_t bind_neg);
  double entropy(int64_t bind_pos, int64_t bind_neg) const { return auec(bind_pos, bind_neg) - clause_area_; }
  bool choose_random(const Literal& random, const Literal& regular, int64_t uncovered,
                     const Coverage& covered_positives) const;

  std::vector<std::unique_ptr<Entry>> heap_;  // min-heap on score
  std::vector<std::unique_ptr<EvaluatedInfo>> saved_;
  int64_t total_uncovered_positive_;
  ClauseCounts clause_;
  double clause_area_;
  const LiteralSet& black_;
  SelectorOptions opt_;
  double max_random_f_ = -1;
  std::unique_ptr<Entry> best_random_;
};

}  // namespace qfl
