// Candidate literal enumeration of the host learner.
//
// Restates learner/CandidateLiteralEnumerator.{hpp,cpp} (learner/CandidateLiteralEnumerator.cpp:23-598)
// and learner/LiteralSearchStats.hpp.  Containers keyed by predicate iterate in predicate-id
// order (the reference iterates pointer-keyed hash maps), which fixes one canonical order for
// every backend.
#pragma once

#include <map>
#include <memory>
#include <unordered_set>
#include <vector>

#include "foil.hpp"

namespace qfl {

using LiteralMap = std::map<int, std::vector<Literal>>;              // predicate id -> literals
using LiteralPtrMap = std::map<int, std::vector<const Literal*>>;   // predicate id -> literals

// learner/LiteralSearchStats.hpp:19-31.
struct LiteralSearchStats {
  std::shared_ptr<LiteralMap> generated;                  // null before the first enumeration
  std::shared_ptr<std::unordered_set<const Literal*>> pruned_by_coverage;  // point into `generated`
};

class Enumerator {
 public:
  explicit Enumerator(const std::vector<const Predicate*>& background) : background_(background) {}

  void enumerate(const Clause& clause, const LiteralSearchStats& last, LiteralMap* entire, LiteralPtrMap* pruned);

  // Exposed for tests (learner/CandidateLiteralEnumerator_test.cpp).
  bool replaceable_duplicate(const Clause& clause, const Literal& literal,
                             const std::map<int, std::vector<const Literal*>>& pred_to_body);
  void clear_canonical() { canonical_.clear(); }

 private:
  void most_general(const Clause& clause, const LiteralSearchStats& last, LiteralMap* entire, LiteralPtrMap* pruned);
  void most_general_from_scratch(const Clause& clause, LiteralMap* entire, LiteralPtrMap* pruned);
  void non_most_general(const Clause& clause, const LiteralSearchStats& last, LiteralMap* entire,
                        LiteralPtrMap* pruned);
  void generate_and_prune(bool has_key, const Clause& clause, const LiteralSearchStats& last, const Predicate* pred,
                          const std::map<int, std::vector<Variable>>& type_to_vars,
                          const std::map<int, std::vector<const Literal*>>& pred_to_body, LiteralMap* entire,
                          LiteralPtrMap* pruned);
  void variable_vector(bool has_key, const std::map<int, std::vector<Variable>>& type_to_vars, const Predicate* pred,
                       const VariableSet& key_vars, std::vector<std::vector<Variable>>* per_arg) const;
  void generate(int filled, const std::vector<std::vector<Variable>>& per_arg, std::vector<Literal>* out) const;
  void generate_from(const Literal& base, int filled, const std::vector<std::vector<Variable>>& per_arg,
                     std::vector<Literal>* out) const;
  const std::vector<std::vector<int>>& canonical_database(const Clause& clause, const Predicate& pred,
                                                          const std::map<int, std::vector<const Literal*>>& pred_to_body);
  template <bool kCheckBound, bool kReturnIfAny>
  bool nested_loops_join(const Literal& l, const std::vector<std::vector<int>>& pred_rel,
                         std::vector<std::vector<int>>* joined) const;

  std::map<int, std::vector<std::vector<int>>> canonical_;  // predicate id -> joined canonical rows
  const std::vector<const Predicate*>& background_;
};

}  // namespace qfl
