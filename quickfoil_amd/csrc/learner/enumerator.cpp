// Candidate literal enumeration (learner/CandidateLiteralEnumerator.cpp).
#include "enumerator.hpp"

#include <cassert>

namespace qfl {

void Enumerator::enumerate(const Clause& clause, const LiteralSearchStats& last, LiteralMap* entire,
                           LiteralPtrMap* pruned) {
  // learner/CandidateLiteralEnumerator.hpp:123-133
  if (clause.body().empty()) most_general(clause, last, entire, pruned);
  else non_most_general(clause, last, entire, pruned);
}

// learner/CandidateLiteralEnumerator.cpp:23-46: reuse the previous enumeration of the most
// general clause minus the literals that covered no positive binding.
void Enumerator::most_general(const Clause& clause, const LiteralSearchStats& last, LiteralMap* entire,
                              LiteralPtrMap* pruned) {
  if (!last.generated) {
    most_general_from_scratch(clause, entire, pruned);
    return;
  }
  for (const auto& kv : *last.generated) {
    std::vector<Literal>& gen = (*entire)[kv.first];
    for (const Literal& l : kv.second)
      if (!last.pruned_by_coverage || last.pruned_by_coverage->count(&l) == 0) gen.push_back(l);
    std::vector<const Literal*>& pr = (*pruned)[kv.first];
    for (const Literal& l : gen) pr.push_back(&l);
  }
}

// learner/CandidateLiteralEnumerator.cpp:48-99.
void Enumerator::most_general_from_scratch(const Clause& clause, LiteralMap* entire, LiteralPtrMap* pruned) {
  std::map<int, std::vector<Variable>> type_to_vars;
  for (const Variable& v : clause.vars()) type_to_vars[v.type].push_back(v);
  for (const Predicate* pred : background_) {
    std::vector<std::vector<Variable>> per_arg;
    variable_vector(false, type_to_vars, pred, VariableSet(), &per_arg);
    std::vector<Literal> lits{Literal(pred)};
    generate(0, per_arg, &lits);
    assert(lits.back().all_unbound());
    lits.pop_back();  // the literal with every variable new
    if (lits.empty()) continue;
    const std::vector<Literal>* imm;
    if (pred != clause.head().predicate()) {
      imm = &entire->emplace(pred->id, std::move(lits)).first->second;
    } else {
      std::vector<Literal>& gen = (*entire)[pred->id];
      for (const Literal& l : lits)
        if (!l.equals(clause.head())) gen.push_back(l);
      imm = &gen;
    }
    std::vector<const Literal*>& pr = (*pruned)[pred->id];
    for (const Literal& l : *imm) pr.push_back(&l);
  }
}

// learner/CandidateLiteralEnumerator.cpp:101-194.
void Enumerator::generate_and_prune(bool has_key, const Clause& clause, const LiteralSearchStats& last,
                                    const Predicate* pred, const std::map<int, std::vector<Variable>>& type_to_vars,
                                    const std::map<int, std::vector<const Literal*>>& pred_to_body,
                                    LiteralMap* entire, LiteralPtrMap* pruned) {
  const std::vector<const Literal*>* body_with_pred = nullptr;
  auto bit = pred_to_body.find(pred->id);
  if (bit != pred_to_body.end()) body_with_pred = &bit->second;

  VariableSet key_vars;
  // A key outside the argument list selects no variable.
  if (has_key && body_with_pred && pred->key < pred->arity())
    for (const Literal* bl : *body_with_pred) key_vars.insert(bl->var(pred->key));

  std::vector<std::vector<Variable>> per_arg;
  variable_vector(has_key && !key_vars.empty(), type_to_vars, pred, key_vars, &per_arg);
  std::vector<Literal> lits{Literal(pred)};
  generate(0, per_arg, &lits);
  assert(lits.back().all_unbound());
  lits.pop_back();
  std::vector<Literal>* ent = &entire->emplace(pred->id, std::move(lits)).first->second;

  auto found = last.generated->find(pred->id);
  if (found != last.generated->end()) {
    for (const Literal& old : found->second) {
      if (last.pruned_by_coverage && last.pruned_by_coverage->count(&old)) continue;
      std::vector<Literal> ext{Literal(pred)};
      generate_from(old, 0, per_arg, &ext);
      ext.pop_back();  // the old literal itself
      ent->insert(ent->end(), ext.begin(), ext.end());
      if (!clause.body().back().equals(old)) ent->push_back(old);
    }
  }

  std::vector<const Literal*>& pr = (*pruned)[pred->id];
  for (const Literal& l : *ent)
    if (!body_with_pred || !replaceable_duplicate(clause, l, pred_to_body)) pr.push_back(&l);
}

// learner/CandidateLiteralEnumerator.cpp:196-245: candidates use the variables the last body
// literal introduced (plus, through the previous enumeration, the older ones).
void Enumerator::non_most_general(const Clause& clause, const LiteralSearchStats& last, LiteralMap* entire,
                                  LiteralPtrMap* pruned) {
  std::map<int, std::vector<Variable>> type_to_vars;
  const Literal& nl = clause.body().back();
  for (const Variable& v : nl.vars())
    if (v.id >= clause.num_vars_without_last()) type_to_vars[v.type].push_back(v);
  std::map<int, std::vector<const Literal*>> pred_to_body;
  for (const Literal& bl : clause.body()) pred_to_body[bl.predicate()->id].push_back(&bl);
  for (const Predicate* pred : background_)
    generate_and_prune(pred->key >= 0, clause, last, pred, type_to_vars, pred_to_body, entire, pruned);
  canonical_.clear();
}

// learner/CandidateLiteralEnumerator.cpp:247-279: per argument, the clause variables of its type
// (minus key variables already used at the predicate's key), then a new variable.
void Enumerator::variable_vector(bool has_key, const std::map<int, std::vector<Variable>>& type_to_vars,
                                 const Predicate* pred, const VariableSet& key_vars,
                                 std::vector<std::vector<Variable>>* per_arg) const {
  for (int i = 0; i < pred->arity(); ++i) {
    auto it = type_to_vars.find(pred->types[i]);
    per_arg->emplace_back();
    if (it != type_to_vars.end()) {
      for (const Variable& v : it->second)
        if (!has_key || key_vars.count(v) == 0) per_arg->back().push_back(v);
    }
    per_arg->back().push_back(Variable(pred->types[i]));
  }
}

// learner/CandidateLiteralEnumerator.cpp:281-331.
void Enumerator::generate_from(const Literal& base, int filled, const std::vector<std::vector<Variable>>& per_arg,
                               std::vector<Literal>* out) const {
  if (base.var(filled).bound()) {
    for (Literal& l : *out) l.add(base.var(filled));
  } else {
    const std::vector<Variable>& cands = per_arg[filled];
    std::vector<Literal> extra;
    for (Literal& l : *out) {
      for (size_t i = 1; i < cands.size(); ++i) {
        extra.push_back(l);
        extra.back().add(cands[i]);
      }
      l.add(cands[0]);
    }
    out->insert(out->end(), extra.begin(), extra.end());
  }
  if (++filled == (int)per_arg.size()) return;
  generate_from(base, filled, per_arg, out);
}

// learner/CandidateLiteralEnumerator.cpp:333-361.
void Enumerator::generate(int filled, const std::vector<std::vector<Variable>>& per_arg,
                          std::vector<Literal>* out) const {
  const std::vector<Variable>& cands = per_arg[filled];
  std::vector<Literal> extra;
  for (Literal& l : *out) {
    for (size_t i = 1; i < cands.size(); ++i) {
      extra.push_back(l);
      extra.back().add(cands[i]);
    }
    l.add(cands[0]);
  }
  out->insert(out->end(), extra.begin(), extra.end());
  if (++filled == (int)per_arg.size()) return;
  generate(filled, per_arg, out);
}

static void add_existing_row(const Literal& l, std::vector<std::vector<int>>* rel) {
  rel->emplace_back();
  for (const Variable& v : l.vars()) rel->back().push_back(v.id);
}

static void add_new_row(const Literal& l, int next_id, std::vector<std::vector<int>>* rel) {
  rel->emplace_back();
  for (const Variable& v : l.vars()) rel->back().push_back(v.bound() ? v.id : next_id++);
}

// learner/CandidateLiteralEnumerator.cpp:363-447: the candidate is a replaceable duplicate of an
// existing literal of its predicate when swapping the two leaves the clause's canonical
// database (one constant per variable) joinable both ways.
bool Enumerator::replaceable_duplicate(const Clause& clause, const Literal& literal,
                                       const std::map<int, std::vector<const Literal*>>& pred_to_body) {
  auto it = pred_to_body.find(literal.predicate()->id);
  if (it == pred_to_body.end() || it->second.empty()) return false;
  const std::vector<const Literal*>& lits = it->second;
  const std::vector<std::vector<int>>& db = canonical_database(clause, *literal.predicate(), pred_to_body);
  std::vector<std::vector<int>> rel_wo;
  for (const Literal* l : lits) add_existing_row(*l, &rel_wo);
  for (size_t li = 0; li < lits.size(); ++li) {
    std::vector<std::vector<int>> joined = db;
    for (size_t i = 0; i < lits.size(); ++i)
      if (i != li) nested_loops_join<false, false>(*lits[i], rel_wo, &joined);
    if (nested_loops_join<true, true>(literal, rel_wo, &joined)) {
      std::vector<std::vector<int>> rel_with(rel_wo.begin(), rel_wo.begin() + li);
      rel_with.insert(rel_with.end(), rel_wo.begin() + li + 1, rel_wo.end());
      add_new_row(literal, clause.num_vars(), &rel_with);
      joined = db;
      for (size_t i = 0; i + 1 < lits.size(); ++i) nested_loops_join<false, false>(*lits[i], rel_with, &joined);
      if (nested_loops_join<false, true>(*lits.back(), rel_with, &joined)) return true;
    }
  }
  return false;
}

// learner/CandidateLiteralEnumerator.cpp:449-526: join of the canonical relations of every body
// predicate but `pred` (and the head).
const std::vector<std::vector<int>>& Enumerator::canonical_database(
    const Clause& clause, const Predicate& pred, const std::map<int, std::vector<const Literal*>>& pred_to_body) {
  auto have = canonical_.find(pred.id);
  if (have != canonical_.end()) return have->second;
  std::vector<std::vector<int>> joined;
  const int nv = clause.num_vars();
  bool joined_head = false;
  for (const auto& kv : pred_to_body) {
    if (kv.first == pred.id) continue;
    std::vector<std::vector<int>> rel;
    for (const Literal* bl : kv.second) add_existing_row(*bl, &rel);
    if (!joined_head && clause.head().predicate()->id == kv.first) {
      joined_head = true;
      add_existing_row(clause.head(), &rel);
    }
    if (joined.empty()) {
      const Literal& first = *kv.second[0];
      for (const std::vector<int>& row : rel) {
        joined.emplace_back(nv, -1);
        for (size_t i = 0; i < row.size(); ++i) joined.back()[first.var((int)i).id] = row[i];
      }
      for (size_t i = 1; i < kv.second.size(); ++i) nested_loops_join<false, false>(*kv.second[i], rel, &joined);
    } else {
      for (const Literal* l : kv.second) nested_loops_join<false, false>(*l, rel, &joined);
    }
  }
  if (!joined_head) {
    if (joined.empty()) {
      joined.emplace_back(nv, -1);
      for (const Variable& v : clause.head().vars()) joined.back()[v.id] = v.id;
    } else {
      std::vector<std::vector<int>> rel;
      add_existing_row(clause.head(), &rel);
      nested_loops_join<false, false>(clause.head(), rel, &joined);
    }
  }
  return canonical_.emplace(pred.id, std::move(joined)).first->second;
}

// learner/CandidateLiteralEnumerator.cpp:528-575.
template <bool kCheckBound, bool kReturnIfAny>
bool Enumerator::nested_loops_join(const Literal& l, const std::vector<std::vector<int>>& pred_rel,
                                   std::vector<std::vector<int>>* joined) const {
  std::vector<std::vector<int>> out;
  for (const std::vector<int>& prow : pred_rel) {
    for (const std::vector<int>& jrow : *joined) {
      size_t i;
      for (i = 0; i < prow.size(); ++i) {
        if (kCheckBound && !l.var((int)i).bound()) continue;
        const int vid = l.var((int)i).id;
        if (jrow[vid] != -1 && jrow[vid] != prow[i]) break;
      }
      if (i == prow.size()) {
        if (kReturnIfAny) return true;
        std::vector<int> row = jrow;
        for (i = 0; i < prow.size(); ++i)
          if (row[l.var((int)i).id] == -1) row[l.var((int)i).id] = prow[i];
        out.push_back(std::move(row));
      }
    }
  }
  *joined = std::move(out);
  return !joined->empty();
}

}  // namespace qfl
