// FOIL schema objects of the host learner: variables, predicates, literals and clauses.
//
// Restates schema/FoilVariable.hpp, schema/FoilPredicate.hpp, schema/FoilLiteral.hpp and
// schema/FoilClause.{hpp,cpp} for the learner's search control.  Binding data is not part of the
// clause here: the learner keeps it beside the clause as a backend table handle.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <unordered_set>
#include <vector>

namespace qfl {

// schema/FoilVariable.hpp:15-50.  id -1 = unbound (a new variable).
struct Variable {
  int id = -1;
  int type = -1;
  Variable() = default;
  explicit Variable(int type_) : id(-1), type(type_) {}
  Variable(int id_, int type_) : id(id_), type(type_) {}
  bool bound() const { return id >= 0; }
  bool equals(const Variable& o) const { return id == o.id; }  // FoilVariable::Equals compares ids only
};

struct VariableHash {
  size_t operator()(const Variable& v) const { return (size_t)v.id; }
};
struct VariableEq {
  bool operator()(const Variable& a, const Variable& b) const { return a.equals(b); }
};
using VariableSet = std::unordered_set<Variable, VariableHash, VariableEq>;

// schema/FoilPredicate.hpp:22-80 (facts live in the backend, addressed by `index`).
struct Predicate {
  int id = 0;     // background predicates 0..n-1, the target n
  int index = 0;  // backend relation index (-1 for the target)
  std::string name;
  int key = -1;
  std::vector<int> types;
  int64_t num_facts = 0;
  int arity() const { return (int)types.size(); }
};

// schema/FoilLiteral.hpp:22-150.
class Literal {
 public:
  explicit Literal(const Predicate* p) : pred_(p) { vars_.reserve(p->arity()); }
  Literal(const Predicate* p, const std::vector<Variable>& vars) : pred_(p), vars_(vars) {
    for (size_t i = 0; i < vars_.size(); ++i) update_join_key((int)i);
  }
  const Predicate* predicate() const { return pred_; }
  const std::vector<Variable>& vars() const { return vars_; }
  const Variable& var(int i) const { return vars_[i]; }
  int num_vars() const { return (int)vars_.size(); }
  void add(const Variable& v) {
    vars_.push_back(v);
    update_join_key((int)vars_.size() - 1);
  }
  bool equals(const Literal& o) const {
    if (pred_ != o.pred_) return false;
    for (int i = 0; i < num_vars(); ++i)
      if (!vars_[i].equals(o.vars_[i])) return false;
    return true;
  }
  int num_unbound() const {
    int n = 0;
    for (const Variable& v : vars_) n += v.bound() ? 0 : 1;
    return n;
  }
  bool is_bound() const { return num_unbound() == 0; }
  bool all_unbound() const { return num_unbound() == num_vars(); }
  int join_key() const { return join_key_; }
  std::string to_string() const {
    std::string s = pred_->name + "(";
    for (int i = 0; i < num_vars(); ++i) {
      if (i) s += ", ";
      s += std::to_string(vars_[i].id);
    }
    return s + ")";
  }
  // FoilLiteral::CreateUnBoundLiteral (schema/FoilLiteral.hpp:148-158).
  Literal unbound_from(int start_unbound_id) const {
    Literal l(pred_);
    for (const Variable& v : vars_) l.add(v.id < start_unbound_id ? v : Variable(v.type));
    return l;
  }

 private:
  // FoilLiteral::UpdateJoinKey (schema/FoilLiteral.hpp:133-140): the bound position with the
  // largest variable type id, first on ties.
  void update_join_key(int pos) {
    if (vars_[pos].bound() && (join_key_ == -1 || vars_[join_key_].type < vars_[pos].type)) join_key_ = pos;
  }
  const Predicate* pred_;
  std::vector<Variable> vars_;
  int join_key_ = -1;
};

// FoilLiteralHash (schema/FoilLiteral.hpp:166-174) with the identity hash of utility/Hash.hpp.
struct LiteralHash {
  size_t operator()(const Literal& l) const {
    size_t seed = (size_t)(uint32_t)l.predicate()->id;
    for (const Variable& v : l.vars()) seed ^= (size_t)(uint32_t)v.id + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    return seed;
  }
};
struct LiteralEq {
  bool operator()(const Literal& a, const Literal& b) const { return a.equals(b); }
};
using LiteralSet = std::unordered_set<Literal, LiteralHash, LiteralEq>;

// schema/FoilClause.{hpp,cpp} without binding data.
class Clause {
 public:
  explicit Clause(const Literal& head) : head_(head), nv_wo_last_(head.num_vars()) {
    for (const Variable& v : head.vars())
      if (v.id == (int)vars_.size()) vars_.push_back(v);
  }
  const Literal& head() const { return head_; }
  const std::vector<Literal>& body() const { return body_; }
  int num_body() const { return (int)body_.size(); }
  const std::vector<Variable>& vars() const { return vars_; }
  int num_vars() const { return (int)vars_.size(); }
  int num_vars_without_last() const { return nv_wo_last_; }
  const std::vector<bool>& random_flags() const { return random_; }
  int num_random() const {
    int n = 0;
    for (bool r : random_) n += r ? 1 : 0;
    return n;
  }
  // FoilClause::AddUnBoundBodyLiteral (schema/FoilClause.cpp:43-61): new variables get the
  // next ids.
  void add_unbound_body_literal(const Literal& l, bool is_random) {
    nv_wo_last_ = (int)vars_.size();
    std::vector<Variable> lv;
    for (const Variable& v : l.vars()) {
      if (!v.bound()) {
        lv.emplace_back((int)vars_.size(), v.type);
        vars_.push_back(lv.back());
      } else {
        lv.push_back(v);
      }
    }
    body_.emplace_back(l.predicate(), lv);
    random_.push_back(is_random);
  }
  // FoilClause::AddBoundBodyLiteral (schema/FoilClause.cpp:63-79), used when a clause is read
  // back from text: variables keep their ids.
  void add_bound_body_literal(const Literal& l, bool is_random) {
    nv_wo_last_ = (int)vars_.size();
    for (const Variable& v : l.vars()) {
      if (v.id >= (int)vars_.size()) {
        vars_.resize(v.id);
        vars_.push_back(v);
      } else if (!vars_[v.id].bound()) {
        vars_[v.id] = v;
      }
    }
    body_.push_back(l);
    random_.push_back(is_random);
  }
  Literal unbound_last_literal() const { return body_.back().unbound_from(nv_wo_last_); }
  // FoilClause::ToString (schema/FoilClause.cpp:31-41).
  std::string to_string() const {
    std::string s = head_.to_string() + " :-  ";
    for (const Literal& l : body_) s += l.to_string() + ", ";
    s.erase(s.end() - 2, s.end());
    return s;
  }

 private:
  Literal head_;
  std::vector<Literal> body_;
  std::vector<Variable> vars_;
  int nv_wo_last_;
  std::vector<bool> random_;
};

}  // namespace qfl
