// Test driver for the learner's candidate enumerator: reads a script on stdin, prints results.
//   PRED <name> <key> <type>...        declare a background predicate (ids in order)
//   CLAUSE <head> :- <body literals>    set the clause (variables by id, FoilParser style)
//   ENUM [<literal> ...]                enumerate with the chained search stats; the listed
//                                        literals (of this enumeration) count as covering no
//                                        positive binding for the next one
//   DUP <literal>                       replaceable-duplicate check against the clause
// Output: "ENTIRE <literals>" and "PRUNED <literals>" after each ENUM, "DUP 0|1".
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "enumerator.hpp"

using namespace qfl;

static std::vector<std::unique_ptr<Predicate>> preds;
static std::map<std::string, const Predicate*> by_name;

static std::vector<std::string> split_literals(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  int depth = 0;
  for (char c : s) {
    if (c == '(') ++depth;
    if (c == ')') --depth;
    if (c == ',' && depth == 0) {
      out.push_back(cur);
      cur.clear();
      continue;
    }
    if (c != ' ' || depth > 0) cur.push_back(c);
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

static Literal parse_literal(const std::string& text) {
  const size_t open = text.find('(');
  std::string name = text.substr(0, open);
  while (!name.empty() && name[0] == ' ') name.erase(0, 1);
  const Predicate* p = by_name.at(name);
  Literal l(p);
  std::stringstream ss(text.substr(open + 1, text.find(')') - open - 1));
  std::string tok;
  int i = 0;
  while (std::getline(ss, tok, ',')) {
    const int id = std::stoi(tok);
    l.add(id >= 0 ? Variable(id, p->types[i]) : Variable(p->types[i]));
    ++i;
  }
  return l;
}

int main() {
  std::unique_ptr<Clause> clause;
  std::vector<const Predicate*> background;
  std::unique_ptr<Enumerator> en;
  LiteralSearchStats stats;
  std::string line;
  while (std::getline(std::cin, line)) {
    std::stringstream ss(line);
    std::string cmd;
    ss >> cmd;
    if (cmd == "PRED") {
      std::unique_ptr<Predicate> p(new Predicate);
      ss >> p->name >> p->key;
      int t;
      while (ss >> t) p->types.push_back(t);
      p->id = (int)preds.size();
      p->index = p->id;
      by_name[p->name] = p.get();
      background.push_back(p.get());
      preds.push_back(std::move(p));
      en.reset(new Enumerator(background));
    } else if (cmd == "CLAUSE") {
      const std::string rest = line.substr(7);
      const size_t sep = rest.find(":-");
      clause.reset(new Clause(parse_literal(rest.substr(0, sep))));
      for (const std::string& lt : split_literals(rest.substr(sep + 2))) clause->add_bound_body_literal(parse_literal(lt), false);
    } else if (cmd == "ENUM") {
      auto entire = std::make_shared<LiteralMap>();
      LiteralPtrMap pruned;
      en->enumerate(*clause, stats, entire.get(), &pruned);
      auto removed = std::make_shared<std::unordered_set<const Literal*>>();
      const std::string rest = line.size() > 5 ? line.substr(5) : "";
      for (const std::string& lt : split_literals(rest)) {
        const Literal q = parse_literal(lt);
        for (const auto& kv : *entire)
          for (const Literal& l : kv.second)
            if (l.equals(q)) removed->insert(&l);
      }
      std::cout << "ENTIRE";
      for (const auto& kv : *entire)
        for (const Literal& l : kv.second) std::cout << " " << l.to_string();
      std::cout << "\nPRUNED";
      for (const auto& kv : pruned)
        for (const Literal* l : kv.second) std::cout << " " << l->to_string();
      std::cout << "\n";
      stats.generated = entire;
      stats.pruned_by_coverage = removed;
    } else if (cmd == "DUP") {
      const Literal q = parse_literal(line.substr(4));
      std::map<int, std::vector<const Literal*>> pred_to_body;
      for (const Literal& bl : clause->body()) pred_to_body[bl.predicate()->id].push_back(&bl);
      en->clear_canonical();
      std::cout << "DUP " << (en->replaceable_duplicate(*clause, q, pred_to_body) ? 1 : 0) << "\n";
    }
  }
  return 0;
}
