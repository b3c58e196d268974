"""A compiled C++ caller of the library (tests/abi_driver.cpp over include/quickfoil_amd.hpp, the
reference's operator names over the C ABI) reproduces the oracle: step evaluation, the two-pass
split, CountAggregator over explicit plans, RadixPartition and both chained-table builds."""
import subprocess

import numpy as np
import pytest

import oracle
from helpers import abi_driver_script, build_abi_driver, parse_driver_output, random_problem, uncle_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    return build_abi_driver(str(tmp_path_factory.mktemp("abi")))


def _plans(literals, lit_ids):
    """One flat PredicateEvaluationPlan per (relation, join key, join variable) group, from the
    oracle's plan builder (learner/CandidateLiteralEvaluator.cpp:105-262)."""
    groups = {}
    for i in lit_ids:
        rel, vars_ = literals[i]
        jk = oracle.literal_join_key(vars_)
        groups.setdefault((rel, jk, vars_[jk]), []).append(i)
    plans = []
    for (rel, jk, var), ids in groups.items():
        arity = len(literals[ids[0]][1])
        plan = oracle.generate_plan(arity, jk, [literals[i][1] for i in ids])
        na = len(plan["atoms"])
        atoms = [(bg, bd, ids[plan["literal"][x]] if plan["literal"][x] >= 0 else -1)
                 for x, (bg, bd) in enumerate(plan["atoms"])]
        conj = [(plan["left"][x], plan["right"][x], ids[plan["literal"][x]] if plan["literal"][x] >= 0 else -1)
                for x in range(na, len(plan["left"]))]
        root = ids[plan["root_literal"]] if plan["root_literal"] >= 0 else -1
        plans.append((rel, jk, var, atoms, conj, root))
    return plans


@pytest.mark.parametrize("case", ["uncle", "random"])
def test_cpp_driver_matches_oracle(driver, case):
    if case == "uncle":
        binding, npos, rels, lits = uncle_problem()
    else:
        binding, npos, rels, lits = random_problem(61, n=700, rel_shapes=((3, 400), (2, 350)), domain=16)
    prob = oracle.Problem(binding, npos, rels, lits)
    want, _ = prob.bruteforce()
    plans = _plans(lits, range(len(lits)))
    script = abi_driver_script(binding, npos, rels, lits, plans)
    print("abi driver: case", case, "literals", len(lits), flush=True)
    r = subprocess.run([driver], input=script, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = parse_driver_output(r.stdout)
    for tag in ("EVAL", "SPLIT", "AGG", "AGG2"):
        got = np.array([out[tag][i] for i in range(len(lits))], dtype=np.uint64)
        np.testing.assert_array_equal(got, want, err_msg=tag)
    for i in range(len(lits)):
        assert out["SIZE"][i] == str(int(want[i][2] + want[i][3]))
    # RadixPartition of binding column 0 at 3 bits, then the chained tables on its partitions.
    pairs, offsets = oracle.radix_partition(binding[0], 3)
    assert out["OFFSETS"] == offsets.tolist()
    assert out["PAIRS"] == pairs.tolist()
    buckets, boff, nxt = oracle.build_hash_table_on_partitions(pairs, offsets, 3)
    assert out["PBUCKETS"] == buckets.tolist()[: boff[-1]]
    assert out["PNEXT"] == nxt.tolist()[: len(binding[0])]
    tb, tn = oracle.build_hash_table_on_table(rels[0])
    assert out["TBUCKETS"] == tb.tolist()
    assert out["TNEXT"] == tn.tolist()
