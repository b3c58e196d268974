"""The learner's candidate enumerator against the expectations of
learner/CandidateLiteralEnumerator_test.cpp (predicate fixture :93-130, duplicates :335-350,
enumeration without and with dynamic pruning :352-450).  The expected literal lists are the
reference test's data; the comparison is by set, as the reference's EQUAL_CHECK does."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "quickfoil_amd", "csrc", "learner")

# name, key, argument types (fixture :93-130; keys follow creation order, so p_3..p_5 have keys
# beyond their arity, which select no variable); background order p_0, p_1, p_5, p_2, p_3, p_4.
PREDICATES = [("p_0", 0, [0]), ("p_1", 1, [0, 1]), ("p_5", 5, [0, 0]), ("p_2", 2, [1, 0, 2]),
              ("p_3", 3, [0, 0, 1]), ("p_4", 4, [1, 0, 1])]
CLAUSES = ["p_5(0, 1) :-", "p_5(0, 1) :- p_5(0, 2)", "p_5(0, 1) :- p_5(0, 2), p_5(2, 1)",
           "p_5(0, 1) :- p_5(0, 2), p_5(2, 1), p_5(0, 3)"]

ENTIRE_NO_PRUNING = """p_0(0) p_0(1) p_0(2) p_0(3) p_1(0,-1) p_1(1,-1) p_1(2,-1) p_1(3,-1) p_2(-1,0,-1) p_2(-1,1,-1)
p_2(-1,2,-1) p_2(-1,3,-1) p_3(-1,0,-1) p_3(-1,1,-1) p_3(-1,2,-1) p_3(-1,3,-1) p_3(0,-1,-1) p_3(0,0,-1) p_3(0,1,-1)
p_3(0,2,-1) p_3(0,3,-1) p_3(1,-1,-1) p_3(1,0,-1) p_3(1,1,-1) p_3(1,2,-1) p_3(1,3,-1) p_3(2,-1,-1) p_3(2,0,-1)
p_3(2,1,-1) p_3(2,2,-1) p_3(2,3,-1) p_3(3,-1,-1) p_3(3,0,-1) p_3(3,1,-1) p_3(3,2,-1) p_3(3,3,-1) p_4(-1,0,-1)
p_4(-1,1,-1) p_4(-1,2,-1) p_4(-1,3,-1) p_5(-1,0) p_5(-1,1) p_5(-1,2) p_5(-1,3) p_5(0,-1) p_5(0,0) p_5(0,2) p_5(0,3)
p_5(1,-1) p_5(1,0) p_5(1,1) p_5(1,2) p_5(1,3) p_5(2,-1) p_5(2,0) p_5(2,2) p_5(2,3) p_5(3,-1) p_5(3,0) p_5(3,1)
p_5(3,2) p_5(3,3)"""
PRUNED_NO_PRUNING = """p_0(0) p_1(0,-1) p_2(-1,0,-1) p_3(-1,0,-1) p_3(0,-1,-1) p_3(0,3,-1) p_3(0,2,-1) p_3(0,1,-1)
p_3(0,0,-1) p_4(-1,0,-1) p_5(-1,0) p_5(0,0) p_0(1) p_1(1,-1) p_2(-1,1,-1) p_3(-1,1,-1) p_3(1,-1,-1) p_3(1,3,-1)
p_3(1,2,-1) p_3(1,1,-1) p_3(1,0,-1) p_4(-1,1,-1) p_5(1,-1) p_5(1,3) p_5(1,2) p_5(1,1) p_5(1,0) p_0(2) p_1(2,-1)
p_2(-1,2,-1) p_3(-1,2,-1) p_3(2,-1,-1) p_3(2,3,-1) p_3(2,2,-1) p_3(2,1,-1) p_3(2,0,-1) p_4(-1,2,-1) p_5(2,2)
p_5(2,0) p_0(3) p_1(3,-1) p_2(-1,3,-1) p_3(-1,3,-1) p_3(3,-1,-1) p_3(3,3,-1) p_3(3,2,-1) p_3(3,1,-1) p_3(3,0,-1)
p_4(-1,3,-1) p_5(3,0)"""
ENTIRE_PRUNING = """p_0(0) p_0(1) p_0(2) p_0(3) p_1(1,-1) p_1(2,-1) p_1(3,-1) p_2(-1,0,-1) p_2(-1,1,-1) p_2(-1,2,-1)
p_2(-1,3,-1) p_3(-1,1,-1) p_3(-1,2,-1) p_3(-1,3,-1) p_3(0,0,-1) p_3(0,1,-1) p_3(1,-1,-1) p_3(1,0,-1) p_3(1,1,-1)
p_3(1,2,-1) p_3(1,3,-1) p_3(2,1,-1) p_3(2,2,-1) p_3(3,-1,-1) p_3(3,1,-1) p_3(3,2,-1) p_3(3,3,-1) p_4(-1,0,-1)
p_4(-1,1,-1) p_4(-1,2,-1) p_4(-1,3,-1) p_5(-1,0) p_5(-1,1) p_5(-1,2) p_5(-1,3) p_5(0,-1) p_5(0,0) p_5(0,2) p_5(0,3)
p_5(1,-1) p_5(1,0) p_5(1,1) p_5(1,2) p_5(1,3) p_5(2,0) p_5(2,2) p_5(3,-1) p_5(3,0) p_5(3,1) p_5(3,2) p_5(3,3)"""
PRUNED_PRUNING = """p_0(0) p_0(1) p_0(2) p_0(3) p_1(1,-1) p_1(2,-1) p_1(3,-1) p_2(-1,0,-1) p_2(-1,1,-1) p_2(-1,2,-1)
p_2(-1,3,-1) p_3(-1,1,-1) p_3(-1,2,-1) p_3(-1,3,-1) p_3(0,0,-1) p_3(0,1,-1) p_3(1,-1,-1) p_3(1,0,-1) p_3(1,1,-1)
p_3(1,2,-1) p_3(1,3,-1) p_3(2,1,-1) p_3(2,2,-1) p_3(3,-1,-1) p_3(3,1,-1) p_3(3,2,-1) p_3(3,3,-1) p_4(-1,0,-1)
p_4(-1,1,-1) p_4(-1,2,-1) p_4(-1,3,-1) p_5(-1,0) p_5(0,0) p_5(1,-1) p_5(1,0) p_5(1,1) p_5(1,2) p_5(1,3) p_5(2,0)
p_5(2,2) p_5(3,0)"""
REMOVED_PRUNING = ["p_1(0, -1), p_3(-1, 0, -1), p_3(0, -1, -1)", "p_3(2, -1, -1), p_5(2, -1)", "", ""]


@pytest.fixture(scope="module")
def driver():
    out = os.path.join(tempfile.mkdtemp(), "enumerator_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + SRC, os.path.join(HERE, "enumerator_driver.cpp"),
                    os.path.join(SRC, "enumerator.cpp"), "-o", out], check=True)
    return out


def run(driver, lines):
    script = "".join("PRED %s %d %s\n" % (n, k, " ".join(map(str, t))) for n, k, t in PREDICATES)
    script += "".join(l + "\n" for l in lines)
    return subprocess.run([driver], input=script, capture_output=True, text=True, check=True).stdout.splitlines()


def as_set(text):
    return {t.replace(" ", "") for t in text.split()}


def parse(line, tag):
    assert line.startswith(tag)
    return [t.replace(" ", "") for t in line[len(tag):].replace(", ", ",").split()]


@pytest.mark.parametrize("clause,literal,expect", [
    ("p_0(0) :- p_1(0, 1)", "p_1(0, -1)", 1),
    ("p_0(0) :- p_1(0, 1)", "p_1(-1, 1)", 0),
    ("p_0(0) :- p_1(0, 1), p_2(1, 0, 2)", "p_2(1, -1, 2)", 0),
    ("p_0(0) :- p_1(0, 1), p_2(1, 0, 2)", "p_1(0, -1)", 0),
    ("p_0(0) :- p_1(0, 1), p_2(1, 0, 2)", "p_2(1, 0, -1)", 1),
    ("p_5(0, 1) :- p_5(0, 2), p_5(2, 1), p_5(0, 3)", "p_5(0, -1)", 1),
    ("p_5(0, 1) :- p_5(0, 2), p_5(2, 1), p_5(0, 3)", "p_5(3, 1)", 1),
    ("p_5(0, 1) :- p_5(0, 2), p_5(2, 1), p_5(0, 3)", "p_5(3, -1)", 1),
    ("p_5(0, 1) :- p_5(0, 2), p_5(2, 1), p_5(0, 3)", "p_5(1, 2)", 0),
])
def test_replaceable_duplicates(driver, clause, literal, expect):
    out = run(driver, ["CLAUSE " + clause, "DUP " + literal])
    assert out == ["DUP %d" % expect]


@pytest.mark.parametrize("removed,entire,pruned", [
    (["", "", "", ""], ENTIRE_NO_PRUNING, PRUNED_NO_PRUNING),
    (REMOVED_PRUNING, ENTIRE_PRUNING, PRUNED_PRUNING),
], ids=["without_dynamic_pruning", "with_dynamic_pruning"])
def test_enumeration_sequence(driver, removed, entire, pruned):
    lines = []
    for c, r in zip(CLAUSES, removed):
        lines += ["CLAUSE " + c, "ENUM " + r]
    out = run(driver, lines)
    got_entire = parse(out[-2], "ENTIRE")
    got_pruned = parse(out[-1], "PRUNED")
    assert len(got_entire) == len(set(got_entire)) and len(got_pruned) == len(set(got_pruned))
    assert set(got_entire) == as_set(entire)
    assert set(got_pruned) == as_set(pruned)
