"""The bench's roofline byte accounting and the C layout of LiteralArray (host only)."""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from quickfoil_amd import _native  # noqa: E402


class _Col:
    def __init__(self, n):
        self.n = n

    def numel(self):
        return self.n


def _workload():
    # 1,000 binding rows x 3 columns; relation 0 (arity 2, 50 rows), relation 1 (arity 3, 70 rows).
    binding = [_Col(1000) for _ in range(3)]
    relations = [[_Col(50), _Col(50)], [_Col(70), _Col(70), _Col(70)]]
    literals = [(0, [0, -1]), (0, [0, 1]), (0, [2, 1]), (1, [-1, 0, 2])]
    return {"binding": binding, "relations": relations, "literals": literals}


def test_probe_bytes_per_join_group():
    total, groups = bench.algorithmic_probe_bytes(_workload())
    # Join groups (relation, join position, variable): (0,0,X0) with 2 literals and atom
    # columns {1} / binding {1}; (0,0,X2) with 1 literal, {1} / {1}; (1,1,X0) with 1 literal, {2} / {2}.
    want = (8 * 50 + 8 * 1000 + 4 * 50 * 1 + 4 * 1000 * 1 + 32 * 2) \
        + (8 * 50 + 8 * 1000 + 4 * 50 * 1 + 4 * 1000 * 1 + 32 * 1) \
        + (8 * 70 + 8 * 1000 + 4 * 70 * 1 + 4 * 1000 * 1 + 32 * 1)
    assert groups == 3
    assert total == want


def test_partition_bytes_per_join_variable():
    # Join variables {X0, X2}; referenced binding columns {X0, X1, X2}: each variable's partition
    # reads 3 columns and writes 3 plus the row id.
    assert bench.algorithmic_partition_bytes(_workload()) == 2 * 1000 * 4 * (3 + 3 + 1)


def test_literal_array_matches_the_c_struct():
    lits = [(3, [0, -1]), (1, [2, 1, -1]), (0, [5, 4, 3, 2, 1, 0])]
    la = _native.LiteralArray(lits)
    assert ctypes.sizeof(_native.QfLiteral) == la.array.strides[0]
    p = la.pointer()
    for i, (rel, vars_) in enumerate(lits):
        assert p[i].relation == rel
        assert list(p[i].vars) == list(vars_) + [-1] * (_native.MAX_ARITY - len(vars_))
    assert len(la) == 3
    assert np.array_equal(la.array[:, 0], [3, 1, 0])
