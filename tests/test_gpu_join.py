"""GPU parity of binding-set extension (row for row) and the coverage semi-joins."""
import itertools

import numpy as np
import pytest

import oracle
from helpers import random_problem
from quickfoil_amd import QfError

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


@pytest.mark.parametrize("vars_", [[0, -1], [0, 1], [1, -1], [-1, 0], [0, 0], [1, 0]])
@pytest.mark.parametrize("rel_rows", [50, 5000])
def test_label_aware_binding_table_row_order(qf_ctx, vars_, rel_rows):
    # rel_rows 50 < labels: hash on the relation; 5000: the relation probes both label tables.
    binding, npos, rels, _ = random_problem(31, n=600, num_cols=2, rel_shapes=((2, rel_rows),), domain=25)
    prob = oracle.Problem(binding, npos, rels, [(0, vars_)])
    counts = prob.bruteforce()[0][0]
    got_rows, want = prob.label_aware_binding_table(0, int(counts[2]), int(counts[3]))
    out = qf_ctx.create_label_aware_binding_table([_t(c) for c in binding], npos, [_t(c) for c in rels[0]], vars_,
                                                  int(counts[2]), int(counts[3]))
    assert len(out) == len(want)
    for g, w in zip(out, want):
        np.testing.assert_array_equal(g.cpu().numpy(), w)


def test_label_aware_binding_table_rejects_wrong_size(qf_ctx):
    binding, npos, rels, _ = random_problem(32, n=200, num_cols=2, rel_shapes=((2, 100),), domain=10)
    counts = oracle.Problem(binding, npos, rels, [(0, [0, -1])]).bruteforce()[0][0]
    with pytest.raises(QfError) as e:
        qf_ctx.create_label_aware_binding_table([_t(c) for c in binding], npos, [_t(c) for c in rels[0]], [0, -1],
                                                int(counts[2]) - 1, int(counts[3]))
    assert e.value.status == 4


@pytest.mark.parametrize("vars_,rel_rows", [([0, 2], 50), ([0, 2], 900), ([1, 0], 900), ([2, 0], 300)])
def test_create_binding_table_row_order(qf_ctx, vars_, rel_rows):
    binding, npos, rels, _ = random_problem(33, n=400, num_cols=2, rel_shapes=((2, rel_rows),), domain=20)
    prob = oracle.Problem(binding, npos, rels, [(0, vars_)])
    want = prob.binding_table(0)
    got = qf_ctx.create_binding_table([_t(c) for c in binding], [_t(c) for c in rels[0]], vars_)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g.cpu().numpy(), w)


def test_literal_without_bound_variable_is_rejected(qf_ctx):
    binding, npos, rels, _ = random_problem(34, n=50, num_cols=2, rel_shapes=((2, 40),), domain=5)
    with pytest.raises(QfError) as e:
        qf_ctx.create_binding_table([_t(c) for c in binding], [_t(c) for c in rels[0]], [2, 3])
    assert e.value.status == 1


@pytest.mark.parametrize("nkeys", [1, 2, 3])
@pytest.mark.parametrize("sizes", [(1000, 300), (300, 1000), (70000, 5000)])
def test_semi_joins(qf_ctx, nkeys, sizes):
    rng = np.random.Generator(np.random.PCG64(nkeys * 13 + sizes[0]))
    pk = [rng.integers(0, 40, sizes[0]).astype(np.int32) for _ in range(nkeys)]
    bk = [rng.integers(0, 40, sizes[1]).astype(np.int32) for _ in range(nkeys)]
    left, ones = qf_ctx.left_semi_join([_t(c) for c in pk], [_t(c) for c in bk])
    want = oracle.left_semi_join(pk, bk)
    np.testing.assert_array_equal(left.cpu().numpy().view(np.uint64), want)
    assert ones == int(sum(bin(int(w)).count("1") for w in want))
    right, ones = qf_ctx.right_semi_join([_t(c) for c in pk], [_t(c) for c in bk])
    want = oracle.right_semi_join(pk, bk)
    np.testing.assert_array_equal(right.cpu().numpy().view(np.uint64), want)


def test_hash_table_after_semi_join(qf_ctx):
    rng = np.random.Generator(np.random.PCG64(44))
    n = 20000
    keys = [rng.integers(0, 300, n).astype(np.int32), rng.integers(0, 4, n).astype(np.int32)]
    bits = oracle.left_semi_join(keys, [rng.integers(0, 300, 500).astype(np.int32),
                                        rng.integers(0, 4, 500).astype(np.int32)])
    import torch
    d_bits = torch.from_numpy(bits.view(np.int64).copy()).cuda()
    buckets, nxt, first = qf_ctx.build_hash_table_after_semi_join([_t(c) for c in keys], d_bits)
    want_first = oracle.dedup_after_semi_join(keys, bits)
    np.testing.assert_array_equal(first.cpu().numpy(), want_first)
    # Chains hold exactly the first occurrences of each bucket, newest first.
    b = buckets.cpu().numpy()
    nx = nxt.cpu().numpy()
    nb = len(b)
    ins = np.nonzero(want_first)[0]
    for bucket in range(nb):
        pos, got = int(b[bucket]) - 1, []
        while pos >= 0:
            got.append(pos)
            pos = int(nx[pos]) - 1
        want = [i for i in ins[::-1] if (oracle.hash_combine(oracle.hash_int(keys[0][i]), keys[1][i]) & (nb - 1)) == bucket]
        assert got == want


@pytest.mark.parametrize("vars_,rows", [([0, -1], (0, 300)), ([1, 2], (100, 400)), ([2, 2], (0, 400)), ([-1, 0], (50, 60))])
def test_example_coverage(qf_ctx, vars_, rows):
    """Examples covered by the binding rows in range that match the literal (brute force)."""
    rng = np.random.Generator(np.random.PCG64(sum(rows) + len(vars_)))
    binding = [rng.integers(0, 12, 400).astype(np.int32) for _ in range(3)]
    rel = [rng.integers(0, 12, 150).astype(np.int32) for _ in range(2)]
    examples = [rng.integers(0, 12, 200).astype(np.int32) for _ in range(2)]
    r0, r1 = rows
    facts = {tuple(int(rel[p][s]) for p in range(2) if vars_[p] >= 0) for s in range(150)}
    heads = set()
    for r in range(r0, r1):
        key = tuple(int(binding[v][r]) for v in vars_ if v >= 0)
        if key in facts:
            heads.add((int(binding[0][r]), int(binding[1][r])))
    want = [(int(examples[0][e]), int(examples[1][e])) in heads for e in range(200)]
    bits, ones = qf_ctx.example_coverage([_t(c) for c in binding], r0, r1, [_t(c) for c in rel], vars_,
                                         [_t(c) for c in examples])
    b = bits.cpu().numpy().view(np.uint64)
    got = [bool((int(b[e >> 6]) >> (e & 63)) & 1) for e in range(200)]
    assert got == want
    assert ones == sum(want)


@pytest.mark.parametrize("seed", [41, 42])
def test_device_resident_binding_lifecycle(qf_ctx, seed):
    """Two refinement steps with the binding set kept on the device: step 1's counts size the
    label-aware extension (built on the device), whose columns feed step 2's evaluation directly;
    both steps' counts and the extension's rows equal the oracle's."""
    binding, npos, rels, lits = random_problem(seed, n=3000, num_cols=2, rel_shapes=((2, 400), (3, 300)),
                                               domain=30)
    dev_rels = [[_t(c) for c in r] for r in rels]
    counts1, _ = qf_ctx.evaluate_literals([_t(c) for c in binding], npos, dev_rels, lits)
    prob1 = oracle.Problem(binding, npos, rels, lits)
    np.testing.assert_array_equal(counts1, prob1.grouped()[0])
    # The literal with the most binding rows that adds a new variable.
    cand = [i for i, (r, v) in enumerate(lits) if -1 in v and counts1[i][2] + counts1[i][3] > 0]
    pick = max(cand, key=lambda i: int(counts1[i][2] + counts1[i][3]))
    rel, vars_ = lits[pick]
    ext = qf_ctx.create_label_aware_binding_table([_t(c) for c in binding], npos, dev_rels[rel], vars_,
                                                  int(counts1[pick][2]), int(counts1[pick][3]))
    _, want_ext = prob1.label_aware_binding_table(pick, int(counts1[pick][2]), int(counts1[pick][3]))
    for g, w in zip(ext, want_ext):
        np.testing.assert_array_equal(g.cpu().numpy(), w)
    # Step 2 on the extension as it lies in device memory (positives first).
    npos2 = int(counts1[pick][2])
    ncols2 = len(ext)
    lits2 = [(r, v) for r, (arity, _) in enumerate(((2, 400), (3, 300)))
             for v in itertools.product(range(-1, ncols2), repeat=arity) if any(x >= 0 for x in v)]
    lits2 = [(r, list(v)) for r, v in lits2]
    counts2, _ = qf_ctx.evaluate_literals(list(ext), npos2, dev_rels, lits2)
    prob2 = oracle.Problem([g.cpu().numpy() for g in ext], npos2, rels, lits2)
    np.testing.assert_array_equal(counts2, prob2.grouped()[0])
