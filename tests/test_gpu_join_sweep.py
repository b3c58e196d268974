"""Seeded sweep of the binding-set extension, CreateBindingTable and the coverage semi-joins over
random shapes (empty and one-row tables, arities 1-4, repeated variables, one-value to large
domains, skew, keys of 1-6 columns), row for row and bit for bit against the oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = 80
MAX_OUT_ROWS = 2_000_000


def _cols(rng, n, k, domain, zipf):
    if zipf:
        return [((rng.zipf(1.5, size=n) - 1) % domain).astype(np.int32) for _ in range(k)]
    return [rng.integers(0, domain, size=n, dtype=np.int64).astype(np.int32) for _ in range(k)]


def join_case(seed):
    rng = np.random.Generator(np.random.PCG64(0xB1D0 + seed))
    nb = int(rng.choice([0, 1, 64, 500, 3000]))
    ncols = int(rng.integers(1, 5))
    domain = int(rng.choice([1, 3, 30, 1000]))
    zipf = bool(rng.integers(0, 2))
    arity = int(rng.integers(1, 5))
    rows = int(rng.choice([0, 1, 40, 600, 4000]))
    binding = _cols(rng, nb, ncols, domain, zipf)
    rel = _cols(rng, rows, arity, domain, zipf)
    while True:
        vars_ = [int(v) for v in rng.integers(-1, ncols, size=arity)]
        if any(v >= 0 for v in vars_):
            break
    return binding, int(rng.integers(0, nb + 1)), rel, vars_


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


@pytest.mark.parametrize("seed", range(CASES))
def test_extension_and_binding_table_sweep(qf_ctx, seed):
    binding, npos, rel, vars_ = join_case(seed)
    prob = oracle.Problem(binding, npos, [rel], [(0, vars_)])
    counts = prob.grouped()[0][0]
    total = int(counts[2] + counts[3])
    print("join case", seed, "rows", len(binding[0]), "cols", len(binding), "relation", (len(rel), len(rel[0])),
          "vars", vars_, "output rows", total, flush=True)
    if total > MAX_OUT_ROWS:
        pytest.skip("output beyond the sweep's size bound")
    _, want = prob.label_aware_binding_table(0, int(counts[2]), int(counts[3]))
    got = qf_ctx.create_label_aware_binding_table([_t(c) for c in binding], npos, [_t(c) for c in rel], vars_,
                                                  int(counts[2]), int(counts[3]))
    assert len(got) == len(want)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g.cpu().numpy(), w)
    want = prob.binding_table(0)
    got = qf_ctx.create_binding_table([_t(c) for c in binding], [_t(c) for c in rel], vars_)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g.cpu().numpy(), w)


@pytest.mark.parametrize("seed", range(CASES))
def test_semi_join_sweep(qf_ctx, seed):
    rng = np.random.Generator(np.random.PCG64(0x5E31 + seed))
    nkeys = int(rng.integers(1, 7))
    domain = int(rng.choice([1, 3, 30, 1000]))
    zipf = bool(rng.integers(0, 2))
    pk = _cols(rng, int(rng.choice([1, 63, 64, 65, 2000, 40000])), nkeys, domain, zipf)
    bk = _cols(rng, int(rng.choice([1, 5, 700, 30000])), nkeys, domain, zipf)
    print("semi-join case", seed, "keys", nkeys, "probe", len(pk[0]), "build", len(bk[0]), flush=True)
    left, ones = qf_ctx.left_semi_join([_t(c) for c in pk], [_t(c) for c in bk])
    want = oracle.left_semi_join(pk, bk)
    np.testing.assert_array_equal(left.cpu().numpy().view(np.uint64), want)
    assert ones == int(sum(bin(int(w)).count("1") for w in want))
    right, _ = qf_ctx.right_semi_join([_t(c) for c in pk], [_t(c) for c in bk])
    np.testing.assert_array_equal(right.cpu().numpy().view(np.uint64), oracle.right_semi_join(pk, bk))
