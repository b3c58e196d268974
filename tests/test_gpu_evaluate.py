"""GPU parity of candidate-literal evaluation: counts and covered bitvectors, bit-exact."""
import numpy as np
import pytest

import oracle
from helpers import random_problem, uncle_problem

pytestmark = pytest.mark.gpu

MODES = ["batched", "plain_group", "plain_literal"]


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


def _gpu_eval(ctx, binding, npos, rels, lits, **kw):
    counts, bits = ctx.evaluate_literals([_t(c) for c in binding], npos, [[_t(c) for c in r] for r in rels], lits, **kw)
    if bits is not None:
        bits = bits.cpu().numpy().view(np.uint64)
    return counts, bits


@pytest.mark.parametrize("mode", MODES)
def test_uncle_sample_data(qf_ctx, mode):
    target, npos, rels, lits = uncle_problem()
    want, want_bits = oracle.Problem(target, npos, rels, lits).bruteforce(semi=True)
    got, bits = _gpu_eval(qf_ctx, target, npos, rels, lits, mode=mode, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("mode", MODES)
def test_random_small_vs_bruteforce(qf_ctx, seed, mode):
    binding, npos, rels, lits = random_problem(seed, n=400, rel_shapes=((2, 300), (3, 200), (1, 40)), domain=15,
                                               zipf=(seed % 2 == 0))
    prob = oracle.Problem(binding, npos, rels, lits)
    want, want_bits = prob.bruteforce(semi=True)
    pipe, _, _ = prob.pipeline()
    np.testing.assert_array_equal(pipe, want)
    got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, mode=mode, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


@pytest.mark.parametrize("radix_bits,tile", [(1, 64), (4, 256), (8, 1024), (12, 2048)])
def test_radix_and_tile_choices_do_not_change_counts(qf_ctx, radix_bits, tile):
    binding, npos, rels, lits = random_problem(21, n=20000, rel_shapes=((2, 9000), (3, 7000)), domain=400,
                                               zipf=True, max_literals=40)
    want, want_bits = oracle.Problem(binding, npos, rels, lits).grouped(semi=True)
    got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, radix_bits=radix_bits, tile_rows=tile, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


def test_heavy_partitions_and_skew(qf_ctx):
    # A relation with one value holding most rows forces HBM tables for its partition.
    rng = np.random.Generator(np.random.PCG64(5))
    n, m = 50000, 60000
    binding = [rng.integers(0, 50, n).astype(np.int32) for _ in range(3)]
    hot = np.where(rng.random(m) < 0.7, 7, rng.integers(0, 50, m)).astype(np.int32)
    rel = [hot, rng.integers(0, 50, m).astype(np.int32), rng.integers(0, 3, m).astype(np.int32)]
    lits = [(0, [0, -1, -1]), (0, [0, 1, -1]), (0, [0, 1, 2]), (0, [-1, 2, -1]), (0, [1, 1, -1]), (0, [2, 0, 1])]
    want, want_bits = oracle.Problem(binding, 20000, [rel], lits).grouped(semi=True)
    for mode in MODES:
        got, bits = _gpu_eval(qf_ctx, binding, 20000, [rel], lits, mode=mode, semi_bits=True, radix_bits=6)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(bits, want_bits)


def test_edge_cases(qf_ctx):
    binding = [np.array([1, 2, 3], np.int32), np.array([4, 5, 6], np.int32)]
    rel = [np.array([9], np.int32), np.array([9], np.int32)]
    lits = [(0, [0, -1]), (0, [0, 1])]
    for npos in (0, 3):  # all negative, all positive
        want, _ = oracle.Problem(binding, npos, [rel], lits).bruteforce()
        got, _ = _gpu_eval(qf_ctx, binding, npos, [rel], lits)
        np.testing.assert_array_equal(got, want)
    # Duplicate rows count separately; repeated variables; negative values.
    binding = [np.array([-5, -5, 7, 7], np.int32), np.array([-5, 1, 7, 7], np.int32)]
    rel = [np.array([-5, -5, 7, 7, 7], np.int32), np.array([-5, -5, 7, 1, 7], np.int32)]
    lits = [(0, [0, 0]), (0, [1, 0]), (0, [0, -1]), (0, [-1, 1])]
    want, want_bits = oracle.Problem(binding, 2, [rel], lits).bruteforce(semi=True)
    got, bits = _gpu_eval(qf_ctx, binding, 2, [rel], lits, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


def test_no_literals_and_empty_binding(qf_ctx):
    binding = [np.array([1, 2], np.int32)]
    rel = [np.array([1], np.int32)]
    got, _ = _gpu_eval(qf_ctx, binding, 1, [rel], [])
    assert got.shape == (0, 4)


def test_count_aggregate_plan_matches_literals(qf_ctx):
    binding, npos, rels, _ = random_problem(8, n=3000, num_cols=3, rel_shapes=((3, 2000),), domain=12)
    lits_vars = [[0, -1, -1], [0, 1, -1], [0, -1, 2], [0, 1, 2], [0, 2, 1], [0, 0, 0]]
    plan = oracle.generate_plan(3, 0, lits_vars)
    conj = [(plan["left"][x], plan["right"][x], plan["literal"][x]) for x in range(len(plan["atoms"]), len(plan["left"]))]
    p = {"relation": 0, "join_column": 0, "binding_column": 0, "atoms": plan["atoms"],
         "atom_literals": plan["literal"][: len(plan["atoms"])], "conjunctions": conj,
         "root_literal": plan["root_literal"]}
    got, _ = qf_ctx.count_aggregate([_t(c) for c in binding], npos, [[_t(c) for c in r] for r in rels], [p],
                                    len(lits_vars))
    want, _ = oracle.Problem(binding, npos, rels, [(0, v) for v in lits_vars]).bruteforce()
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dist", ["uniform", "zipf"])
def test_microbench_shape_scaled(qf_ctx, dist):
    """The microbench workload at 1/64 scale against the grouped-count closed form."""
    from quickfoil_amd import datagen
    w = datagen.microbench(dist, rows=262144, rel_rows=65536, num_values=16384)
    want, want_bits = oracle.Problem(w["binding"], w["num_positive"], w["relations"], w["literals"]).grouped(semi=True)
    for mode in MODES:
        got, bits = _gpu_eval(qf_ctx, w["binding"], w["num_positive"], w["relations"], w["literals"], mode=mode,
                              semi_bits=True)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(bits, want_bits)


@pytest.mark.parametrize("partition_rows,tile,semi", [(0, 0, False), (128, 512, False), (256, 768, True),
                                                     (512, 2048, False)])
def test_refinement_step_scaled_golden(qf_ctx, partition_rows, tile, semi):
    """The refinement-step workload at 1/256 scale (2,080 literals over 20 relations of arity 2 and
    3: key subsets of two and three columns, whose parents' hits gate them) against the oracle's
    committed counts; covered bitvectors, where asked for, popcount to the covered counts."""
    import os
    from quickfoil_amd import datagen
    want = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "step_1_256_counts.npy"))
    w = datagen.refinement_step(scale=1 / 256)
    got, bits = _gpu_eval(qf_ctx, w["binding"], w["num_positive"], w["relations"], w["literals"],
                          partition_rows=partition_rows, tile_rows=tile, semi_bits=semi)
    np.testing.assert_array_equal(got, want)
    if semi:
        n, npos = len(w["binding"][0]), w["num_positive"]
        rows = np.unpackbits(bits.view(np.uint8), axis=1, bitorder="little")[:, :n]
        np.testing.assert_array_equal(rows[:, :npos].sum(axis=1), want[:, 0])
        np.testing.assert_array_equal(rows[:, npos:].sum(axis=1), want[:, 1])


def _covered_from_bits(bits, n, npos):
    rows = np.unpackbits(bits.view(np.uint8), axis=1, bitorder="little")[:, :n]
    return rows[:, :npos].sum(axis=1), rows[:, npos:].sum(axis=1)


@pytest.mark.parametrize("dist", ["uniform", "zipf"])
def test_microbench_full_size(qf_ctx, dist):
    """The operator microbench at full size (16,777,216 x 3 bindings, 4,194,304 x 2 facts over 1 M
    keys) against the oracle: its materialising pipeline for uniform keys, the grouped-count closed
    form for Zipf keys (5.6e11 matching pairs, beyond what a pipeline can hold); the covered
    bitvectors popcount to the covered counts."""
    from quickfoil_amd import datagen
    w = datagen.microbench(dist)
    prob = oracle.Problem(w["binding"], w["num_positive"], w["relations"], w["literals"])
    want = prob.pipeline()[0] if dist == "uniform" else prob.grouped()[0]
    got, bits = _gpu_eval(qf_ctx, w["binding"], w["num_positive"], w["relations"], w["literals"], semi_bits=True)
    np.testing.assert_array_equal(got, want)
    cov_pos, cov_neg = _covered_from_bits(bits, len(w["binding"][0]), w["num_positive"])
    np.testing.assert_array_equal(cov_pos, want[:, 0])
    np.testing.assert_array_equal(cov_neg, want[:, 1])
    qf_ctx.trim()


def test_refinement_step_full_size_batched_equals_plain(qf_ctx):
    """The refinement step at full size (8,388,608 x 4 bindings, 20 relations of up to 26.7 M facts,
    2,080 literals), where no CPU oracle finishes in test time: the batched form (one radix-bit
    count for every relation, batches of up to 32 units, graph replay) equals the plain per-group
    form (each relation at its own radix bits, one launch per group), and the counts obey
    covered <= binding, covered positives <= positives, and monotonicity: binding one more
    position of a literal never raises any of its four counts."""
    import torch
    from quickfoil_amd import datagen
    w = datagen.refinement_step(device=torch.device("cuda:0"))
    args = (w["binding"], w["num_positive"], w["relations"], w["literals"])
    batched, _ = qf_ctx.evaluate_literals(*args, reuse_partitions=True, use_graph=True)
    plain, _ = qf_ctx.evaluate_literals(*args, reuse_partitions=True, mode="plain_group")
    np.testing.assert_array_equal(batched, plain)
    c = batched.astype(np.int64)
    n, npos = int(w["binding"][0].numel()), w["num_positive"]
    assert (c[:, 0] <= c[:, 2]).all() and (c[:, 1] <= c[:, 3]).all()
    assert (c[:, 0] <= npos).all() and (c[:, 1] <= n - npos).all()
    assert c.sum() > 0
    index = {(r, tuple(v)): i for i, (r, v) in enumerate(w["literals"])}
    pairs = 0
    for i, (r, v) in enumerate(w["literals"]):
        for p, x in enumerate(v):
            if x >= 0:
                continue
            for var in range(len(w["binding"])):
                j = index[(r, tuple(v[:p]) + (var,) + tuple(v[p + 1:]))]
                assert (c[j] <= c[i]).all(), (w["literals"][i], w["literals"][j])
                pairs += 1
    assert pairs > 1000
    qf_ctx.trim()


def test_graph_replay_matches_eager(qf_ctx):
    binding, npos, rels, lits = random_problem(41, n=30000, rel_shapes=((2, 20000), (3, 9000)), domain=300,
                                               zipf=True, max_literals=60)
    want, _ = oracle.Problem(binding, npos, rels, lits).grouped()
    dev_b = [_t(c) for c in binding]
    dev_r = [[_t(c) for c in r] for r in rels]
    eager, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, lits)
    np.testing.assert_array_equal(eager, want)
    for i in range(3):
        got, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, lits, use_graph=True, reuse_partitions=True)
        np.testing.assert_array_equal(got, want)
        # Eager work between replays.
        again, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, lits, mode="plain_group")
        np.testing.assert_array_equal(again, want)
    qf_ctx.trim()


def test_graph_replay_follows_changed_calls(qf_ctx):
    """A repeated call replays the captured step straight from its raw arguments; any change of
    literal list, num_positive or binding contents (same columns, rewritten in place) gives that
    call's own counts, and returning to the first call gives the first counts again."""
    binding, npos, rels, lits = random_problem(43, n=20000, rel_shapes=((2, 12000), (3, 7000)), domain=250,
                                               zipf=True, max_literals=50)
    dev_b = [_t(c) for c in binding]
    dev_r = [[_t(c) for c in r] for r in rels]
    kw = dict(use_graph=True, reuse_partitions=True)
    want, _ = oracle.Problem(binding, npos, rels, lits).grouped()
    for _ in range(2):
        got, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, lits, **kw)
        np.testing.assert_array_equal(got, want)
    fewer = lits[: len(lits) // 2]
    got, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, fewer, **kw)
    np.testing.assert_array_equal(got, oracle.Problem(binding, npos, rels, fewer).grouped()[0])
    got, _ = qf_ctx.evaluate_literals(dev_b, npos // 2, dev_r, lits, **kw)
    np.testing.assert_array_equal(got, oracle.Problem(binding, npos // 2, rels, lits).grouped()[0])
    got, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, lits, **kw)
    np.testing.assert_array_equal(got, want)
    shifted = [np.roll(c, 777) for c in binding]
    for d, c in zip(dev_b, shifted):
        d.copy_(_t(c))
    got, _ = qf_ctx.evaluate_literals(dev_b, npos, dev_r, lits, **kw)
    np.testing.assert_array_equal(got, oracle.Problem(shifted, npos, rels, lits).grouped()[0])
    qf_ctx.trim()


def test_table_set_matches_column_lists(qf_ctx):
    """Relations handed over once as a TableSet count exactly as the same column lists, eagerly
    and through graph replays."""
    import quickfoil_amd
    binding, npos, rels, lits = random_problem(47, n=9000, rel_shapes=((2, 6000), (3, 4000), (1, 300)), domain=120,
                                               zipf=True, max_literals=40)
    want, _ = oracle.Problem(binding, npos, rels, lits).grouped()
    dev_b = [_t(c) for c in binding]
    dev_r = [[_t(c) for c in r] for r in rels]
    ts = quickfoil_amd.TableSet(dev_r)
    assert len(ts) == len(rels)
    got, _ = qf_ctx.evaluate_literals(dev_b, npos, ts, lits)
    np.testing.assert_array_equal(got, want)
    for _ in range(2):
        got, _ = qf_ctx.evaluate_literals(dev_b, npos, ts, lits, use_graph=True, reuse_partitions=True)
        np.testing.assert_array_equal(got, want)
    qf_ctx.trim()


def test_autotuned_step_matches_golden(qf_ctx):
    """autotune=1 times its (facts per partition, tile) grid on the first call for a workload shape
    and runs the fastest; the counts of the 1/256 refinement step equal the committed golden ones,
    on the tuning call and on the cached choice."""
    import os
    from quickfoil_amd import datagen
    want = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "step_1_256_counts.npy"))
    w = datagen.refinement_step(scale=1 / 256)
    b = [_t(c) for c in w["binding"]]
    r = [[_t(c) for c in rel] for rel in w["relations"]]
    for _ in range(2):
        got, _ = qf_ctx.evaluate_literals(b, w["num_positive"], r, w["literals"], autotune=True, reuse_partitions=True)
        np.testing.assert_array_equal(got, want)
        choice = qf_ctx.autotune_choice()
        assert choice["partition_rows"] in (128, 256, 512)
        assert choice["tile_rows"] in (512, 768, 1024, 1536, 2048)
        assert choice["best_ms"] > 0
    qf_ctx.trim()


def test_count_range_is_reported_with_exact_counts(qf_ctx):
    """A binding count above int32 (60,000 positive rows x 40,000 facts of one join value: 2.4e9
    pairs) returns QF_ERR_COUNT_RANGE when asked for, and the uint64 counts are exact either way (the
    reference's int32 size_type would wrap there, learner/CandidateLiteralInfo.hpp:21-25)."""
    from quickfoil_amd._native import QfError
    n, m, npos = 70000, 40000, 60000
    binding = [np.zeros(n, np.int32), np.arange(n, dtype=np.int32)]
    rel = [[np.zeros(m, np.int32), np.arange(m, dtype=np.int32) + 5]]
    lits = [(0, [0, -1]), (0, [0, 1])]
    want, _ = oracle.Problem(binding, npos, rel, lits).grouped()
    assert int(want[0][2]) == npos * m > 2 ** 31 and int(want[0][3]) == (n - npos) * m
    got, _ = _gpu_eval(qf_ctx, binding, npos, rel, lits)
    np.testing.assert_array_equal(got, want)
    with pytest.raises(QfError) as e:
        _gpu_eval(qf_ctx, binding, npos, rel, lits, check_range=True)
    assert e.value.status == 5
    qf_ctx.trim()


@pytest.mark.parametrize("mode", MODES)
def test_arg_types_join_key_rule(qf_ctx, mode):
    """Argument types choose each literal's join key (schema/FoilLiteral.hpp:133-144: the bound
    position with the largest type id, first on ties), which regroups the literals; the counts are
    those of the brute-force definition either way."""
    binding, npos, rels, lits = random_problem(31, n=500, rel_shapes=((2, 300), (3, 250)), domain=14)
    arg_types = [[0, 2], [1, 0, 3]]
    prob = oracle.Problem(binding, npos, rels, lits, arg_types=arg_types)
    want, want_bits = prob.bruteforce(semi=True)
    np.testing.assert_array_equal(prob.pipeline()[0], want)
    # The rule picks a later position for some literals, so the grouping differs from the default.
    assert any(oracle.literal_join_key(v, arg_types[r]) != oracle.literal_join_key(v) for r, v in lits)
    got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, arg_types=arg_types, mode=mode, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("npos", [170, 400, 0])
def test_split_positive_and_negative_tables(qf_ctx, mode, npos):
    """Separate positive and negative binding tables (the restart clause's two passes,
    learner/CandidateLiteralEvaluator.cpp:345-405) give the counts of the concatenated table, with
    each table's covered bitvector over its own rows; a table may be empty."""
    binding, _, rels, lits = random_problem(51, n=400, rel_shapes=((2, 300), (3, 200)), domain=12, zipf=True)
    want, want_bits = oracle.Problem(binding, npos, rels, lits).bruteforce(semi=True)
    pos = [_t(c[:npos]) for c in binding]
    neg = [_t(c[npos:]) for c in binding]
    got, bp, bn = qf_ctx.evaluate_literals_split(pos, neg, [[_t(c) for c in r] for r in rels], lits, mode=mode,
                                                 semi_bits=True)
    np.testing.assert_array_equal(got, want)
    rows = np.unpackbits(want_bits.view(np.uint8), axis=1, bitorder="little")[:, :400]
    for part, lo, hi in ((bp, 0, npos), (bn, npos, 400)):
        got_rows = np.unpackbits(part.cpu().numpy().view(np.uint8), axis=1, bitorder="little")[:, : hi - lo]
        np.testing.assert_array_equal(got_rows, rows[:, lo:hi])


@pytest.mark.parametrize("mode", MODES)
def test_wide_relations_vs_bruteforce(qf_ctx, mode):
    # Arity 4-6: keys of up to six columns (the kernel's generic key width), fact columns beyond
    # the three carried in registers between units, and up to 32 key subsets in one unit.
    binding, npos, rels, lits = random_problem(71, n=400, num_cols=3, rel_shapes=((4, 300), (5, 200), (6, 160)),
                                               domain=4, max_literals=700)
    assert max(len(v) for _, v in lits) == 6
    want, want_bits = oracle.Problem(binding, npos, rels, lits).bruteforce(semi=True)
    got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, mode=mode, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


def test_wide_relations_on_hbm_tables(qf_ctx):
    # One radix bit: every partition holds about half of a 2,000-row relation, more than an LDS
    # table takes, so every key subset of widths 1-6 is counted in HBM tables.
    binding, npos, rels, lits = random_problem(72, n=3000, num_cols=3, rel_shapes=((4, 2000), (6, 1800)),
                                               domain=4, max_literals=500)
    want, want_bits = oracle.Problem(binding, npos, rels, lits).grouped(semi=True)
    for mode in MODES:
        got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, mode=mode, semi_bits=True, radix_bits=1)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(bits, want_bits)


def _plan_counts_bruteforce(binding, npos, rel, keys):
    """Counts of literals given as key lists [(relation column, binding column), ...] (a relation
    column may appear twice), by direct comparison of every (binding row, fact) pair."""
    b = [np.asarray(c, dtype=np.int64) for c in binding]
    r = [np.asarray(c, dtype=np.int64) for c in rel]
    out = np.zeros((len(keys), 4), dtype=np.uint64)
    for i, key in enumerate(keys):
        m = np.ones((len(b[0]), len(r[0])), dtype=bool)
        for pc, bc in key:
            m &= b[bc][:, None] == r[pc][None, :]
        cnt = m.sum(axis=1)
        pos = np.arange(len(cnt)) < npos
        out[i] = [np.sum((cnt > 0) & pos), np.sum((cnt > 0) & ~pos), cnt[pos].sum(), cnt[~pos].sum()]
    return out


@pytest.mark.parametrize("mode", MODES)
def test_plan_with_repeated_relation_columns(qf_ctx, mode):
    # A plan may bind one relation column to two binding columns (CountAggregator's atoms are
    # (background column, binding column) pairs, operations/CountAggregator.hpp:25-29); the
    # library keeps one key position and checks the two binding columns for equality in the row.
    binding, npos, rels, _ = random_problem(9, n=3000, num_cols=3, rel_shapes=((3, 2000),), domain=6)
    atoms = [(1, 1), (1, 2), (0, 2)]
    conj = [(0, 1, 3), (0, 2, 4)]
    p = {"relation": 0, "join_column": 0, "binding_column": 0, "atoms": atoms, "atom_literals": [0, 1, 2],
         "conjunctions": conj, "root_literal": 5}
    keys = [[(0, 0), (1, 1)], [(0, 0), (1, 2)], [(0, 0), (0, 2)], [(0, 0), (1, 1), (1, 2)],
            [(0, 0), (1, 1), (0, 2)], [(0, 0)]]
    want = _plan_counts_bruteforce(binding, npos, rels[0], keys)
    assert want[3, 2] > 0 and want[4, 2] > 0  # the equality literals match some rows
    got, _ = qf_ctx.count_aggregate([_t(c) for c in binding], npos, [[_t(c) for c in r] for r in rels], [p], 6,
                                    mode=mode)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("mode", ["batched", "plain_group"])
def test_twelve_binding_columns(qf_ctx, mode):
    # Twelve clause variables: twelve staged columns plus the row id in a work item, up to twelve
    # join variables in one unit, and binding partitions done in batches of six data columns.
    binding, npos, rels, lits = random_problem(73, n=2000, num_cols=12, rel_shapes=((2, 900), (3, 700)), domain=7,
                                               max_literals=400)
    assert len({v for _, vs in lits for v in vs if v >= 0}) == 12
    want, want_bits = oracle.Problem(binding, npos, rels, lits).bruteforce(semi=True)
    got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, mode=mode, semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)


def test_thirteen_binding_columns_are_refused(qf_ctx):
    binding, npos, rels, _ = random_problem(74, n=100, num_cols=13, rel_shapes=((2, 50),), domain=5)
    lits = [(0, [v, (v + 1) % 13]) for v in range(13)]  # every variable referenced
    with pytest.raises(Exception, match="too many binding columns"):
        _gpu_eval(qf_ctx, binding, npos, rels, lits, mode="batched")


@pytest.mark.parametrize("zipf", [False, True])
def test_single_key_units(qf_ctx, zipf):
    # Forty relations of arity 1 and one of arity 2 keyed on its second column: 41 units whose
    # literals key on the join column alone, counted from value runs in the batched form (two
    # launches of at most 32 units), over six join variables (two passes of four), with skewed
    # binding values (cells of many 256-row chunks) and relations holding one value many times.
    rng = np.random.Generator(np.random.PCG64(81 + zipf))
    n, domain = 30000, 300

    def col(size, s=1.3):
        if zipf:
            return (rng.zipf(s, size=size) % domain).astype(np.int32)
        return rng.integers(0, domain, size=size, dtype=np.int64).astype(np.int32)

    binding = [col(n) for _ in range(6)]
    rels = [[col(int(rng.choice([1, 20, 900, 20000])))] for _ in range(40)]
    rels.append([col(5000), col(5000)])
    lits = [(r, [v]) for r in range(40) for v in range(6) if rng.random() < 0.7]
    lits += [(40, [-1, v]) for v in range(6)] + [(3, [2])]  # a repeated literal shares its counts
    npos = 11000
    want, want_bits = oracle.Problem(binding, npos, rels, lits).grouped(semi=True)
    got, _ = _gpu_eval(qf_ctx, binding, npos, rels, lits, mode="batched")
    np.testing.assert_array_equal(got, want)
    # With covered bitvectors the same units go through the general kernel.
    got, bits = _gpu_eval(qf_ctx, binding, npos, rels, lits, mode="batched", semi_bits=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(bits, want_bits)
