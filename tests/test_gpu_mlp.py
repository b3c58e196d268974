"""The estimator MLP on MFMA against its CPU oracle (forward and gradients), its training on
joins the engine records, and presized joins with exact fallback."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle  # noqa: E402
import oracle  # noqa: E402
from quickfoil_amd import estimator  # noqa: E402

pytestmark = pytest.mark.gpu


def _random(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.normal(0, 1, (n, 40)).astype(np.float32), rng.normal(3, 2, n).astype(np.float32)


@pytest.mark.parametrize("n", [1, 64, 300])
def test_forward_matches_oracle(qf_ctx, n):
    X, _ = _random(n, n)
    est = estimator.Estimator(qf_ctx, estimator.init_params(seed=5))
    try:
        got = est.predict_raw(X)
        want, _ = mlp_oracle.forward(est.params(), X)
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4 * (1 + np.abs(want).max()))
    finally:
        est.close()


@pytest.mark.parametrize("n", [17, 128, 300])
def test_gradients_match_oracle(qf_ctx, n):
    X, y = _random(n, 100 + n)
    est = estimator.Estimator(qf_ctx, estimator.init_params(seed=6))
    try:
        g, loss = est.gradients(X, y)
        want, want_loss = mlp_oracle.gradients(est.params(), X, y)
        assert abs(loss - want_loss) <= 1e-4 * want_loss
        blocks = [(mlp_oracle.OFF_W1, mlp_oracle.OFF_B1), (mlp_oracle.OFF_B1, mlp_oracle.OFF_W2),
                  (mlp_oracle.OFF_W2, mlp_oracle.OFF_B2), (mlp_oracle.OFF_B2, mlp_oracle.OFF_W3),
                  (mlp_oracle.OFF_W3, mlp_oracle.OFF_B3), (mlp_oracle.OFF_B3, mlp_oracle.NUM_PARAMS)]
        for a, b in blocks:
            err = np.linalg.norm(g[a:b] - want[a:b])
            assert err <= 2e-3 * np.linalg.norm(want[a:b]) + 1e-7, (a, b, err)
    finally:
        est.close()


@pytest.fixture(scope="module")
def recorded(qf_ctx):
    train = [estimator.record_step(qf_ctx, 1.0 / 128, seed)[:2] for seed in (11, 12, 13, 14)]
    X = np.concatenate([t[0] for t in train])
    y = np.concatenate([t[1] for t in train])
    Xh, yh, w = estimator.record_step(qf_ctx, 1.0 / 128, 15)
    return X, y, Xh, yh, w


def test_training_beats_independence_estimate(qf_ctx, recorded):
    X, y, Xh, yh, _ = recorded
    est = estimator.Estimator(qf_ctx)
    try:
        hist = est.fit(X, y, epochs=40, batch=512, lr=3e-3)
        assert hist[-1] < 0.5 * hist[0]
        q_mlp = estimator.q_error(est.predict(Xh), yh)
        q_ind = estimator.q_error(estimator.independence_estimate(Xh), yh)
        assert np.median(q_mlp) < np.median(q_ind)
        assert np.percentile(q_mlp, 90) < np.percentile(q_ind, 90)
    finally:
        est.close()


def test_presized_join_is_exact(qf_ctx, recorded):
    import torch
    X, y, Xh, yh, w = recorded
    est = estimator.Estimator(qf_ctx)
    try:
        est.fit(X, y, epochs=20, batch=512, lr=3e-3)
        pred = est.predict(Xh)
        dev = torch.device("cuda")
        binding = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in w["binding"]]
        fallbacks = 0
        picks = list(range(0, len(w["literals"]), 97))[:20]
        for i in picks:
            rel, vars_ = w["literals"][i]
            relation = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in w["relations"][rel]]
            want = qf_ctx.create_binding_table(binding, relation, vars_)
            got, fell_back = qf_ctx.create_binding_table_presized(binding, relation, vars_,
                                                                  estimator.presize_rows(pred[i]))
            fallbacks += int(fell_back)
            assert len(got) == len(want)
            for g, e in zip(got, want):
                assert torch.equal(g, e)
        # A capacity of one row forces the exact fallback.
        rel, vars_ = w["literals"][picks[0]]
        relation = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in w["relations"][rel]]
        want = qf_ctx.create_binding_table(binding, relation, vars_)
        got, fell_back = qf_ctx.create_binding_table_presized(binding, relation, vars_, 1)
        assert fell_back == (want[0].numel() > 1)
        for g, e in zip(got, want):
            assert torch.equal(g, e)
    finally:
        est.close()


def test_batched_presized_joins_and_forced_fallback(qf_ctx, recorded):
    """The estimator in the loop: CreateBindingTable for many literals in one call, capacities from
    the network's predictions, largest prediction first; joins whose rows exceed their capacity
    write nothing and are rerun at their exact size.  With capacities forced far below the
    predictions (under-prediction) every non-empty join takes the fallback: the overflow flag is
    raised, the retry writes, and every output equals the two-pass join and the oracle's rows."""
    import torch
    X, y, Xh, yh, w = recorded
    est = estimator.Estimator(qf_ctx)
    try:
        est.fit(X, y, epochs=20, batch=512, lr=3e-3)
        dev = torch.device("cuda")
        binding = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in w["binding"]]
        rels = [[torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in r] for r in w["relations"]]
        picks = [w["literals"][i] for i in range(0, len(w["literals"]), 53)][:36]
        for margin in (1.0, -40.0):
            got, stats = estimator.presized_joins(qf_ctx, est, binding, rels, picks, margin=margin)
            nonempty = int(sum(1 for r in stats["rows"] if r > 0))
            if margin < 0:
                assert stats["fallbacks"] == sum(1 for r in stats["rows"] if r > 16)
                assert stats["fallbacks"] > 0
            assert nonempty > 0
            for (rel, vars_), g, rows in zip(picks, got, stats["rows"]):
                want = qf_ctx.create_binding_table(binding, rels[rel], vars_)
                assert want[0].numel() == rows
                for a, b in zip(g, want):
                    assert torch.equal(a, b)
        # Against the oracle's CreateBindingTable as a multiset, on a few literals.
        sub = picks[:6]
        prob = oracle.Problem(w["binding"], w["num_positive"], w["relations"], sub)
        outs, rows, fits = qf_ctx.create_binding_tables_presized(binding, rels, sub, [1 << 22] * len(sub))
        assert fits.all()
        for i in range(len(sub)):
            want = prob.binding_table(i)
            assert int(rows[i]) == len(want[0])
            got_rows = sorted(zip(*[c.cpu().numpy().tolist() for c in outs[i]]))
            assert got_rows == sorted(zip(*[c.tolist() for c in want]))
    finally:
        est.close()
