"""The estimator's CPU oracle: exact-arithmetic gradients agree with central finite differences
(which fixes the backward formulas), and BF16 rounding changes them only slightly."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle  # noqa: E402


def _problem(seed=3, n=64):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(mlp_oracle.NUM_PARAMS, np.float64)
    p[:] = rng.normal(0, 0.2, mlp_oracle.NUM_PARAMS)
    W1 = p[mlp_oracle.OFF_W1:mlp_oracle.OFF_B1].reshape(64, 48)
    W1[:, 40:] = 0
    return p, rng.normal(0, 1, (n, 40)), rng.normal(0, 1, n)


def test_exact_gradients_match_finite_differences():
    p, X, y = _problem()
    g, _ = mlp_oracle.gradients(p.astype(np.float32).astype(np.float64), X, y, rounding=False)
    rng = np.random.Generator(np.random.PCG64(7))
    idx = list(rng.choice(np.arange(mlp_oracle.OFF_W1, mlp_oracle.OFF_W1 + 64 * 40), 8, replace=False))
    idx += list(rng.choice(np.arange(mlp_oracle.OFF_W2, mlp_oracle.OFF_B2), 8, replace=False))
    idx += [mlp_oracle.OFF_B1 + 5, mlp_oracle.OFF_B2 + 9, mlp_oracle.OFF_W3 + 11, mlp_oracle.OFF_B3]
    eps = 1e-4
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = mlp_oracle.gradients(pp, X, y, rounding=False)[1]
        lm = mlp_oracle.gradients(pm, X, y, rounding=False)[1]
        assert abs(g[i] - (lp - lm) / (2 * eps)) < 1e-5 + 1e-4 * abs(g[i]), i


def test_padding_columns_have_no_gradient():
    p, X, y = _problem()
    g, _ = mlp_oracle.gradients(p, X, y)
    assert not np.any(g[mlp_oracle.OFF_W1:mlp_oracle.OFF_B1].reshape(64, 48)[:, 40:])


def test_bf16_rounding_is_close_to_exact():
    p, X, y = _problem()
    ge, le = mlp_oracle.gradients(p, X, y, rounding=False)
    gr, lr = mlp_oracle.gradients(p, X, y, rounding=True)
    assert abs(le - lr) < 0.02 * le
    assert np.linalg.norm(ge - gr) < 0.05 * np.linalg.norm(ge)


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-3], np.float32)
    got = mlp_oracle.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0).
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125 and got[3] == -2.5
    assert abs(got[4] - 3.0e-3) < 3.0e-3 * 2 ** -8
