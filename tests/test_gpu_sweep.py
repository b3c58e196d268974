"""Seeded sweep over random problem shapes: binding columns, empty and one-row tables, relation
arities 1-4, tiny and large value domains, skew, radix bits, tiles, all three evaluation forms
and covered bitvectors, each checked bit-exactly against the oracle's grouped closed form (which
the CPU tests pin to brute force and to the restated pipeline)."""
import itertools

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = 200


def sweep_case(seed):
    rng = np.random.Generator(np.random.PCG64(0xC0FFEE + seed))
    nb = int(rng.choice([0, 1, 2, 63, 64, 65, 700, 3000]))
    ncols = int(rng.integers(1, 7))
    domain = int(rng.choice([1, 2, 5, 40, 1000]))
    zipf = bool(rng.integers(0, 2))

    def col(size):
        if zipf:
            return ((rng.zipf(1.5, size=size) - 1) % domain).astype(np.int32)
        return rng.integers(0, domain, size=size, dtype=np.int64).astype(np.int32)

    binding = [col(nb) for _ in range(ncols)]
    shapes = [(int(rng.integers(1, 5)), int(rng.choice([0, 1, 5, 300, 2500]))) for _ in range(int(rng.integers(1, 4)))]
    rels = [[col(rows) for _ in range(arity)] for arity, rows in shapes]
    lits = [(r, list(v)) for r, (arity, _) in enumerate(shapes)
            for v in itertools.product(range(-1, ncols), repeat=arity) if any(x >= 0 for x in v)]
    if len(lits) > 60:
        lits = [lits[i] for i in sorted(rng.choice(len(lits), size=60, replace=False))]
    opts = {"mode": str(rng.choice(["batched", "plain_group", "plain_literal"])),
            "radix_bits": int(rng.choice([0, 1, 4, 9])), "tile_rows": int(rng.choice([0, 64, 320, 768]))}
    npos = int(rng.integers(0, nb + 1))
    return binding, npos, rels, lits, opts, bool(rng.integers(0, 2))


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


@pytest.mark.parametrize("seed", range(CASES))
def test_sweep_matches_oracle(qf_ctx, seed):
    binding, npos, rels, lits, opts, semi = sweep_case(seed)
    print("sweep case", seed, "rows", len(binding[0]), "cols", len(binding), "relations",
          [(len(r), len(r[0])) for r in rels], "literals", len(lits), opts, "semi", semi, flush=True)
    want, want_bits = oracle.Problem(binding, npos, rels, lits).grouped(semi=semi)
    got, bits = qf_ctx.evaluate_literals([_t(c) for c in binding], npos, [[_t(c) for c in r] for r in rels], lits,
                                         semi_bits=semi, **opts)
    np.testing.assert_array_equal(got, want)
    if semi:
        np.testing.assert_array_equal(bits.cpu().numpy().view(np.uint64), want_bits)
