"""GPU parity of the reference-layout operators: RadixPartition and the chained hash tables."""
import numpy as np
import pytest

import oracle
from helpers import radix_spec

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


SPEC = [c for c in radix_spec() if len(c["values"]) in (10, 100, 1000)]


@pytest.mark.parametrize("case", SPEC, ids=lambda c: "%s-r%d-n%d" % (c["name"], c["radix_bits"], len(c["values"])))
def test_radix_partition_reference_cases(qf_ctx, case):
    v = np.array(case["values"], dtype=np.int32)
    pairs, offsets = qf_ctx.radix_partition(_t(v), case["radix_bits"])
    pairs = pairs.cpu().numpy().reshape(-1, 2)
    offsets = offsets.cpu().numpy()
    for p in range(1 << case["radix_bits"]):
        got = pairs[offsets[p]:offsets[p + 1]].tolist()
        assert got == sorted(case["partitions"].get(str(p), []), key=lambda t: t[1])


@pytest.mark.parametrize("bits", [1, 3, 5, 7, 8, 9, 12, 16])
@pytest.mark.parametrize("n", [0, 1, 63, 2048, 2049, 100000, 1 << 20])
def test_radix_partition_matches_oracle(qf_ctx, bits, n):
    rng = np.random.Generator(np.random.PCG64(bits * 1000 + n))
    v = rng.integers(-(2 ** 31), 2 ** 31, size=n, dtype=np.int64).astype(np.int32)
    pairs, offsets = qf_ctx.radix_partition(_t(v), bits)
    want_pairs, want_off = oracle.radix_partition(v, bits)
    np.testing.assert_array_equal(offsets.cpu().numpy(), want_off)
    np.testing.assert_array_equal(pairs.cpu().numpy(), want_pairs)


def test_radix_partition_skewed(qf_ctx):
    rng = np.random.Generator(np.random.PCG64(77))
    v = (rng.zipf(1.3, size=300000) % 5000).astype(np.int32)
    for bits in (5, 11):
        pairs, offsets = qf_ctx.radix_partition(_t(v), bits)
        want_pairs, want_off = oracle.radix_partition(v, bits)
        np.testing.assert_array_equal(offsets.cpu().numpy(), want_off)
        np.testing.assert_array_equal(pairs.cpu().numpy(), want_pairs)


@pytest.mark.parametrize("bits", [1, 5, 10])
@pytest.mark.parametrize("n,domain", [(1000, 50), (200000, 1 << 30), (200000, 300)])
def test_build_hash_table_on_partitions_bit_exact(qf_ctx, bits, n, domain):
    rng = np.random.Generator(np.random.PCG64(n + bits))
    v = rng.integers(0, domain, size=n, dtype=np.int64).astype(np.int32)
    pairs, offsets = qf_ctx.radix_partition(_t(v), bits)
    buckets, boff, nxt = qf_ctx.build_hash_table_on_partitions(pairs, offsets, bits)
    opairs, ooff = oracle.radix_partition(v, bits)
    wb, wboff, wn = oracle.build_hash_table_on_partitions(opairs, ooff, bits)
    np.testing.assert_array_equal(boff, wboff)
    np.testing.assert_array_equal(buckets.cpu().numpy(), wb)
    np.testing.assert_array_equal(nxt.cpu().numpy(), wn)


@pytest.mark.parametrize("nkeys", [1, 2, 3, 6])
@pytest.mark.parametrize("n", [1, 17, 5000, 300000])
def test_build_hash_table_on_table_bit_exact(qf_ctx, nkeys, n):
    rng = np.random.Generator(np.random.PCG64(nkeys * 7 + n))
    cols = [rng.integers(-40, 40, size=n, dtype=np.int64).astype(np.int32) for _ in range(nkeys)]
    buckets, nxt = qf_ctx.build_hash_table_on_table([_t(c) for c in cols])
    wb, wn = oracle.build_hash_table_on_table(cols)
    np.testing.assert_array_equal(buckets.cpu().numpy(), wb)
    np.testing.assert_array_equal(nxt.cpu().numpy(), wn)
