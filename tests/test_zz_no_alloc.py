"""qf_radix_partition_noalloc (temporaries in the library's fixed device array, no allocation)
gives exactly qf_radix_partition's pairs and offsets, through the C ABI."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(ctx, fn, values, bits):
    import torch
    n = values.numel()
    pairs = torch.full((2 * max(1, n),), -7, dtype=torch.int32, device=values.device)
    offsets = torch.full(((1 << bits) + 1,), -7, dtype=torch.int64, device=values.device)
    st = fn(ctx.h, ctypes.c_void_p(values.data_ptr()), ctypes.c_int64(n), bits, ctypes.c_void_p(pairs.data_ptr()),
            ctypes.c_void_p(offsets.data_ptr()))
    return st, pairs[: 2 * n].cpu().numpy(), offsets.cpu().numpy()


def test_noalloc_partition_equals_allocating_entry_point(qf_ctx):
    import torch
    lib = qf_ctx.lib
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    # Two calls in a row, the second smaller (the array keeps the first call's contents), then a
    # one-pass digit; sizes up to the maximum.
    for n, bits in [(4_194_304, 16), (300_001, 13), (70_000, 5), (0, 3)]:
        v = torch.randint(-(2 ** 31), 2 ** 31 - 1, (max(1, n),), generator=g, device="cuda", dtype=torch.int64)
        v = v.to(torch.int32)[:n]
        st_a, pairs_a, off_a = _run(qf_ctx, lib.qf_radix_partition, v, bits)
        st_b, pairs_b, off_b = _run(qf_ctx, lib.qf_radix_partition_noalloc, v, bits)
        assert st_a == 0 and st_b == 0, (n, bits, st_a, st_b)
        np.testing.assert_array_equal(pairs_b, pairs_a)
        np.testing.assert_array_equal(off_b, off_a)
        # Known answer of the partition itself (operations/RadixPartition_test.cpp:106-168):
        # partition p holds the values whose low bits are p, in row order.
        if n:
            p = pairs_b.reshape(-1, 2)
            vals = v.cpu().numpy()
            np.testing.assert_array_equal(vals[p[:, 1]], p[:, 0])
            digit = (p[:, 0].astype(np.int64) & 0xFFFFFFFF) & ((1 << bits) - 1)
            assert np.all(np.diff(digit) >= 0)
            assert off_b[-1] == n


def test_noalloc_partition_rejects_more_rows_than_its_array_holds(qf_ctx):
    import torch
    n = 4_194_305
    v = torch.zeros(n, dtype=torch.int32, device="cuda")
    st, _, _ = _run(qf_ctx, qf_ctx.lib.qf_radix_partition_noalloc, v, 8)
    assert st == 1  # QF_ERR_INVALID, before any device work
    assert "QF_NOALLOC_MAX_ROWS" in qf_ctx.lib.qf_last_error(qf_ctx.h).decode()
