"""The relation-file loader (main/main.cpp:49-109 format) on CPU."""
import numpy as np
import pytest

from helpers import load_relation as py_load
from quickfoil_amd import QfError
from quickfoil_amd.loader import load_relation, parse_relation_text
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uncle")


@pytest.mark.parametrize("name", ["uncle", "brother", "parent", "sister"])
def test_sample_files(name):
    got = load_relation(os.path.join(GOLDEN, name))
    want = py_load(os.path.join(GOLDEN, name))
    assert len(got) == len(want)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


def test_rules():
    text = "# comment\n1|2\n\n -3|+4\n2147483647|-2147483648\r\n5|6"
    cols = parse_relation_text(text, 2)
    assert cols[0].tolist() == [1, -3, 2147483647, 5]
    assert cols[1].tolist() == [2, 4, -2147483648, 6]
    cols = parse_relation_text(text, 2, keep=[False, True])
    assert len(cols) == 1 and cols[0].tolist() == [2, 4, -2147483648, 6]


@pytest.mark.parametrize("bad", ["1|2|3\n", "1\n", "1|x\n", "2147483648|1\n"])
def test_malformed(bad):
    with pytest.raises(QfError):
        parse_relation_text("7|8\n" + bad, 2)


def test_exact_capacity_is_checked():
    import ctypes
    from quickfoil_amd._native import load_library
    lib = load_library()
    data = b"1|2\n3|4\n"
    a = np.zeros(1, dtype=np.int32)
    b = np.zeros(1, dtype=np.int32)
    ptrs = (ctypes.c_void_p * 2)(a.ctypes.data, b.ctypes.data)
    rows = ctypes.c_int64(0)
    st = lib.qf_parse_relation_text(data, len(data), 2, None, ptrs, 1, ctypes.byref(rows), None)
    assert st == 4 and rows.value == 1
