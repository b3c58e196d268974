"""The C-ABI library loads and exports every function include/quickfoil_amd.h declares."""
import os
import re

import quickfoil_amd
from quickfoil_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "quickfoil_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(qf_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert "qf_evaluate_literals" in names and "qf_radix_partition" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = quickfoil_amd.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_signatures_cover_header():
    assert set(declared_functions()) == set(_native.SIGNATURES)


def test_bucket_layout_host_helper():
    import ctypes
    import numpy as np
    lib = quickfoil_amd.load_library()
    off = np.array([0, 3, 3, 8, 9], dtype=np.int64)
    boff = np.zeros(5, dtype=np.int64)
    total = lib.qf_partition_bucket_layout(off.ctypes.data_as(ctypes.c_void_p), 4,
                                           boff.ctypes.data_as(ctypes.c_void_p))
    # P2(3)=4, empty -> 0, P2(5)=8, P2(1)=1 (storage/FoilHashTable.cpp:34-49).
    assert total == 13
    assert boff.tolist() == [0, 4, 4, 12, 13]


def test_cpp_adapter_driver_compiles_and_links(tmp_path):
    """tests/abi_driver.cpp, a C++ caller of the adapter header include/quickfoil_amd.hpp (the
    reference's operator names over the C ABI), compiles and links against the library."""
    from helpers import build_abi_driver
    path = build_abi_driver(str(tmp_path))
    assert os.path.exists(path)
