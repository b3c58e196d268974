"""Relation files in the reference's pipe-delimited format, parsed by the C ABI's host loader."""
import ctypes

import numpy as np

from ._native import QfError, load_library, STATUS_NAMES


def parse_relation_text(text, num_fields, keep=None):
    """Returns the kept fields of every tuple as int32 numpy columns."""
    lib = load_library()
    data = text.encode() if isinstance(text, str) else bytes(text)
    keep_arr = None
    if keep is not None:
        keep_arr = (ctypes.c_int * num_fields)(*[1 if k else 0 for k in keep])
    rows = ctypes.c_int64(0)
    err = ctypes.c_int64(-1)
    st = lib.qf_parse_relation_text(data, len(data), num_fields, keep_arr, None, 0, ctypes.byref(rows),
                                    ctypes.byref(err))
    if st != 0:
        raise QfError(st, "line %d is malformed" % err.value)
    ncols = num_fields if keep is None else sum(1 for k in keep if k)
    cols = [np.zeros(max(1, rows.value), dtype=np.int32) for _ in range(ncols)]
    ptrs = (ctypes.c_void_p * max(1, ncols))(*[c.ctypes.data for c in cols])
    st = lib.qf_parse_relation_text(data, len(data), num_fields, keep_arr, ptrs, rows.value, ctypes.byref(rows),
                                    ctypes.byref(err))
    if st != 0:
        raise QfError(st, "%s at line %d" % (STATUS_NAMES.get(st, st), err.value))
    return [c[: rows.value] for c in cols]


def load_relation(path, num_fields=None, keep=None):
    with open(path, "rb") as f:
        data = f.read()
    if num_fields is None:
        first = next((l for l in data.splitlines() if l.strip() and not l.startswith(b"#")), b"")
        num_fields = first.count(b"|") + 1
    return parse_relation_text(data, num_fields, keep)
