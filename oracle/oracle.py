"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.  The
product (quickfoil_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAX_ARITY = 6


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.orc_hash_int.restype = ctypes.c_uint32
        _lib.orc_hash_int.argtypes = [ctypes.c_int32]
        _lib.orc_hash_combine.restype = ctypes.c_uint32
        _lib.orc_hash_combine.argtypes = [ctypes.c_uint32, ctypes.c_int32]
        _lib.orc_next_pow2.restype = ctypes.c_uint32
        _lib.orc_next_pow2.argtypes = [ctypes.c_uint32]
        _lib.orc_partition_bucket_offsets.restype = ctypes.c_int64
        _lib.orc_evaluate_pipeline.restype = ctypes.c_int
        _lib.orc_create_label_aware_binding_table.restype = ctypes.c_int64
        _lib.orc_create_binding_table.restype = ctypes.c_int64
        _lib.orc_generate_plan.restype = ctypes.c_int
        _lib.orc_literal_join_key.restype = ctypes.c_int
        _lib.orc_relation_cache_create.restype = ctypes.c_void_p
        _lib.orc_relation_cache_create.argtypes = []
        _lib.orc_relation_cache_destroy.restype = None
        _lib.orc_relation_cache_destroy.argtypes = [ctypes.c_void_p]
        _lib.orc_relation_cache_fill.restype = None
        _lib.orc_evaluate_pipeline_cached.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def hash_int(v):
    return lib().orc_hash_int(int(v))


def hash_combine(seed, v):
    return lib().orc_hash_combine(seed, int(v))


def radix_partition(values, radix_bits):
    v = _i32(values)
    n = v.size
    pairs = np.zeros(2 * max(1, n), dtype=np.int32)
    offsets = np.zeros((1 << radix_bits) + 1, dtype=np.int64)
    lib().orc_radix_partition(_p(v), ctypes.c_int64(n), radix_bits, _p(pairs), _p(offsets))
    return pairs[: 2 * n], offsets


def build_hash_table_on_partitions(pairs, offsets, radix_bits):
    pairs = _i32(pairs)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nparts = 1 << radix_bits
    boff = np.zeros(nparts + 1, dtype=np.int64)
    total = lib().orc_partition_bucket_offsets(_p(offsets), nparts, _p(boff))
    buckets = np.zeros(max(1, total), dtype=np.int32)
    nxt = np.zeros(max(1, pairs.size // 2), dtype=np.int32)
    lib().orc_build_hash_table_on_partitions(_p(pairs), _p(offsets), radix_bits, _p(boff), _p(buckets), _p(nxt))
    return buckets[:total], boff, nxt[: pairs.size // 2]


class _Cols:
    def __init__(self, cols):
        self.arrs = [_i32(c) for c in cols]
        self.ptrs = (ctypes.c_void_p * max(1, len(self.arrs)))(*[a.ctypes.data for a in self.arrs])


def build_hash_table_on_table(key_columns):
    kc = _Cols(key_columns)
    n = kc.arrs[0].size
    nb = lib().orc_next_pow2(n) if n > 0 else 0
    buckets = np.zeros(max(1, nb), dtype=np.int32)
    nxt = np.zeros(max(1, n), dtype=np.int32)
    lib().orc_build_hash_table_on_table(kc.ptrs, len(kc.arrs), ctypes.c_int64(n), _p(buckets), _p(nxt))
    return buckets[:nb], nxt[:n]


class EvalInput(ctypes.Structure):
    _fields_ = [
        ("binding_columns", ctypes.c_void_p),
        ("num_binding_columns", ctypes.c_int),
        ("num_bindings", ctypes.c_int64),
        ("num_positive", ctypes.c_int64),
        ("rel_columns", ctypes.c_void_p),
        ("rel_col_start", ctypes.c_void_p),
        ("rel_arity", ctypes.c_void_p),
        ("rel_rows", ctypes.c_void_p),
        ("arg_types", ctypes.c_void_p),
        ("num_relations", ctypes.c_int),
        ("lit_rel", ctypes.c_void_p),
        ("lit_vars", ctypes.c_void_p),
        ("num_literals", ctypes.c_int),
    ]


class Problem:
    """Binding table + relations + literals in the oracle's flat input form."""

    def __init__(self, binding_columns, num_positive, relations, literals, arg_types=None):
        self.b = _Cols(binding_columns)
        self.n = self.b.arrs[0].size if self.b.arrs else 0
        self.num_positive = int(num_positive)
        flat = [c for r in relations for c in r]
        self.r = _Cols(flat)
        self.col_start = np.array(np.cumsum([0] + [len(r) for r in relations])[:-1], dtype=np.int32)
        self.arity = np.array([len(r) for r in relations], dtype=np.int32)
        self.rows = np.array([len(r[0]) for r in relations], dtype=np.int64)
        self.types = None if arg_types is None else np.ascontiguousarray(
            np.concatenate([np.asarray(t, dtype=np.int32) for t in arg_types]))
        self.lit_rel = np.array([l[0] for l in literals], dtype=np.int32)
        lv = np.full((max(1, len(literals)), MAX_ARITY), -1, dtype=np.int32)
        for i, (_, vars_) in enumerate(literals):
            lv[i, : len(vars_)] = vars_
        self.lit_vars = lv
        self.L = len(literals)
        self.inp = EvalInput(ctypes.cast(self.b.ptrs, ctypes.c_void_p).value, len(self.b.arrs), self.n,
                             self.num_positive, ctypes.cast(self.r.ptrs, ctypes.c_void_p).value,
                             self.col_start.ctypes.data, self.arity.ctypes.data, self.rows.ctypes.data,
                             self.types.ctypes.data if self.types is not None else None, len(relations),
                             self.lit_rel.ctypes.data, self.lit_vars.ctypes.data, self.L)

    def _out(self, semi):
        counts = np.zeros((max(1, self.L), 4), dtype=np.uint64)
        bits = np.zeros((max(1, self.L), (self.n + 63) // 64), dtype=np.uint64) if semi else None
        return counts, bits

    def pipeline(self, radix_bits=5, chunk_size=32768, semi=False):
        counts, bits = self._out(semi)
        st = lib().orc_evaluate_pipeline(ctypes.byref(self.inp), radix_bits, chunk_size, _p(counts),
                                         _p(bits) if semi else None)
        return counts[: self.L], (bits[: self.L] if semi else None), st

    def pipeline_cached(self, cache, radix_bits=5, chunk_size=32768, semi=False):
        """The pipeline over a RelationCache's background partitions (made outside this call)."""
        counts, bits = self._out(semi)
        st = lib().orc_evaluate_pipeline_cached(ctypes.byref(self.inp), ctypes.c_void_p(cache.h), radix_bits,
                                                chunk_size, _p(counts), _p(bits) if semi else None)
        return counts[: self.L], (bits[: self.L] if semi else None), st

    def bruteforce(self, semi=False):
        counts, bits = self._out(semi)
        lib().orc_evaluate_bruteforce(ctypes.byref(self.inp), _p(counts), _p(bits) if semi else None)
        return counts[: self.L], (bits[: self.L] if semi else None)

    def grouped(self, semi=False):
        counts, bits = self._out(semi)
        lib().orc_evaluate_grouped(ctypes.byref(self.inp), _p(counts), _p(bits) if semi else None)
        return counts[: self.L], (bits[: self.L] if semi else None)

    def label_aware_binding_table(self, literal, num_binding_pos, num_binding_neg):
        ncols = len(self.b.arrs) + int(np.sum(self.lit_vars[literal][: self.arity[self.lit_rel[literal]]] < 0))
        rows = int(num_binding_pos + num_binding_neg)
        outs = [np.zeros(max(1, rows), dtype=np.int32) for _ in range(ncols)]
        ptrs = (ctypes.c_void_p * ncols)(*[o.ctypes.data for o in outs])
        got = lib().orc_create_label_aware_binding_table(ctypes.byref(self.inp), literal,
                                                         ctypes.c_int64(num_binding_pos),
                                                         ctypes.c_int64(num_binding_neg), ptrs)
        return got, [o[:rows] for o in outs]

    def binding_table(self, literal):
        rows = lib().orc_create_binding_table(ctypes.byref(self.inp), literal, None)
        if rows < 0:
            raise ValueError("literal has no bound variable")
        arity = self.arity[self.lit_rel[literal]]
        vars_ = self.lit_vars[literal][:arity]
        ncols = len(self.b.arrs) + int(np.sum((vars_ < 0) | (vars_ >= len(self.b.arrs))))
        outs = [np.zeros(max(1, rows), dtype=np.int32) for _ in range(ncols)]
        ptrs = (ctypes.c_void_p * ncols)(*[o.ctypes.data for o in outs])
        lib().orc_create_binding_table(ctypes.byref(self.inp), literal, ptrs)
        return [o[:rows] for o in outs]


def literal_join_key(vars_, arg_types=None):
    v = np.full(MAX_ARITY, -1, dtype=np.int32)
    v[: len(vars_)] = vars_
    t = None if arg_types is None else np.ascontiguousarray(arg_types, dtype=np.int32)
    return lib().orc_literal_join_key(len(vars_), _p(v), _p(t) if t is not None else None)


def generate_plan(arity, join_key, lit_vars, max_atoms=64, max_nodes=256):
    n = len(lit_vars)
    lv = np.full((max(1, n), MAX_ARITY), -1, dtype=np.int32)
    for i, v in enumerate(lit_vars):
        lv[i, : len(v)] = v
    na = ctypes.c_int(0)
    root = ctypes.c_int(-1)
    abg = np.zeros(max_atoms, dtype=np.int32)
    abd = np.zeros(max_atoms, dtype=np.int32)
    nl = np.zeros(max_nodes, dtype=np.int32)
    nr = np.zeros(max_nodes, dtype=np.int32)
    nlit = np.zeros(max_nodes, dtype=np.int32)
    nn = lib().orc_generate_plan(arity, join_key, _p(lv), n, max_atoms, max_nodes, ctypes.byref(na), _p(abg),
                                 _p(abd), ctypes.byref(root), _p(nl), _p(nr), _p(nlit))
    if nn < 0:
        raise RuntimeError("plan buffers too small")
    a = na.value
    return {
        "atoms": list(zip(abg[:a].tolist(), abd[:a].tolist())),
        "root_literal": root.value,
        "left": nl[:nn].tolist(),
        "right": nr[:nn].tolist(),
        "literal": nlit[:nn].tolist(),
    }


def _bits(a, n):
    return np.ascontiguousarray(a, dtype=np.uint64)


def left_semi_join(probe_keys, build_keys):
    p, b = _Cols(probe_keys), _Cols(build_keys)
    n = p.arrs[0].size
    out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
    lib().orc_left_semi_join(p.ptrs, ctypes.c_int64(n), b.ptrs, ctypes.c_int64(b.arrs[0].size), len(p.arrs), _p(out))
    return out[: (n + 63) // 64]


def right_semi_join(probe_keys, build_keys):
    p, b = _Cols(probe_keys), _Cols(build_keys)
    n = b.arrs[0].size
    out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
    lib().orc_right_semi_join(p.ptrs, ctypes.c_int64(p.arrs[0].size), b.ptrs, ctypes.c_int64(n), len(p.arrs), _p(out))
    return out[: (n + 63) // 64]


def dedup_after_semi_join(key_columns, bits):
    k = _Cols(key_columns)
    n = k.arrs[0].size
    b = np.ascontiguousarray(bits, dtype=np.uint64)
    out = np.zeros(max(1, n), dtype=np.uint8)
    lib().orc_dedup_after_semi_join(k.ptrs, len(k.arrs), ctypes.c_int64(n), _p(b), _p(out))
    return out[:n]


class RelationCache:
    """Background partitions of every (relation, join key) a problem's literals use, made once
    (learner/CandidateLiteralEvaluator.cpp:291-299 keeps them on the fact tables); valid for
    problems over the same relation arrays."""

    def __init__(self, problem, radix_bits=5):
        self.h = lib().orc_relation_cache_create()
        self.radix_bits = radix_bits
        lib().orc_relation_cache_fill(ctypes.c_void_p(self.h), ctypes.byref(problem.inp), radix_bits)

    def close(self):
        if self.h:
            lib().orc_relation_cache_destroy(ctypes.c_void_p(self.h))
            self.h = None

    def __del__(self):
        self.close()
