"""ctypes binding of libquickfoil_amd.so (the C ABI declared in include/quickfoil_amd.h).

Device memory is passed as torch tensors on the GPU; torch is used only to own device memory and
to order work on its current stream.  Every call fails loudly if the HIP library is missing:
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libquickfoil_amd.so")
MAX_ARITY = 6

QF_OK = 0
STATUS_NAMES = {
    0: "QF_OK",
    1: "QF_ERR_INVALID",
    2: "QF_ERR_DEVICE",
    3: "QF_ERR_OOM",
    4: "QF_ERR_OVERFLOW",
    5: "QF_ERR_COUNT_RANGE",
    6: "QF_ERR_UNSUPPORTED",
}

MODE_BATCHED = 0
MODE_PLAIN_GROUP = 1
MODE_PLAIN_LITERAL = 2
MODES = {"batched": MODE_BATCHED, "plain_group": MODE_PLAIN_GROUP, "plain_literal": MODE_PLAIN_LITERAL}


class QfError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


class QfTable(ctypes.Structure):
    _fields_ = [
        ("columns", ctypes.POINTER(ctypes.c_void_p)),
        ("num_columns", ctypes.c_int),
        ("num_rows", ctypes.c_int64),
    ]


class QfLiteral(ctypes.Structure):
    _fields_ = [("relation", ctypes.c_int), ("vars", ctypes.c_int * MAX_ARITY)]


class LiteralArray:
    """Literals converted once to the C layout (qf_literal: relation, then MAX_ARITY variables),
    for callers that evaluate the same literal list repeatedly."""

    def __init__(self, literals):
        rows = [[rel] + list(vars_) + [-1] * (MAX_ARITY - len(vars_)) for rel, vars_ in literals]
        self.array = np.ascontiguousarray(np.array(rows, dtype=np.int32).reshape(-1, 1 + MAX_ARITY))
        self.count = len(literals)

    def __len__(self):
        return self.count

    def pointer(self):
        return self.array.ctypes.data_as(ctypes.POINTER(QfLiteral))


class QfEvalOptions(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("radix_bits", ctypes.c_int),
        ("tile_rows", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("reuse_partitions", ctypes.c_int),
        ("partition_rows", ctypes.c_int),
        ("autotune", ctypes.c_int),
    ]


class QfJoinGroupPlan(ctypes.Structure):
    _fields_ = [
        ("relation", ctypes.c_int),
        ("join_column", ctypes.c_int),
        ("binding_column", ctypes.c_int),
        ("num_atoms", ctypes.c_int),
        ("atom_bg_col", ctypes.POINTER(ctypes.c_int)),
        ("atom_binding_col", ctypes.POINTER(ctypes.c_int)),
        ("num_nodes", ctypes.c_int),
        ("node_left", ctypes.POINTER(ctypes.c_int)),
        ("node_right", ctypes.POINTER(ctypes.c_int)),
        ("node_literal", ctypes.POINTER(ctypes.c_int)),
        ("root_literal", ctypes.c_int),
    ]


# Every symbol the header declares, with (restype, argtypes).
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
SIGNATURES = {
    "qf_context_create": (_I, [_I, ctypes.POINTER(_P)]),
    "qf_context_destroy": (None, [_P]),
    "qf_context_set_stream": (_I, [_P, _P]),
    "qf_synchronize": (_I, [_P]),
    "qf_last_error": (ctypes.c_char_p, [_P]),
    "qf_context_trim": (_I, [_P]),
    "qf_context_enable_timing": (_I, [_P, _I]),
    "qf_context_reset_timing": (_I, [_P]),
    "qf_context_phase_times": (_I, [_P, _P, _P, _P, _I]),
    "qf_radix_partition": (_I, [_P, _P, _I64, _I, _P, _P]),
    "qf_radix_partition_noalloc": (_I, [_P, _P, _I64, _I, _P, _P]),
    "qf_partition_bucket_layout": (_I64, [_P, _I, _P]),
    "qf_build_hash_table_on_partitions": (_I, [_P, _P, _P, _I, _P, _I64, _I64, _P, _P]),
    "qf_build_hash_table_on_table": (_I, [_P, ctypes.POINTER(QfTable), _P, _P]),
    "qf_eval_options_default": (None, [ctypes.POINTER(QfEvalOptions)]),
    "qf_autotune_choice": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)]),
    "qf_evaluate_literals": (_I, [_P, ctypes.POINTER(QfTable), _I64, ctypes.POINTER(QfTable), _P, _I,
                                  ctypes.POINTER(QfLiteral), _I, ctypes.POINTER(QfEvalOptions), _P, _P]),
    "qf_evaluate_literals_split": (_I, [_P, ctypes.POINTER(QfTable), ctypes.POINTER(QfTable), ctypes.POINTER(QfTable),
                                        _P, _I, ctypes.POINTER(QfLiteral), _I, ctypes.POINTER(QfEvalOptions), _P, _P,
                                        _P]),
    "qf_count_aggregate": (_I, [_P, ctypes.POINTER(QfTable), _I64, ctypes.POINTER(QfTable), _I,
                                ctypes.POINTER(QfJoinGroupPlan), _I, _I, ctypes.POINTER(QfEvalOptions), _P, _P]),
    "qf_create_label_aware_binding_table": (_I, [_P, ctypes.POINTER(QfTable), _I64, ctypes.POINTER(QfTable),
                                                 ctypes.POINTER(QfLiteral), _I64, _I64, _P]),
    "qf_create_binding_table": (_I, [_P, ctypes.POINTER(QfTable), ctypes.POINTER(QfTable),
                                     ctypes.POINTER(QfLiteral), _P, ctypes.POINTER(_I64)]),
    "qf_create_binding_tables_presized": (_I, [_P, ctypes.POINTER(QfTable), ctypes.POINTER(QfTable), _I,
                                               ctypes.POINTER(QfLiteral), _I, _P, _P, _P, _P]),
    "qf_left_semi_join": (_I, [_P, ctypes.POINTER(QfTable), ctypes.POINTER(QfTable), _P, ctypes.POINTER(_I64)]),
    "qf_right_semi_join": (_I, [_P, ctypes.POINTER(QfTable), ctypes.POINTER(QfTable), _P, ctypes.POINTER(_I64)]),
    "qf_build_hash_table_after_semi_join": (_I, [_P, ctypes.POINTER(QfTable), _P, _P, _P, _P]),
    "qf_example_coverage": (_I, [_P, ctypes.POINTER(QfTable), _I64, _I64, ctypes.POINTER(QfTable),
                                 ctypes.POINTER(QfLiteral), ctypes.POINTER(QfTable), _P, ctypes.POINTER(_I64)]),
    "qf_parse_relation_text": (_I, [ctypes.c_char_p, _I64, _I, _P, _P, _I64, ctypes.POINTER(_I64),
                                    ctypes.POINTER(_I64)]),
    "qf_create_binding_table_presized": (_I, [_P, ctypes.POINTER(QfTable), ctypes.POINTER(QfTable),
                                              ctypes.POINTER(QfLiteral), _I64, _P, ctypes.POINTER(_I64)]),
    "qf_mlp_num_params": (_I, []),
    "qf_mlp_create": (_I, [_P, _P, ctypes.POINTER(_P)]),
    "qf_mlp_destroy": (None, [_P]),
    "qf_mlp_get_params": (_I, [_P, _P, _P]),
    "qf_mlp_forward": (_I, [_P, _P, _P, _I64, _P]),
    "qf_mlp_gradients": (_I, [_P, _P, _P, _P, _I64, _P, ctypes.POINTER(ctypes.c_float)]),
    "qf_mlp_train_step": (_I, [_P, _P, _P, _P, _I64, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libquickfoil_amd.so is not built (run quickfoil_amd.build.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class _TableHolder:
    """Keeps a QfTable and the pointer array it refers to alive."""

    def __init__(self, columns, num_rows=None):
        import torch
        cols = []
        for c in columns:
            if not (c.dtype == torch.int32 and c.is_cuda and c.is_contiguous()):
                raise ValueError("columns must be contiguous int32 GPU tensors")
            cols.append(c)
        if num_rows is None:
            num_rows = int(cols[0].numel()) if cols else 0
        for c in cols:
            if c.numel() != num_rows:
                raise ValueError("columns must have equal length")
        self.cols = cols
        self.arr = (ctypes.c_void_p * max(1, len(cols)))(*[c.data_ptr() for c in cols])
        self.t = QfTable(ctypes.cast(self.arr, ctypes.POINTER(ctypes.c_void_p)), len(cols), num_rows)


class TableSet:
    """Device tables (each a list of contiguous int32 GPU column tensors) converted once to the C
    layout, an array of qf_table, for callers that pass the same tables on every call (a learner's
    relations).  The tensors stay referenced; their contents may change, their storage may not."""

    def __init__(self, tables):
        self.holders = [_TableHolder(t) for t in tables]
        self.array = (QfTable * max(1, len(self.holders)))(*[h.t for h in self.holders])

    def __len__(self):
        return len(self.holders)


class Context:
    """A qf_context bound to torch's current stream on one GPU."""

    def __init__(self, device=0):
        import torch
        self.lib = load_library()
        self.torch = torch
        self.device = device
        h = ctypes.c_void_p()
        self._check(self.lib.qf_context_create(device, ctypes.byref(h)), ctx=None)
        self.h = h
        stream = torch.cuda.current_stream(device).cuda_stream
        self._check(self.lib.qf_context_set_stream(self.h, ctypes.c_void_p(stream)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.qf_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status, ctx=True):
        if status != QF_OK:
            msg = self.lib.qf_last_error(self.h).decode() if ctx else "context creation failed"
            raise QfError(status, msg)

    def synchronize(self):
        self._check(self.lib.qf_synchronize(self.h))

    def trim(self):
        self._check(self.lib.qf_context_trim(self.h))

    def enable_timing(self, on=True):
        self._check(self.lib.qf_context_enable_timing(self.h, int(on)))

    def reset_timing(self):
        self._check(self.lib.qf_context_reset_timing(self.h))

    def phase_times(self):
        """{phase: (milliseconds, launches)} accumulated since the last reset."""
        n = 16
        names = (ctypes.c_char_p * n)()
        ms = np.zeros(n, dtype=np.float64)
        launches = np.zeros(n, dtype=np.int64)
        k = self.lib.qf_context_phase_times(self.h, names, ms.ctypes.data_as(ctypes.c_void_p),
                                            launches.ctypes.data_as(ctypes.c_void_p), n)
        if k < 0:
            self._check(-k)
        return {names[i].decode(): (float(ms[i]), int(launches[i])) for i in range(min(k, n))}

    # ---- reference-layout operators --------------------------------------------------------
    def radix_partition(self, values, radix_bits):
        t = self.torch
        n = values.numel()
        pairs = t.empty(2 * max(1, n), dtype=t.int32, device=values.device)
        offsets = t.empty((1 << radix_bits) + 1, dtype=t.int64, device=values.device)
        self._check(self.lib.qf_radix_partition(self.h, _ptr(values), n, radix_bits, _ptr(pairs), _ptr(offsets)))
        return pairs[: 2 * n], offsets

    def build_hash_table_on_partitions(self, pairs, offsets, radix_bits):
        t = self.torch
        h_off = offsets.cpu().numpy().astype(np.int64)
        nparts = 1 << radix_bits
        h_boff = np.zeros(nparts + 1, dtype=np.int64)
        total = self.lib.qf_partition_bucket_layout(h_off.ctypes.data_as(ctypes.c_void_p), nparts,
                                                    h_boff.ctypes.data_as(ctypes.c_void_p))
        if total < 0:
            raise QfError(-total, "bad partition offsets")
        n = pairs.numel() // 2
        d_boff = t.from_numpy(h_boff).to(pairs.device)
        buckets = t.empty(max(1, total), dtype=t.int32, device=pairs.device)
        nxt = t.empty(max(1, n), dtype=t.int32, device=pairs.device)
        self._check(self.lib.qf_build_hash_table_on_partitions(self.h, _ptr(pairs), _ptr(offsets), radix_bits,
                                                               _ptr(d_boff), total, n, _ptr(buckets), _ptr(nxt)))
        return buckets[:total], h_boff, nxt[:n]

    def build_hash_table_on_table(self, key_columns):
        t = self.torch
        tab = _TableHolder(key_columns)
        n = tab.t.num_rows
        nb = 1 << (max(n, 1) - 1).bit_length() if n > 0 else 0
        buckets = t.empty(max(1, nb), dtype=t.int32, device=key_columns[0].device)
        nxt = t.empty(max(1, n), dtype=t.int32, device=key_columns[0].device)
        self._check(self.lib.qf_build_hash_table_on_table(self.h, ctypes.byref(tab.t), _ptr(buckets), _ptr(nxt)))
        return buckets[:nb], nxt[:n]

    # ---- evaluation ------------------------------------------------------------------------
    @staticmethod
    def _options(mode, radix_bits, tile_rows, use_graph, reuse, partition_rows=0, autotune=False):
        o = QfEvalOptions()
        o.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        o.radix_bits = radix_bits
        o.tile_rows = tile_rows
        o.use_graph = int(use_graph)
        o.reuse_partitions = int(reuse)
        o.partition_rows = int(partition_rows)
        o.autotune = int(autotune)
        return o

    def autotune_choice(self):
        pr, tr, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
        self._check(self.lib.qf_autotune_choice(self.h, ctypes.byref(pr), ctypes.byref(tr), ctypes.byref(ms)))
        return {"partition_rows": pr.value, "tile_rows": tr.value, "best_ms": ms.value}

    def evaluate_literals(self, binding_columns, num_positive, relations, literals, arg_types=None,
                          mode="batched", radix_bits=0, tile_rows=0, use_graph=False, reuse_partitions=False,
                          semi_bits=False, check_range=False, partition_rows=0, autotune=False):
        """Returns (counts[L,4] uint64 ndarray, bits tensor or None).

        literals: list of (relation index, [var per position]) with -1 for new variables, or a
        LiteralArray made from one; relations: a list of column lists, or a TableSet made from one.
        Counts are {covered_pos, covered_neg, binding_pos, binding_neg}.
        """
        t = self.torch
        b = _TableHolder(binding_columns)
        if isinstance(relations, TableSet):
            rels, rel_arr = relations.holders, relations.array
        else:
            rels = [_TableHolder(r) for r in relations]
            rel_arr = (QfTable * max(1, len(rels)))(*[r.t for r in rels])
        if not isinstance(literals, LiteralArray):
            literals = LiteralArray(literals)
        L = len(literals)
        lit_arr = literals.pointer() if L > 0 else (QfLiteral * 1)()
        types = None
        types_ptr = ctypes.c_void_p(0)
        if arg_types is not None:
            types = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int32) for x in arg_types]))
            types_ptr = types.ctypes.data_as(ctypes.c_void_p)
        counts = np.zeros((max(1, L), 4), dtype=np.uint64)
        bits = None
        n = b.t.num_rows
        if semi_bits:
            words = (n + 63) // 64
            bits = t.empty(max(1, L * words), dtype=t.int64, device=binding_columns[0].device)
        o = self._options(mode, radix_bits, tile_rows, use_graph, reuse_partitions, partition_rows, autotune)
        st = self.lib.qf_evaluate_literals(self.h, ctypes.byref(b.t), num_positive, rel_arr, types_ptr, len(rels),
                                           lit_arr, L, ctypes.byref(o), counts.ctypes.data_as(ctypes.c_void_p),
                                           _ptr(bits))
        if st != QF_OK and not (st == 5 and not check_range):
            self._check(st)
        if bits is not None:
            bits = bits[: L * ((n + 63) // 64)].view(L, (n + 63) // 64)
        return counts[:L], bits

    def evaluate_literals_split(self, positive_columns, negative_columns, relations, literals, arg_types=None,
                                mode="batched", radix_bits=0, tile_rows=0, reuse_partitions=False, semi_bits=False,
                                check_range=False, partition_rows=0):
        """Counts over a clause whose positive and negative bindings are separate tables (the two
        passes); returns (counts[L,4], positive bits or None, negative bits or None)."""
        t = self.torch
        bp, bn = _TableHolder(positive_columns), _TableHolder(negative_columns)
        rels = [_TableHolder(r) for r in relations]
        rel_arr = (QfTable * max(1, len(rels)))(*[r.t for r in rels])
        if not isinstance(literals, LiteralArray):
            literals = LiteralArray(literals)
        L = len(literals)
        lit_arr = literals.pointer() if L > 0 else (QfLiteral * 1)()
        types = None
        types_ptr = ctypes.c_void_p(0)
        if arg_types is not None:
            types = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int32) for x in arg_types]))
            types_ptr = types.ctypes.data_as(ctypes.c_void_p)
        counts = np.zeros((max(1, L), 4), dtype=np.uint64)
        bits = [None, None]
        words = [(bp.t.num_rows + 63) // 64, (bn.t.num_rows + 63) // 64]
        if semi_bits:
            dev = positive_columns[0].device
            bits = [t.empty(max(1, L * w), dtype=t.int64, device=dev) for w in words]
        o = self._options(mode, radix_bits, tile_rows, False, reuse_partitions, partition_rows)
        st = self.lib.qf_evaluate_literals_split(self.h, ctypes.byref(bp.t), ctypes.byref(bn.t), rel_arr, types_ptr,
                                                 len(rels), lit_arr, L, ctypes.byref(o),
                                                 counts.ctypes.data_as(ctypes.c_void_p), _ptr(bits[0]), _ptr(bits[1]))
        if st != QF_OK and not (st == 5 and not check_range):
            self._check(st)
        if semi_bits:
            bits = [b[: L * w].view(L, w) for b, w in zip(bits, words)]
        return counts[:L], bits[0], bits[1]

    def count_aggregate(self, binding_columns, num_positive, relations, plans, num_slots, mode="batched",
                        radix_bits=0, tile_rows=0, semi_bits=False):
        """plans: list of dicts with relation, join_column, binding_column, atoms [(bg, bind)],
        nodes [(left, right, literal)] for conjunctions (atoms' literals in atom_literals), root_literal."""
        t = self.torch
        b = _TableHolder(binding_columns)
        rels = [_TableHolder(r) for r in relations]
        rel_arr = (QfTable * max(1, len(rels)))(*[r.t for r in rels])
        keep = []
        parr = (QfJoinGroupPlan * max(1, len(plans)))()
        for i, p in enumerate(plans):
            na = len(p["atoms"])
            nn = na + len(p.get("conjunctions", []))
            bg = (ctypes.c_int * max(1, na))(*[a[0] for a in p["atoms"]])
            bd = (ctypes.c_int * max(1, na))(*[a[1] for a in p["atoms"]])
            left = [-1] * na + [c[0] for c in p.get("conjunctions", [])]
            right = [-1] * na + [c[1] for c in p.get("conjunctions", [])]
            lit = list(p.get("atom_literals", [-1] * na)) + [c[2] for c in p.get("conjunctions", [])]
            la = (ctypes.c_int * max(1, nn))(*left)
            ra = (ctypes.c_int * max(1, nn))(*right)
            ta = (ctypes.c_int * max(1, nn))(*lit)
            keep += [bg, bd, la, ra, ta]
            parr[i] = QfJoinGroupPlan(p["relation"], p["join_column"], p["binding_column"], na,
                                      ctypes.cast(bg, ctypes.POINTER(ctypes.c_int)),
                                      ctypes.cast(bd, ctypes.POINTER(ctypes.c_int)), nn,
                                      ctypes.cast(la, ctypes.POINTER(ctypes.c_int)),
                                      ctypes.cast(ra, ctypes.POINTER(ctypes.c_int)),
                                      ctypes.cast(ta, ctypes.POINTER(ctypes.c_int)), p.get("root_literal", -1))
        counts = np.zeros((max(1, num_slots), 4), dtype=np.uint64)
        n = b.t.num_rows
        bits = None
        if semi_bits:
            bits = t.empty(max(1, num_slots * ((n + 63) // 64)), dtype=t.int64, device=binding_columns[0].device)
        o = self._options(mode, radix_bits, tile_rows, False, False)
        self._check(self.lib.qf_count_aggregate(self.h, ctypes.byref(b.t), num_positive, rel_arr, len(rels), parr,
                                                len(plans), num_slots, ctypes.byref(o),
                                                counts.ctypes.data_as(ctypes.c_void_p), _ptr(bits)))
        if bits is not None:
            bits = bits[: num_slots * ((n + 63) // 64)].view(num_slots, (n + 63) // 64)
        return counts[:num_slots], bits

    # ---- binding-set extension and coverage ------------------------------------------------
    @staticmethod
    def _literal(rel, vars_):
        lit = QfLiteral()
        lit.relation = 0 if rel is None else rel
        vv = list(vars_) + [-1] * (MAX_ARITY - len(vars_))
        for k in range(MAX_ARITY):
            lit.vars[k] = vv[k]
        return lit

    def create_label_aware_binding_table(self, binding_columns, num_positive, relation_columns, vars_,
                                         num_binding_pos, num_binding_neg, out=None):
        """The extended binding table (binding columns, then one per new variable), positives then
        negatives; `out`: int32 device buffers to write into (one per output column, long enough)."""
        t = self.torch
        b = _TableHolder(binding_columns)
        r = _TableHolder(relation_columns)
        lit = self._literal(0, vars_)
        n_new = sum(1 for v in vars_ if v < 0)
        rows = num_binding_pos + num_binding_neg
        dev = binding_columns[0].device
        ncols = len(binding_columns) + n_new
        if out is not None:
            assert len(out) >= ncols and all(o.numel() >= rows for o in out[:ncols]), "output buffers too small"
            outs = list(out[:ncols])
        else:
            outs = [t.empty(max(1, rows), dtype=t.int32, device=dev) for _ in range(ncols)]
        arr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        self._check(self.lib.qf_create_label_aware_binding_table(self.h, ctypes.byref(b.t), num_positive,
                                                                 ctypes.byref(r.t), ctypes.byref(lit),
                                                                 num_binding_pos, num_binding_neg, arr))
        return [o[:rows] for o in outs]

    def create_binding_table(self, binding_columns, relation_columns, vars_, out=None):
        """CreateBindingTable in two calls: the row count, then the rows (into `out`, a list of
        int32 device buffers, one per output column, when they are long enough)."""
        t = self.torch
        b = _TableHolder(binding_columns)
        r = _TableHolder(relation_columns)
        lit = self._literal(0, vars_)
        rows = ctypes.c_int64(0)
        self._check(self.lib.qf_create_binding_table(self.h, ctypes.byref(b.t), ctypes.byref(r.t), ctypes.byref(lit),
                                                     None, ctypes.byref(rows)))
        n_new = sum(1 for v in vars_ if v < 0 or v >= len(binding_columns))
        dev = binding_columns[0].device
        ncols = len(binding_columns) + n_new
        if out is not None and len(out) >= ncols and all(o.numel() >= max(1, rows.value) for o in out[:ncols]):
            outs = list(out[:ncols])
        else:
            outs = [t.empty(max(1, rows.value), dtype=t.int32, device=dev) for _ in range(ncols)]
        arr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        self._check(self.lib.qf_create_binding_table(self.h, ctypes.byref(b.t), ctypes.byref(r.t), ctypes.byref(lit),
                                                     arr, ctypes.byref(rows)))
        return [o[: rows.value] for o in outs]

    def create_binding_tables_presized(self, binding_columns, relations, literals, capacities):
        """CreateBindingTable of every literal ((relation, vars) over `relations`) into outputs of
        the given row capacities, in one call.  Returns (outputs: per literal its columns, or None
        when its rows exceeded the capacity and nothing was written; exact row counts; fits)."""
        t = self.torch
        b = _TableHolder(binding_columns)
        rels = [_TableHolder(r) for r in relations]
        rel_arr = (QfTable * max(1, len(rels)))(*[r.t for r in rels])
        lits = LiteralArray(literals)
        C = len(binding_columns)
        ncols = [C + sum(1 for v in vars_ if v < 0 or v >= C) for _, vars_ in literals]
        caps = np.ascontiguousarray(np.asarray(capacities, dtype=np.int64).reshape(-1))
        assert caps.size == len(literals) and (caps >= 0).all()
        total = int(sum(max(1, int(c)) * n for c, n in zip(caps, ncols)))
        arena = t.empty(max(1, total), dtype=t.int32, device=binding_columns[0].device)
        cols, col_arrays, at = [], [], 0
        for c, n in zip(caps, ncols):
            rows = max(1, int(c))
            these = [arena[at + k * rows: at + (k + 1) * rows] for k in range(n)]
            at += n * rows
            cols.append(these)
            col_arrays.append((ctypes.c_void_p * n)(*[x.data_ptr() for x in these]))
        ptrs = (ctypes.c_void_p * max(1, len(col_arrays)))(*[ctypes.cast(a, ctypes.c_void_p).value
                                                             for a in col_arrays])
        out_rows = np.zeros(max(1, len(literals)), dtype=np.int64)
        fits = np.zeros(max(1, len(literals)), dtype=np.uint8)
        lit_arr = lits.pointer() if len(literals) else (QfLiteral * 1)()
        st = self.lib.qf_create_binding_tables_presized(self.h, ctypes.byref(b.t), rel_arr, len(rels), lit_arr,
                                                        len(literals), caps.ctypes.data_as(ctypes.c_void_p), ptrs,
                                                        out_rows.ctypes.data_as(ctypes.c_void_p),
                                                        fits.ctypes.data_as(ctypes.c_void_p))
        if st not in (QF_OK, 4):
            self._check(st)
        outs = [[x[: int(r)] for x in these] if f else None for these, r, f in zip(cols, out_rows, fits)]
        return outs[: len(literals)], out_rows[: len(literals)], fits[: len(literals)].astype(bool)

    def create_binding_table_presized(self, binding_columns, relation_columns, vars_, capacity, out=None):
        """CreateBindingTable into buffers of `capacity` rows (the first `capacity` rows of `out`
        when given); on overflow the exact size is used for a second call.  Returns (columns,
        fell_back)."""
        t = self.torch
        b = _TableHolder(binding_columns)
        r = _TableHolder(relation_columns)
        lit = self._literal(0, vars_)
        n_new = sum(1 for v in vars_ if v < 0 or v >= len(binding_columns))
        dev = binding_columns[0].device
        ncols = len(binding_columns) + n_new
        if out is not None and len(out) >= ncols and all(o.numel() >= max(1, capacity) for o in out[:ncols]):
            outs = list(out[:ncols])
        else:
            outs = [t.empty(max(1, capacity), dtype=t.int32, device=dev) for _ in range(ncols)]
        arr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        rows = ctypes.c_int64(0)
        st = self.lib.qf_create_binding_table_presized(self.h, ctypes.byref(b.t), ctypes.byref(r.t),
                                                       ctypes.byref(lit), capacity, arr, ctypes.byref(rows))
        if st == 0:
            return [o[: rows.value] for o in outs], False
        if st != 4:  # anything but the overflow that asks for the exact size
            self._check(st)
        if not (out is not None and len(out) >= ncols and all(o.numel() >= max(1, rows.value) for o in out[:ncols])):
            outs = [t.empty(max(1, rows.value), dtype=t.int32, device=dev) for _ in range(ncols)]
        arr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        self._check(self.lib.qf_create_binding_table(self.h, ctypes.byref(b.t), ctypes.byref(r.t), ctypes.byref(lit),
                                                     arr, ctypes.byref(rows)))
        return [o[: rows.value] for o in outs], True

    def _semi(self, fn, probe_keys, build_keys, nbits):
        t = self.torch
        p = _TableHolder(probe_keys)
        bt = _TableHolder(build_keys)
        words = (nbits + 63) // 64
        bits = t.empty(max(1, words), dtype=t.int64, device=probe_keys[0].device)
        ones = ctypes.c_int64(0)
        self._check(fn(self.h, ctypes.byref(p.t), ctypes.byref(bt.t), _ptr(bits), ctypes.byref(ones)))
        return bits[:words], ones.value

    def left_semi_join(self, probe_keys, build_keys):
        return self._semi(self.lib.qf_left_semi_join, probe_keys, build_keys, probe_keys[0].numel())

    def right_semi_join(self, probe_keys, build_keys):
        return self._semi(self.lib.qf_right_semi_join, probe_keys, build_keys, build_keys[0].numel())

    def example_coverage(self, binding, row_begin, row_end, relation, vars_, examples):
        """Bits over example rows covered by the binding rows [row_begin, row_end) that match the
        literal (relation columns, variables) -- the learner's coverage semi-join chain."""
        t = self.torch
        b, r, e = _TableHolder(binding), _TableHolder(relation), _TableHolder(examples)
        lit = QfLiteral()
        lit.relation = 0
        for i in range(MAX_ARITY):
            lit.vars[i] = vars_[i] if i < len(vars_) else -1
        ne = e.t.num_rows
        bits = t.zeros(max(1, (ne + 63) // 64), dtype=t.int64, device=examples[0].device)
        ones = _I64(0)
        self._check(self.lib.qf_example_coverage(self.h, ctypes.byref(b.t), row_begin, row_end, ctypes.byref(r.t),
                                                 ctypes.byref(lit), ctypes.byref(e.t), _ptr(bits), ctypes.byref(ones)))
        return bits[: (ne + 63) // 64], ones.value

    def build_hash_table_after_semi_join(self, key_columns, bits):
        t = self.torch
        k = _TableHolder(key_columns)
        n = k.t.num_rows
        nb = 1 << (max(n, 1) - 1).bit_length() if n > 0 else 0
        dev = key_columns[0].device
        buckets = t.empty(max(1, nb), dtype=t.int32, device=dev)
        nxt = t.empty(max(1, n), dtype=t.int32, device=dev)
        first = t.empty(max(1, n), dtype=t.uint8, device=dev)
        self._check(self.lib.qf_build_hash_table_after_semi_join(self.h, ctypes.byref(k.t), _ptr(bits),
                                                                 _ptr(buckets), _ptr(nxt), _ptr(first)))
        return buckets[:nb], nxt[:n], first[:n]
