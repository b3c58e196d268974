"""Host learner binding: QuickFoil's search control (qf_learn, include/quickfoil_learner.h) over an
evaluation backend -- the device backend on this library's operators, or any other that fills
the same function table.

A problem is a target relation (positives first) plus background relations, as the reference's
JSON configuration describes them (main/Configuration.cpp:20-117, main/main.cpp:100-160).
"""
import ctypes
import json
import os
import re

import numpy as np

from . import _native
from .loader import load_relation

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64


class LearnerBackend(ctypes.Structure):
    _fields_ = [("self", _P), ("upload", _P), ("download", _P), ("release", _P), ("concat", _P),
                ("evaluate", _P), ("extend", _P), ("coverage", _P), ("join", _P)]


class LearnerOptions(ctypes.Structure):
    _fields_ = [
        ("positive_threshold", ctypes.c_double),
        ("max_iterations", _I),
        ("minimum_inflated_precision", ctypes.c_double),
        ("minimum_true_precision", ctypes.c_double),
        ("minimum_f_score", ctypes.c_double),
        ("maximum_clause_length", _I),
        ("maximum_random_literals", _I),
        ("maximum_random_trials", _I),
        ("minimum_coverage_for_tied_literal", ctypes.c_double),
        ("num_saved_literals", _I),
        ("random_margin", ctypes.c_double),
    ]


class LearnerProblem(ctypes.Structure):
    _fields_ = [
        ("num_background", _I),
        ("names", ctypes.POINTER(ctypes.c_char_p)),
        ("arities", ctypes.POINTER(_I)),
        ("argument_types", ctypes.POINTER(ctypes.POINTER(_I))),
        ("keys", ctypes.POINTER(_I)),
        ("num_facts", ctypes.POINTER(_I64)),
        ("target_columns", ctypes.POINTER(_P)),
        ("num_positive", _I64),
    ]


def _bind(lib):
    lib.qf_learner_options_default.restype = None
    lib.qf_learner_options_default.argtypes = [ctypes.POINTER(LearnerOptions)]
    lib.qf_learn.restype = _I
    lib.qf_learn.argtypes = [ctypes.POINTER(LearnerProblem), ctypes.POINTER(LearnerOptions),
                             ctypes.POINTER(LearnerBackend), ctypes.c_char_p, _I64, ctypes.POINTER(_I)]
    lib.qf_learner_device_backend_create.restype = _I
    lib.qf_learner_device_backend_create.argtypes = [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.POINTER(_P)),
                                                     ctypes.POINTER(_I64), ctypes.POINTER(ctypes.POINTER(_I)),
                                                     ctypes.POINTER(LearnerBackend)]
    lib.qf_learner_test.restype = _I
    lib.qf_learner_test.argtypes = [ctypes.POINTER(LearnerBackend), ctypes.POINTER(_P), _I, _I64,
                                    ctypes.POINTER(_native.QfLiteral), ctypes.POINTER(_I), _I, ctypes.POINTER(_I64)]
    lib.qf_learner_device_backend_destroy.restype = None
    lib.qf_learner_device_backend_destroy.argtypes = [ctypes.POINTER(LearnerBackend)]
    return lib


class Relation:
    def __init__(self, name, columns, types, key=-1):
        self.name = name
        self.columns = [np.ascontiguousarray(c, dtype=np.int32) for c in columns]
        self.types = list(types)
        self.key = key
        assert len(self.columns) == len(self.types)

    @property
    def rows(self):
        return int(self.columns[0].size) if self.columns else 0


class Problem:
    """Background relations (relation index = position) and the target with its first
    num_positive facts positive."""

    def __init__(self, target, num_positive, background):
        self.target = target
        self.num_positive = int(num_positive)
        self.background = list(background)

    def _arrays(self):
        rels = self.background + [self.target]
        self._names = (ctypes.c_char_p * len(rels))(*[r.name.encode() for r in rels])
        self._arities = (_I * len(rels))(*[len(r.types) for r in rels])
        self._types = [(_I * len(r.types))(*r.types) for r in rels]
        self._type_ptrs = (ctypes.POINTER(_I) * len(rels))(*[ctypes.cast(t, ctypes.POINTER(_I)) for t in self._types])
        self._keys = (_I * len(rels))(*[r.key for r in rels])
        self._facts = (_I64 * len(rels))(*[r.rows for r in rels])
        self._target_cols = (_P * len(self.target.columns))(*[c.ctypes.data for c in self.target.columns])
        p = LearnerProblem()
        p.num_background = len(self.background)
        p.names = self._names
        p.arities = self._arities
        p.argument_types = self._type_ptrs
        p.keys = self._keys
        p.num_facts = self._facts
        p.target_columns = self._target_cols
        p.num_positive = self.num_positive
        return p

    def relation_arrays(self):
        """ctypes arrays describing the background relations for a backend constructor."""
        n = len(self.background)
        cols = [(_P * len(r.columns))(*[c.ctypes.data for c in r.columns]) for r in self.background]
        col_ptrs = (ctypes.POINTER(_P) * max(1, n))(*[ctypes.cast(c, ctypes.POINTER(_P)) for c in cols])
        arities = (_I * max(1, n))(*[len(r.columns) for r in self.background])
        rows = (_I64 * max(1, n))(*[r.rows for r in self.background])
        types = [(_I * len(r.types))(*r.types) for r in self.background]
        type_ptrs = (ctypes.POINTER(_I) * max(1, n))(*[ctypes.cast(t, ctypes.POINTER(_I)) for t in types])
        keep = (cols, types)
        return n, arities, col_ptrs, rows, type_ptrs, keep


def default_options(**overrides):
    lib = _bind(_native.load_library())
    o = LearnerOptions()
    lib.qf_learner_options_default(ctypes.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


def learn(problem, backend, options=None):
    """Runs QuickFoil::Learn with `backend` (a LearnerBackend); returns the clause strings."""
    lib = _bind(_native.load_library())
    opts = options if options is not None else default_options()
    p = problem._arrays()
    cap = 1 << 20
    out = ctypes.create_string_buffer(cap)
    n = _I(0)
    st = lib.qf_learn(ctypes.byref(p), ctypes.byref(opts), ctypes.byref(backend), out, cap, ctypes.byref(n))
    if st != 0:
        raise _native.QfError(st, out.value.decode(errors="replace"))
    text = out.value.decode()
    return [l for l in text.split("\n") if l]


class DeviceBackend:
    """The device backend over a Context (relations uploaded once)."""

    def __init__(self, ctx, problem):
        self.lib = _bind(_native.load_library())
        self.ctx = ctx
        n, arities, col_ptrs, rows, type_ptrs, self._keep = problem.relation_arrays()
        self.backend = LearnerBackend()
        st = self.lib.qf_learner_device_backend_create(ctx.h, n, arities, col_ptrs, rows, type_ptrs,
                                                       ctypes.byref(self.backend))
        if st != 0:
            raise _native.QfError(st, "device backend")

    def close(self):
        if self.backend.self:
            self.lib.qf_learner_device_backend_destroy(ctypes.byref(self.backend))


def _strip_comments(text):
    # The reference's configurations carry // and /* */ comments (folly::json::stripComments).
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return "\n".join(re.sub(r"//.*$", "", l) for l in text.splitlines())


def load_configuration(path, data_dir=None):
    """Reads a reference JSON configuration (main/Configuration.cpp:20-117).  Relation files are
    looked up by base name in `data_dir` (default: the configuration's directory).  Returns the
    Problem and the test setting (file path, num_positive) or None."""
    with open(path) as f:
        conf = json.loads(_strip_comments(f.read()))
    base = data_dir or os.path.dirname(os.path.abspath(path))
    background_names = list(conf.get("background", []))
    rels = {}
    target = None
    for r in conf["relations"]:
        attrs = r["attributes"]
        keep = [not a.get("skip", False) for a in attrs]
        types = [a["domain_type"] for a, k in zip(attrs, keep) if k]
        cols = load_relation(os.path.join(base, os.path.basename(r["file"])), len(attrs), keep)
        rel = Relation(r["name"], cols, types, r.get("key", -1))
        if r["name"] == conf["target"]:
            target = (rel, r["num_positive"])
        elif r["name"] in background_names:
            rels[r["name"]] = rel
    # Background predicates in the configuration's relation order (main/main.cpp:217-224).
    background = [rels[r["name"]] for r in conf["relations"] if r["name"] in rels]
    test = conf.get("test")
    test_setting = None
    if test is not None:
        test_setting = (os.path.join(base, os.path.basename(test["file"])), test["num_positive"])
    return Problem(target[0], target[1], background), test_setting


def parse_clause(text, problem):
    """FoilClause::ToString text -> list of body literals [(relation index, [variable ids])]."""
    index = {r.name: i for i, r in enumerate(problem.background)}
    body = text.split(":-", 1)[1].strip()
    lits = []
    for m in re.finditer(r"([^\s(),]+)\(([^)]*)\)", body):
        lits.append((index[m.group(1)], [int(v) for v in m.group(2).split(",")]))
    return lits


def run_test(problem, backend, clauses, test_columns):
    """QuickFoilTestRunner::RunTest (learner/QuickFoilTestRunner.cpp:135-160): the number of test
    examples no clause covers."""
    lib = _bind(_native.load_library())
    cols = [np.ascontiguousarray(c, dtype=np.int32) for c in test_columns]
    body, lengths = [], []
    for c in clauses:
        lits = parse_clause(c, problem)
        lengths.append(len(lits))
        for rel, vars_ in lits:
            q = _native.QfLiteral()
            q.relation = rel
            for i in range(_native.MAX_ARITY):
                q.vars[i] = vars_[i] if i < len(vars_) else -1
            body.append(q)
    arr = (_native.QfLiteral * max(1, len(body)))(*body)
    lens = (_I * max(1, len(lengths)))(*lengths)
    ptrs = (_P * len(cols))(*[c.ctypes.data for c in cols])
    out = _I64(0)
    st = lib.qf_learner_test(ctypes.byref(backend), ptrs, len(cols), int(cols[0].size), arr, lens, len(lengths),
                             ctypes.byref(out))
    if st != 0:
        raise _native.QfError(st, "test run")
    return out.value


def test_metrics(problem, backend, clauses, test_columns, num_test_positive):
    """Covered test positives/negatives, precision and recall as main/main.cpp:264-290 reports."""
    pos = [c[:num_test_positive] for c in test_columns]
    neg = [c[num_test_positive:] for c in test_columns]
    cov_pos = num_test_positive - run_test(problem, backend, clauses, pos)
    n_neg = int(neg[0].size)
    cov_neg = n_neg - run_test(problem, backend, clauses, neg) if n_neg else 0
    precision = 0.0 if cov_pos == 0 and cov_neg == 0 else cov_pos / (cov_pos + cov_neg)
    return {"covered_test_positive": cov_pos, "covered_test_negative": cov_neg, "precision": precision,
            "recall": cov_pos / num_test_positive}
