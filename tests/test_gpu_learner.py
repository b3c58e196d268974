"""The device backend learns exactly the clauses the oracle backend learns."""
import pytest

from learner_problems import PROBLEMS_BY_NAME, bongard_test_file, cpu_backend, release_cpu_backend
from quickfoil_amd import learner

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["uncle", "bongard_sample", "bongard_sample_1000", "kinship", "bongard_synthetic",
                                  "bongard_synthetic_200k"])
def test_identical_clauses(qf_ctx, name):
    make = PROBLEMS_BY_NAME[name]
    problem, opts = make()
    options = learner.default_options(**opts)
    be, keep = cpu_backend(problem)
    try:
        want = learner.learn(problem, be, options)
    finally:
        release_cpu_backend(be)
    dev = learner.DeviceBackend(qf_ctx, problem)
    try:
        got = learner.learn(problem, dev.backend, options)
    finally:
        dev.close()
    assert want, "no clause learnt"
    assert got == want


@pytest.mark.parametrize("name", ["bongard_sample", "kinship"])
def test_test_runner_identical(qf_ctx, name):
    """QuickFoilTestRunner counts agree between the backends (tests = the target facts)."""
    problem, opts = PROBLEMS_BY_NAME[name]()
    options = learner.default_options(**opts)
    be, keep = cpu_backend(problem)
    dev = learner.DeviceBackend(qf_ctx, problem)
    try:
        clauses = learner.learn(problem, be, options)
        tests = problem.target.columns
        want = learner.test_metrics(problem, be, clauses, tests, problem.num_positive)
        got = learner.test_metrics(problem, dev.backend, clauses, tests, problem.num_positive)
        assert got == want
        assert want["covered_test_positive"] > 0
    finally:
        release_cpu_backend(be)
        dev.close()


@pytest.mark.parametrize("fold,name", [("bongard100", "bongard_sample"), ("bongard1000", "bongard_sample_1000")])
def test_test_runner_on_test_file(qf_ctx, fold, name):
    """The fold's own test file (main/main.cpp:264-290 reads it after learning): counts agree
    between the backends."""
    problem, opts = PROBLEMS_BY_NAME[name]()
    options = learner.default_options(**opts)
    tests, num_test_positive = bongard_test_file(fold)
    be, keep = cpu_backend(problem)
    dev = learner.DeviceBackend(qf_ctx, problem)
    try:
        clauses = learner.learn(problem, be, options)
        want = learner.test_metrics(problem, be, clauses, tests, num_test_positive)
        got = learner.test_metrics(problem, dev.backend, clauses, tests, num_test_positive)
        assert got == want
        assert want["covered_test_positive"] > 0
    finally:
        release_cpu_backend(be)
        dev.close()


def test_backend_evaluate_keeps_counts_above_int32(qf_ctx):
    """The device backend's evaluate returns success with exact uint64 counts when a candidate's
    binding count exceeds int32 (60,000 positive rows x 40,000 facts of one join value), so one
    skewed candidate does not end a learning run."""
    import ctypes
    import numpy as np
    n, m, npos = 70000, 40000, 60000
    rel = learner.Relation("r", [np.zeros(m, np.int32), np.arange(m, dtype=np.int32)], [0, 0])
    target = learner.Relation("t", [np.arange(n, dtype=np.int32), np.zeros(n, np.int32)], [0, 0])
    problem = learner.Problem(target, npos, [rel])
    dev = learner.DeviceBackend(qf_ctx, problem)
    try:
        be = dev.backend
        upload = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                  ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p))(be.upload)
        evaluate = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)(be.evaluate)
        release = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)(be.release)
        cols = (ctypes.c_void_p * 2)(*[c.ctypes.data for c in target.columns])
        h = ctypes.c_void_p()
        assert upload(be.self, cols, 2, n, ctypes.byref(h)) == 0
        from quickfoil_amd._native import LiteralArray
        lits = LiteralArray([(0, [1, -1]), (0, [1, 0])])
        counts = np.zeros((2, 4), dtype=np.uint64)
        st = evaluate(be.self, h, 2, n, npos, lits.array.ctypes.data, 2, counts.ctypes.data)
        release(be.self, h)
        assert st == 0
        assert counts[0].tolist() == [npos, n - npos, npos * m, (n - npos) * m]
        assert counts[1].tolist() == [m, 0, m, 0]  # row r < m meets fact (0, r) once
    finally:
        dev.close()
