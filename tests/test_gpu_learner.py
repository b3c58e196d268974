"""The device backend learns exactly the clauses the oracle backend learns."""
import pytest

from learner_problems import PROBLEMS_BY_NAME
from quickfoil_amd import learner

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["uncle", "bongard_sample", "kinship", "bongard_synthetic"])
def test_identical_clauses(qf_ctx, name):
    make = PROBLEMS_BY_NAME[name]
    problem, opts = make()
    options = learner.default_options(**opts)
    from learner_problems import cpu_backend, release_cpu_backend
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
    from learner_problems import cpu_backend, release_cpu_backend
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
