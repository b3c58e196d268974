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
