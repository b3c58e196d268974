"""Learning problems shared by the learner tests: the reference's sample data (uncle, Bongard
th3/100 and th3/1000 fold 0) and the synthetic kinship and Bongard generators."""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))

import oracle  # noqa: E402
from quickfoil_amd import datagen, learner  # noqa: E402
from quickfoil_amd.loader import load_relation  # noqa: E402

GOLDEN = os.path.join(HERE, "golden")


def uncle():
    # sample_data/uncle/uncle.json:14-40: target uncle (4 positives), background brother and
    # parent, every argument of domain type 1.
    def rel(name):
        return learner.Relation(name, load_relation(os.path.join(GOLDEN, "uncle", name), 2), [1, 1])
    return learner.Problem(rel("uncle"), 4, [rel("brother"), rel("parent")]), {}


# sample_data/bongard/th3/100/fold_0/bongard.json:4-126 (th3/1000/fold_0/bongard.json:1-118 lists
# the same relations and types): relation name, file, domain types.
_BONGARD = [
    ("element_train", "element_merge", [0, 1]),
    ("circle_train", "circle_merge", [1]),
    ("rectangle_train", "rectangle_merge", [1]),
    ("triangle_train", "triangle_merge", [1]),
    ("triangle_up_train", "triangle_up_merge", [1]),
    ("triangle_down_train", "triangle_down_merge", [1]),
    ("east_train", "east_merge", [1, 1]),
    ("north_train", "north_merge", [1, 1]),
    ("inside_train", "inside_merge", [1, 1]),
]


# Positives of each fold's training target and test file: th3/100/fold_0/bongard.json:15-25
# (test num_positive 10, 88 training positives) and th3/1000/fold_0/bongard.json:14-21 (100; 901).
_BONGARD_FOLDS = {"bongard100": (88, 10), "bongard1000": (901, 100)}


def bongard_sample(fold="bongard100"):
    d = os.path.join(GOLDEN, fold)
    bg = [learner.Relation(n, load_relation(os.path.join(d, f), len(t)), t) for n, f, t in _BONGARD]
    target = learner.Relation("bongard", load_relation(os.path.join(d, "bongard_train_5"), 1), [0])
    return learner.Problem(target, _BONGARD_FOLDS[fold][0], bg), {}


def bongard_test_file(fold):
    """The fold's test examples (positives first) and their number of positives."""
    return load_relation(os.path.join(GOLDEN, fold, "bongard_test"), 1), _BONGARD_FOLDS[fold][1]


def kinship(num_persons=2000):
    k = datagen.kinship(num_persons=num_persons)
    r = k["relations"]
    names = ["parent", "married", "male", "female", "brother", "sister"]
    bg = [learner.Relation(n, r[n], [0] * len(r[n])) for n in names]
    return (learner.Problem(learner.Relation("uncle", k["target"], [0, 0]), k["num_positive"], bg),
            {"maximum_clause_length": 3})


def bongard_synthetic(num_pictures):
    b = datagen.bongard(num_pictures)
    bg = [learner.Relation(n, c, t) for n, (c, t) in b["relations"].items()]
    return (learner.Problem(learner.Relation("picture", b["target"], [0]), b["num_positive"], bg),
            {"maximum_clause_length": 4})


def cpu_backend(problem):
    """The oracle's backend (test infrastructure): returns (backend, handle keeping arrays alive)."""
    lib = oracle.lib()
    lib.orc_learner_backend_create.restype = ctypes.c_int
    n, ar, cols, rows, types, keep = problem.relation_arrays()
    be = learner.LearnerBackend()
    st = lib.orc_learner_backend_create(n, ar, cols, rows, types, ctypes.byref(be))
    assert st == 0
    return be, keep


def release_cpu_backend(be):
    lib = oracle.lib()
    lib.orc_learner_backend_destroy.restype = None
    lib.orc_learner_backend_destroy(ctypes.byref(be))


PROBLEMS_BY_NAME = {
    "uncle": uncle,
    "bongard_sample": bongard_sample,
    "kinship": kinship,
    "bongard_sample_1000": lambda: bongard_sample("bongard1000"),
    "bongard_synthetic": lambda: bongard_synthetic(5000),
    "bongard_synthetic_200k": lambda: bongard_synthetic(200000),
}
