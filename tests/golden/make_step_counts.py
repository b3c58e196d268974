"""Writes step_1_256_counts.npy: the oracle's counts (uint64 [2080, 4]: covered positive, covered
negative, binding positive, binding negative per literal, in literal order) for the batched
refinement-step workload at 1/256 scale (quickfoil_amd.datagen.refinement_step(scale=1/256):
32,768 binding rows over four variables, 20 relations of 47 to 104,123 rows, 2,080 literals),
computed with the grouped-count closed form.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
OUT = os.path.join(HERE, "step_1_256_counts.npy")
SCALE = 1.0 / 256


def workload():
    from quickfoil_amd import datagen
    return datagen.refinement_step(scale=SCALE)


def main():
    import oracle
    w = workload()
    counts, _ = oracle.Problem(w["binding"], w["num_positive"], w["relations"], w["literals"]).grouped(semi=False)
    np.save(OUT, np.asarray(counts, dtype=np.uint64))
    print("wrote", OUT, counts.shape, int(counts.sum()))


if __name__ == "__main__":
    main()
