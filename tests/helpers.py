"""Shared inputs and checks for the oracle and GPU parity tests."""
import itertools
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_relation(path):
    """Pipe-delimited int32 rows (the reference loader's format, main/main.cpp:49-109)."""
    rows = []
    with open(path) as f:
        for line in f.read().splitlines():
            line = line.strip()
            if line:
                rows.append([int(x) for x in line.split("|")])
    a = np.array(rows, dtype=np.int32)
    return [np.ascontiguousarray(a[:, i]) for i in range(a.shape[1])]


def uncle_problem():
    """uncle/2 target (4 positives first) with brother, parent and sister as background."""
    d = os.path.join(GOLDEN, "uncle")
    target = load_relation(os.path.join(d, "uncle"))
    rels = [load_relation(os.path.join(d, n)) for n in ("brother", "parent", "sister")]
    literals = []
    for r in range(3):
        for vars_ in itertools.product(range(-1, 2), repeat=2):
            if any(v >= 0 for v in vars_):
                literals.append((r, list(vars_)))
    return target, 4, rels, literals


def radix_spec():
    with open(os.path.join(GOLDEN, "radix_partition_spec.json")) as f:
        return json.load(f)


def random_problem(seed, n=300, num_cols=3, rel_shapes=((2, 200), (3, 150), (1, 50)), domain=20, pos_frac=0.4,
                   zipf=False, max_literals=None):
    """Random binding table and relations over a small domain, with every literal placement."""
    rng = np.random.Generator(np.random.PCG64(seed))

    def col(size):
        if zipf:
            return (rng.zipf(1.6, size=size) % domain).astype(np.int32)
        return rng.integers(0, domain, size=size, dtype=np.int64).astype(np.int32)

    binding = [col(n) for _ in range(num_cols)]
    rels = [[col(rows) for _ in range(arity)] for arity, rows in rel_shapes]
    literals = []
    for r, (arity, _) in enumerate(rel_shapes):
        for vars_ in itertools.product(range(-1, num_cols), repeat=arity):
            if any(v >= 0 for v in vars_):
                literals.append((r, list(vars_)))
    if max_literals is not None and len(literals) > max_literals:
        idx = rng.choice(len(literals), size=max_literals, replace=False)
        literals = [literals[i] for i in sorted(idx)]
    return binding, int(n * pos_frac), rels, literals


def partitions_as_lists(pairs, offsets):
    pairs = np.asarray(pairs).reshape(-1, 2)
    out = []
    for p in range(len(offsets) - 1):
        out.append(pairs[offsets[p]:offsets[p + 1]].tolist())
    return out


def chain(buckets, nxt, b):
    """Positions on bucket b's chain (1-based links, 0 ends)."""
    out = []
    pos = int(buckets[b]) - 1
    while pos >= 0:
        out.append(pos)
        pos = int(nxt[pos]) - 1
    return out


def brute_extended_join(binding, num_positive, relation, vars_):
    """Rows of the extended binding table as a sorted list (positives then negatives kept apart)."""
    n = len(binding[0])
    m = len(relation[0])
    C = len(binding)
    unbound = [i for i, v in enumerate(vars_) if v < 0]
    pos_rows, neg_rows = [], []
    for r in range(n):
        for s in range(m):
            if all(relation[i][s] == binding[v][r] for i, v in enumerate(vars_) if v >= 0):
                row = tuple([int(binding[c][r]) for c in range(C)] + [int(relation[u][s]) for u in unbound])
                (pos_rows if r < num_positive else neg_rows).append(row)
    return sorted(pos_rows), sorted(neg_rows)


def build_abi_driver(out_dir):
    """Compiles tests/abi_driver.cpp (a C++ caller of include/quickfoil_amd.hpp) with g++ against
    the built library; returns the executable's path."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib_dir = os.path.join(root, "quickfoil_amd")
    out = os.path.join(out_dir, "abi_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(root, "include"), os.path.join(root, "tests", "abi_driver.cpp"),
                    "-L" + lib_dir, "-lquickfoil_amd", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + lib_dir + ":/opt/rocm/lib", "-o", out], check=True)
    return out


def abi_driver_script(binding, num_positive, relations, literals, plans=()):
    """The driver's stdin for a problem (see tests/abi_driver.cpp); plans: (relation, join column,
    binding column, atoms [(bg, bind, literal)], conjunctions [(left, right, literal)], root)."""
    lines = ["BINDING %d %d %d" % (len(binding), len(binding[0]), num_positive)]
    lines += [" ".join(map(str, c.tolist())) for c in binding]
    for r in relations:
        lines.append("REL %d %d" % (len(r), len(r[0])))
        lines += [" ".join(map(str, c.tolist())) for c in r]
    for rel, vars_ in literals:
        lines.append("LIT %d %s" % (rel, " ".join(map(str, list(vars_) + [-1] * (6 - len(vars_))))))
    for rel, jc, bc, atoms, conj, root in plans:
        lines.append("PLAN %d %d %d %d %d %d" % (rel, jc, bc, len(atoms), len(conj), root))
        lines += ["%d %d %d" % a for a in atoms]
        lines += ["%d %d %d" % c for c in conj]
    lines.append("END")
    return "\n".join(lines) + "\n"


def parse_driver_output(text):
    """{tag: {index: [values]}} for counted lines, {tag: [values]} for arrays."""
    out = {}
    for line in text.splitlines():
        parts = line.split()
        tag = parts[0]
        if tag in ("EVAL", "SPLIT", "AGG", "AGG2"):
            out.setdefault(tag, {})[int(parts[1])] = [int(x) for x in parts[2:]]
        elif tag == "SIZE":
            out.setdefault(tag, {})[int(parts[1])] = parts[2]
        else:
            out[tag] = [int(x) for x in parts[1:]]
    return out
