"""Seeded synthetic workloads for the candidate-literal evaluation path.

* microbench():   16,777,216 x 3 binding rows (20 % positive) against a 4,194,304 x 2 relation,
                  join keys over 1,000,000 values, uniform or Zipf(s=1.0); three literals
                  P(X0,_), P(X0,X1), P(X0,X2).
* refinement_step(): 8,388,608 x 4 binding rows (30 % positive) against 20 relations of
                  arity 2-3 with log-uniform sizes 10K..50M and Zipf s in {0, 0.8, 1.2};
                  every literal placement with at least one bound variable (2,080 literals).
* kinship():      a seeded family forest (2,000 persons, 6 generations) with parent, married,
                  male, female, brother, sister and the uncle/2 target.

Zipf values are ranks k = 1..D drawn with P(k) ~ k^-s, mapped to values through a seeded
permutation of [0, D): one permutation for every column of the microbench (its frequent join
values coincide on both sides), one per column elsewhere.  Generation runs on the GPU when a torch device is given (fast), else on
the host with numpy; the same seed gives the same data for a given device kind.
"""
import itertools
import math

import numpy as np

MICRO_ROWS = 16_777_216
MICRO_REL_ROWS = 4_194_304
MICRO_KEYS = 1_000_000
STEP_ROWS = 8_388_608
STEP_KEYS = 1_000_000


def _zipf_cdf(num_values, s):
    k = np.arange(1, num_values + 1, dtype=np.float64)
    w = k ** (-float(s)) if s > 0 else np.ones_like(k)
    c = np.cumsum(w)
    return c / c[-1]


def column_values(n, num_values, s, seed, device=None, perm_seed=None):
    """n int32 values over [0, num_values): uniform when s == 0, else Zipf(s) ranks mapped to values
    through a seeded permutation (`perm_seed`, default derived from `seed`; columns that share it
    share their frequent values)."""
    ps = (int(seed) * 7919 + 17) if perm_seed is None else int(perm_seed)
    if device is not None:
        import torch
        g = torch.Generator(device=device)
        g.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
        if s == 0:
            return torch.randint(0, num_values, (n,), generator=g, device=device, dtype=torch.int64).to(torch.int32)
        cdf = torch.from_numpy(_zipf_cdf(num_values, s)).to(device)
        u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
        rank = torch.clamp(torch.searchsorted(cdf, u), max=num_values - 1)
        gp = torch.Generator(device="cpu")
        gp.manual_seed(ps & 0x7FFFFFFFFFFFFFFF)
        perm = torch.randperm(num_values, generator=gp).to(device=device, dtype=torch.int32)
        return perm[rank]
    rng = np.random.Generator(np.random.PCG64(seed))
    if s == 0:
        return rng.integers(0, num_values, size=n, dtype=np.int64).astype(np.int32)
    cdf = _zipf_cdf(num_values, s)
    rank = np.minimum(np.searchsorted(cdf, rng.random(n)), num_values - 1)
    perm = np.random.Generator(np.random.PCG64(ps)).permutation(num_values).astype(np.int32)
    return perm[rank]


def all_placements(arity, num_vars):
    """Every assignment of clause variables (or -1) to positions with at least one bound."""
    out = []
    for combo in itertools.product(range(-1, num_vars), repeat=arity):
        if any(v >= 0 for v in combo):
            out.append(list(combo))
    return out


def microbench(distribution="uniform", rows=MICRO_ROWS, rel_rows=MICRO_REL_ROWS, num_values=MICRO_KEYS,
               device=None, seed=0x51F00000):
    s = 0.0 if distribution == "uniform" else 1.0
    # One permutation for every column on both sides: the frequent join values coincide (the
    # skewed-join case: about 5.6e11 matching pairs at full size).
    binding = [column_values(rows, num_values, s, seed + c, device, perm_seed=seed) for c in range(3)]
    relation = [column_values(rel_rows, num_values, s, seed + 16 + c, device, perm_seed=seed) for c in range(2)]
    num_positive = int(rows * 0.2)
    literals = [(0, [0, -1]), (0, [0, 1]), (0, [0, 2])]
    return {"binding": binding, "num_positive": num_positive, "relations": [relation], "literals": literals,
            "name": "microbench-%s" % distribution}


def step_relation_specs(num_relations=20, min_rows=10_000, max_rows=50_000_000, seed=0x5EED0003):
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = []
    for i in range(num_relations):
        arity = 2 if i < 4 else 3
        rows = int(round(10 ** rng.uniform(math.log10(min_rows), math.log10(max_rows))))
        specs.append({"arity": arity, "rows": rows, "s": [0.0, 0.8, 1.2][i % 3]})
    return specs


def refinement_step(rows=STEP_ROWS, num_vars=4, num_values=STEP_KEYS, specs=None, device=None, seed=0x5EED0000,
                    scale=1.0):
    """The batched refinement-step workload; `scale` shrinks every size for tests."""
    if specs is None:
        specs = step_relation_specs()
    rows = max(1, int(rows * scale))
    binding = [column_values(rows, num_values, 0.0, seed + c, device) for c in range(num_vars)]
    relations = []
    for i, sp in enumerate(specs):
        n = max(1, int(sp["rows"] * scale))
        relations.append([column_values(n, num_values, sp["s"], seed + 1000 * (i + 1) + c, device)
                          for c in range(sp["arity"])])
    literals = []
    for i, sp in enumerate(specs):
        for vars_ in all_placements(sp["arity"], num_vars):
            literals.append((i, vars_))
    return {"binding": binding, "num_positive": int(rows * 0.3), "relations": relations, "literals": literals,
            "name": "refinement-step", "specs": specs}


# ---------------------------------------------------------------------------------------------
# Kinship forest.

def kinship(num_persons=2000, generations=6, children_mean=2.5, seed=0x4B1E5, negative_ratio=4):
    """Returns dict of relations (lists of int32 numpy columns) plus uncle positives/negatives."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sex = []        # 0 = male, 1 = female
    gen_of = []
    parents = []    # (parent, child)
    married = []
    next_id = [0]

    def new_person(g, sx=None):
        pid = next_id[0]
        next_id[0] += 1
        sex.append(int(rng.integers(0, 2)) if sx is None else sx)
        gen_of.append(g)
        return pid

    # Founders: an eighth of the persons; with Poisson(children_mean) children per couple and
    # spouses joining each generation the forest reaches num_persons within the generations.
    founders = max(4, num_persons // 8)
    current = [new_person(0) for _ in range(founders)]
    for g in range(generations):
        if next_id[0] >= num_persons:
            break
        # Couples: each person of this generation marries someone of the other sex (from the
        # generation if available, else a newcomer of the same generation).
        rng.shuffle(current)
        males = [p for p in current if sex[p] == 0]
        females = [p for p in current if sex[p] == 1]
        couples = []
        for i in range(max(len(males), len(females))):
            if next_id[0] >= num_persons:
                break
            m = males[i] if i < len(males) else new_person(g, 0)
            f = females[i] if i < len(females) else new_person(g, 1)
            couples.append((m, f))
            married.append((m, f))
            married.append((f, m))
        if g == generations - 1:
            break
        nxt = []
        for m, f in couples:
            for _ in range(int(rng.poisson(children_mean))):
                if next_id[0] >= num_persons:
                    break
                c = new_person(g + 1)
                parents.append((m, c))
                parents.append((f, c))
                nxt.append(c)
        current = nxt
    n = next_id[0]
    sex = np.array(sex, dtype=np.int32)
    par_of = {}
    for p, c in parents:
        par_of.setdefault(c, []).append(p)
    # Siblings: same (full) parent set.
    by_parents = {}
    for c, ps in par_of.items():
        by_parents.setdefault(tuple(sorted(ps)), []).append(c)
    brother, sister = [], []
    for kids in by_parents.values():
        for x in kids:
            for y in kids:
                if x != y:
                    (brother if sex[x] == 0 else sister).append((x, y))
    spouse = {}
    for a, b in married:
        spouse[a] = b
    uncle = set()
    for c, ps in par_of.items():
        for p in ps:
            for x, y in brother:
                if y == p:
                    uncle.add((x, c))
            for x, y in sister:
                if y == p and x in spouse and sex[spouse[x]] == 0:
                    uncle.add((spouse[x], c))
    pos = sorted(uncle)
    neg = set()
    while len(neg) < negative_ratio * len(pos):
        a, b = int(rng.integers(0, n)), int(rng.integers(0, n))
        if (a, b) not in uncle:
            neg.add((a, b))
    neg = sorted(neg)

    def cols(pairs, k=2):
        if not pairs:
            return [np.zeros(0, dtype=np.int32) for _ in range(k)]
        a = np.array(pairs, dtype=np.int32).reshape(-1, k)
        return [np.ascontiguousarray(a[:, i]) for i in range(k)]

    rel = {
        "parent": cols(parents),
        "married": cols(married),
        "male": [np.nonzero(sex == 0)[0].astype(np.int32)],
        "female": [np.nonzero(sex == 1)[0].astype(np.int32)],
        "brother": cols(brother),
        "sister": cols(sister),
    }
    examples = cols(pos + neg)
    return {"relations": rel, "target": examples, "num_positive": len(pos), "num_persons": n}


def bongard(num_pictures=200000, objects_mean=6.0, inside_density=0.15, noise=0.02, seed=0xB0A6D):
    """Bongard-style synthetic pictures: element(picture, object) with Poisson(objects_mean)
    objects per picture; circle/square/triangle and small/large on objects; inside(a, b) for 15 %
    of same-picture object pairs.  The seeded rule `element(P, A), circle(A), small(A)` (length 3)
    holds for half the pictures (planted, and kept false for the rest), then 2 % of the labels
    flip.  Picture ids have type 0, object ids type 1.  Returns relations, the target column
    (positives first) and the positive count."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(num_pictures)
    counts = np.maximum(1, rng.poisson(objects_mean, n)).astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)])
    m = int(starts[-1])
    pic_of = np.repeat(np.arange(n, dtype=np.int64), counts)
    shape = rng.integers(0, 3, m)          # 0 circle, 1 square, 2 triangle
    size = rng.integers(0, 2, m)           # 0 small, 1 large
    rule_true = rng.random(n) < 0.5
    # Plant the rule in the first object of rule-true pictures; break it everywhere else.
    first = starts[:-1]
    shape[first[rule_true]] = 0
    size[first[rule_true]] = 0
    small_circle = (shape == 0) & (size == 0)
    bad = small_circle & ~rule_true[pic_of]
    size[bad] = 1
    labels = rule_true ^ (rng.random(n) < noise)
    obj = np.arange(m, dtype=np.int64) + n  # object ids follow picture ids
    # inside: a fraction of the ordered same-picture pairs.
    ins_a, ins_b = [], []
    k = counts
    pairs = k * (k - 1)
    draws = rng.binomial(pairs, inside_density)
    for p in np.nonzero(draws)[0]:
        c = int(k[p])
        idx = rng.choice(c * c, size=int(draws[p]) + c, replace=False)
        a, b = idx // c, idx % c
        keep = a != b
        a, b = a[keep][: draws[p]], b[keep][: draws[p]]
        ins_a.append(obj[starts[p] + a])
        ins_b.append(obj[starts[p] + b])
    ins_a = np.concatenate(ins_a) if ins_a else np.zeros(0, np.int64)
    ins_b = np.concatenate(ins_b) if ins_b else np.zeros(0, np.int64)
    i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
    rel = {
        "element": ([i32(pic_of), i32(obj)], [0, 1]),
        "circle": ([i32(obj[shape == 0])], [1]),
        "square": ([i32(obj[shape == 1])], [1]),
        "triangle": ([i32(obj[shape == 2])], [1]),
        "small": ([i32(obj[size == 0])], [1]),
        "large": ([i32(obj[size == 1])], [1]),
        "inside": ([i32(ins_a), i32(ins_b)], [1, 1]),
    }
    pics = np.arange(n, dtype=np.int32)
    target = np.concatenate([pics[labels], pics[~labels]])
    return {"relations": rel, "target": [i32(target)], "num_positive": int(labels.sum())}
