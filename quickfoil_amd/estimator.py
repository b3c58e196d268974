"""Learned join-cardinality estimator (the MLP of csrc/mlp.hip): features, training on recorded
joins, batched inference, and presizing of binding-table joins with exact fallback.

A join instance is a binding table joined with one literal: the output has
binding_pos + binding_neg rows (the evaluator's exact counts give the training targets).
Features (40): log2 sizes and composite distinct-key estimates of both sides, arity, new-variable
and key counts, the independence estimate, and 16-bin key-frequency histograms of a row sample of
each side.  The independence estimate is the textbook one for equi-joins under uniform,
independent columns: N_b N_p / prod over key columns k of max(D_b(k), D_p(k)), with per-column
distinct counts D estimated from the samples.

In the loop (presized_joins): the network's predictions size the outputs of CreateBindingTable
joins (the joins whose size no count gives beforehand), which run in one call without a host round
trip per join; a join whose rows exceed its capacity writes nothing and is rerun at its exact size,
so every output is exact.
"""
import ctypes
import math

import numpy as np

from . import _native

NUM_FEATURES = 40
SAMPLE = 1024


def _rows(c):
    return int(c.numel()) if hasattr(c, "numel") else int(np.asarray(c).size)


def _host(c):
    return c.cpu().numpy() if hasattr(c, "cpu") else np.asarray(c)


def _sample(cols, sample):
    """An evenly spaced row sample of the columns (numpy arrays or device tensors: only the
    sampled rows leave the device)."""
    n = _rows(cols[0])
    if n <= sample:
        return [_host(c) for c in cols], n
    idx = np.linspace(0, n - 1, sample).astype(np.int64)
    if hasattr(cols[0], "cpu"):  # a torch tensor
        import torch
        ti = torch.from_numpy(idx).to(cols[0].device)
        return [c[ti].cpu().numpy() for c in cols], n
    return [np.asarray(c)[idx] for c in cols], n


def _key_stats(cols, sample=SAMPLE):
    """Distinct-key estimate (GEE: sqrt(N/S) f1 + sum_{j>=2} f_j) and the fraction of sampled
    rows whose key occurs 2^b..2^(b+1)-1 times in the sample, b = 0..15."""
    s, n = _sample(cols, sample)
    if n == 0:
        return 0.0, np.zeros(16)
    key = np.zeros(s[0].size, dtype=np.int64)
    for c in s:
        key = key * 1000003 + c.astype(np.int64)
    _, counts = np.unique(key, return_counts=True)
    S = key.size
    f1 = np.count_nonzero(counts == 1)
    distinct = math.sqrt(n / S) * f1 + np.count_nonzero(counts >= 2)
    bins = np.minimum(15, np.floor(np.log2(counts)).astype(np.int64))
    hist = np.bincount(bins, weights=counts, minlength=16)[:16] / S
    return float(distinct), hist


def join_features(binding_keys, relation_keys, arity, num_new, cache=None, cache_keys=(None, None),
                  column_keys=None):
    """Features of joining binding rows (key columns) with relation rows (key columns).
    cache_keys name the two composite keys, column_keys the (binding, relation) name of each key
    column, for caching their sample statistics across joins."""
    def stats(cols, ck):
        if cache is not None and ck is not None and ck in cache:
            return cache[ck]
        v = _key_stats(cols)
        if cache is not None and ck is not None:
            cache[ck] = v
        return v
    nb, nr = _rows(binding_keys[0]), _rows(relation_keys[0])
    db, hb = stats(binding_keys, cache_keys[0])
    dr, hr = stats(relation_keys, cache_keys[1])
    lb, lr = math.log2(1 + nb), math.log2(1 + nr)
    # Independence: N_b N_p / prod_k max(D_b(k), D_p(k)) over the key columns.
    log_rows = math.log2(max(1, nb)) + math.log2(max(1, nr))
    for k in range(len(binding_keys)):
        ckb, ckr = column_keys[k] if column_keys is not None else (None, None)
        d_b = stats([binding_keys[k]], ckb)[0]
        d_r = stats([relation_keys[k]], ckr)[0]
        log_rows -= math.log2(max(1.0, d_b, d_r))
    ind = math.log2(1.0 + 2.0 ** min(62.0, log_rows)) if nb and nr else 0.0
    f = [lb, lr, math.log2(1 + db), math.log2(1 + dr), arity, num_new, len(binding_keys), ind]
    return np.concatenate([np.array(f), hb, hr]).astype(np.float32)


def independence_estimate(features):
    """The estimator without the network: log2(1 + output rows) under uniform, independent key
    columns (feature 7)."""
    return np.maximum(0.0, features[:, 7])


def literal_instances(binding, relations, literals, counts, cache, tag):
    """Features and log2(1 + output rows) targets of every literal of an evaluated step."""
    X, y = [], []
    for (rel, vars_), c in zip(literals, counts):
        bound = [i for i, v in enumerate(vars_) if v >= 0]
        bkeys = [binding[vars_[i]] for i in bound]
        rkeys = [relations[rel][i] for i in bound]
        num_new = len(vars_) - len(bound)
        ck = ((tag, "b", tuple(vars_[i] for i in bound)), (tag, "r", rel, tuple(bound)))
        col_keys = [((tag, "b", (vars_[i],)), (tag, "r", rel, (i,))) for i in bound]
        X.append(join_features(bkeys, rkeys, len(vars_), num_new, cache, ck, col_keys))
        y.append(math.log2(1 + float(c[2]) + float(c[3])))
    return np.stack(X), np.array(y, dtype=np.float32)


def init_params(seed=0):
    """He-initialised parameters in the kernel layout."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w1 = np.zeros((64, 48), np.float32)
    w1[:, :NUM_FEATURES] = rng.normal(0, math.sqrt(2.0 / NUM_FEATURES), (64, NUM_FEATURES))
    w2 = rng.normal(0, math.sqrt(2.0 / 64), (64, 64)).astype(np.float32)
    w3 = rng.normal(0, math.sqrt(1.0 / 64), 64).astype(np.float32)
    return np.concatenate([w1.ravel(), np.zeros(64), w2.ravel(), np.zeros(64), w3, [0.0]]).astype(np.float32)


class Estimator:
    """The MLP on a context: forward, gradients, Adam training.  Inputs are standardised with the
    training set's mean and deviation (kept beside the weights)."""

    def __init__(self, ctx, params=None, mean=None, std=None):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.lib = ctx.lib
        self.n_params = self.lib.qf_mlp_num_params()
        p = np.ascontiguousarray(init_params() if params is None else params, dtype=np.float32)
        assert p.size == self.n_params
        self.h = ctypes.c_void_p()
        ctx._check(self.lib.qf_mlp_create(ctx.h, p.ctypes.data_as(ctypes.c_void_p), ctypes.byref(self.h)))
        self.mean = np.zeros(NUM_FEATURES, np.float32) if mean is None else mean
        self.std = np.ones(NUM_FEATURES, np.float32) if std is None else std

    def close(self):
        if self.h:
            self.lib.qf_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx._check(self.lib.qf_mlp_get_params(self.ctx.h, self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def _dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()

    def normalise(self, X):
        return (X - self.mean) / self.std

    def predict_raw(self, X):
        d = self._dev(X)
        out = self.torch.empty(max(1, d.shape[0]), dtype=self.torch.float32, device=d.device)
        self.ctx._check(self.lib.qf_mlp_forward(self.ctx.h, self.h, _native._ptr(d), d.shape[0], _native._ptr(out)))
        return out[: d.shape[0]].cpu().numpy()

    def predict(self, X):
        return self.predict_raw(self.normalise(X))

    def gradients(self, X, y):
        dx, dy = self._dev(X), self._dev(y)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float(0)
        self.ctx._check(self.lib.qf_mlp_gradients(self.ctx.h, self.h, _native._ptr(dx), _native._ptr(dy), dx.shape[0],
                                                  g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss)))
        return g, loss.value

    def fit(self, X, y, epochs=60, batch=512, lr=2e-3, seed=0, log=None):
        self.mean = X.mean(0).astype(np.float32)
        self.std = (X.std(0) + 1e-6).astype(np.float32)
        Xn = self.normalise(X).astype(np.float32)
        dx, dy = self._dev(Xn), self._dev(y)
        rng = np.random.Generator(np.random.PCG64(seed))
        n = Xn.shape[0]
        loss = ctypes.c_float(0)
        history = []
        for ep in range(epochs):
            perm = self.torch.from_numpy(rng.permutation(n)).cuda()
            xs, ys = dx[perm].contiguous(), dy[perm].contiguous()
            tot = 0.0
            for b0 in range(0, n, batch):
                xb, yb = xs[b0:b0 + batch], ys[b0:b0 + batch]
                self.ctx._check(self.lib.qf_mlp_train_step(self.ctx.h, self.h, _native._ptr(xb), _native._ptr(yb),
                                                           xb.shape[0], ctypes.c_float(lr), ctypes.byref(loss)))
                tot += loss.value * xb.shape[0]
            history.append(tot / n)
            if log and (ep % 10 == 0 or ep == epochs - 1):
                log("epoch %d loss %.4f" % (ep, history[-1]))
        return history


def q_error(pred_log2, true_log2):
    """max(est/true, true/est) of the row counts (log2(1 + rows) in, rows >= 1 assumed)."""
    return np.exp2(np.abs(np.asarray(pred_log2) - np.asarray(true_log2)))


def presize_rows(pred_log2, margin=1.0):
    """Buffer rows for a prediction: 2^(pred + margin) (margin in log2 units)."""
    return int(min(2 ** 31 - 1, max(16, math.ceil(2.0 ** (float(pred_log2) + margin)))))


def record_step(ctx, scale, seed, num_values=None, log=None, on_device=False):
    """Join instances of one refinement step (every literal of the step workload at `scale`,
    keys from `seed`): evaluates the step on the device and returns (features, targets, workload).
    on_device: the workload is generated on the GPU and stays there (features read row samples)."""
    import torch
    from . import datagen
    kw = {} if num_values is None else {"num_values": num_values}
    specs = datagen.step_relation_specs(seed=seed + 3)
    dev = torch.device("cuda")
    w = datagen.refinement_step(scale=scale, seed=seed, specs=specs, device=dev if on_device else None, **kw)
    if on_device:
        b, r = w["binding"], w["relations"]
    else:
        b = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in w["binding"]]
        r = [[torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in rel] for rel in w["relations"]]
    counts, _ = ctx.evaluate_literals(b, w["num_positive"], r, w["literals"])
    X, y = literal_instances(w["binding"], w["relations"], w["literals"], counts, {}, seed)
    if log:
        log("recorded %d joins (scale %g, seed %d)" % (len(y), scale, seed))
    return X, y, w


def record_micro(ctx, scale, seed, distribution, log=None):
    """Join instances of the operator microbench at `scale` (one relation of arity 2 against three
    binding variables, every literal placement), with the counts of the device evaluator."""
    import torch
    from . import datagen
    rows = max(1024, int(datagen.MICRO_ROWS * scale))
    w = datagen.microbench(distribution, rows=rows, rel_rows=max(256, int(datagen.MICRO_REL_ROWS * scale)),
                           num_values=max(64, int(datagen.MICRO_KEYS * scale)), seed=seed)
    literals = [(0, v) for v in datagen.all_placements(2, 3)]
    dev = torch.device("cuda")
    b = [torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in w["binding"]]
    r = [[torch.from_numpy(np.ascontiguousarray(c)).to(dev) for c in rel] for rel in w["relations"]]
    counts, _ = ctx.evaluate_literals(b, w["num_positive"], r, literals)
    X, y = literal_instances(w["binding"], w["relations"], literals, counts, {}, ("micro", seed))
    if log:
        log("recorded %d joins (microbench %s, scale %g, seed %d)" % (len(y), distribution, scale, seed))
    return X, y


def binding_join_features(binding, relations, literals, cache=None):
    """Features of CreateBindingTable joins (operations/MultiColumnHashJoin.cpp:26-115): binding
    table x literal, where variables below the binding's column count are join keys and the
    others new columns.  Columns may be device tensors (only samples are copied)."""
    C = len(binding)
    cache = {} if cache is None else cache
    X = []
    for rel, vars_ in literals:
        bound = [i for i, v in enumerate(vars_) if 0 <= v < C]
        bkeys = [binding[vars_[i]] for i in bound]
        rkeys = [relations[rel][i] for i in bound]
        ck = (("b", tuple(vars_[i] for i in bound)), ("r", rel, tuple(bound)))
        col_keys = [(("b", (vars_[i],)), ("r", rel, (i,))) for i in bound]
        X.append(join_features(bkeys, rkeys, len(vars_), len(vars_) - len(bound), cache, ck, col_keys))
    return np.stack(X) if X else np.zeros((0, NUM_FEATURES), np.float32)


def presized_joins(ctx, est, binding, relations, literals, margin=1.0, cache=None, budget_bytes=2 << 30):
    """CreateBindingTable for every literal with the estimator in the loop: predicted sizes give
    the capacities (2^(pred + margin) rows, at most the product of the input sizes), the joins run
    largest prediction first, as few calls as the output budget allows, and those whose rows
    exceed their capacity are rerun at their exact size.  Returns (outputs in literal order,
    stats)."""
    X = binding_join_features(binding, relations, literals, cache)
    pred = est.predict(X) if len(literals) else np.zeros(0, np.float32)
    C = len(binding)
    nb = _rows(binding[0])
    ncols = [C + sum(1 for v in vars_ if v < 0 or v >= C) for _, vars_ in literals]
    caps = [min(presize_rows(p, margin), max(16, nb * _rows(relations[rel][0])), max(16, budget_bytes // (8 * n)))
            for p, (rel, _), n in zip(pred, literals, ncols)]
    order = sorted(range(len(literals)), key=lambda i: -float(pred[i]))
    result = [None] * len(literals)
    exact = np.zeros(len(literals), np.int64)
    retry = []
    pos = 0
    while pos < len(order):  # calls of at most budget_bytes of outputs
        chunk, used = [], 0
        while pos < len(order) and (not chunk or used + 4 * caps[order[pos]] * ncols[order[pos]] <= budget_bytes):
            used += 4 * caps[order[pos]] * ncols[order[pos]]
            chunk.append(order[pos])
            pos += 1
        outs, rows, fits = ctx.create_binding_tables_presized(binding, relations, [literals[i] for i in chunk],
                                                              [caps[i] for i in chunk])
        for k, i in enumerate(chunk):
            exact[i] = int(rows[k])
            if fits[k]:
                result[i] = outs[k]
            else:
                retry.append(i)
    for i in retry:  # the exact fallback, one join at a time
        outs, rows, fits = ctx.create_binding_tables_presized(binding, relations, [literals[i]], [int(exact[i])])
        assert fits[0], "an exact-size retry must fit"
        result[i] = outs[0]
    return result, {"predicted_log2": pred, "capacities": caps, "rows": exact, "fallbacks": len(retry)}
