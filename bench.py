"""Benchmark of candidate-literal evaluation on one MI355X.

Workload (default): the batched refinement step -- 8,388,608 binding rows x 4 int32 variable
columns (30 % positive) against 20 background relations of arity 2-3, sizes log-uniform
10K..50M rows, Zipf s in {0, 0.8, 1.2}; all 2,080 candidate literals (every placement with a
bound variable).  A step = evaluating every literal's four counts from device-resident tables
(relation partitions cached, as the reference caches them on the fact tables).  `value` is
literals/sec of the batched form; the plain form (one launch per literal) is timed beside it.

`--workload micro` runs the operator microbench instead: 16,777,216 x 3 binding rows against
4,194,304 x 2 facts, keys over 1M values (uniform or Zipf 1.0), three literals.

Prints one JSON line (see the keys below).  --gpus above 1 is refused: this build runs on one
GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_probe_bytes(workload):
    """Per step: for each join group (relation, join column, clause variable), the relation's
    rows (8 B: join value + row id), the binding rows (8 B), 4 B per relation atom column and per
    binding atom column, and 32 B of counts per literal (the per-group formula used for the
    roofline; per-match random gathers are not counted)."""
    from quickfoil_amd import datagen  # noqa: F401
    rels = workload["relations"]
    nb = int(workload["binding"][0].numel())
    groups = {}
    for rel, vars_ in workload["literals"]:
        bound = [i for i, v in enumerate(vars_) if v >= 0]
        jk = bound[0]
        key = (rel, jk, vars_[jk])
        g = groups.setdefault(key, {"bg": set(), "bind": set(), "lits": 0})
        g["lits"] += 1
        for i in bound:
            if i != jk:
                g["bg"].add(i)
                g["bind"].add(vars_[i])
    total = 0
    for (rel, jk, var), g in groups.items():
        npr = int(rels[rel][0].numel())
        total += 8 * npr + 8 * nb + 4 * npr * len(g["bg"]) + 4 * nb * len(g["bind"]) + 32 * g["lits"]
    return total, len(groups)


def measured_traffic(workload):
    """HBM bytes per k_eval launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    this workload (profiles/k_eval_traffic.json), or None when none was recorded."""
    path = os.path.join(ROOT, "profiles", "k_eval_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return rec.get("traffic_bytes_per_launch") if rec else None


def cpu_baseline(workload, fraction, log_fn):
    """The oracle's restatement of the reference pipeline (1 thread) on the first `fraction`
    of the binding rows with every relation and literal."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nb = int(workload["binding"][0].numel())
    rows = max(64, int(nb * fraction))
    binding = [c[:rows].cpu().numpy() for c in workload["binding"]]
    npos = min(rows, int(rows * workload["num_positive"] / nb))
    rels = [[c.cpu().numpy() for c in r] for r in workload["relations"]]
    prob = oracle.Problem(binding, npos, rels, workload["literals"])
    t0 = time.perf_counter()
    prob.pipeline(radix_bits=5, chunk_size=32768)
    dt = time.perf_counter() - t0
    log_fn("cpu baseline: %d literals over %d binding rows in %.2f s" % (len(workload["literals"]), rows, dt))
    return dt, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["step", "micro"], default="step")
    ap.add_argument("--dist", choices=["uniform", "zipf"], default="zipf")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-fraction", type=float, default=1.0 / 256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, help="MiB of a device copy to run once")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import quickfoil_amd
    from quickfoil_amd import datagen

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    if args.workload == "step":
        w = datagen.refinement_step(device=dev)
        metric = "candidate literals evaluated/sec (batched refinement step)"
        workload_name = "refinement-step: 8388608x4 bindings (30% pos), 20 relations 10K-50M rows, 2080 literals"
    else:
        w = datagen.microbench(args.dist, device=dev)
        metric = "candidate literals evaluated/sec (operator microbench, %s)" % args.dist
        workload_name = "microbench-%s: 16777216x3 bindings (20%% pos), 4194304x2 facts, 1M keys, 3 literals" % args.dist
    torch.cuda.synchronize()
    rel_rows = [int(r[0].numel()) for r in w["relations"]]
    log("data ready in %.1f s: %d binding rows, %d relations (%d rows), %d literals" % (
        time.perf_counter() - t0, w["binding"][0].numel(), len(rel_rows), sum(rel_rows), len(w["literals"])))

    ctx = quickfoil_amd.Context(0)
    ctx.enable_timing(True)
    L = len(w["literals"])

    def step(mode):
        counts, _ = ctx.evaluate_literals(w["binding"], w["num_positive"], w["relations"], w["literals"], mode=mode,
                                          reuse_partitions=True, radix_bits=args.radix_bits, tile_rows=args.tile_rows,
                                          use_graph=(mode == "batched" and not args.no_graph))
        return counts

    # Warm-up (first step partitions and caches the relations: the cold cost).
    tc = time.perf_counter()
    ref_counts = step("batched")
    torch.cuda.synchronize()
    cold_s = time.perf_counter() - tc
    log("cold batched step: %.3f s" % cold_s)
    for _ in range(max(0, args.warmup - 1)):
        step("batched")
    torch.cuda.synchronize()
    ctx.reset_timing()

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        counts = step("batched")
        log("batched step %d done" % (i + 1))
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t_start) / args.steps
    assert np.array_equal(counts, ref_counts), "batched counts changed between steps"
    # Per-kernel times: HIP events around every kernel launch of the same steps run eagerly
    # (events cannot be timed inside a graph replay); the rocprof trace of this command
    # cross-checks them.
    ctx.reset_timing()
    for i in range(args.steps):
        eager_counts, _ = ctx.evaluate_literals(w["binding"], w["num_positive"], w["relations"], w["literals"],
                                                mode="batched", reuse_partitions=True, radix_bits=args.radix_bits,
                                                tile_rows=args.tile_rows)
    torch.cuda.synchronize()
    assert np.array_equal(eager_counts, ref_counts), "eager and graph counts differ"
    phases = ctx.phase_times()

    # Plain form: one launch per literal, same partitions.
    plain_s = None
    if args.plain_steps > 0:
        ctx.reset_timing()
        step("plain_literal")  # warm its code path
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for i in range(args.plain_steps):
            plain_counts = step("plain_literal")
            log("plain step %d done" % (i + 1))
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t_start) / args.plain_steps
        assert np.array_equal(plain_counts, ref_counts), "plain and batched counts differ"

    if args.copy_calibration:
        # A plain device copy of known size, for scaling rocprof FETCH_SIZE / WRITE_SIZE.
        src = torch.empty(args.copy_calibration * (1 << 20) // 4, dtype=torch.int32, device=dev).fill_(1)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB" % args.copy_calibration)

    algo_bytes, num_groups = algorithmic_probe_bytes(w)
    probe_ms, probe_launches = phases.get("probe_count", (0.0, 0))
    probe_ms_per_step = probe_ms / max(1, args.steps)
    achieved = algo_bytes / (probe_ms_per_step * 1e-3) / 1e9 if probe_ms_per_step > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        dt, rows = cpu_baseline(w, args.cpu_fraction, log)
        nb = int(w["binding"][0].numel())
        cpu = {"value": L / (dt * nb / rows), "unit": "literals/s", "cores": 1, "kind": "port",
               "sample": "oracle pipeline (reference algorithm, 1 thread) on the first %d of %d binding rows with "
                         "all relations and literals: %.2f s; value scales that time by %d/%d rows"
                         % (rows, nb, dt, nb, rows)}

    out = {
        "metric": metric,
        "value": L / batched_s,
        "unit": "literals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": workload_name, "state": "warm (relation partitions cached)",
                   "literals": L, "join_groups": num_groups},
        "plain_literals_per_s": (L / plain_s) if plain_s else None,
        "plain_ms_per_step": (plain_s * 1e3) if plain_s else None,
        "batched_over_plain": (plain_s / batched_s) if plain_s else None,
        "cold_step_ms": cold_s * 1e3,
        "phase_ms_per_step": {k: v[0] / max(1, args.steps) for k, v in phases.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload),
                     "kernel": "k_eval (probe and count; HIP events around each launch on the context stream)",
                     "algorithmic_bytes_per_step": algo_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
