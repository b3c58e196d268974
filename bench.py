"""Benchmark of candidate-literal evaluation on one MI355X.

Workload (default): the batched refinement step -- 8,388,608 binding rows x 4 int32 variable
columns (30 % positive) against 20 background relations of arity 2-3, sizes log-uniform
10K..50M rows, Zipf s in {0, 0.8, 1.2}; all 2,080 candidate literals (every placement with a
bound variable).  A step = evaluating every literal's four counts from device-resident tables
(relation partitions cached, as the reference caches them on the fact tables).  `value` is
literals/sec of the batched form; the plain form (one launch per literal) is timed beside it.

`--workload micro` runs the operator microbench instead: 16,777,216 x 3 binding rows against
4,194,304 x 2 facts, keys over 1M values (uniform or Zipf 1.0), three literals.

`--workload learn-kinship` / `learn-bongard` time end-to-end rule learning (host learner over the
device backend) on the synthetic kinship forest (2,000 persons, clause length 3) and Bongard
pictures (200,000, clause length 4), with the oracle backend timed beside it and the learnt
clauses compared.  `--workload estimator` records join instances from refinement steps, trains
the cardinality MLP on the device and reports q-error, presize fallback rate and the time of
presized joins against the two-pass joins.

Prints one JSON line (see the keys below).  --gpus above 1 is refused: the library runs on one
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


def algorithmic_partition_bytes(workload):
    """Per step: the binding rows are partitioned once per clause join variable, reading every
    referenced binding column once and writing it plus the row id once (4 B each)."""
    nb = int(workload["binding"][0].numel())
    join_vars, cols = set(), set()
    for _, vars_ in workload["literals"]:
        bound = [v for v in vars_ if v >= 0]
        join_vars.add(bound[0])
        cols.update(bound)
    return len(join_vars) * nb * 4 * (2 * len(cols) + 1)


def measured_traffic(workload, kernel="k_stream"):
    """HBM bytes per launch of `kernel` (per step for "binding_partition") from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload (profiles/<kernel>_traffic.json), or
    None when none was recorded."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.json" % kernel)
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return rec.get("traffic_bytes_per_launch") if rec else None


def cpu_baseline(workload, fractions, reps, log_fn, grouped=False):
    """The oracle's restatement of the reference pipeline on one pinned host core, warm: the
    relations' partitions are made before the timer starts, as the reference keeps them on the
    fact tables across refinement steps (learner/CandidateLiteralEvaluator.cpp:291-299).  Each
    of `fractions` (of the binding rows, every relation and literal) is timed `reps` times and
    the median kept.  With two fractions the full step is extrapolated as fixed + slope * rows
    (the relation rows' probes do not depend on the binding rows, the matches grow with them);
    with one fraction of 1.0 the median is the full step.  `grouped`: the oracle's grouped-count
    closed form instead, for inputs whose matching pairs the materialising pipeline cannot hold.
    Returns (seconds per full step, description)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nb = int(workload["binding"][0].numel())
    rels = [[c.cpu().numpy() for c in r] for r in workload["relations"]]
    binding_all = [c.cpu().numpy() for c in workload["binding"]]
    probs, rows_list = [], []
    for f in fractions:
        rows = min(nb, max(64, int(nb * f)))
        npos = min(rows, int(rows * workload["num_positive"] / nb))
        probs.append(oracle.Problem([c[:rows] for c in binding_all], npos, rels, workload["literals"]))
        rows_list.append(rows)
    # One pinned host core, as the reference runs single-threaded.
    saved = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(saved)})
    try:
        t0 = time.perf_counter()
        cache = None if grouped else oracle.RelationCache(probs[0])
        part_s = time.perf_counter() - t0
        medians = []
        for prob, rows in zip(probs, rows_list):
            times = []
            for k in range(reps):
                t0 = time.perf_counter()
                if grouped:
                    prob.grouped()
                else:
                    prob.pipeline_cached(cache)
                times.append(time.perf_counter() - t0)
                log_fn("cpu baseline: %d binding rows, run %d/%d: %.2f s" % (rows, k + 1, reps, times[-1]))
            medians.append(float(np.median(times)))
    finally:
        os.sched_setaffinity(0, saved)
        if cache is not None:
            cache.close()
    method = ("oracle grouped-count closed form (the reference's materialising pipeline cannot hold this input's "
              "5.6e11 matching pairs)" if grouped else "oracle pipeline (reference algorithm)")
    if len(fractions) == 1:
        full = medians[0] * nb / rows_list[0]
        desc = ("%s, 1 thread pinned to one core, relation partitions made before timing (%.2f s, not counted); "
                "median of %d runs on %d of %d binding rows with all relations and literals: %.2f s"
                % (method, part_s, reps, rows_list[0], nb, medians[0]))
        if rows_list[0] != nb:
            desc += "; scaled by %d/%d rows" % (nb, rows_list[0])
        return full, desc
    slope = (medians[1] - medians[0]) / (rows_list[1] - rows_list[0])
    fixed = medians[0] - slope * rows_list[0]
    full = fixed + slope * nb
    desc = ("%s, 1 thread pinned to one core, relation partitions made before timing (%.2f s, not counted); "
            "medians of %d runs with all relations and literals on %d and %d of %d binding rows: %.2f s and %.2f s; "
            "full step extrapolated as fixed + slope * rows = %.2f s + %.3e s/row * %d = %.1f s"
            % (method, part_s, reps, rows_list[0], rows_list[1], nb, medians[0], medians[1], fixed, slope, nb, full))
    return full, desc


def host_cpu():
    """The host's CPU model and the number of CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0))}


def copy_ceiling(torch, dev, mib=1024, reps=5):
    """GB/s of a plain device-to-device copy (read + write bytes), best of `reps`."""
    src = torch.empty(mib * (1 << 20) // 4, dtype=torch.int32, device=dev).fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = 0.0
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = max(best, 2 * src.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del src, dst
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["step", "micro", "learn-kinship", "learn-bongard", "estimator"],
                    default="step")
    ap.add_argument("--dist", choices=["uniform", "zipf"], default="zipf")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--fresh-steps", type=int, default=3,
                    help="steps on a new binding table each (relations cached, host planning and eager launch)")
    ap.add_argument("--extend-literals", type=int, default=4,
                    help="literals whose binding-set extension (CreateLabelAwareBindingTables) is timed")
    ap.add_argument("--cpu-fractions", type=float, nargs="+", default=None,
                    help="binding-row fractions the CPU baseline is timed at (default: 1/256 and 1/64 of the "
                         "refinement step, extrapolated to the full step; all of the microbench)")
    ap.add_argument("--cpu-reps", type=int, default=5, help="runs per CPU fraction (the median is kept)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="fixed 512 facts per partition and 1024-row tiles instead of the autotuner's choice")
    ap.add_argument("--copy-calibration", type=int, default=0, help="MiB of a device copy to run once")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (the library has no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import quickfoil_amd
    from quickfoil_amd import datagen

    if args.workload.startswith("learn-"):
        return bench_learner(args)
    if args.workload == "estimator":
        return bench_estimator(args)

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
    # The literal list and the relations' table descriptors in the C layout, made once (a host
    # learner hands the library such arrays).
    lits = quickfoil_amd.LiteralArray(w["literals"])
    rels = quickfoil_amd.TableSet(w["relations"])

    # The batched form takes its facts per partition and tile from the autotuner (it times a grid
    # of both on the first call for this workload shape); the plain form reuses the choice.
    tune = {"radix_bits": args.radix_bits, "tile_rows": args.tile_rows, "partition_rows": 0}

    def step(mode):
        auto = mode == "batched" and not args.no_autotune and not args.radix_bits and not args.tile_rows
        counts, _ = ctx.evaluate_literals(w["binding"], w["num_positive"], rels, lits, mode=mode,
                                          reuse_partitions=True, radix_bits=tune["radix_bits"],
                                          tile_rows=0 if auto else tune["tile_rows"],
                                          partition_rows=0 if auto else tune["partition_rows"],
                                          use_graph=(mode == "batched" and not args.no_graph), autotune=auto)
        return counts

    # Cold step: nothing cached, default facts per partition and tile (partitions the relations).
    tc = time.perf_counter()
    ref_counts, _ = ctx.evaluate_literals(w["binding"], w["num_positive"], rels, lits,
                                          mode="batched", reuse_partitions=True, radix_bits=args.radix_bits,
                                          tile_rows=args.tile_rows)
    torch.cuda.synchronize()
    cold_s = time.perf_counter() - tc
    log("cold batched step: %.3f s" % cold_s)
    autotune_s = None
    if not args.no_autotune and not args.radix_bits and not args.tile_rows:
        ta = time.perf_counter()
        tuned_counts = step("batched")  # times the autotuner's grid, then runs the choice
        torch.cuda.synchronize()
        autotune_s = time.perf_counter() - ta
        assert np.array_equal(tuned_counts, ref_counts), "autotuned counts differ"
        choice = ctx.autotune_choice()
        pr, tr, best_ms = choice["partition_rows"], choice["tile_rows"], choice["best_ms"]
        tune.update(partition_rows=pr, tile_rows=tr)
        log("autotuner: %d facts per partition, %d-row tiles (%.2f ms)" % (pr, tr, best_ms))
    # Eager calls (no graph) with the configuration the graph replays use: the autotuned one
    # (looked up in the context's autotuner cache, which also holds whether the streamed form won)
    # or the one given.
    auto_cfg = not args.no_autotune and not args.radix_bits and not args.tile_rows
    eager_kw = (dict(autotune=True) if auto_cfg else
                dict(radix_bits=tune["radix_bits"], tile_rows=tune["tile_rows"], partition_rows=tune["partition_rows"]))
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
        eager_counts, _ = ctx.evaluate_literals(w["binding"], w["num_positive"], rels, lits,
                                                mode="batched", reuse_partitions=True, **eager_kw)
    torch.cuda.synchronize()
    assert np.array_equal(eager_counts, ref_counts), "eager and graph counts differ"
    phases = ctx.phase_times()

    # A learner's step: a new binding table (new device columns) against cached relations, so
    # host planning, the binding partition and an eager launch run every time (no graph replay).
    fresh_s = None
    if args.fresh_steps > 0:
        times = []
        for i in range(args.fresh_steps):
            fresh = [c.clone() for c in w["binding"]]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fresh_counts, _ = ctx.evaluate_literals(fresh, w["num_positive"], rels, lits, mode="batched",
                                                    reuse_partitions=True, **eager_kw)
            times.append(time.perf_counter() - t0)
            assert np.array_equal(fresh_counts, ref_counts), "counts on a new binding table differ"
            del fresh
        fresh_s = float(np.median(times))
        log("new-binding-table step: %.2f ms (median of %d)" % (fresh_s * 1e3, len(times)))

    # Plain form: one launch per literal, same partitions; its kernels' own time from the phase
    # events beside the wall time (which includes one host planning pass per literal group).
    plain_s = None
    plain_kernel_ms = None
    if args.plain_steps > 0:
        step("plain_literal")  # warm its code path
        torch.cuda.synchronize()
        ctx.reset_timing()
        t_start = time.perf_counter()
        for i in range(args.plain_steps):
            plain_counts = step("plain_literal")
            log("plain step %d done" % (i + 1))
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t_start) / args.plain_steps
        plain_kernel_ms = sum(v[0] for v in ctx.phase_times().values()) / args.plain_steps
        assert np.array_equal(plain_counts, ref_counts), "plain and batched counts differ"

    # Binding-set extension (CreateLabelAwareBindingTables) for a few literals whose new binding
    # tables are large, into preallocated outputs.
    extension = None
    if args.extend_literals > 0 and args.workload == "step":
        sizes = ref_counts[:, 2].astype(np.int64) + ref_counts[:, 3].astype(np.int64)
        cand = [i for i in np.argsort(-sizes) if sizes[i] <= 40_000_000][: args.extend_literals]
        rows_max = int(max(sizes[i] for i in cand)) if cand else 0
        arena = [torch.empty(max(1, rows_max), dtype=torch.int32, device=dev) for _ in range(len(w["binding"]) + 3)]
        ext_ms, ext_rows = 0.0, 0
        for i in cand:
            rel, vars_ = w["literals"][i]
            args_ext = (w["binding"], w["num_positive"], w["relations"][rel], vars_, int(ref_counts[i, 2]),
                        int(ref_counts[i, 3]))
            ctx.create_label_aware_binding_table(*args_ext, out=arena)  # warm
            torch.cuda.synchronize()
            reps = []
            for _ in range(3):
                t0 = time.perf_counter()
                ctx.create_label_aware_binding_table(*args_ext, out=arena)
                torch.cuda.synchronize()
                reps.append(time.perf_counter() - t0)
            ext_ms += float(np.median(reps)) * 1e3
            ext_rows += int(sizes[i])
        del arena
        extension = {"literals": len(cand), "rows": ext_rows, "ms": ext_ms,
                     "rows_per_s": ext_rows / (ext_ms * 1e-3) if ext_ms > 0 else None}
        log("extension: %d literals, %d rows in %.2f ms" % (len(cand), ext_rows, ext_ms))

    if args.copy_calibration:
        # A plain device copy of known size, for scaling rocprof FETCH_SIZE / WRITE_SIZE.
        src = torch.empty(args.copy_calibration * (1 << 20) // 4, dtype=torch.int32, device=dev).fill_(1)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB" % args.copy_calibration)

    copy_gbs = copy_ceiling(torch, dev)
    log("device copy ceiling: %.0f GB/s" % copy_gbs)
    algo_bytes, num_groups = algorithmic_probe_bytes(w)
    part_bytes = algorithmic_partition_bytes(w)
    part_ms_per_step = phases.get("partition_binding", (0.0, 0))[0] / max(1, args.steps)
    part_achieved = part_bytes / (part_ms_per_step * 1e-3) / 1e9 if part_ms_per_step > 0 else 0.0
    # The probe phase: k_stream counts every unit in the cells whose binding rows fit its work
    # items, k_eval the rest (cells of heavily repeated join values); the roofline takes the
    # step's algorithmic bytes over both kernels' time.
    stream_ms_per_step = phases.get("probe_count_stream", (0.0, 0))[0] / max(1, args.steps)
    eval_ms_per_step = phases.get("probe_count", (0.0, 0))[0] / max(1, args.steps)
    probe_ms_per_step = stream_ms_per_step + eval_ms_per_step
    achieved = algo_bytes / (probe_ms_per_step * 1e-3) / 1e9 if probe_ms_per_step > 0 else 0.0
    traffic = measured_traffic(args.workload)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_stream + k_eval (probe and count of all %d join groups; HIP events around each "
                          "launch on the context stream)" % num_groups,
                "algorithmic_bytes_per_step": algo_bytes, "k_stream_ms_per_step": stream_ms_per_step,
                "k_eval_ms_per_step": eval_ms_per_step}
    if traffic:
        # The per-join-group formula counts the binding rows once per join group (224 here); the
        # streamed form stages each binding cell once for all of them, so its measured HBM traffic
        # (rocprofv3 FETCH_SIZE + WRITE_SIZE per k_stream launch) is far below the formula and
        # `frac` can exceed 1.  The bandwidth the kernel actually draws is traffic / time.
        roofline["traffic_GBs"] = traffic / (probe_ms_per_step * 1e-3) / 1e9 if probe_ms_per_step > 0 else 0.0
        roofline["traffic_frac"] = roofline["traffic_GBs"] / HBM_PEAK_GBS

    part_traffic = measured_traffic(args.workload, kernel="binding_partition")
    roofline_part = {"bound": "hbm", "achieved": part_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": part_achieved / HBM_PEAK_GBS, "traffic": part_traffic,
                     "kernel": "binding radix partition of the clause join variables (histograms, scans, scatter "
                               "passes, offsets; HIP events around the phase)",
                     "algorithmic_bytes_per_step": part_bytes}
    if part_traffic:
        # rocprofv3 FETCH_SIZE + WRITE_SIZE summed over the partition's kernels of one step: both LSD
        # passes and their histograms, against the formula's single read and write per variable.
        roofline_part["traffic_GBs"] = part_traffic / (part_ms_per_step * 1e-3) / 1e9 if part_ms_per_step > 0 else 0.0
        roofline_part["traffic_frac"] = roofline_part["traffic_GBs"] / HBM_PEAK_GBS

    cpu = None
    if not args.no_cpu:
        fracs = args.cpu_fractions or ([1.0 / 256, 1.0 / 64] if args.workload == "step" else [1.0])
        grouped = args.workload == "micro" and args.dist == "zipf"
        full_s, desc = cpu_baseline(w, fracs, max(1, args.cpu_reps), log, grouped=grouped)
        cpu = {"value": L / full_s, "unit": "literals/s", "cores": 1, "kind": "port", **host_cpu(),
               "seconds_per_step": full_s, "sample": desc}

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
                   "literals": L, "join_groups": num_groups,
                   "facts_per_partition": tune["partition_rows"] or 256, "tile_rows": tune["tile_rows"] or 1024},
        "plain_literals_per_s": (L / plain_s) if plain_s else None,
        "plain_ms_per_step": (plain_s * 1e3) if plain_s else None,
        "plain_kernel_ms_per_step": plain_kernel_ms,
        "plain_steps": args.plain_steps,
        "batched_over_plain": (plain_s / batched_s) if plain_s else None,
        "fresh_binding_ms_per_step": (fresh_s * 1e3) if fresh_s else None,
        "fresh_binding_literals_per_s": (L / fresh_s) if fresh_s else None,
        "extension": extension,
        "cold_step_ms": cold_s * 1e3,
        "autotune_ms": (autotune_s * 1e3) if autotune_s is not None else None,
        "copy_ceiling_GBs": copy_gbs,
        "phase_ms_per_step": {k: v[0] / max(1, args.steps) for k, v in phases.items()},
        "roofline": roofline,
        "probe_phase_ms_per_step": probe_ms_per_step,
        "roofline_partition": roofline_part,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_learner(args):
    """End-to-end QuickFoil learning on a synthetic problem, device backend vs the oracle backend
    (one host thread); clauses must be identical."""
    import ctypes
    import quickfoil_amd
    from quickfoil_amd import datagen, learner
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if args.workload == "learn-kinship":
        k = datagen.kinship()
        r = k["relations"]
        bg = [learner.Relation(n, r[n], [0] * len(r[n])) for n in ["parent", "married", "male", "female", "brother",
                                                                  "sister"]]
        problem = learner.Problem(learner.Relation("uncle", k["target"], [0, 0]), k["num_positive"], bg)
        opts, name = {"maximum_clause_length": 3}, "kinship: 2000 persons, uncle/2, clause length 3"
    else:
        b = datagen.bongard(200000)
        bg = [learner.Relation(n, c, t) for n, (c, t) in b["relations"].items()]
        problem = learner.Problem(learner.Relation("picture", b["target"], [0]), b["num_positive"], bg)
        opts, name = {"maximum_clause_length": 4}, "bongard: 200000 pictures, 7 relations, clause length 4"
    options = learner.default_options(**opts)
    ctx = quickfoil_amd.Context(0)
    dev = learner.DeviceBackend(ctx, problem)
    learner.learn(problem, dev.backend, options)  # warm-up (code paths, pools)
    times = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        clauses = learner.learn(problem, dev.backend, options)
        times.append(time.perf_counter() - t0)
    dev.close()
    lib = oracle.lib()
    n, ar, cols, rows, types, keep = problem.relation_arrays()
    cpu = learner.LearnerBackend()
    lib.orc_learner_backend_create(n, ar, cols, rows, types, ctypes.byref(cpu))
    t0 = time.perf_counter()
    cpu_clauses = learner.learn(problem, cpu, options)
    cpu_s = time.perf_counter() - t0
    gpu_s = min(times)
    out = {"metric": "end-to-end rule learning time", "value": gpu_s, "unit": "s", "n_gpus": 1, "steps": len(times),
           "warmup": 1, "higher_is_better": False, "dtype": "int32", "data": "synthetic (seeded generators)",
           "config": {"workload": name}, "clauses": clauses, "identical_to_oracle_backend": clauses == cpu_clauses,
           "cpu_baseline": {"value": cpu_s, "unit": "s", "cores": 1, "kind": "port",
                            "sample": "the same learner over the oracle backend (reference pipeline restated, "
                                      "1 thread), full problem"}}
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_estimator(args):
    """c5: train the cardinality MLP on joins recorded from refinement steps (1/64 scale, varied
    seeds) and operator microbenches, report its q-error on held-out steps beside the textbook
    independence estimate, then use it in the loop at the full refinement step: CreateBindingTable
    joins of the step's literals presized by the predictions in one call (exact fallback for
    under-predictions) against the two-pass joins, outputs asserted identical."""
    import torch
    import quickfoil_amd
    from quickfoil_amd import datagen, estimator
    ctx = quickfoil_amd.Context(0)
    t0 = time.perf_counter()
    # Steps at four scales (1/64 to the full 8,388,608-row step, six seeds each: 49,920 joins) and
    # operator microbenches at two scales and both key distributions.
    scales = [1.0 / 64, 1.0 / 16, 1.0 / 4, 1.0]
    train = [estimator.record_step(ctx, sc, 100 + 10 * k + i, log=log, on_device=True)[:2]
             for k, sc in enumerate(scales) for i in range(6)]
    train += [estimator.record_micro(ctx, sc, 0x51F00000 + 97 * i, d, log=log)
              for i, sc in enumerate([1.0 / 64, 1.0 / 16] * 4) for d in ("uniform", "zipf")]
    X = np.concatenate([t[0] for t in train])
    y = np.concatenate([t[1] for t in train])
    torch.cuda.empty_cache()
    held = [estimator.record_step(ctx, sc, 900 + i, log=log, on_device=True)[:2]
            for i, sc in enumerate([1.0 / 64, 1.0 / 4])]
    Xh = np.concatenate([h[0] for h in held])
    yh = np.concatenate([h[1] for h in held])
    torch.cuda.empty_cache()
    rec_s = time.perf_counter() - t0
    est = estimator.Estimator(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    est.fit(X, y, epochs=60, batch=512, lr=3e-3, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    q = estimator.q_error(est.predict(Xh), yh)
    qi = estimator.q_error(estimator.independence_estimate(Xh), yh)

    # In the loop, full refinement step: the step's exact counts (one batched evaluation) pick
    # which literals' binding tables to build and judge the predictions; the joins themselves
    # see only the estimator.
    dev = torch.device("cuda:0")
    w = datagen.refinement_step(device=dev)
    lits = w["literals"]
    counts, _ = ctx.evaluate_literals(w["binding"], w["num_positive"], w["relations"], lits, reuse_partitions=True)
    rows_of = counts[:, 2].astype(np.int64) + counts[:, 3].astype(np.int64)
    mid = [i for i in range(len(lits)) if (1 << 10) <= rows_of[i] <= (1 << 22)]
    wide = [i for i in range(len(lits)) if sum(1 for v in lits[i][1] if v >= 0) >= 2]
    picks = mid[:: max(1, len(mid) // 36)][:36] + wide[:: max(1, len(wide) // 12)][:12]
    joins = [lits[i] for i in picks]
    log("in-loop joins: %d literals, %d output rows" % (len(joins), int(rows_of[picks].sum())))
    binding, rels = w["binding"], w["relations"]
    # Warm both code paths.
    estimator.presized_joins(ctx, est, binding, rels, joins[:2])
    ctx.create_binding_table(binding, rels[joins[0][0]], joins[0][1])
    torch.cuda.synchronize()
    arena_rows = int(max(1, rows_of[picks].max()))
    arena = [torch.empty(arena_rows, dtype=torch.int32, device=dev) for _ in range(len(binding) + 3)]
    t0 = time.perf_counter()
    for rel, vars_ in joins:
        ctx.create_binding_table(binding, rels[rel], vars_, out=arena)
    torch.cuda.synchronize()
    two_pass_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    got, stats = estimator.presized_joins(ctx, est, binding, rels, joins)
    torch.cuda.synchronize()
    assisted_s = time.perf_counter() - t0
    identical = True
    for (rel, vars_), g in zip(joins, got):
        want = ctx.create_binding_table(binding, rels[rel], vars_)
        identical = identical and len(g) == len(want) and all(torch.equal(a, b) for a, b in zip(g, want))
    assert identical, "presized joins differ from the two-pass joins"
    assert np.array_equal(stats["rows"], rows_of[picks]), "join sizes differ from the evaluator's counts"
    y_loop = np.log2(1.0 + rows_of[picks].astype(np.float64))
    q_loop = estimator.q_error(stats["predicted_log2"], y_loop)
    q_loop_ind = estimator.q_error(estimator.independence_estimate(
        estimator.binding_join_features(binding, rels, joins)), y_loop)
    out = {"metric": "join-cardinality estimator q-error (held-out joins) and in-loop presized joins",
           "value": float(np.median(q)), "unit": "q-error p50", "n_gpus": 1, "higher_is_better": False,
           "dtype": "bf16 (fp32 accumulate)",
           "data": "synthetic (joins recorded from refinement steps and operator microbenches; in-loop joins on "
                   "the full refinement step, whose seed no training step uses)",
           "config": {"workload": "estimator: MLP 40-64-64-1, %d training joins (24 refinement steps at 1/64 to full "
                                  "scale, 16 microbenches), %d held-out (steps at 1/64 and 1/4 scale); %d in-loop "
                                  "CreateBindingTable joins of the full refinement step"
                                  % (len(y), len(yh), len(joins))},
           "q_error_p50": float(np.median(q)), "q_error_p90": float(np.percentile(q, 90)),
           "q_error_p99": float(np.percentile(q, 99)),
           "independence_q_error_p50": float(np.median(qi)), "independence_q_error_p90": float(np.percentile(qi, 90)),
           "independence_q_error_p99": float(np.percentile(qi, 99)),
           "in_loop": {"joins": len(joins), "output_rows": int(rows_of[picks].sum()),
                       "fallbacks": stats["fallbacks"], "fallback_rate": stats["fallbacks"] / max(1, len(joins)),
                       "q_error_p50": float(np.median(q_loop)), "q_error_p99": float(np.percentile(q_loop, 99)),
                       "independence_q_error_p50": float(np.median(q_loop_ind)),
                       "independence_q_error_p99": float(np.percentile(q_loop_ind, 99)),
                       "two_pass_s": two_pass_s, "assisted_s": assisted_s,
                       "speedup": two_pass_s / assisted_s if assisted_s > 0 else None,
                       "identical_outputs": identical},
           "record_s": rec_s, "train_s": train_s}
    print(json.dumps(out), flush=True)
    est.close()
    ctx.close()


if __name__ == "__main__":
    main()
