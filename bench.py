"""Benchmark of the MI355X triangle-counting path; prints ONE JSON line.

Headline (``value``): TC edges/s on RMAT scale 22, edge factor 16
(a,b,c = .57,.19,.19), seed 1, whole graph on one GPU -- one step = one
tc_count call (degree binning + persistent count kernel + u64 copy back) on
the DAG already resident in HBM.  Beside it:

* roofline: the count kernels' algorithmic bytes -- what the partner-scheme
  count reads with no reuse: per work unit its 16-byte row record, 16 bytes of
  row offsets and the owner's list, per DAG edge an 8-byte list reference and
  the shorter list 4 min(d+(u), d+(v)) -- divided by their average duration
  from HIP events on the library's stream, against 8 TB/s; the reference
  algorithm's edge-iterator bytes 8(R+1) + 20 m_R + 4 sum d+(v) over the same
  time are reported beside it, and so is the device's practical read ceiling
  (a streaming read of 4 GiB, measured in the same run);
* cpu_baseline: the CPU oracle (OpenMP merge count, the reference's CPU
  algorithm) on the same graph -- the whole graph when a run fits the CPU
  budget, else a random sample of pivot rows -- in a child process with
  OMP_PROC_BIND=close, median of 3 runs;
* extra legs: the plain one-wave-per-pivot kernel, the 4096-graph batched
  workload (one captured persistent launch, default and autotuned, vs one
  plain launch per graph), the serving loop under Poisson arrivals, and
  partition 0 of RMAT scale 24 split 4 ways with its halo cache, with every
  part of the reference's vertex split and of a partner-balanced split.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


# Time and text of the last progress line (read by the side-leg watchdog).
_progress = [time.monotonic(), "start"]
_side_legs_done = threading.Event()


def log(msg: str) -> None:
    _progress[0], _progress[1] = time.monotonic(), msg
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def start_watchdog(line: dict, limit_s: float) -> None:
    """Once the headline measurement is in `line`: if the side legs print
    nothing for limit_s seconds, print the line (the headline figures, with
    extra.side_legs_incomplete naming the last step reached) and end the
    process, so a stalled side leg does not take the headline record with it."""
    def watch():
        while True:
            time.sleep(5.0)
            silent = time.monotonic() - _progress[0]
            if silent > limit_s and not _side_legs_done.is_set():
                line.setdefault("extra", {})["side_legs_incomplete"] = {
                    "after": _progress[1], "silent_s": round(silent, 1)}
                print(f"[bench] side legs silent for {silent:.0f} s after: {_progress[1]}; "
                      f"printing the headline line", file=sys.stderr, flush=True)
                print(json.dumps(line), flush=True)
                os._exit(0)

    threading.Thread(target=watch, daemon=True).start()


def device_sync() -> None:
    """hipDeviceSynchronize: the library's calls already end in a stream sync."""
    try:
        lib = ctypes.CDLL("libamdhip64.so")
        lib.hipDeviceSynchronize()
    except OSError:
        pass


def cores_available() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_steps(fn, steps: int, warmup: int):
    for _ in range(warmup):
        fn()
    device_sync()
    t0 = time.perf_counter()
    outs = [fn() for _ in range(steps)]
    device_sync()
    return (time.perf_counter() - t0) * 1e3, outs


def physical_cores() -> int:
    """Distinct (socket, core) pairs of the host in /proc/cpuinfo."""
    pairs, phys = set(), None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    phys = line.split(":", 1)[1].strip()
                elif line.startswith("core id"):
                    pairs.add((phys, line.split(":", 1)[1].strip()))
    except OSError:
        pass
    return len(pairs) or (os.cpu_count() or 1)


def cpu_baseline_child(scale: int, seed: int, target_s: float, runs: int = 3, sample_seed: int = 7) -> dict:
    """Runs in a child process started with OMP_PROC_BIND=close (so the
    binding never reaches the GPU process's server and client threads): the
    oracle's OpenMP merge count -- the reference's CPU algorithm,
    graph.cc:376-391 inside the tc_omp_base loop, dynamic schedule of 1024
    pivots -- on the same RMAT DAG, median of `runs` timed runs.  The whole
    graph when one run fits target_s, else a random sample of pivot rows
    sized to it (dynamic schedule of 64 rows)."""
    # the thread count first: with OMP_PROC_BIND set, the first OpenMP region
    # (graph generation below) binds this thread to one CPU, and the
    # affinity mask no longer shows the CPUs the process may use
    threads = cores_available()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the timed CPU baseline
    from hetegeneous_parallel_programming_amd import Graph

    g = Graph.rmat_scale(scale, 16, seed)
    g.orientation()
    rp, ci = g.out_rowptr(), g.out_colidx()
    n = g.V()
    rng = np.random.default_rng(sample_seed)
    frac = 0.01
    rows = np.sort(rng.choice(n, max(1, int(n * frac)), replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    oracle.tc_count_rows(rp, ci, rows, threads=threads)
    est_whole = (time.perf_counter() - t0) / frac
    times = []
    if est_whole <= target_s:
        edges, what = g.E(), f"the whole RMAT-{scale} DAG ({n} rows, {g.E()} DAG edges)"
        count = None
        for _ in range(runs):
            t0 = time.perf_counter()
            count = oracle.tc_count(rp, ci, chunk=1024, threads=threads)
            times.append(time.perf_counter() - t0)
    else:
        frac = target_s / est_whole
        rows = np.sort(rng.choice(n, max(1, int(n * frac)), replace=False)).astype(np.int32)
        edges = int((rp[rows + 1] - rp[rows]).sum())
        what = f"{rows.size} random pivot rows ({frac:.4%} of {n}, {edges} DAG edges) of the RMAT-{scale} DAG"
        count = None
        for _ in range(runs):
            t0 = time.perf_counter()
            count = oracle.tc_count_rows(rp, ci, rows, chunk=64, threads=threads)
            times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    return {"value": edges / t, "unit": "edges/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "physical_cores_host": physical_cores(),
            "logical_cpus_host": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "runs_s": times, "count": int(count), "whole_graph": est_whole <= target_s,
            "sample": f"{what}: oracle OpenMP merge count on {threads} threads (the CPUs this process may "
                      f"use; OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}), median of {runs} runs = {t:.2f} s"}


def cpu_baseline(args) -> dict:
    """The CPU baseline in a child process (see cpu_baseline_child); the
    child builds the same graph from the same seed and never touches the GPU."""
    import subprocess

    env = dict(os.environ, OMP_PROC_BIND="close", OMP_NUM_THREADS=str(cores_available()))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", "--scale", str(args.scale),
           "--seed", str(args.seed), "--cpu-seconds", str(args.cpu_seconds)]
    p = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({p.returncode}): {p.stderr[-2000:]}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def leg_er(ctx, seed, steps, warmup):
    """ER G(n=16384, avg degree 16): tc_count (which takes the single plain
    launch at this size), the binned count (its plan kept after the first
    call) and the plain kernel on the resident DAG, beside the CPU oracle's
    count of the same DAG on all host cores (median of 3); and TCSolver's
    one-shot call (tc_count_csr: copy in, one launch, count back).  A
    ~131k-edge graph is launch-bound on the GPU: each is timed over at least
    50 calls."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, and here the timed CPU figure beside the GPU one
    from hetegeneous_parallel_programming_amd import Graph

    g = Graph.er(16384, 16, seed).oriented()
    rp, ci = g.out_rowptr(), g.out_colidx()
    threads = cores_available()
    cpu = []
    for _ in range(3):
        t0 = time.perf_counter()
        want = oracle.tc_count(rp, ci, threads=threads)
        cpu.append(time.perf_counter() - t0)
    cpu_s = sorted(cpu)[1]
    res = {"vertices": g.V(), "dag_edges": g.E(), "count": want,
           "cpu": {"ms": cpu_s * 1e3, "edges_per_s": g.E() / cpu_s, "cores": threads}}
    from hetegeneous_parallel_programming_amd.solver import TCSolver

    dg = ctx.upload(g)
    exact = True
    for kind, fn in (("tc_count", ctx.count), ("binned", ctx.count_binned), ("plain", ctx.count_plain)):
        ms_total, outs = time_steps(lambda: fn(dg), steps, warmup)
        exact = exact and all(o[0] == want for o in outs)
        k_avg = float(np.mean([o[1]["count_ms"] for o in outs]))
        res[kind] = {"ms_per_step": ms_total / steps, "kernel_ms_avg": k_avg,
                     "edges_per_s": g.E() / (ms_total / steps / 1e3),
                     "single_launch": bool(outs[-1][1]["single_launch"])}
    dg.free()
    # each one-shot call timed on its own: a few calls take ~1-4 ms (host
    # side), so the median is quoted beside the mean
    for _ in range(warmup):
        TCSolver(g, ctx=ctx)
    one, outs = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        outs.append(TCSolver(g, ctx=ctx))
        one.append((time.perf_counter() - t0) * 1e3)
    exact = exact and all(o == want for o in outs)
    res["tcsolver_one_shot_ms"] = float(np.mean(one))
    res["tcsolver_one_shot_median_ms"] = float(np.median(one))
    res["exact_vs_cpu_oracle"] = exact
    log(f"er16384: count={want} tc_count step={res['tc_count']['ms_per_step']:.3f} ms "
        f"(single launch: {res['tc_count']['single_launch']}), binned {res['binned']['ms_per_step']:.3f} ms "
        f"(kernels {res['binned']['kernel_ms_avg']:.3f}), plain {res['plain']['ms_per_step']:.3f} ms, "
        f"TCSolver one-shot {res['tcsolver_one_shot_median_ms']:.3f} ms median "
        f"({res['tcsolver_one_shot_ms']:.3f} mean), cpu {cpu_s * 1e3:.2f} ms on {threads} cores, "
        f"exact={exact}")
    return res


def local_counts(ctx, g, dg, total, steps=3, warmup=1):
    """tc_count_local on the resident graph: per-edge, per-vertex and per-PE
    attribution, timed beside the binned count, with the reference push/pull
    per-PE sums of a 4-way vertex split and of a partner-balanced split; then
    the same without vertex counts (edge and part counts only), and the part
    counts alone.
    Size-independent checks: edge counts sum to the total, vertex counts to
    three times it, part counts to it; the edges-only outputs equal the full
    call's."""
    from hetegeneous_parallel_programming_amd import PartitionedGraph

    out = {}
    # the first local call rebuilds the partner structure with each entry's
    # edge position and writes the per-part pass's row words; no host outputs here
    t0 = time.perf_counter()
    ctx.count_local(dg, edges=False, vertices=False)
    first_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    ctx.count_local(dg, edges=False, vertices=False)
    next_ms = (time.perf_counter() - t0) * 1e3
    out["first_call_ms"], out["next_call_ms"] = first_ms, next_ms
    log(f"local: first call {first_ms:.1f} ms (partner rebuild with edge positions, per-part row words), "
        f"next {next_ms:.1f} ms (no host outputs)")
    for balance in ("vertices", "partner"):
        bounds = PartitionedGraph(g, 4, balance).bounds()
        ms_total, outs = time_steps(lambda: ctx.count_local(dg, bounds=bounds), steps, warmup)
        o = outs[-1][1]
        exact = (all(x[0] == total for x in outs)
                 and int(o["edge_counts"].sum(dtype=np.uint64)) == total
                 and int(o["vertex_counts"].sum(dtype=np.uint64)) == 3 * total
                 and int(o["part_counts"].sum()) == total)
        k_avg = float(np.mean([x[1]["stats"]["count_ms"] for x in outs]))
        out[balance] = {"bounds": [int(x) for x in bounds],
                        "part_counts": [int(x) for x in o["part_counts"]],
                        "kernel_ms_avg": k_avg, "ms_per_step": ms_total / steps, "exact_sums": exact}
        log(f"local ({balance} split): kernel={k_avg:.3f} ms, per-PE push/pull sums "
            f"{out[balance]['part_counts']}, sums exact={exact}")
        # the reference's own attribution only (per-edge counts and per-PE
        # sums, no per-vertex counts): the kernels keep no slot hit counters
        ms_e, outs_e = time_steps(lambda: ctx.count_local(dg, vertices=False, bounds=bounds), steps, warmup)
        oe = outs_e[-1][1]
        exact_e = (all(x[0] == total for x in outs_e)
                   and np.array_equal(oe["edge_counts"], o["edge_counts"])
                   and np.array_equal(oe["part_counts"], o["part_counts"]))
        ke_avg = float(np.mean([x[1]["stats"]["count_ms"] for x in outs_e]))
        out[balance]["edges_only"] = {"kernel_ms_avg": ke_avg, "ms_per_step": ms_e / steps, "exact": exact_e}
        log(f"local ({balance} split), edge and part counts only: kernel={ke_avg:.3f} ms, "
            f"same as with vertex counts={exact_e}")
        # the per-PE sums alone (what the reference's push/pull kernel
        # reports): each partner row's hits go to its owner's part
        ms_p, outs_p = time_steps(lambda: ctx.count_local(dg, edges=False, vertices=False, bounds=bounds),
                                  steps, warmup)
        op = outs_p[-1][1]
        exact_p = (all(x[0] == total for x in outs_p)
                   and np.array_equal(op["part_counts"], o["part_counts"]))
        kp_avg = float(np.mean([x[1]["stats"]["count_ms"] for x in outs_p]))
        out[balance]["parts_only"] = {"kernel_ms_avg": kp_avg, "ms_per_step": ms_p / steps, "exact": exact_p}
        log(f"local ({balance} split), part counts only: kernel={kp_avg:.3f} ms, "
            f"same as with vertex counts={exact_p}")
    out["max_vertex_triangles"] = int(o["vertex_counts"].max())
    return out


def leg_single(ctx, g, steps, warmup, name, plain=True, local=False):
    """Binned count (partner scheme: each edge streams the shorter list; its
    algorithmic bytes are work['partner_bytes']) and the plain kernel (the
    reference's edge iterator: N+(v) streamed for every edge, work['alg_bytes'])."""
    work = g.work()
    t0 = time.perf_counter()
    dg = ctx.upload(g)
    upload_s = time.perf_counter() - t0
    info = dg.info()
    log(f"{name}: upload {upload_s * 1e3:.0f} ms incl. partner build {info['build_ms']:.2f} ms on the "
        f"device; {info['partner_rows']} partner rows")
    res = {"upload": {"ms": upload_s * 1e3, "partner_build_ms": info["build_ms"],
                      "partner_rows": info["partner_rows"], "device_bytes": info["bytes"]}}
    kinds = [("binned", ctx.count)] + ([("plain", ctx.count_plain)] if plain else [])
    alg = {"binned": work["partner_bytes"], "plain": work["alg_bytes"]}
    for kind, fn in kinds:
        ms_total, outs = time_steps(lambda: fn(dg), steps, warmup)
        counts = {o[0] for o in outs}
        if len(counts) != 1:
            raise RuntimeError(f"{name} {kind}: counts differ across steps: {counts}")
        kms = sorted(o[1]["count_ms"] for o in outs)
        pms = sorted(o[1]["plan_ms"] for o in outs)
        k_avg = float(np.mean(kms))
        res[kind] = {"count": counts.pop(), "ms_per_step": ms_total / steps,
                     "kernel_ms_avg": k_avg, "kernel_ms_min": kms[0],
                     "plan_ms_avg": float(np.mean(pms)),
                     "edges_per_s": work["edges"] / (ms_total / steps / 1e3),
                     "alg_bytes": alg[kind], "alg_GBps": alg[kind] / (k_avg / 1e3) / 1e9,
                     "edge_iterator_GBps": work["alg_bytes"] / (k_avg / 1e3) / 1e9,
                     "bins": outs[-1][1]["bin_pivots"],
                     "hash_fallbacks": outs[-1][1]["hash_fallbacks"],
                     "single_launch": outs[-1][1]["single_launch"]}
        log(f"{name} {kind}: count={res[kind]['count']} step={res[kind]['ms_per_step']:.3f} ms "
            f"kernel={k_avg:.3f} ms alg={res[kind]['alg_GBps']:.0f} GB/s")
    if local:
        res["local"] = local_counts(ctx, g, dg, res["binned"]["count"])
    dg.free()
    return work, res


def golden_big() -> dict:
    """Oracle counts at full size (tests/golden/big.json), when present."""
    p = os.path.join(ROOT, "tests", "golden", "big.json")
    return json.load(open(p)) if os.path.exists(p) else {}


def per_graph_tc_count(ctx, graphs):
    """Second baseline: every graph uploaded on its own (tc_graph_upload) and
    counted by tc_count, one call per graph (the single plain launch for a
    small graph, the binned count with its kept plan otherwise); per-graph
    host time of the call (each ends in a stream sync), after one untimed
    pass.  Returns (counts, ms per graph)."""
    dgs = [ctx.upload(g) for g in graphs]
    try:
        for dg in dgs:
            ctx.count(dg)
        counts = np.zeros(len(graphs), np.uint64)
        ms = np.zeros(len(graphs))
        for i, dg in enumerate(dgs):
            t0 = time.perf_counter()
            counts[i] = ctx.count(dg)[0]
            ms[i] = (time.perf_counter() - t0) * 1e3
    finally:
        for dg in dgs:
            dg.free()
    return counts, ms


def leg_batch(ctx, graphs, seed, reps, tuner=None):
    """4096-graph mix: one captured persistent launch, first with the default
    launch setting, then with the autotuned one (tuned on a different batch of
    the same mix), against two per-graph baselines -- one plain launch per
    graph (memset + plain kernel + copy back + sync, tc_batch_count_each), and
    tc_count per separately uploaded graph -- with the speed-up also quoted
    against the faster of the two for every graph.  The batch's own creation
    (host-to-device copies and the partner build of all graphs) is reported
    beside it, with a build-inclusive rate."""
    n_graphs = len(graphs)
    edges = sum(g.E() for g in graphs)
    # partner-scheme bytes of the whole batch, as for the headline roofline
    alg_bytes = sum(g.work()["partner_bytes"] for g in graphs)
    t0 = time.perf_counter()
    b = ctx.batch(graphs)
    create_ms = (time.perf_counter() - t0) * 1e3
    build_ms = b.info()["build_ms"]
    log(f"batch: created in {create_ms:.0f} ms (partner build {build_ms:.1f} ms)")
    each, ms_each, st_each = b.count_each()
    each2, ms_each2, st_each2 = b.count_each()
    base_s = st_each2["total_ms"] / 1e3
    log(f"batch: plain launch per graph {base_s * 1e3:.0f} ms for all")
    tcc, ms_tcc = per_graph_tc_count(ctx, graphs)
    log(f"batch: tc_count per graph {ms_tcc.sum():.0f} ms for all")
    if not np.array_equal(tcc, each):
        raise RuntimeError("tc_count per graph differs from the plain launch per graph")
    best_s = float(np.minimum(ms_each2, ms_tcc).sum()) / 1e3

    def timed():
        b.count()  # capture + first replay
        ms_total, outs = time_steps(lambda: b.count(), reps, 1)
        for o in outs:
            if not np.array_equal(o[0], each):
                raise RuntimeError("batched counts differ from per-graph counts")
        return ms_total / reps / 1e3, [o[1]["count_ms"] for o in outs]

    default_s, kms_default = timed()
    log(f"batch: captured launch, default setting {default_s * 1e3:.2f} ms")
    res = {"graphs": n_graphs, "edges": edges, "sum_count": int(each.sum()),
           "baseline_graphs_per_s": n_graphs / base_s,
           "baseline_p50_ms": float(np.percentile(ms_each2, 50)),
           "baseline_p99_ms": float(np.percentile(ms_each2, 99)),
           "baseline_tc_count_graphs_per_s": n_graphs / (ms_tcc.sum() / 1e3),
           "baseline_tc_count_p99_ms": float(np.percentile(ms_tcc, 99)),
           "baseline_best_per_graph_graphs_per_s": n_graphs / best_s,
           "create_ms": create_ms, "partner_build_ms": build_ms,
           "default_tune": ctx.tune, "default_graphs_per_s": n_graphs / default_s,
           "default_speedup": base_s / default_s}
    batched_s, kms = default_s, kms_default
    if tuner is not None:
        tune, report = tuner(ctx)
        res["tune"], res["autotune"] = tune, report
        batched_s, kms = timed()
    res.update({"batched_graphs_per_s": n_graphs / batched_s, "batched_ms": batched_s * 1e3,
                "batched_p99_ms": float(np.percentile(kms, 99)), "speedup": base_s / batched_s,
                "speedup_vs_best_per_graph": best_s / batched_s,
                "build_inclusive_graphs_per_s": n_graphs / (batched_s + build_ms / 1e3),
                "build_inclusive_speedup_vs_best_per_graph": best_s / (batched_s + build_ms / 1e3),
                "alg_bytes": alg_bytes, "alg_GBps": alg_bytes / batched_s / 1e9})
    gold = golden_big().get(f"batch_seed{seed}_{n_graphs}")
    res["exact_vs_cpu_oracle"] = (bool(np.array_equal(each, np.array(gold["triangles"], np.uint64)))
                                  if gold else None)
    log(f"batch: baselines {res['baseline_graphs_per_s']:.0f} graphs/s (plain launch per graph), "
        f"{res['baseline_tc_count_graphs_per_s']:.0f} (tc_count per graph), "
        f"{res['baseline_best_per_graph_graphs_per_s']:.0f} (faster per graph); batched default "
        f"{res['default_graphs_per_s']:.0f} graphs/s, batched tuned {res['batched_graphs_per_s']:.0f} "
        f"graphs/s ({res['batched_ms']:.2f} ms), speedup {res['speedup']:.2f}x / "
        f"{res['speedup_vs_best_per_graph']:.2f}x vs the faster baseline; creation {create_ms:.0f} ms "
        f"(partner build {build_ms:.1f} ms), build-inclusive {res['build_inclusive_graphs_per_s']:.0f} graphs/s "
        f"({res['build_inclusive_speedup_vs_best_per_graph']:.2f}x); exact vs oracle: {res['exact_vs_cpu_oracle']}")
    b.free()
    return res, each


def batch_tuner(args):
    """Autotune on a 1024-graph batch of the same mix drawn with another seed
    (or load the setting saved for this device from --tune-file)."""
    def tune(ctx):
        from hetegeneous_parallel_programming_amd import Graph

        key = "batch_mix"
        if args.tune_file and os.path.exists(args.tune_file) and ctx.load_tune(args.tune_file, key):
            log(f"autotune: loaded {ctx.tune} from {args.tune_file}")
            return ctx.tune, {"source": args.tune_file}
        tg = Graph.batch(args.seed + 1000, 0, 1024)
        tb = ctx.batch(tg)
        t, rep = ctx.autotune(batch=tb, reps=3, budget_ms=args.tune_seconds * 1e3)
        tb.free()
        rep["tuned_on"] = f"1024 graphs of the same mix, seed {args.seed + 1000}"
        log(f"autotune: {t} after {rep['trials']} settings in {rep['elapsed_ms'] / 1e3:.1f} s "
            f"({rep['start_ms']:.2f} -> {rep['best_ms']:.2f} ms on the tuning batch)")
        if args.tune_file:
            ctx.save_tune(args.tune_file, key)
        return t, rep
    return tune


def serving_mix_capacity(ctx, req, want, heavy, k=2000, reps=3):
    """Batched capacity of the serving mix: k requests of the run's draw with
    its heavy share (the first round(k * share) RMAT-20 requests and the first
    light ones, so a short sample does not under- or over-represent the 1%
    heavy requests) counted by one captured batched launch.  Replay-only
    graphs/s, and with the batch's partner build included (a served graph
    arrives without its partner structure)."""
    k = min(k, len(req))
    hidx = np.nonzero(heavy)[0]
    lidx = np.nonzero(~heavy)[0]
    nh = min(len(hidx), int(round(k * len(hidx) / len(req))))
    pick = np.concatenate([hidx[:nh], lidx[:k - nh]])
    b = ctx.batch([req[i] for i in pick])
    try:
        build_ms = b.info()["build_ms"]
        b.count()
        ms_total, outs = time_steps(lambda: b.count(), reps, 1)
        exact = all(np.array_equal(o[0], want[pick]) for o in outs)
    finally:
        b.free()
    replay_s = ms_total / reps / 1e3
    return {"requests": k, "heavy_requests": int(nh), "replay_ms": replay_s * 1e3, "partner_build_ms": build_ms,
            "replay_graphs_per_s": k / replay_s, "graphs_per_s": k / (replay_s + build_ms / 1e3),
            "exact": exact}


def leg_serving(ctx, graphs, counts, heavy, heavy_count, n_requests, seed, clients=8, slots=3,
                saturated_only=False, fill_graphs=0, fill_wait_us=0, max_ids=0, heavy_frac=0.01):
    """Poisson TC requests from the batched mix plus 1% RMAT-20 through the
    continuous-batching server: saturated throughput; then Poisson arrivals at
    50% and 90% of the batched capacity of the same mix (build included), and
    at 50% and 90% of the server's own saturated throughput."""
    from hetegeneous_parallel_programming_amd.solver import Server, clock_ms

    rng = np.random.default_rng(seed)
    picks = np.where(rng.random(n_requests) < heavy_frac, -1, rng.integers(0, len(graphs), n_requests))
    want = np.array([heavy_count if k < 0 else counts[k] for k in picks], np.uint64)
    req = [heavy if k < 0 else graphs[k] for k in picks]
    out = {"requests": n_requests, "heavy_fraction": float((picks < 0).mean()), "slots": slots}
    cap = serving_mix_capacity(ctx, req, want, picks < 0)
    out["batched_capacity"] = cap
    log(f"serving mix, batched: {cap['graphs_per_s']:.0f} graphs/s with the partner build "
        f"({cap['replay_graphs_per_s']:.0f} replay only; {cap['requests']} requests, {cap['heavy_requests']} "
        f"RMAT-20), exact={cap['exact']}")
    with Server(ctx, hbm_budget=16 << 30, slots=slots, max_batch_graphs=2048, max_batch_rows=16 << 20,
                max_queue=1 << 20, host_staging=4 << 30, fill_graphs=fill_graphs, fill_wait_us=fill_wait_us,
                max_batch_ids=max_ids) as srv:
        for i in range(64):  # warm-up
            srv.submit(req[i], i)
        srv.drain()
        srv.poll(max_results=4096)

        def run(arrivals, clients):
            """Requests i = c, c + clients, ... go out from client thread c,
            each at its scheduled arrival (or back to back when saturating);
            the main thread collects results.  Latency = completion - arrival."""
            got = {}
            t0 = clock_ms() + 50.0
            sched = t0 + arrivals if arrivals is not None else None
            errors = []

            def client(c):
                try:
                    for i in range(c, n_requests, clients):
                        if sched is not None:
                            srv.submit_at(req[i], i, sched[i])
                        else:
                            srv.submit(req[i], i)
                except Exception as e:  # surfaced after the run
                    errors.append(e)

            st0 = srv.stats()
            start = clock_ms()
            ths = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
            for t in ths:
                t.start()
            while len(got) < n_requests:
                rs = srv.poll(timeout_ms=200)
                for r in rs:
                    got[r["tag"]] = r
                if not rs and not any(t.is_alive() for t in ths):
                    srv.drain()  # everything submitted: collect the rest
                    rs = srv.poll()
                    while rs:
                        for r in rs:
                            got[r["tag"]] = r
                        rs = srv.poll()
                    break
            for t in ths:
                t.join()
            if errors:
                raise errors[0]
            done = np.array([got[i]["t_done_ms"] for i in range(n_requests)])
            base = sched if sched is not None else np.array([got[i]["t_submit_ms"] for i in range(n_requests)])
            lat = done - base
            cnt = np.array([got[i]["count"] for i in range(n_requests)], np.uint64)
            ok = bool(np.array_equal(cnt, want)) and all(got[i]["status"] == 0 for i in range(n_requests))
            span = (done.max() - (sched[0] if sched is not None else start)) / 1e3
            light = lat[picks >= 0]
            heavy_lat = lat[picks < 0]
            st1 = srv.stats()
            # where the time went during this run: ms per second of the run's span
            where = {k: (st1[k] - st0[k]) / span for k in
                     ("submit_ms", "retire_ms", "admit_ms", "launch_ms", "wait_busy_ms", "wait_idle_ms", "device_ms")}
            where["launches"] = st1["launches"] - st0["launches"]
            return {"graphs_per_s": n_requests / span, "p50_ms": float(np.percentile(lat, 50)),
                    "p99_ms": float(np.percentile(lat, 99)),
                    "light_p50_ms": float(np.percentile(light, 50)),
                    "light_p99_ms": float(np.percentile(light, 99)),
                    "heavy_p50_ms": float(np.percentile(heavy_lat, 50)) if heavy_lat.size else None,
                    "heavy_p99_ms": float(np.percentile(heavy_lat, 99)) if heavy_lat.size else None,
                    "exact": ok, "clients": clients, "time_per_s": where,
                    "mean_batch": float(np.mean([got[i]["batch_graphs"] for i in range(n_requests)]))}

        sat = run(None, clients)
        out["saturated"] = sat
        log(f"serving saturated: {sat['graphs_per_s']:.0f} graphs/s, p99 {sat['p99_ms']:.1f} ms, "
            f"exact={sat['exact']}; ms per s: " + ", ".join(f"{k} {v:.0f}" for k, v in sat["time_per_s"].items()))
        # load_50 / load_90: fractions of the batched capacity of this mix;
        # load_50_of_saturated / ...: of the server's own saturated rate
        points = () if saturated_only else (
            ("load_50", 0.5, cap["graphs_per_s"]), ("load_90", 0.9, cap["graphs_per_s"]),
            ("load_50_of_saturated", 0.5, sat["graphs_per_s"]), ("load_90_of_saturated", 0.9, sat["graphs_per_s"]))
        for key, frac, base in points:
            lam = frac * base
            arr = np.cumsum(rng.exponential(1.0 / lam, n_requests)) * 1e3
            r = run(arr, clients)
            r["offered_per_s"] = lam
            out[key] = r
            log(f"serving {key} ({lam:.0f}/s offered): {r['graphs_per_s']:.0f} graphs/s, "
                f"p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms (mix graphs p50 {r['light_p50_ms']:.2f} / "
                f"p99 {r['light_p99_ms']:.2f} ms, RMAT-20 p50 {r['heavy_p50_ms'] or float('nan'):.2f} ms), "
                f"exact={r['exact']}, mean batch {r['mean_batch']:.1f}")
        out["server_stats"] = srv.stats()
    out["exact_all"] = cap["exact"] and all(
        out[k]["exact"] for k in ("saturated", "load_50", "load_90", "load_50_of_saturated", "load_90_of_saturated")
        if k in out)
    return out


def leg_partition(ctx, seed, steps, warmup, split=True):
    from hetegeneous_parallel_programming_amd import Graph, PartitionedGraph

    t0 = time.perf_counter()
    g = Graph.rmat_scale(24, 16, seed)
    g.orientation()
    part = PartitionedGraph(g, 4).halo_partition(0)
    work = part.work()
    log(f"partition: RMAT-24 + halo built in {time.perf_counter() - t0:.1f} s; owned edges "
        f"{work['edges']}, halo lists {part.halo_ids.size}")
    if not split:
        del g
    dg = ctx.upload_partition(part)
    info = dg.info()
    ms_total, outs = time_steps(lambda: ctx.count(dg), steps, warmup)
    counts = {o[0] for o in outs}
    if len(counts) != 1:
        raise RuntimeError("partition counts differ across steps")
    k_avg = float(np.mean([o[1]["count_ms"] for o in outs]))
    dg.free()
    res = {"count": counts.pop(), "owned_edges": work["edges"], "alg_bytes": work["partner_bytes"],
           "alg_bytes_edge_iterator": work["alg_bytes"],
           "halo_lists": int(part.halo_ids.size), "halo_ids": int(part.halo_colidx.size),
           "partner_rows": info["partner_rows"], "partner_build_ms": info["build_ms"],
           "ms_per_step": ms_total / steps, "kernel_ms_avg": k_avg,
           "edges_per_s": work["edges"] / (ms_total / steps / 1e3),
           "alg_GBps": work["partner_bytes"] / (k_avg / 1e3) / 1e9,
           "edge_iterator_GBps": work["alg_bytes"] / (k_avg / 1e3) / 1e9}
    res["roofline_frac"] = res["alg_GBps"] / PEAK_HBM_GBS
    log(f"partition: count={res['count']} kernel={k_avg:.3f} ms frac={res['roofline_frac']:.3f}")
    if split:
        res["split4"] = split_parts(ctx, g, res, steps=3, warmup=1)
    return res


def split_parts(ctx, g, p0, steps, warmup):
    """Every part of the 4-way split, under the reference's vertex split and
    the partner-balanced split: per-part count time, and the slowest part
    (what a 4-GPU run would wait for).  Both splits must sum to the same total."""
    from hetegeneous_parallel_programming_amd import PartitionedGraph

    out = {}
    for balance in ("vertices", "partner"):
        pg = PartitionedGraph(g, 4, balance)
        parts = []
        for i in range(4):
            if balance == "vertices" and i == 0:
                parts.append({"rows": [0, p0_rows(pg)], "edges": p0["owned_edges"],
                              "count": p0["count"], "kernel_ms": p0["kernel_ms_avg"]})
                continue
            part = pg.halo_partition(i)
            dg = ctx.upload_partition(part)
            _, outs = time_steps(lambda: ctx.count(dg), steps, warmup)
            dg.free()
            counts = {o[0] for o in outs}
            if len(counts) != 1:
                raise RuntimeError("partition counts differ across steps")
            parts.append({"rows": [part.row_begin, part.row_end], "edges": part.work()["edges"],
                          "count": counts.pop(),
                          "kernel_ms": float(np.mean([o[1]["count_ms"] for o in outs]))})
            del part
        out[balance] = {"parts": parts, "total": sum(x["count"] for x in parts),
                        "max_kernel_ms": max(x["kernel_ms"] for x in parts),
                        "sum_kernel_ms": sum(x["kernel_ms"] for x in parts)}
        log(f"split4 {balance}: kernel ms per part "
            f"{[round(x['kernel_ms'], 2) for x in parts]}, total {out[balance]['total']}")
    out["exact"] = out["vertices"]["total"] == out["partner"]["total"]
    out["critical_path_speedup"] = (out["vertices"]["max_kernel_ms"]
                                    / out["partner"]["max_kernel_ms"])
    return out


def p0_rows(pg):
    return int(pg.bounds()[1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-partition", action="store_true")
    ap.add_argument("--no-split", action="store_true",
                    help="skip timing the other parts of the RMAT-24 4-way splits")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU time per baseline run (3 runs); the whole graph when it fits")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-rmat", action="store_true", help="skip the headline RMAT leg")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain kernel")
    ap.add_argument("--no-er", action="store_true", help="skip the ER n=16384 leg")
    ap.add_argument("--no-local", action="store_true",
                    help="skip the per-edge / per-vertex / per-PE attribution count")
    ap.add_argument("--tune", default="", help="launch parameters, e.g. t_wave=8192,stage_ratio16=24")
    ap.add_argument("--requests", type=int, default=10000, help="requests per serving run")
    ap.add_argument("--clients", type=int, default=8, help="client threads submitting requests")
    ap.add_argument("--serving-slots", type=int, default=3, help="concurrent launches of the server")
    ap.add_argument("--serving-saturated-only", action="store_true", help="serving leg: the saturated run only")
    ap.add_argument("--serving-fill", type=int, default=0, help="server fill_graphs (0: its default)")
    ap.add_argument("--serving-fill-wait-us", type=int, default=0, help="server fill_wait_us (0: its default)")
    ap.add_argument("--serving-max-ids", type=int, default=0, help="server max_batch_ids (0: its default)")
    ap.add_argument("--serving-heavy-frac", type=float, default=0.01, help="share of RMAT-20 requests")
    ap.add_argument("--no-autotune", action="store_true", help="time the batch with the default setting only")
    ap.add_argument("--side-leg-limit", type=float, default=150.0,
                    help="seconds without progress in the side legs before the headline line is printed "
                         "alone (0: no limit)")
    ap.add_argument("--tune-seconds", type=float, default=12.0, help="autotuner wall-time budget")
    ap.add_argument("--tune-file", default="", help="load the batch setting from / save it to this file")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy (counter calibration) and a plain read of this many bytes")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        print(json.dumps(cpu_baseline_child(args.scale, args.seed, args.cpu_seconds)), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this triangle-counting build has no multi-GPU "
              "path (run with --gpus 1).", file=sys.stderr)
        return 2

    from hetegeneous_parallel_programming_amd import Graph
    from hetegeneous_parallel_programming_amd.solver import Context

    if args.copy_probe:
        ctx = Context(0)
        ms = ctx.probe_copy(args.copy_probe, max(1, args.steps))
        ms_read = ctx.probe_read(args.copy_probe, max(1, args.steps))
        ctx.close()
        print(json.dumps({"copy_bytes": args.copy_probe, "iters": max(1, args.steps), "ms": ms,
                          "GBps_read_plus_write": 2 * args.copy_probe / (ms / 1e3) / 1e9,
                          "read_ms": ms_read,
                          "GBps_read_only": args.copy_probe / (ms_read / 1e3) / 1e9}), flush=True)
        return 0
    ctx = Context(0)
    log(f"device: {ctx.name}")
    if args.tune:
        ctx.tune = {k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(","))}
    log(f"tune: {ctx.tune}")
    line = {"metric": "TC edges/s, %HBM peak; batched graphs/s + p99 vs one-launch-per-graph",
            "value": None, "extra": {}}
    if not args.no_rmat:
        line = leg_headline(ctx, args)
    if args.side_leg_limit > 0:
        start_watchdog(line, args.side_leg_limit)
    run_side_legs(ctx, args, line)
    _side_legs_done.set()
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


def traffic_of(workload: str):
    """Measured HBM bytes per step of the six count kernels, from the committed
    rocprofv3 summary profiles/count_kernel_<workload>.md (its line
    `traffic_bytes_per_step = N`: FETCH_SIZE and WRITE_SIZE scaled by the
    copy-kernel calibration), when there is one."""
    p = os.path.join(ROOT, "profiles", f"count_kernel_{workload}.md")
    if not os.path.exists(p):
        return None, None
    m = re.search(r"^traffic_bytes_per_step = (\d+)\s*$", open(p).read(), re.M)
    if not m:
        return None, None
    return int(m.group(1)), os.path.relpath(p, ROOT)


def leg_headline(ctx, args) -> dict:
    from hetegeneous_parallel_programming_amd import Graph

    t0 = time.perf_counter()
    g = Graph.rmat_scale(args.scale, 16, args.seed)
    g.orientation()
    log(f"RMAT-{args.scale} generated+oriented in {time.perf_counter() - t0:.1f} s: "
        f"n={g.V()} DAG edges={g.E()} max out={g.get_max_degree()}")
    work, single = leg_single(ctx, g, args.steps, args.warmup, f"rmat{args.scale}",
                              plain=not args.no_plain, local=not args.no_local)
    best = single["binned"]
    traffic, traffic_src = traffic_of(f"rmat{args.scale}")
    # the practical read ceiling of this device, measured in the same run: a
    # streaming read of 4 GiB (well beyond the 256 MB Infinity Cache)
    read_bytes = 1 << 32
    read_GBps = read_bytes / (ctx.probe_read(read_bytes, 10) / 1e3) / 1e9
    log(f"read probe: 4 GiB streaming read {read_GBps:.0f} GB/s; count at "
        f"{best['alg_GBps'] / read_GBps:.2f} of it")
    line = {
        "metric": "TC edges/s, %HBM peak; batched graphs/s + p99 vs one-launch-per-graph",
        "value": best["edges_per_s"], "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"rmat{args.scale}_ef16_seed{args.seed}_whole_graph",
                   "graph": f"RMAT scale {args.scale}, edge factor 16, (a,b,c)=(.57,.19,.19), "
                            "symmetrised, deduplicated, degree-oriented DAG",
                   "vertices": g.V(), "dag_edges": work["edges"], "count": best["count"],
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": best["alg_GBps"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": best["alg_GBps"] / PEAK_HBM_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "count_kernel<bin> (one persistent launch per non-empty bin per step; four bin groups on four streams)",
                     "alg_bytes": work["partner_bytes"],
                     "alg_bytes_note": "bytes the partner-scheme count reads with no reuse: per partner row "
                                       "32 + 4 d+(x), per edge 8 + 4 min(d+(u), d+(v))",
                     "alg_bytes_edge_iterator": work["alg_bytes"],
                     "edge_iterator_GBps": best["edge_iterator_GBps"],
                     "kernel_ms_avg": best["kernel_ms_avg"],
                     "read_probe_GBps": read_GBps,
                     "frac_of_read_probe": best["alg_GBps"] / read_GBps},
        "cpu_baseline": None,
        "extra": {"plain_kernel": single.get("plain"), "binned": best, "upload": single["upload"],
                  "local_counts": single.get("local")},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
        if line["cpu_baseline"]["whole_graph"] and line["cpu_baseline"]["count"] != best["count"]:
            raise RuntimeError("CPU baseline count differs from the GPU count")
        log(f"cpu baseline: {line['cpu_baseline']['value']:.3e} edges/s on "
            f"{line['cpu_baseline']['cores']} cores")
    gold = golden_big().get(f"rmat{args.scale}_seed{args.seed}")
    line["config"]["exact_vs_cpu_oracle"] = (gold["triangles"] == best["count"]) if gold else None
    return line


def side_leg(line, name, fn) -> None:
    """Runs one side leg; an exception in it is logged with its traceback and
    recorded in extra.side_leg_errors[name], and the remaining legs and the
    JSON line (with the headline measurement) still follow."""
    try:
        fn()
    except Exception as e:  # noqa: BLE001 -- recorded in the JSON line, not hidden
        import traceback
        log(f"side leg {name} failed: {e!r}\n{traceback.format_exc()}")
        line["extra"].setdefault("side_leg_errors", {})[name] = repr(e)


def run_side_legs(ctx, args, line) -> None:
    if not args.no_er:
        side_leg(line, "er16384", lambda: line["extra"].__setitem__(
            "er16384", leg_er(ctx, args.seed, max(50, args.steps), max(5, args.warmup))))
    # the partition legs before the batch and serving legs: the longest side
    # legs come last
    if not args.no_partition:
        side_leg(line, "partition_rmat24_p0", lambda: line["extra"].__setitem__(
            "partition_rmat24_p0", leg_partition(ctx, args.seed, max(3, args.steps // 4),
                                                 args.warmup, split=not args.no_split)))
    if not args.no_batch or not args.no_serving:
        side_leg(line, "batched_serving", lambda: batch_and_serving_legs(ctx, args, line))


def batch_and_serving_legs(ctx, args, line) -> None:
    from hetegeneous_parallel_programming_amd import Graph

    t0 = time.perf_counter()
    graphs = Graph.batch(args.seed, 0, args.batch_graphs)
    log(f"batch: generated {len(graphs)} graphs in {time.perf_counter() - t0:.1f} s")
    default = ctx.tune
    res, counts = leg_batch(ctx, graphs, args.seed, max(3, args.steps // 4),
                            None if args.no_autotune else batch_tuner(args))
    if not args.no_batch:
        line["extra"]["batched"] = res
    if not args.no_serving:
        heavy = Graph.rmat_scale(20, 16, args.seed)
        heavy.orientation()
        dh = ctx.upload(heavy)
        heavy_count = ctx.count(dh)[0]
        heavy_build_ms = dh.info()["build_ms"]
        dh.free()
        log(f"rmat20: partner build {heavy_build_ms:.2f} ms on the device")
        line["extra"]["serving"] = leg_serving(ctx, graphs, counts, heavy, heavy_count,
                                               args.requests, args.seed + 100, args.clients,
                                               args.serving_slots, args.serving_saturated_only,
                                               args.serving_fill, args.serving_fill_wait_us,
                                               args.serving_max_ids, args.serving_heavy_frac)
        line["extra"]["serving"]["rmat20_build_ms"] = heavy_build_ms
        gold20 = golden_big().get(f"rmat20_seed{args.seed}")
        line["extra"]["serving"]["rmat20_exact_vs_cpu_oracle"] = (
            gold20["triangles"] == heavy_count if gold20 else None)
        del heavy
    ctx.tune = default
    del graphs


if __name__ == "__main__":
    sys.exit(main())
