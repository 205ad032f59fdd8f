"""Benchmark: triangle TEPS on one MI355X, exact counts.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat]
                  [--scale 26] [--schedule degree|learned] [--tune SECONDS]
                  [--cpu-seconds 15] [--no-cpu]

Workloads (one JSON line each; the default is the headline):
  rmat      R-MAT scale S (default 26), edge factor 16, (a,b,c,d) =
            (.57,.19,.19,.05), seed 1.  One step = one full count of the
            graph resident in HBM: planning kernels (tier split or learned
            plan) + the counting launches, timed with hipEvents on the launch
            stream.  Three schedules run on the same graph and must agree: the
            plain degree-threshold split (hub threshold 100) at the default
            launch, the same split at the reference README's launch (1024 x
            1024, chunk 1), and the learned schedule (headline), whose cost
            model is trained on this graph first (reported separately); its
            feature, inference and queue kernels run inside every timed step.
  chung-lu  Chung-Lu n = 3.8M, 16.5M pairs, gamma 2.5, max expected degree
            800, seed 1; degree split with autotuned launch parameters.
  batch     10,000 graphs, half Erdos-Renyi, half Barabasi-Albert, n
            log-uniform in [1k, 200k], mean degree uniform in [8, 64], seed 1;
            one count of the whole batch per step, replayed as a captured
            hipGraph, learned schedule with autotuned launch parameters.
            Every graph's count is checked against the CPU oracle.  Reports
            graphs/s and p50/p99 per-graph completion latency.
  tiles     R-MAT scale S hash-partitioned into n x n tiles (the reference's
            partition layout), all n^3 tile jobs streamed through the GPU with
            asynchronous tile copies; the total must equal the unpartitioned
            count.  One step = one streamed count (host-to-device copies
            included, tiles in pinned-free pageable host memory).
  ingest    host only, no GPU: an R-MAT edge list (--ingest-scale, default
            21) written as text, then the input tools timed -- the text path
            (trust_tool undirect + preprocess, each --order) and the binary
            fast path (txt2bin + preprocess-bin) and the tile writer
            (trust_tool partition, --parts n) -- beside the reference's own
            tools built from source in oracle/_ref when present, with every
            output compared byte for byte and each tool's peak anonymous
            memory sampled from /proc.

--order plain|indegree (rmat) counts the graph in the row order of the
reference's otherPreprocess tools instead of preprocess_outdegree's.

Generation, preprocessing and the initial host-to-device copy happen before
the timed region.  `value` is the whole job's throughput on one GPU.
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
METRICS = {
    "rmat": "triangle TEPS on 1 MI355X, learned vs degree-split schedule, exact count",
    "chung-lu": "triangle TEPS on 1 MI355X, autotuned degree split, exact count",
    "batch": "graphs/s on 1 MI355X, 10k-graph ER/BA batch, hipGraph replay, exact per-graph counts",
    "tiles": "triangle TEPS on 1 MI355X, 2D-partition tiles streamed from host, exact count",
    "ingest": "edge-list ingest (text -> begin.bin/adjacent.bin) on the host, byte-identical "
              "to the reference tools",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ helpers
def cpu_baseline(beg, adj, E, target_seconds, gpu_count):
    """Time the oracle's hashed count on a strided vertex sample and scale
    to the whole graph by probe share (the oracle's cost is ~linear in
    probes).  When the sample is the whole graph (stride 1) its count must
    equal the GPU's.  Returns the cpu_baseline object."""
    import oracle

    # the oracle walks the reference's side of every edge: scale by its W
    W = oracle.probes(beg, adj)
    threads = oracle.max_threads()
    stride = 4096
    t0 = time.perf_counter()
    _, p = oracle.count_hash_sample(beg, adj, 0, stride)
    dt = max(time.perf_counter() - t0, 1e-3)
    rate = max(p, 1) / dt
    want_probes = rate * target_seconds
    stride = int(max(1, min(4096, W / max(want_probes, 1.0))))
    t0 = time.perf_counter()
    cnt, p = oracle.count_hash_sample(beg, adj, 0, stride)
    dt = time.perf_counter() - t0
    est_full = dt * (W / max(p, 1))
    if stride == 1:
        if cnt != gpu_count:
            raise SystemExit(f"oracle count {cnt} != GPU count {gpu_count}")
        log(f"cpu baseline covered the whole graph: oracle count {cnt} == GPU count")
        what = f"the whole graph: {p} probes in {dt:.2f}s (count {cnt} equals the GPU's)"
    else:
        what = (f"every {stride}-th vertex: {p} of {W} probes in {dt:.2f}s, scaled by probe "
                f"share to an estimated {est_full:.1f}s for the whole graph")
    return {
        "value": E / est_full / 1e9,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": "oracle hashed count (C/OpenMP restatement of the reference loop) over " + what,
    }


def check_fixture(name, beg, adj, triangles):
    """Compares a count with the oracle's count of the same workload committed
    under tests/golden/large_counts.json (written on the host by
    tests/golden/make_large_counts.py).  The graph must be the one the
    fixture was made from (same V, E and xxh64 of adjacent.bin and
    begin.bin).  Returns a note for the JSON line, or None when no fixture
    covers the workload."""
    import xxhash

    path = os.path.join(ROOT, "tests", "golden", "large_counts.json")
    if not os.path.exists(path):
        return None
    fx = json.load(open(path)).get(name)
    if fx is None:
        return None
    got = {"V": len(beg) - 1, "E": len(adj),
           "adj_xxh64": xxhash.xxh64(memoryview(np.ascontiguousarray(adj))).hexdigest(),
           "beg_xxh64": xxhash.xxh64(memoryview(np.ascontiguousarray(beg))).hexdigest()}
    for k, v in got.items():
        if fx[k] != v:
            raise SystemExit(f"{name}: the generated graph differs from the fixture's ({k})")
    if int(fx["triangles"]) != int(triangles):
        raise SystemExit(f"{name}: GPU count {triangles} != oracle fixture {fx['triangles']}")
    log(f"{name}: count {triangles} equals the oracle's committed count")
    return f"equals the oracle's count of the same CSR (tests/golden/large_counts.json:{name})"


def committed_traffic(workload):
    """HBM bytes per counting launch from the committed rocprof summary of this
    workload (FETCH_SIZE and WRITE_SIZE, calibrated on the copy kernel), or
    None when no summary exists for it."""
    path = os.path.join(ROOT, "profiles", f"tc_count_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_steps(fn, warmup, steps, label, E):
    """warmup + timed calls of fn() -> Result; counts must be identical."""
    ref = None
    for _ in range(warmup):
        r = fn()
        ref = r.triangles if ref is None else ref
        if r.triangles != ref:
            raise SystemExit(f"{label}: count changed between runs")
    res = []
    t_start = time.perf_counter()
    for _ in range(steps):
        res.append(fn())
    t_end = time.perf_counter()
    counts = {r.triangles for r in res} | ({ref} if ref is not None else set())
    if len(counts) != 1:
        raise SystemExit(f"{label}: counts differ across steps: {counts}")
    dev_s = sum(r.seconds for r in res) / len(res)
    cnt_s = sum(r.count_seconds for r in res) / len(res)
    idx_s = sum(r.index_seconds for r in res) / len(res)
    plan_s = sum(r.plan_seconds for r in res) / len(res)
    dense_s = sum(r.dense_seconds for r in res) / len(res)
    core = (f", dense core {res[0].core_vertices} vertices: {res[0].dense_edges} edges in "
            f"{dense_s * 1e3:.2f} ms of the count" if res[0].core_vertices else "")
    log(f"{label}: {res[0].triangles} triangles, {dev_s * 1e3:.2f} ms/step "
        f"(index {idx_s * 1e3:.2f} ms, plan {plan_s * 1e3:.2f} ms, count "
        f"kernels {cnt_s * 1e3:.2f} ms{core}), {E / dev_s / 1e9:.3f} GTEPS")
    return dict(r=res[0], dev_s=dev_s, cnt_s=cnt_s, idx_s=idx_s, dense_s=dense_s,
                wall_ms=(t_end - t_start) / steps * 1e3)


def core_row_bytes(nc):
    """Bytes of bit rows one dense edge reads: the core's rows are nc bits
    rounded up to S slices of a multiple of 4 words (core.hip)."""
    S = 16 if nc > 16384 else 8
    wd = max(4, ((nc + 32 * S - 1) // (32 * S) + 3) & ~3)
    return 4 * wd * S


def core_config(T, r, dense_s):
    """The dense core's share of a count (None when it had none), with the
    L2s' measured read rate its bit rows are read at."""
    if not r.core_vertices:
        return None
    rb = core_row_bytes(r.core_vertices)
    gbs = r.dense_edges * rb / dense_s / 1e9 if dense_s > 0 else None
    l2 = T.membench_read(2 << 20) / 1e9
    log(f"dense core: {gbs:.0f} GB/s of bit rows; an L2-resident 2 MiB read: {l2:.0f} GB/s")
    return {"vertices": r.core_vertices, "dense_edges": r.dense_edges,
            "dense_probes": r.dense_probes, "W_probes_short_walk": r.probes + r.dense_probes,
            "dense_ms": dense_s * 1e3, "row_bytes_per_edge": rb, "bit_row_GBs": gbs,
            "l2_read_GBs": l2, "frac_of_l2_read": gbs / l2 if gbs else None}


def roofline(T, ctx, r0, cnt_s, kw, workload, dense_s=0.0):
    """Roofline object for the hashed counting launches: algorithmic bytes /
    time, against HBM peak, plus the measured load-path references (the same
    walk with hashing replaced by a checksum, and a streaming copy).  The
    dense core's launch (dense_s of the count's time, bound by the L2s that
    hold its bit rows) is left out of this kernel's time."""
    cnt_s = cnt_s - dense_s
    achieved = r0.alg_bytes / cnt_s / 1e9
    gkw = dict(kw, schedule=T.SCHED_DEGREE)
    ctx.gather(**gkw)
    gs = [ctx.gather(**gkw).count_seconds for _ in range(3)]
    gather_gbs = r0.alg_bytes / (sum(gs) / len(gs)) / 1e9
    copy_bytes = 2 << 30
    copy_s = T.membench_copy(copy_bytes, reps=5)
    copy_gbs = 2 * copy_bytes / copy_s / 1e9
    log(f"count kernels {cnt_s * 1e3:.2f} ms ({achieved:.0f} GB/s alg); gather-only "
        f"{sum(gs) / len(gs) * 1e3:.2f} ms ({gather_gbs:.0f} GB/s); copy {copy_gbs:.0f} GB/s")
    return {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": committed_traffic(workload),
        "alg_bytes_per_launch": r0.alg_bytes,
        "gather_only_GBs": gather_gbs,
        "frac_of_gather_only": achieved / gather_gbs,
        "copy_GBs": copy_gbs,
    }


def autotune(ctx, seconds, base):
    if seconds <= 0:
        return {}, None
    tuned, sec, n = ctx.autotune(reps=2, max_seconds=seconds, **base)
    log(f"autotune ({n} candidates, {seconds:.0f}s budget): {tuned} -> {sec * 1e3:.2f} ms/count")
    return tuned, dict(params=tuned, seconds_per_count=sec, candidates=n, budget_s=seconds)


def line_head(args, workload, value, unit, ms, dtype="int32", higher=True):
    return {
        "metric": METRICS[workload],
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": higher,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
    }


# ---------------------------------------------------------------- workloads
def bench_rmat(args, T):
    t0 = time.perf_counter()
    src, dst = T.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    t1 = time.perf_counter()
    log(f"generated {len(src)} R-MAT pairs in {t1 - t0:.1f}s; preprocessing ({args.order} order)")
    beg, adj = T.preprocess_edges(src, dst, order=args.order)
    del src, dst
    t2 = time.perf_counter()
    V, E = len(beg) - 1, len(adj)
    log(f"rmat{args.scale}: V={V} E={E} gen {t1 - t0:.1f}s preprocess {t2 - t1:.1f}s")

    ctx = T.Context(0)
    h2d = ctx.load(beg, adj)
    log(f"h2d {h2d:.2f}s ({(beg.nbytes + adj.nbytes) / h2d / 1e9:.1f} GB/s incl. pageable staging)")
    base_kw = dict(threads=args.threads, blocks=args.blocks, chunk=args.chunk)
    fresh = T.FLAG_FRESH_INDEX

    # The first count of a graph decides its walk and builds that walk's
    # per-graph index (the short-side walk's in-walk rows, or the reference
    # walk's start index).  A fresh-graph step (FLAG_FRESH_INDEX) rebuilds
    # it inside its timed device work; the headline counts fresh graphs.
    first = ctx.count(**base_kw, schedule=T.SCHED_DEGREE)
    index_s, walk = first.index_seconds, first.walk_used
    log(f"walk: {'short-side' if walk else 'reference'}; per-graph index {index_s * 1e3:.1f} ms")
    if args.only_head:
        # the headline's own steps: the same training and judge, then the
        # AUTO schedule on fresh graphs (a kernel trace of this run sees the
        # judge's four counts first, then warmup + steps headline counts)
        hkw = dict(base_kw, schedule=T.SCHED_DEGREE, flags=fresh)
        if args.schedule == "learned":
            ctx.train_cost_model(epochs=15, max_samples=200000, seed=1, **base_kw)
            hkw["schedule"] = T.SCHED_AUTO
        head = run_steps(lambda: ctx.count(**hkw), args.warmup, args.steps,
                         "headline only, fresh graph", E)
        r0 = head["r"]
        line = line_head(args, "rmat", E / head["dev_s"] / 1e9, "GTEPS", head["dev_s"] * 1e3)
        line["config"] = {"workload": f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}",
                          "V": V, "E": E, "triangles": r0.triangles, "W_probes": r0.probes,
                          "only_head": True, "index_ms": head["idx_s"] * 1e3,
                          "plan_ms": r0.plan_seconds * 1e3,
                          "count_kernel_ms": head["cnt_s"] * 1e3,
                          "core": core_config(T, r0, head["dense_s"]),
                          "schedule_used": "learned" if r0.schedule_used == T.SCHED_LEARNED
                          else "degree",
                          "head_counts": args.warmup + args.steps,
                          "alg_bytes_per_count": r0.alg_bytes}
        ctx.close()
        # one 2 GiB copy: the known-traffic kernel that calibrates byte counters
        line["config"]["copy_bytes"] = 2 << 30
        line["config"]["copy_s"] = T.membench_copy(2 << 30, reps=1)
        return line
    deg = run_steps(lambda: ctx.count(**base_kw, schedule=T.SCHED_DEGREE), args.warmup,
                    args.steps, "degree split, index kept", E)
    dkw = dict(base_kw, schedule=T.SCHED_DEGREE, flags=fresh)
    deg_fresh = run_steps(lambda: ctx.count(**dkw), args.warmup, args.steps,
                          "degree split, fresh graph", E)
    # the reference's walk: every v in N+(u), its offsets looked up per entry
    # (no per-graph index)
    ref_flags = T.FLAG_REFERENCE_WALK | T.FLAG_OFFSET_LOOKUP
    rw_kw = dict(base_kw, schedule=T.SCHED_DEGREE, flags=ref_flags)
    ref_walk = run_steps(lambda: ctx.count(**rw_kw), 1, max(1, args.steps // 2),
                         "degree split, reference walk", E)
    # the same walk reading the offsets from this build's start index
    rs_kw = dict(base_kw, schedule=T.SCHED_DEGREE, flags=T.FLAG_REFERENCE_WALK | fresh)
    ref_starts = run_steps(lambda: ctx.count(**rs_kw), 1, max(1, args.steps // 2),
                           "degree split, reference walk with the start index, fresh graph", E)
    # the reference's own launch (README: 1024 x 1024, chunk 1) and walk
    ref_kw = dict(schedule=T.SCHED_DEGREE, threads=1024, blocks=1024, chunk=1, flags=ref_flags)
    ref_launch = run_steps(lambda: ctx.count(**ref_kw), 1, max(1, args.steps // 2),
                           "degree split, reference launch and walk", E)
    train = tune = forced = None
    if args.schedule == "learned":
        mse, train_s = ctx.train_cost_model(epochs=15, max_samples=200000, seed=1, **base_kw)
        decision, jl, jd = ctx.schedule_judgement()
        log(f"cost model trained: mse {mse:.5f} (log-cycle units/32), {train_s:.2f}s device time; "
            f"judge: learned {jl * 1e3:.2f} ms vs degree {jd * 1e3:.2f} ms -> "
            f"{'learned' if decision == T.SCHED_LEARNED else 'degree'}")
        train = dict(mse=mse, seconds=train_s, judge_learned_ms=jl * 1e3, judge_degree_ms=jd * 1e3,
                     verdict="learned" if decision == T.SCHED_LEARNED else "degree")
        head_kw = dict(base_kw, schedule=T.SCHED_AUTO)
        tuned, tune = autotune(ctx, args.tune, head_kw)
        head_kw.update(tuned)
        lkw = dict(head_kw, schedule=T.SCHED_LEARNED, flags=fresh)
        forced = run_steps(lambda: ctx.count(**lkw), 1, max(1, args.steps // 2),
                           "learned order (forced), fresh graph", E)
        kept = run_steps(lambda: ctx.count(**head_kw), 1, max(1, args.steps // 2),
                         "auto, index kept", E)
        head_kw["flags"] = fresh
        head = run_steps(lambda: ctx.count(**head_kw), args.warmup, args.steps,
                         "auto (learned, judged by the degree split), fresh graph", E)
        if head["r"].schedule_used != decision:
            raise SystemExit("the schedule that ran is not the judge's verdict")
    else:
        head_kw = dict(base_kw, schedule=T.SCHED_DEGREE)
        tuned, tune = autotune(ctx, args.tune, head_kw)
        head_kw.update(tuned)
        kept = run_steps(lambda: ctx.count(**head_kw), 1, max(1, args.steps // 2),
                         "degree, index kept", E)
        head_kw["flags"] = fresh
        head = run_steps(lambda: ctx.count(**head_kw), args.warmup, args.steps,
                         "degree (head), fresh graph", E)
    runs = [head, kept, deg, deg_fresh, ref_launch, ref_walk, ref_starts] + ([forced] if forced else [])
    if len({x["r"].triangles for x in runs}) != 1:
        raise SystemExit("schedules disagree on the count")
    r0, dev_s, cnt_s = head["r"], head["dev_s"], head["cnt_s"]
    W = r0.probes
    workload = f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}"
    if args.order != "outdegree":
        workload += f"_{args.order}"
    fixture = check_fixture(f"rmat{args.scale}", beg, adj, r0.triangles) \
        if args.order == "outdegree" and args.edge_factor == 16 and args.seed == 1 else None
    line = line_head(args, "rmat", E / dev_s / 1e9, "GTEPS", dev_s * 1e3)
    line["ms_per_step_wall"] = head["wall_ms"]
    line["config"] = {
        "workload": workload,
        "generator": "R-MAT a=0.57 b=0.19 c=0.19 d=0.05, reference preprocess contract",
        "row_order": args.order,
        "V": V, "E": E, "W_probes": W, "triangles": r0.triangles,
        "oracle_parity": fixture,
        "step": "one count of a freshly loaded graph resident in HBM: walk statistics, the "
                "walk's per-graph index, plan and counting launches (hipEvents, one stream)",
        "schedule": "auto (learned cost model, judged against the degree split)"
        if args.schedule == "learned" else "degree",
        "schedule_used": "learned" if r0.schedule_used == T.SCHED_LEARNED else "degree",
        "threads": r0.threads, "blocks": r0.blocks,
        "tiers_spill_hub_wave": [r0.n_spill, r0.n_hub, r0.n_wave],
        "parallelism": "single-gpu",
        "walk": "short-side" if r0.walk_used else "reference",
        "index_ms": head["idx_s"] * 1e3,
        "plan_ms": (dev_s - cnt_s - head["idx_s"]) * 1e3,
        "count_kernel_ms": cnt_s * 1e3,
        "core": core_config(T, r0, head["dense_s"]),
        "gteps_index_kept": E / kept["dev_s"] / 1e9,
        "W_probes_reference_walk": ref_walk["r"].probes,
        "degree_split_gteps": E / deg_fresh["dev_s"] / 1e9,
        "degree_split_index_kept_gteps": E / deg["dev_s"] / 1e9,
        "degree_split_reference_walk_gteps": E / ref_walk["dev_s"] / 1e9,
        "degree_split_reference_walk_start_index_gteps": E / ref_starts["dev_s"] / 1e9,
        "degree_split_reference_launch_and_walk_gteps": E / ref_launch["dev_s"] / 1e9,
        "speedup_vs_degree_split": deg_fresh["dev_s"] / dev_s,
        "speedup_vs_degree_split_reference_walk": ref_walk["dev_s"] / dev_s,
        "speedup_vs_reference_launch_and_walk": ref_launch["dev_s"] / dev_s,
        "learned_forced_gteps": E / forced["dev_s"] / 1e9 if forced else None,
        "learned_forced_vs_degree_split": deg_fresh["dev_s"] / forced["dev_s"] if forced else None,
        "cost_model_training": train,
        "autotune": tune,
        "h2d_s": h2d, "preprocess_s": t2 - t1,
    }
    line["roofline"] = roofline(T, ctx, r0, cnt_s, dict(head_kw, flags=0), workload,
                                head["dense_s"])
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(beg, adj, E, args.cpu_seconds, r0.triangles)
    ctx.close()
    return line


def bench_chung_lu(args, T):
    n, m = 3_800_000, 16_500_000
    t0 = time.perf_counter()
    beg, adj = T.preprocess_edges(*T.gen_chung_lu(n, m, 2.5, 800.0, seed=args.seed))
    V, E = len(beg) - 1, len(adj)
    log(f"chung-lu: V={V} E={E} max d+ {np.diff(beg).max()} in {time.perf_counter() - t0:.1f}s")
    ctx = T.Context(0)
    ctx.load(beg, adj)
    fresh = T.FLAG_FRESH_INDEX
    base_kw = dict(schedule=T.SCHED_DEGREE)
    deg = run_steps(lambda: ctx.count(**base_kw, flags=fresh), args.warmup, args.steps,
                    "degree split, fresh graph", E)
    ref_launch = run_steps(lambda: ctx.count(schedule=T.SCHED_DEGREE, threads=1024, blocks=1024,
                                             chunk=1, flags=fresh), 1, args.steps,
                           "reference launch, fresh graph", E)
    tuned, tune = autotune(ctx, args.tune if args.tune > 0 else 30.0, base_kw)
    head_kw = dict(base_kw, **tuned)
    kept = run_steps(lambda: ctx.count(**head_kw), 1, args.steps, "autotuned, index kept", E)
    head = run_steps(lambda: ctx.count(**head_kw, flags=fresh), args.warmup, args.steps,
                     "autotuned, fresh graph", E)
    if len({x["r"].triangles for x in (deg, ref_launch, kept, head)}) != 1:
        raise SystemExit("launches disagree on the count")
    r0, dev_s, cnt_s = head["r"], head["dev_s"], head["cnt_s"]
    fixture = check_fixture("chunglu3.8M", beg, adj, r0.triangles) if args.seed == 1 else None
    workload = f"chunglu3.8M_m16.5M_g2.5_dmax800_seed{args.seed}"
    line = line_head(args, "chung-lu", E / dev_s / 1e9, "GTEPS", dev_s * 1e3)
    line["config"] = {
        "workload": workload, "V": V, "E": E, "W_probes": r0.probes, "triangles": r0.triangles,
        "oracle_parity": fixture,
        "step": "one count of a freshly loaded graph resident in HBM: walk statistics, the "
                "walk's per-graph index, plan and counting launches",
        "walk": "short-side" if r0.walk_used else "reference (start index)",
        "parallelism": "single-gpu",
        "index_ms": head["idx_s"] * 1e3, "plan_ms": (dev_s - cnt_s - head["idx_s"]) * 1e3,
        "count_kernel_ms": cnt_s * 1e3,
        "core": core_config(T, r0, head["dense_s"]),
        "gteps_index_kept": E / kept["dev_s"] / 1e9,
        "default_launch_gteps": E / deg["dev_s"] / 1e9,
        "reference_launch_gteps": E / ref_launch["dev_s"] / 1e9,
        "speedup_vs_default_launch": deg["dev_s"] / dev_s,
        "speedup_vs_reference_launch": ref_launch["dev_s"] / dev_s,
        "autotune": tune,
    }
    rl = roofline(T, ctx, r0, cnt_s, head_kw, workload, head["dense_s"])
    # The graph (adj 66 MB, offsets 15 MB, start index 132 MB) is resident in
    # the 256 MB Infinity Cache, so the bound is the cache's bandwidth,
    # measured here with the same copy kernel on a cache-resident buffer.
    mall_bytes = 32 << 20
    mall_s = T.membench_copy(mall_bytes, reps=50)
    mall_gbs = 2 * mall_bytes / mall_s / 1e9
    log(f"cache-resident copy ({2 * mall_bytes >> 20} MB moved): {mall_gbs:.0f} GB/s")
    rl.update({"bound": "mall", "peak": mall_gbs, "frac": rl["achieved"] / mall_gbs,
               "peak_source": f"measured: copy kernel on a {mall_bytes >> 20} MiB buffer "
                              f"(source + destination resident in the Infinity Cache)",
               "hbm_peak": HBM_PEAK_GBS, "frac_of_hbm_peak": rl["achieved"] / HBM_PEAK_GBS})
    line["roofline"] = rl
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(beg, adj, E, args.cpu_seconds, r0.triangles)
    ctx.close()
    return line


def bench_batch(args, T):
    import oracle

    ng = args.graphs
    t0 = time.perf_counter()
    beg, adj, vb = T.batch_generate(ng, seed=args.seed, nmin=args.nmin, nmax=args.nmax,
                                    dmin=8.0, dmax=64.0)
    V, E = len(beg) - 1, len(adj)
    log(f"batch: {ng} graphs V={V} E={E} generated in {time.perf_counter() - t0:.1f}s")
    ctx = T.Context(0)
    t1 = time.perf_counter()
    ctx.load_batch(beg, adj, vb)
    log(f"loaded in {time.perf_counter() - t1:.1f}s")
    mse, train_s = ctx.train_cost_model(epochs=15, max_samples=200000, seed=1)
    decision, jl, jd = ctx.schedule_judgement()
    verdict = "learned" if decision == T.SCHED_LEARNED else "degree"
    log(f"cost model trained: mse {mse:.5f}, {train_s:.2f}s device time; judge: learned "
        f"{jl * 1e3:.1f} ms vs degree {jd * 1e3:.1f} ms -> {verdict}")
    kw = dict(schedule=T.SCHED_AUTO)
    tuned, tune = autotune(ctx, args.tune if args.tune > 0 else 60.0, kw)
    kw.update(tuned)
    counts, done, r = ctx.count_batch(**kw)
    if r.schedule_used != decision:
        raise SystemExit("the schedule that ran is not the judge's verdict")
    log(f"eager batch count: {r.triangles} triangles in {r.seconds * 1e3:.1f} ms")
    fx_path = os.path.join(ROOT, "tests", "golden", "large_counts.json")
    fx = json.load(open(fx_path)).get("batch10000") if os.path.exists(fx_path) else None
    same = (fx is not None and (ng, args.seed, args.nmin, args.nmax) == (10000, 1, 1000, 200000)
            and (fx["V"], fx["E"]) == (V, E))
    if same:
        import xxhash

        if xxhash.xxh64(memoryview(adj)).hexdigest() != fx["adj_xxh64"]:
            raise SystemExit("the generated batch differs from the fixture's")
        want = np.load(os.path.join(ROOT, "tests", "golden", fx["per_graph_counts"]))
        parity = f"all {ng} graphs equal the oracle's committed counts (tests/golden)"
    else:  # every graph against the oracle, in slices so progress shows
        t2 = time.perf_counter()
        want = np.zeros(ng, np.uint64)
        step = max(1, ng // 10)
        for g0 in range(0, ng, step):
            g1 = min(ng, g0 + step)
            lo, hi = int(vb[g0]), int(vb[g1])
            b = beg[lo: hi + 1] - beg[lo]
            want[g0:g1] = oracle.count_hash_batch(b, adj[beg[lo]: beg[hi]] - np.int32(lo),
                                                  vb[g0: g1 + 1] - lo)
            log(f"oracle graphs [{g0}, {g1}) done, {time.perf_counter() - t2:.1f}s")
        parity = f"all {ng} graphs equal the oracle (computed in this run)"
    bad = np.nonzero(counts != want)[0]
    if len(bad):
        raise SystemExit(f"per-graph counts differ from the oracle at graphs {bad[:10].tolist()}")
    log(f"all {ng} per-graph counts match the oracle")
    gkw = dict(kw, flags=T.FLAG_HIP_GRAPH)
    for _ in range(args.warmup):
        ctx.count_batch(**gkw)
    secs, dones = [], []
    t3 = time.perf_counter()
    for _ in range(args.steps):
        c, d, rr = ctx.count_batch(**gkw)
        if not np.array_equal(c, want):
            raise SystemExit("a hipGraph replay changed a per-graph count")
        secs.append(rr.seconds)
        dones.append(d)
    wall = (time.perf_counter() - t3) / args.steps
    dev_s = float(np.mean(secs))
    lat = np.concatenate(dones)
    p50, p99 = np.percentile(lat, [50, 99])
    log(f"hipGraph replay: {dev_s * 1e3:.1f} ms/batch, {ng / dev_s:.0f} graphs/s, "
        f"{E / dev_s / 1e9:.3f} GTEPS, latency p50 {p50 * 1e3:.2f} ms p99 {p99 * 1e3:.2f} ms")
    eager = run_steps(lambda: ctx.count_batch(**kw)[2], 0, max(1, args.steps // 2),
                      "eager batch", E)
    lkw = dict(kw, schedule=T.SCHED_LEARNED, flags=T.FLAG_HIP_GRAPH)
    ctx.count_batch(**lkw)
    forced = run_steps(lambda: ctx.count_batch(**lkw)[2], 0, max(1, args.steps // 2),
                       "learned order (forced), hipGraph", E)
    ekw = dict(kw, schedule=T.SCHED_LEARNED)
    forced_eager = run_steps(lambda: ctx.count_batch(**ekw)[2], 0, 2, "learned order (forced), eager", E)
    # the plain degree split on the same batch, same launch parameters
    dkw = dict(kw, schedule=T.SCHED_DEGREE, flags=T.FLAG_HIP_GRAPH)
    ctx.count_batch(**dkw)
    deg = run_steps(lambda: ctx.count_batch(**dkw)[2], 0, max(1, args.steps // 2),
                    "degree split, hipGraph", E)
    _, ddone, _ = ctx.count_batch(**dkw)
    dp50, dp99 = np.percentile(ddone, [50, 99])
    workload = f"batch{ng}_er_ba_n{args.nmin}-{args.nmax}_d8-64_seed{args.seed}"
    line = line_head(args, "batch", ng / dev_s, "graphs/s", dev_s * 1e3)
    line["ms_per_step_wall"] = wall * 1e3
    line["config"] = {
        "workload": workload, "graphs": ng, "V": V, "E": E, "W_probes": r.probes,
        "triangles": r.triangles, "parallelism": "single-gpu",
        "schedule": "auto (learned cost model, judged against the degree split)",
        "schedule_used": verdict,
        "launch": "hipGraph replay (plan + count captured once)",
        "gteps": E / dev_s / 1e9,
        "latency_p50_ms": p50 * 1e3, "latency_p99_ms": p99 * 1e3,
        "latency_definition": "per graph: first workgroup start to the end of the graph's "
                              "last vertex, device wall clock, pooled over the timed replays",
        "eager_ms_per_batch": eager["dev_s"] * 1e3,
        "eager_plan_ms": (eager["dev_s"] - eager["cnt_s"]) * 1e3,
        "degree_split_graphs_per_s": ng / deg["dev_s"],
        "degree_split_latency_p50_ms": dp50 * 1e3, "degree_split_latency_p99_ms": dp99 * 1e3,
        "learned_forced_graphs_per_s": ng / forced["dev_s"],
        "learned_forced_eager_plan_ms": (forced_eager["dev_s"] - forced_eager["cnt_s"]) * 1e3,
        "auto_vs_degree_split": deg["dev_s"] / dev_s,
        "per_graph_parity": parity,
        "cost_model_training": dict(mse=mse, seconds=train_s, judge_learned_ms=jl * 1e3,
                                    judge_degree_ms=jd * 1e3, verdict=verdict),
        "autotune": tune,
    }
    line["roofline"] = roofline(T, ctx, r, r.count_seconds, kw, workload)
    if not args.no_cpu:
        # the oracle on every stride-th graph (parallel over graphs), scaled
        # to the batch by edge share
        stride = max(1, int(args.cpu_graph_stride))
        gs = np.arange(0, ng, stride)
        t2 = time.perf_counter()
        # the sampled graphs as one batch with their own global ids
        sv = np.zeros(len(gs) + 1, np.int64)
        sv[1:] = np.cumsum(vb[gs + 1] - vb[gs])
        sb_parts, sa_parts, e0 = [np.zeros(1, np.int64)], [], 0
        for k, g in enumerate(gs):
            b, a = T.batch_graph(beg, adj, vb, int(g))
            sb_parts.append(b[1:] + e0)
            sa_parts.append(a + np.int32(sv[k]))
            e0 += len(a)
        sb, sa = np.concatenate(sb_parts), np.concatenate(sa_parts).astype(np.int32)
        t3 = time.perf_counter()
        got = oracle.count_hash_batch(sb, sa, sv)
        oracle_s = time.perf_counter() - t3
        if not np.array_equal(got, want[gs]):
            raise SystemExit("the oracle's sampled graphs differ from the reference counts")
        frac = len(sa) / E
        line["cpu_baseline"] = {
            "value": len(gs) / oracle_s, "unit": "graphs/s", "cores": oracle.max_threads(),
            "kind": "port",
            "sample": f"oracle hashed count of every {stride}-th graph ({len(gs)} graphs, "
                      f"{frac * 100:.1f}% of the edges) in {oracle_s:.1f}s (parallel over "
                      f"graphs; {t3 - t2:.1f}s to cut them out); the whole batch would take "
                      f"about {oracle_s / frac:.0f}s",
        }
    ctx.close()
    return line


def bench_tiles(args, T):
    import oracle  # noqa: F401  (checker only: the totals are compared below)

    n = args.parts
    t0 = time.perf_counter()
    src, dst = T.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    log(f"generated {len(src)} R-MAT pairs in {time.perf_counter() - t0:.1f}s; preprocessing")
    t0 = time.perf_counter()
    beg, adj = T.preprocess_edges(src, dst)
    del src, dst
    V, E = len(beg) - 1, len(adj)
    log(f"rmat{args.scale}: V={V} E={E} preprocessed in {time.perf_counter() - t0:.1f}s; "
        f"partitioning")
    t0 = time.perf_counter()
    tiles = T.partition_csr(beg, adj, n)
    log(f"{n}x{n} tiles in {time.perf_counter() - t0:.1f}s")
    ctx = T.Context(0)
    ctx.load(beg, adj)
    whole = run_steps(lambda: ctx.count(flags=T.FLAG_FRESH_INDEX), 1, 2,
                      "unpartitioned, fresh graph", E)
    ctx.close()  # the tile driver keeps its own device buffers
    fixture = check_fixture(f"rmat{args.scale}", beg, adj, whole["r"].triangles) \
        if args.edge_factor == 16 and args.seed == 1 else None
    part = run_steps(lambda: T.count_tiles(tiles, n), 1, args.steps, f"{n}x{n} tiles", E)
    rt = part["r"]
    log(f"tiles: stream {rt.seconds * 1e3:.1f} ms = counting launches {rt.count_seconds * 1e3:.1f} ms"
        f" + tile walk builds {rt.index_seconds * 1e3:.1f} ms + copies not hidden; host time in "
        f"uploads {rt.h2d_seconds * 1e3:.1f} ms")
    if part["r"].triangles != whole["r"].triangles:
        raise SystemExit(f"partitioned {part['r'].triangles} != unpartitioned "
                         f"{whole['r'].triangles}")
    dev_s = part["dev_s"]
    workload = f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}_tiles{n}x{n}"
    tile_bytes = sum(b.nbytes + a.nbytes for b, a in tiles)
    line = line_head(args, "tiles", E / dev_s / 1e9, "GTEPS", dev_s * 1e3)
    rt = part["r"]
    line["config"] = {
        "workload": workload, "V": V, "E": E, "triangles": rt.triangles,
        "oracle_parity": fixture,
        "jobs": n ** 3, "tile_bytes": tile_bytes, "parallelism": "single-gpu",
        "walk": "short-side (global out-degree)" if rt.walk_used else "reference",
        "W_probes": rt.probes,
        "unpartitioned_gteps": E / whole["dev_s"] / 1e9,
        "timing": "streamed count incl. asynchronous host-to-device tile copies and the tile "
                  "walks' builds",
        "count_launches_ms": rt.count_seconds * 1e3,
        "tile_walk_builds_ms": rt.index_seconds * 1e3,
        "host_upload_ms": rt.h2d_seconds * 1e3,
        "overlap_note": "the stream's time against its parts: counting launches, walk builds "
                        "(synchronous) and host time inside the tile uploads (waiting for a "
                        "free pinned staging buffer, then a parallel copy into it), which "
                        "overlaps the counting of the previous job",
    }
    r0 = part["r"]
    achieved = r0.alg_bytes / dev_s / 1e9
    line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": committed_traffic(workload),
                        "alg_bytes_per_launch": r0.alg_bytes}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(beg, adj, E, args.cpu_seconds, whole["r"].triangles)
    return line


def bench_ingest(args, T):
    """Host-only timing of the input tools against the reference's tools."""
    import shutil
    import subprocess
    import tempfile

    import oracle

    tool = os.path.join(T.BIN_DIR, "trust_tool")
    src, dst = T.gen_rmat(args.ingest_scale, args.edge_factor, seed=args.seed)
    m = len(src)
    work = tempfile.mkdtemp(prefix="trust_ingest_")
    try:
        text = os.path.join(work, "e.txt")
        np.savetxt(text, np.stack([src, dst], 1), fmt="%d", delimiter=" ")
        del src, dst
        log(f"ingest: {m} pairs, {os.path.getsize(text) / 1e6:.0f} MB of text in {work}")

        peaks = {}

        def run(cmd, cwd, outputs, ok_abort=False, key=None):
            # wall time, and the peak of the child's anonymous resident memory
            t = time.perf_counter()
            p = subprocess.Popen(cmd, cwd=cwd, stdout=subprocess.DEVNULL,
                                 stderr=subprocess.PIPE)
            peak = 0
            while p.poll() is None:
                try:
                    with open(f"/proc/{p.pid}/status") as f:
                        for ln in f:
                            if ln.startswith("RssAnon:"):
                                peak = max(peak, int(ln.split()[1]))
                except OSError:
                    pass
                time.sleep(0.02)
            dt = time.perf_counter() - t
            err = p.stderr.read().decode(errors="replace")
            if p.returncode != 0 and not (ok_abort and all(
                    os.path.exists(os.path.join(cwd, f)) for f in outputs)):
                raise SystemExit(f"{cmd} failed ({p.returncode}): {err[-400:]}")
            if key:
                peaks[key] = peak / 2**20  # GiB
            return dt

        steps = [  # (label, ours, reference tool, outputs, dir)
            ("undirect", [tool, "undirect", text, "1.mmio"], ["fromDirectToUndirect", text],
             ["1.mmio"], "txt"),
            ("preprocess", [tool, "preprocess", "1.mmio", "."], ["preprocess"],
             ["begin.bin", "adjacent.bin"], "txt"),
            ("preprocess_plain", [tool, "preprocess", "../txt/1.mmio", ".", "--order", "plain"],
             ["preprocess_plain"], ["begin.bin", "adjacent.bin", "edge"], "plain"),
            ("preprocess_indegree", [tool, "preprocess", "../txt/1.mmio", ".", "--order", "indegree"],
             ["preprocess_indegree"], ["begin.bin", "adjacent.bin", "edge"], "indegree"),
            ("txt2bin", [tool, "txt2bin", text, str(m)], ["TransformTxt2Binary", text, str(m)],
             ["prop.txt", "degree", "BinEdgelist"], "bin"),
            ("preprocess_bin", [tool, "preprocess-bin", "."], ["preprocess_out"],
             ["begin.bin", "adjacent.bin"], "bin"),
        ]
        have_ref = all(oracle.ref_tool(s[2][0]) for s in steps)
        times, ref_times, identical = {}, {}, {}
        for side in ("ours", "ref") if have_ref else ("ours",):
            for label, ours, ref, outputs, sub in steps:
                d = os.path.join(work, side, sub)
                os.makedirs(d, exist_ok=True)
                if sub in ("plain", "indegree") and side == "ref":
                    shutil.copy(os.path.join(work, side, "txt", "1.mmio"), d)
                cmd = ours if side == "ours" else [oracle.ref_tool(ref[0])] + ref[1:]
                dt = run(cmd, d, outputs, ok_abort=side == "ref", key=f"{side} {label}")
                (times if side == "ours" else ref_times)[label] = dt
                log(f"{side} {label}: {dt:.2f}s, peak anon {peaks[f'{side} {label}']:.2f} GiB")
        # tile writer (Preprocess/partition.cpp) on the preprocessed graph
        n = args.parts
        tile_files = [os.path.join(f"partition{i}_{j}", f) for i in range(n) for j in range(n)
                      for f in ("begin.bin", "adjacent.bin", "map", "division")]
        for side in ("ours", "ref") if have_ref and n <= 10 else ("ours",):
            d = os.path.join(work, side, "tiles")
            os.makedirs(d)
            for f in ("begin.bin", "adjacent.bin"):
                os.symlink(os.path.join(work, "ours", "txt", f), os.path.join(d, f))
            if side == "ref":  # the reference tool needs the folders to exist
                for i in range(n):
                    for j in range(n):
                        os.makedirs(os.path.join(d, f"partition{i}_{j}"))
            cmd = [tool, "partition", str(n), "."] if side == "ours" else \
                [oracle.ref_tool("partition"), str(n)]
            dt = run(cmd, d, tile_files, key=f"{side} partition")
            (times if side == "ours" else ref_times)["partition"] = dt
            log(f"{side} partition n={n}: {dt:.2f}s, peak anon {peaks[f'{side} partition']:.2f} GiB")
        if have_ref:
            for label, _, _, outputs, sub in steps + [("partition", 0, 0, tile_files, "tiles")]:
                if label in ref_times:
                    identical[label] = all(
                        open(os.path.join(work, "ours", sub, f), "rb").read() ==
                        open(os.path.join(work, "ref", sub, f), "rb").read() for f in outputs)
            if not all(identical.values()):
                raise SystemExit(f"outputs differ from the reference tools: {identical}")
        beg = np.fromfile(os.path.join(work, "ours", "txt", "begin.bin"), np.int64)
        V, E = len(beg) - 1, int(beg[-1])
    finally:
        shutil.rmtree(work, ignore_errors=True)
    text_s = times["undirect"] + times["preprocess"]
    line = line_head(args, "ingest", m / text_s / 1e6, "Mpairs/s", text_s * 1e3, dtype="int32")
    line["n_gpus"] = 0
    line["steps"], line["warmup"] = 1, 0
    line["config"] = {
        "workload": f"ingest_rmat{args.ingest_scale}_ef{args.edge_factor}_seed{args.seed}",
        "pairs": m, "V": V, "E": E, "host_threads": os.cpu_count(),
        "partition_parts": args.parts,
        "seconds": times,
        "reference_seconds": ref_times or None,
        "peak_anon_gib": {k: round(v, 3) for k, v in peaks.items()},
        "speedup_vs_reference": {k: ref_times[k] / times[k] for k in times} if ref_times else None,
        "byte_identical": identical or None,
        "parallelism": "host (OpenMP)",
    }
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(METRICS), default="rmat")
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--schedule", choices=["degree", "learned"], default="learned")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--blocks", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--tune", type=float, default=0.0,
                    help="autotune budget in seconds (0: rmat off; chung-lu 30; batch 60)")
    ap.add_argument("--graphs", type=int, default=10000)
    ap.add_argument("--nmin", type=int, default=1000)
    ap.add_argument("--nmax", type=int, default=200000)
    ap.add_argument("--parts", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-graph-stride", type=int, default=5,
                    help="batch: the cpu baseline counts every n-th graph")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-head", action="store_true",
                    help="rmat: time the headline steps alone (no comparison legs, no cpu "
                         "baseline): what a kernel profile of the headline should see")
    ap.add_argument("--order", choices=["outdegree", "plain", "indegree"], default="outdegree")
    ap.add_argument("--ingest-scale", type=int, default=21)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X only: this build runs the count on a single GPU "
              "(multi-GPU sharding is a design note, not implemented). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import trust_amd as T

    log(f"host: {os.cpu_count()} cpus, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}")
    fn = {"rmat": bench_rmat, "chung-lu": bench_chung_lu, "batch": bench_batch,
          "tiles": bench_tiles, "ingest": bench_ingest}[args.workload]
    print(json.dumps(fn(args, T)), flush=True)


if __name__ == "__main__":
    main()
