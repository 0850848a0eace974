"""PHOLD events/s at 2^20 LPs on one MI355X.

A step is one complete PHOLD simulation of the workload: 2^20 LPs, 32 initial
events per LP (2^25 events in flight), end time 50, float32 timestamps
ts' = (ts + 0.1) + Exp(1), destination self with p = 0.5 else uniform, Philox
streams from seed 42.  The queue is populated outside the timed region (the
reference times the simulation loop only, src/des.cu:148-215); the timed region
of a step is the whole step loop, driven by hipGraph replays.

Beside the JSON line's value:
  roofline      the dominant kernel's algorithmic bytes per launch / its mean
                launch duration (HIP events on the queue's stream), against 8 TB/s;
                traffic = HBM bytes per launch from the rocprofv3 counter summary
                committed in profiles/ (null when absent)
  cpu_baseline  the sequential std::priority_queue oracle on a bounded sample of
                the same workload (end time 0.5), one host core
  phold_batch   32 PHOLD instances of uneven size (2^U[12,18] LPs) in one pool, against
                each alone with a host sync per step
  astar_single  the 4096 x 4096 grid query: tile search, and 65536 open lists in a local heap queue
  astar_batch   512 synthetic grid queries (side 2^U[8,11], density U[0.1,0.3],
                8-connected) in one persistent launch: queries/s, p50/p99 completion
                time, against the plain one-query-per-launch path (host sync per
                expansion step) on the first queries, and the oracle's A* on one core
  serving       a Poisson trace of mixed PHOLD / A* requests served by continuous
                batching: requests/s, p50/p99 latency, against serving them one at a
                time with a host sync per step
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
ASTAR_SEED = 2024
ASTAR_QUERIES = 512
RECORD = 16          # bytes per event record
METRIC = "PHOLD events/s @2^20 LPs; batched A* queries/s + p99 latency vs plain"

WORKLOADS = {
    "phold_2p20_k32_t50": dict(num_lps=1 << 20, init_per_lp=32, end_time=50.0),
    "phold_1024_k16_t100": dict(num_lps=1024, init_per_lp=16, end_time=100.0),
}
QUEUES = {
    "laq": dict(capacity_per_lp=128, stage_per_lp=16),
    "lhq": dict(capacity_per_lp=128, stage_per_lp=16),
    "lsq": dict(capacity_per_lp=128, stage_per_lp=16),
    "gaq": dict(capacity_per_lp=64, stage_per_lp=2),
    "hq": dict(capacity_per_lp=40, stage_per_lp=1),
}
DRAIN = False
FAR = 0            # LAQ two-tier mode: far-store records per LP (0 = one sorted ring per LP)
FAR_STEP = 0.0     # LAQ two-tier: near-horizon step (0 = 4 x lookahead)
CAL = 0            # LAQ calendar mode: windows in the calendar (0 = off)
CAL_SLOTS = 8      # LAQ calendar mode: a window's log records per LP


def make(gpq, kind, wl, stats=False, checks=False):
    heaps = 4 if kind == "hq4" else 1      # hq4: four heaps, each owning a quarter of the LPs
    kind = "hq" if kind == "hq4" else kind
    args = dict(QUEUES[kind])
    if DRAIN:
        args["stage_per_lp"] = 32      # a drained window can send ~3 events to one LP on average
    if kind == "hq":
        # one node holds one safe event per LP of its heap
        args.update(node_size=max(-(-wl["num_lps"] // heaps), 32), heaps=heaps)
    if kind == "laq" and FAR:
        args.update(capacity_per_lp=64, far_per_lp=FAR, far_step=FAR_STEP)
    if kind == "laq" and CAL and DRAIN == "fused":
        args.update(capacity_per_lp=CAL_SLOTS, stage_per_lp=32, far_per_lp=FAR or 128, calendar_windows=CAL)
    if kind in SORT_TILES:
        args["sort_tile"] = SORT_TILES[kind]
    q = gpq.Queue(kind, num_lps=wl["num_lps"], lookahead=0.1, horizon=wl["end_time"], block_threads=256,
                  collect_stats=stats, **args)
    m = gpq.Phold(q, init_per_lp=wl["init_per_lp"], seed=42, end_time=wl["end_time"], checks=checks, drain=DRAIN)
    m.model_init()
    return q, m


def roofline(gpq, kind, wl, log):
    """Mean launch durations of the model step kernel and the queue post kernel from a
    HIP-event-instrumented run, and their algorithmic bytes from the device work
    counters of a second (identical, deterministic) run.

    The end-to-end PHOLD bytes are structure-independent: per committed event the record's
    lifecycle (written to staging, read from staging, written into the queue, read at
    dequeue: 4 x 16 B) and the LP's state (12 B read + 12 B write), 88 B in all, plus 4 B per
    LP and step for the window scan.  Bytes a kernel moves beyond that because of the queue's
    structure (records displaced by a sorted insert, far-store moves, LP words beyond the
    head) are reported as overhead, not as achieved bandwidth."""
    q, m = make(gpq, kind, wl)
    prof = m.profile(batch=64)
    m.model_finish(); q.queue_finish()
    q, m = make(gpq, kind, wl, stats=True)
    before = q.stats()
    rep = gpq.des_run(m, mode="graph", batch=64)
    after = q.stats()
    m.model_finish(); q.queue_finish()
    d = {k: after[k] - before[k] for k in after}
    L = wl["num_lps"]
    events = rep["events"]
    if prof["events"] != events:
        raise RuntimeError("roofline: profiled run committed %d events, counted run %d" % (prof["events"], events))
    steps = prof["steps"]                     # launches of the model step kernel in the profiled run
    launches = max(d["post_launches"], 1)     # post (merge) launches of the counted run
    scan = 12 if FAR else 4                   # LP words the merge scans: head (+ far minimum and count)
    # structure-dependent bytes per launch, from the kernels' own counters
    overhead = {"displaced_records": RECORD * 2 * d["displaced"] / launches,
                "far_store_moves": d["aux_bytes"] / launches,
                "lp_words_beyond_head": (scan - 4) * d["scanned_lps"] / launches}
    path_bytes = 88.0 * events + 4.0 * L * steps       # over all `steps` launches
    if DRAIN == "fused":
        # one kernel merges the staged events and commits the window: the whole path's bytes
        # over the same launch set whose durations were measured
        kernels = {
            "merge_commit_step": {"us": prof["step_kernel_us"], "alg_bytes": path_bytes / steps,
                                  "launches": steps, "overhead_bytes": overhead},
        }
        if not CAL:           # the calendar step's last block moves the window: no publish launch
            kernels["window_publish"] = {"us": prof["window_kernel_us"], "alg_bytes": 24.0, "launches": steps}
    else:
        # queue post: the sorted insert of k records displacing d (16*(2k+2d)), 4 B per scanned
        # head and per block partial; model step: 88 B per committed event + 4 B per LP
        post_bytes = (RECORD * (2 * d["inserted"] + 2 * d["displaced"]) + 4 * d["scanned_lps"]) / launches \
            + 4 * ((L + 255) // 256)
        kernels = {
            "queue_post": {"us": prof["post_kernel_us"], "alg_bytes": post_bytes, "launches": launches,
                           "overhead_bytes": {"far_store_moves": overhead["far_store_moves"],
                                              "lp_words_beyond_head": overhead["lp_words_beyond_head"]}},
            "model_step": {"us": prof["step_kernel_us"], "alg_bytes": path_bytes / steps, "launches": steps},
            "window_publish": {"us": prof["window_kernel_us"], "alg_bytes": 4 * ((L + 255) // 256) + 24,
                               "launches": steps},
        }
    for k in kernels.values():
        k["gbps"] = k["alg_bytes"] / (k["us"] * 1e-6) / 1e9 if k["us"] > 0 else 0.0
    dom = max(kernels, key=lambda k: kernels[k]["us"])
    log("profile: %s" % json.dumps({"prof": prof, "counters": d, "kernels": kernels}))
    return dom, kernels, d


def host_cpu():
    """Model name and logical core count of the host CPU (the lscpu "Model name" field)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, os.cpu_count(), avail


def cpu_baseline(wl, log):
    """The oracle on the host: the sequential std::priority_queue simulation on one core
    (the value), and the windowed drain on every core this process may use (at most 16,
    the share of one GPU's host cores), each on a bounded prefix of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle()
    sample_T = 0.5
    res, _, _ = o.phold("seq", wl["num_lps"], wl["init_per_lp"], sample_T)
    log("cpu_baseline: %d events in %.2fs" % (res.events, res.seconds))
    model, ncpu, avail = host_cpu()
    out = {"value": res.events / res.seconds, "unit": "events/s", "cores": 1, "kind": "port",
           "sample": "sequential std::priority_queue oracle, %d LPs x %d events, end time %.2f "
                     "(%d events), event loop only" % (wl["num_lps"], wl["init_per_lp"], sample_T, res.events),
           "cpu_model": model, "host_logical_cpus": ncpu}
    threads = max(1, min(16, avail))
    par_T = 2.0
    pres, _, _ = o.phold("parallel", wl["num_lps"], wl["init_per_lp"], par_T, threads=threads)
    log("cpu_baseline all cores: %d events in %.2fs on %d threads" % (pres.events, pres.seconds, threads))
    out["all_cores"] = {"value": pres.events / pres.seconds, "unit": "events/s", "cores": threads,
                        "sample": "windowed drain over %d host threads, end time %.1f (%d events), event loop only"
                                  % (threads, par_T, pres.events)}
    return out




def pct(v, p):
    v = sorted(v)
    k = max(0, min(len(v) - 1, int(-(-p * len(v) // 100)) - 1))
    return v[k]


def astar_bench(gpq, log, plain_n=ASTAR_QUERIES, cpu=True):
    specs = [gpq.query_synthetic(ASTAR_SEED, q) for q in range(ASTAR_QUERIES)]
    t0 = time.perf_counter()
    tuning, tuned_s = gpq.autotune_batch(specs, sample=64)      # threads per query x staging, on 64 queries
    tune_wall = time.perf_counter() - t0
    log("astar autotune: %s (%.1f ms on the sample, %.2f s to tune)" % (tuning, 1e3 * tuned_s, tune_wall))
    threads = tuning["threads_per_query"]
    with gpq.AstarBatch(specs, **tuning) as b:
        b.run("graph")                                  # capture + warm up
        t = b.run("graph")
        start, finish = b.timing()
        costs, expanded = b.results()
    # the plain one-query-per-launch path on the same queries (every query, one at a time, a
    # host sync per expansion step), and the batched launch on exactly those queries
    with gpq.AstarBatch(specs[:plain_n], **tuning) as b2:
        tp = b2.run("plain")
        cp, _ = b2.results()
        b2.run("graph")
        tb = b2.run("graph")
    if list(cp) != list(costs[:plain_n]):
        raise RuntimeError("A*: plain and batched costs differ")
    log("astar plain: %d queries in %.2fs; batched on the same queries %.4fs" % (plain_n, tp, tb))
    out = {"queries": ASTAR_QUERIES, "seed": ASTAR_SEED, "threads_per_query": threads,
           "tuning": dict(tuning, autotune_seconds=tune_wall),
           "queries_per_s": ASTAR_QUERIES / t, "seconds": t,
           "expanded_per_s": float(expanded.sum()) / t,
           "p50_completion_ms": pct(finish, 50) / 1000.0, "p99_completion_ms": pct(finish, 99) / 1000.0,
           "plain_queries": plain_n, "plain_queries_per_s": plain_n / tp,
           "batched_same_queries_per_s": plain_n / tb,
           "batched_vs_plain": tp / tb, "batched_vs_plain_basis": "the same %d queries, both timed" % plain_n,
           "unreachable": int((costs < 0).sum())}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle
        o = Oracle()
        n_cpu, t_cpu = 0, 0.0
        for q in range(ASTAR_QUERIES):
            s = specs[q]
            grid = o.grid(s.side, s.side, ASTAR_SEED, 1, q, s.threshold)
            t0 = time.perf_counter()
            c, _ = o.astar(grid, s.side, s.side, 8, (s.sx, s.sy), (s.tx, s.ty))
            t_cpu += time.perf_counter() - t0
            n_cpu += 1
            if c != costs[q]:
                raise RuntimeError("A*: query %d cost %d, oracle %d" % (q, costs[q], c))
            if t_cpu > 20.0:
                break
        out["cpu_queries_per_s"] = n_cpu / t_cpu
        out["cpu_sample"] = "oracle A* (binary heap), 1 core, first %d queries, search only" % n_cpu
        out["costs_vs_cpu"] = "match (%d queries)" % n_cpu
    log("astar_batch: %s" % json.dumps(out))
    return out


def per_queue_bench(gpq, wl, log, drive="graph"):
    """The same workload through every queue kind behind the one API (one timed run each);
    drive "plain" launches every step eagerly (a host sync per step), for kernel tracing."""
    global DRAIN, CAL
    out = {}
    saved, saved_cal = DRAIN, CAL
    runs = [("laq", False), ("lhq", True), ("lhq", False), ("lsq", True), ("gaq", False), ("hq", False),
            ("hq4", False)]
    if saved == "fused":
        runs.insert(0, ("laq", True))             # the two-tier LAQ with a separate post kernel
        if saved_cal:
            runs.insert(0, ("laq", "fused"))      # the two-tier sorted ring with the merge-and-commit step
    for kind, drain in runs:
        DRAIN = drain
        CAL = saved_cal if drain == "fused" and not saved_cal else 0
        q, m = make(gpq, kind, wl)
        q.sync()
        t0 = time.perf_counter()
        rep = gpq.des_run(m, mode=drive, batch=64)
        dt = time.perf_counter() - t0
        h = m.trace_hash()
        m.model_finish()
        q.queue_finish()
        key = kind + ("_ring_fused" if drain == "fused" else "_drain" if drain else "_one_event")
        out[key] = {"events_per_s": rep["events"] / dt, "seconds": dt, "windows": rep["windows"],
                    "trace_hash": "%016x" % h}
        log("per_queue %s: %.3e events/s, %d windows, hash %016x" % (key, rep["events"] / dt, rep["windows"], h))
    DRAIN, CAL = saved, saved_cal
    return out


# Kernel groups of the other queue kinds, with the structure-independent byte model of the
# operation each implements (per record the queue's work counters report).
QUEUE_GROUPS = {
    "gaq": ["sort", "merge", "select"],
    "hq": ["sort", "merge", "select"],
    "lhq": ["post"],
    "lsq": ["post"],
}
GROUP_KERNELS = {
    "sort": ("tile_sort_kernel", "merge_pass_kernel"),
    "merge": ("merge_tasks_kernel",),
    "select": ("select_",),
}


def queue_roofline(gpq, kind, wl, log, counts):
    """One eager run (a launch sequence per step) of the workload through `kind`. With
    `counts`, the queue's work counters are on and the line carries each kernel group's
    algorithmic bytes; without, the run is the one a kernel trace times (the counters add
    launches and atomics). The runs are deterministic, so both see the same work."""
    global DRAIN, CAL
    DRAIN = kind in ("lhq", "lsq")       # the local heaps drain each LP's window per step
    CAL = 0
    q, m = make(gpq, kind, wl, stats=counts)
    q.sync()
    t0 = time.perf_counter()
    rep = gpq.des_run(m, mode="plain")
    dt = time.perf_counter() - t0
    st = q.stats()
    h = m.trace_hash()
    m.model_finish()
    q.queue_finish()
    out = {"queue_roofline": kind, "workload": wl, "drive": "plain (eager)", "events": rep["events"],
           "windows": rep["windows"], "seconds": dt, "trace_hash": "%016x" % h}
    if counts:
        L = wl["num_lps"]
        g = {}
        if kind in ("gaq", "hq"):
            g["sort"] = {"alg_bytes": 2.0 * RECORD * st["sorted"], "records": st["sorted"],
                         "model": "sort of N records: 2 x 16 B per record (merge passes are overhead)"}
            g["merge"] = {"alg_bytes": 2.0 * RECORD * st["merged"], "records": st["merged"],
                          "model": "merge of runs a and b: 2 x 16 B x (a + b)"}
            g["select"] = {"alg_bytes": 2.0 * RECORD * st["selected"] + st["selected"] / 8.0,
                           "records": st["selected"],
                           "model": "stable compaction of the scanned prefix of N records: 2 x 16 B x N + N/8"}
        else:
            moved = st["displaced"] if kind == "lhq" else 0
            g["post"] = {"alg_bytes": RECORD * (2.0 * st["inserted"] + 2.0 * moved) + 4.0 * st["scanned_lps"]
                         + 4.0 * ((L + 255) // 256) * st["post_launches"],
                         "records": st["inserted"], "moved": moved,
                         "splay_visits": st["displaced"] if kind == "lsq" else 0,
                         "model": "insert of k records moving d (heap sift): 16 B x (2k + 2d), plus 4 B per "
                                  "scanned LP head and per block partial; splay visits are overhead"}
        for k, e in g.items():
            e["kernels"] = list(GROUP_KERNELS.get(k, ("%s_post_kernel" % kind,)))
        out["groups"] = g
    log("queue_roofline: %s" % json.dumps(out))
    return out


def _short_kernel(name):
    """`void gpq::(anonymous namespace)::merge_tasks_kernel<256>(...)` -> `merge_tasks_kernel<256>`."""
    head = name.replace("(anonymous namespace)::", "")
    if head.startswith("void "):
        head = head[5:]
    depth, start, end = 0, 0, len(head)
    for i, ch in enumerate(head):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif depth == 0 and ch == ":" and head[i + 1:i + 2] == ":":
            start = i + 2
        elif ch == "(" and depth == 0:
            end = i
            break
    return head[start:end].strip()


def roofline_table(stats_csv, counts_json, fetch_csv=None, write_csv=None, scale=(2.0, 1.0)):
    """Joins a rocprofv3 --kernel-trace --stats summary (kernel time) with the work counts of a
    --queue-roofline --roofline-counts run (algorithmic bytes) and, optionally, FETCH_SIZE and
    WRITE_SIZE counter runs (HBM bytes, scaled by the copy calibration), as a markdown table."""
    import csv
    with open(counts_json) as f:
        c = [json.loads(ln) for ln in f if ln.startswith("{")][-1]
    with open(stats_csv) as f:
        rows = list(csv.DictReader(f))

    def counter(path):
        per = {}
        if not path:
            return per
        with open(path) as f:
            for r in csv.DictReader(f):
                k = _short_kernel(r["Kernel_Name"])
                tot, n = per.get(k, (0.0, 0))
                per[k] = (tot + float(r["Counter_Value"]), n + 1)
        return per

    fetch, write = counter(fetch_csv), counter(write_csv)
    print("| group | kernels | launches | kernel time (ms) | mean (us) | algorithmic bytes | achieved (GB/s) | "
          "of 8 TB/s | HBM bytes (counters) |")
    print("|---|---|---|---|---|---|---|---|---|")
    for gname, e in c["groups"].items():
        ns, calls, names = 0.0, 0, set()
        for r in rows:
            k = _short_kernel(r["Name"])
            if any(k.startswith(p) for p in e["kernels"]):
                ns += float(r["TotalDurationNs"])
                calls += int(r["Calls"])
                names.add(k)
        gbps = e["alg_bytes"] / ns if ns else 0.0
        hbm = "—"
        if fetch and write:
            kb = sum(fetch[k][0] * scale[0] + write.get(k, (0.0, 0))[0] * scale[1] for k in fetch
                     if any(k.startswith(p) for p in e["kernels"]))
            hbm = "%.3g (%.2fx)" % (kb * 1024, kb * 1024 / e["alg_bytes"]) if e["alg_bytes"] else "%.3g" % (kb * 1024)
        print("| %s | %s | %d | %.1f | %.1f | %.4g | %.0f | %.1f%% | %s |"
              % (gname, ", ".join("`%s`" % n for n in sorted(names)), calls, ns / 1e6, ns / 1e3 / max(calls, 1),
                 e["alg_bytes"], gbps, 100.0 * gbps / HBM_PEAK_GBPS, hbm))
    print()
    print("Byte models: " + "; ".join("%s: %s" % (g, e["model"]) for g, e in c["groups"].items()) + ".")
    grouped = tuple(p for e in c["groups"].values() for p in e["kernels"])
    rest = sorted(((float(r["TotalDurationNs"]), int(r["Calls"]), _short_kernel(r["Name"])) for r in rows
                   if not any(_short_kernel(r["Name"]).startswith(p) for p in grouped)), reverse=True)[:8]
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    print()
    print("Other kernels of the run (largest first): " + "; ".join(
        "`%s` %d calls, %.1f ms (%.1f%%)" % (n, k, t / 1e6, 100.0 * t / total) for t, k, n in rest) + ".")


def phold_autotune_bench(gpq, wl, log, sample_end_time=5.0):
    """The PHOLD layout autotuner on the workload's first `sample_end_time` time units: every
    candidate layout's rate, the pick, and whether the benchmarked layout is the pick."""
    t0 = time.perf_counter()
    kw, drain, cands = gpq.phold_autotune(wl["num_lps"], init_per_lp=wl["init_per_lp"], seed=42,
                                          end_time=wl["end_time"], sample_end_time=sample_end_time)
    dt = time.perf_counter() - t0
    if CAL and DRAIN == "fused":
        bench_layout = (CAL_SLOTS, 32, FAR or 128, DRAIN, CAL)
    else:
        bench_layout = ((64, 32, FAR, DRAIN) if FAR else (QUEUES["laq"]["capacity_per_lp"], 32, 0, DRAIN)) + (0,)
    pick = (kw["capacity_per_lp"], kw["stage_per_lp"], kw["far_per_lp"], drain, kw["calendar_windows"])
    out = {"sample_end_time": sample_end_time, "seconds": dt, "pick": dict(kw, drain=drain),
           "benchmarked_layout_is_pick": bench_layout == pick and (kw["far_step"] == FAR_STEP or not FAR or CAL),
           "candidates": cands}
    log("phold_autotune: %s" % json.dumps(out))
    return out


SORT_TILES = {}    # sort tile of the global queues picked by the sort autotuner, by kind


def sort_autotune_bench(gpq, wl, log, sample_end_time=1.0):
    """The sort-tile autotuner of the global queues (GAQ, HQ) on the workload's first
    `sample_end_time` time units; the per-queue runs then use the picks."""
    out = {}
    for kind in ("gaq", "hq"):
        kw = dict(QUEUES[kind])
        if kind == "hq":
            kw["node_size"] = max(wl["num_lps"], 32)
        t0 = time.perf_counter()
        tile, cands = gpq.autotune_sort(kind, wl["num_lps"], init_per_lp=wl["init_per_lp"], seed=42,
                                        end_time=wl["end_time"], sample_end_time=sample_end_time, **kw)
        SORT_TILES[kind] = tile
        out[kind] = {"pick": tile, "seconds": time.perf_counter() - t0, "candidates": cands}
    log("sort_autotune: %s" % json.dumps(out))
    return out


def phold_batch_bench(gpq, log, instances=32, seed=5, plain=True):
    """Instances of uneven size (LPs 2^U[12,18], 16 events per LP, end time 20) advanced together
    in one pool vs each alone through its own queue with a host sync per step (plain=False: the
    pool only, for kernel tracing)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    inst = [(int(round(2 ** rng.uniform(12, 18))), 16, int(rng.integers(1, 2 ** 31)), 20.0) for _ in range(instances)]

    def one(kw, layout):
        with gpq.PholdBatch(inst, **kw) as b:
            b.run("graph", batch=16)              # warm up (graph capture)
            tg = b.run("graph", batch=16)
            hg, eg = b.results()
            if plain:
                tp = b.run("plain")
                hp, ep = b.results()
        if plain and (list(hg) != list(hp) or list(eg) != list(ep)):
            raise RuntimeError("batched PHOLD (%s): batched and plain results differ" % layout)
        ev = int(eg.sum())
        r = {"layout": layout, "events": ev, "events_per_s": ev / tg, "seconds": tg}
        if plain:
            r.update(plain_events_per_s=ev / tp, plain_seconds=tp, batched_vs_plain=tp / tg,
                     results_batched_vs_plain="match")
        return r, hg, eg

    # calendar: every instance in one calendar LAQ on one window sequence; the plain path runs
    # each instance alone in its own calendar LAQ (the same step kind)
    if CAL:
        main, hc, ec = one(dict(calendar_windows=CAL, capacity_per_lp=CAL_SLOTS, stage_per_lp=32, far_per_lp=32),
                           "calendar LAQ (%d windows, logs of %d records per LP), one window sequence" % (CAL, CAL_SLOTS))
    ring_kw = dict(capacity_per_lp=64, far_per_lp=FAR or 128)
    ring, hr, er = one(ring_kw, "two-tier LAQ (ring 64, far 128), post kernel + drain step, a window per instance")
    if not CAL:
        main, hc, ec = ring, hr, er
    elif list(hc) != list(hr) or list(ec) != list(er):
        raise RuntimeError("batched PHOLD: calendar and two-tier results differ")
    out = {"instances": instances, "lps": "2^U[12,18] x 16, end time 20", "total_lps": sum(i[0] for i in inst)}
    out.update(main)
    if CAL:
        out["two_tier_ring"] = ring
    log("phold_batch: %s" % json.dumps(out))
    return out


def astar_single_bench(gpq, log, cpu=True, tile=32, drive="graph"):
    """One 4096 x 4096, 4-connected, Bernoulli(0.2) seed 7 query, (0,0) -> (4095,4095): the tile
    search (one workgroup per tile x tile square) and, for comparison, 65536 open lists in a
    local heap queue; against the oracle's A* on one core."""
    w = h = 4096
    inf = float("inf")

    def one(lists, tile_side, decrease_key=False, two_tier=False):
        if decrease_key:      # one entry per cell in a LAQ, cheaper paths update it
            q = gpq.Queue("laq", num_lps=lists, capacity_per_lp=64, stage_per_lp=16, lookahead=inf, horizon=inf,
                          update_ids=w * h)
        elif two_tier:        # a 32-record ring per list backed by its far store
            q = gpq.Queue("laq", num_lps=lists, capacity_per_lp=32, stage_per_lp=16, far_per_lp=256, lookahead=inf,
                          horizon=inf)
        else:
            q = gpq.Queue("lhq", num_lps=lists, capacity_per_lp=64, stage_per_lp=16, lookahead=inf, horizon=inf)
        a = gpq.Astar(q, w, h, connectivity=4, seed=7, threshold=0.2, tile=tile_side)
        a.model_init()
        a.run(drive, 32)                                 # warm up (graph capture)
        a.model_init()
        rep = a.run(drive, 32)
        grid = a.grid() if cpu else None
        a.model_finish()
        q.queue_finish()
        return rep, grid

    rep, grid = one(64, tile)
    rep_q, _ = one(1 << 16, 0)
    rep_u, _ = one(1 << 16, 0, decrease_key=True)
    rep_t, _ = one(1 << 16, 0, two_tier=True)
    out = {"grid": "4096x4096, 4-connected, Bernoulli(0.2), seed 7", "search": "tiles %dx%d" % (tile, tile),
           "drive": drive,
           "cost": rep["cost"], "seconds": rep["seconds"], "steps": rep["steps"], "expanded": rep["expanded"],
           "closed": rep["closed"], "expanded_per_s": rep["expanded"] / rep["seconds"] if rep["seconds"] > 0 else 0.0,
           "queue_lists": {"open_lists": 1 << 16, "queue": "lhq", "cost": rep_q["cost"], "seconds": rep_q["seconds"],
                           "steps": rep_q["steps"], "expanded": rep_q["expanded"]},
           "queue_lists_decrease_key": {"open_lists": 1 << 16, "queue": "laq, decrease-key", "cost": rep_u["cost"],
                                        "seconds": rep_u["seconds"], "steps": rep_u["steps"],
                                        "expanded": rep_u["expanded"]},
           "queue_lists_two_tier": {"open_lists": 1 << 16, "queue": "laq two-tier (ring 32, far 256)",
                                    "cost": rep_t["cost"], "seconds": rep_t["seconds"], "steps": rep_t["steps"],
                                    "expanded": rep_t["expanded"]}}
    if not rep_q["cost"] == rep_u["cost"] == rep_t["cost"] == rep["cost"]:
        raise RuntimeError("A* single query: tile search %d, queue lists %d, decrease-key lists %d, two-tier %d" %
                           (rep["cost"], rep_q["cost"], rep_u["cost"], rep_t["cost"]))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle
        o = Oracle()
        t0 = time.perf_counter()
        c, e = o.astar(grid, w, h, 4, (0, 0), (w - 1, h - 1))
        dt = time.perf_counter() - t0
        if c != rep["cost"]:
            raise RuntimeError("A* single query: cost %d, oracle %d" % (rep["cost"], c))
        out.update(cpu_seconds=dt, cpu_expanded=e, cost_vs_cpu="match", speedup_vs_cpu=dt / rep["seconds"])
    log("astar_single: %s" % json.dumps(out))
    return out


def serving_bench(gpq, log, requests=64, rates=(50.0, 100.0, 200.0, 400.0)):
    """The Poisson trace at each arrival rate in `rates` (latency under load), the same trace with
    every request arriving at once (saturated throughput), and the plain one-request-at-a-time path."""
    kw = dict(seed=7, phold_fraction=0.5, hbm_budget=8 << 30)
    rep_p, plain = gpq.serve_run(requests, mode="plain", rate=0.0, **kw)
    res = [r["result"] for r in plain]
    # the plain path with every PHOLD request's queue and model created before the clock, as the
    # served pool is: the like-for-like comparison (no allocation inside either clock)
    rep_pa, plain_a = gpq.serve_run(requests, mode="plain", rate=0.0, preallocated=True, **kw)
    if [r["result"] for r in plain_a] != res:
        raise RuntimeError("serving: plain preallocated and plain results differ")
    keys = ("requests_per_second", "p50_ms", "p99_ms", "phold_p99_ms", "astar_p99_ms", "peak_inflight", "peak_bytes",
            "phold_peak_bytes")
    swept = []
    for rate in rates:
        rep_s, served = gpq.serve_run(requests, mode="graph", rate=rate, **kw)
        if [r["result"] for r in served] != res:
            raise RuntimeError("serving at %g/s: served and plain results differ" % rate)
        swept.append(dict({"rate_per_s": rate}, **{k: rep_s[k] for k in keys}))
    rep_x, sat = gpq.serve_run(requests, mode="graph", rate=0.0, **kw)
    if [r["result"] for r in sat] != res:
        raise RuntimeError("serving saturated: served and plain results differ")
    out = {"requests": requests, "phold_lps": "2^U[12,18] x 16, end time 20", "astar": "side 2^U[8,11], 8-connected",
           "served_at_rate": swept,
           "served_saturated": {k: rep_x[k] for k in keys},
           "plain": {k: rep_p[k] for k in ("requests_per_second", "p50_ms", "p99_ms")},
           "plain_preallocated": {k: rep_pa[k] for k in ("requests_per_second", "p50_ms", "p99_ms")},
           "served_vs_plain_throughput": rep_x["requests_per_second"] / rep_p["requests_per_second"],
           "served_vs_plain_preallocated_throughput":
               rep_x["requests_per_second"] / rep_pa["requests_per_second"],
           "pool_bytes": rep_x["pool_bytes"], "hbm_budget": kw["hbm_budget"],
           "results_served_vs_plain": "match"}
    log("serving: %s" % json.dumps(out))
    return out


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path)).get(workload, {})
    for name, e in d.items():
        if name == kernel or name.startswith(kernel + "<"):
            return e.get("bytes_per_launch")
    return None


def copy_calibration():
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 28                                         # 1 GiB of float32
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queue", default="laq", choices=sorted(QUEUES))
    ap.add_argument("--workload", default="phold_2p20_k32_t50", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=64, help="steps per hipGraph replay")
    ap.add_argument("--step", default="fused", choices=("fused", "drain", "one-event"),
                    help="fused: the two-tier LAQ merges the staged events and commits the window in one kernel "
                         "per step (windows one lookahead apart); drain: a post kernel, then each LP commits its "
                         "safe window; one-event: one event per LP per window (the reference's step).  The "
                         "committed trace is the same")
    ap.add_argument("--one-event", action="store_true", help="same as --step one-event")
    ap.add_argument("--far", type=int, default=128,
                    help="LAQ two-tier mode: far-store records per LP (0 = single sorted ring)")
    ap.add_argument("--far-step", type=float, default=0.0, help="LAQ two-tier: near-horizon step (0 = 4 x lookahead)")
    ap.add_argument("--calendar", type=int, default=64,
                    help="LAQ calendar mode for the merge-and-commit step: windows in the calendar (0 = off: the "
                         "two-tier sorted ring)")
    ap.add_argument("--cal-slots", type=int, default=8,
                    help="LAQ calendar mode: a window's log records per LP (capacity_per_lp)")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-astar", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-queues", action="store_true", help="skip the same workload through the other queue kinds")
    ap.add_argument("--no-autotune", action="store_true", help="skip the PHOLD layout autotuner")
    ap.add_argument("--batch-pool-only", action="store_true",
                    help="batched PHOLD: time the pool only, not the plain per-instance path (kernel tracing)")
    ap.add_argument("--per-queue-drive", choices=["graph", "plain"], default="graph",
                    help="drive of the per-queue runs (plain: eager launches, for kernel tracing)")
    ap.add_argument("--astar-drive", default="graph", choices=("graph", "plain"),
                    help="single A* query: hipGraph replays, or one launch sequence per step (eager, for tracing)")
    ap.add_argument("--serve-rates", default="50,100,200,400", help="Poisson arrival rates (requests/s) to sweep")
    ap.add_argument("--skip-phold", action="store_true", help="profiling aid: A* and serving parts only")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy five times (counter calibration)")
    ap.add_argument("--queue-roofline", choices=sorted(QUEUE_GROUPS),
                    help="only run the workload once, eagerly, through this queue kind (for a kernel trace)")
    ap.add_argument("--roofline-counts", action="store_true",
                    help="with --queue-roofline: work counters on; print each kernel group's algorithmic bytes")
    ap.add_argument("--end-time", type=float, default=0.0, help="override the workload's end time (0: keep)")
    ap.add_argument("--roofline-table", nargs="+", metavar="CSV",
                    help="KERNEL_STATS_CSV COUNTS_JSON [FETCH_CSV WRITE_CSV]: print the queue-kernel roofline "
                         "table of a --queue-roofline run (no GPU)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.roofline_table:
        roofline_table(*args.roofline_table)
        return
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; --gpus %d is not provided "
              "(the multi-GPU sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    import gpupq_amd as gpq
    global DRAIN, FAR
    step = "one-event" if args.one_event else args.step
    if step == "fused" and not (args.queue == "laq" and args.far):
        step = "drain"                   # the merge-and-commit step exists for the two-tier LAQ only
    DRAIN = False if step == "one-event" or args.queue not in ("laq", "lhq", "lsq") else \
        ("fused" if step == "fused" else True)
    FAR = args.far
    global FAR_STEP, CAL, CAL_SLOTS
    FAR_STEP = args.far_step
    CAL, CAL_SLOTS = args.calendar, args.cal_slots
    wl = dict(WORKLOADS[args.workload])
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "phold_golden.json")
    if os.path.exists(gpath):
        golden = json.load(open(gpath)).get(args.workload)
    if args.end_time > 0 and args.end_time != wl["end_time"]:
        wl["end_time"] = args.end_time
        golden = None                    # the goldens hold the full workloads only

    if args.queue_roofline:
        out = queue_roofline(gpq, args.queue_roofline, wl, log, args.roofline_counts)
        if golden is not None:
            out["trace_hash_vs_golden"] = "match" if out["trace_hash"] == golden["trace_hash"] else "MISMATCH"
        print(json.dumps(out), flush=True)
        return

    if args.skip_phold:
        line = {"metric": METRIC, "value": None, "unit": "events/s", "n_gpus": 1, "note": "PHOLD part skipped"}
        if not args.no_batch:
            line["phold_batch"] = phold_batch_bench(gpq, log, plain=not args.batch_pool_only)
        if not args.no_astar:
            line["astar_batch"] = astar_bench(gpq, log, cpu=not args.no_cpu)
            line["astar_single"] = astar_single_bench(gpq, log, cpu=not args.no_cpu, drive=args.astar_drive)
        if not args.no_serving:
            line["serving"] = serving_bench(gpq, log, rates=[float(r) for r in args.serve_rates.split(",")])
        print(json.dumps(line), flush=True)
        return
    total_events, total_s, hashes = 0, 0.0, set()
    for i in range(args.warmup + args.steps):
        q, m = make(gpq, args.queue, wl)
        q.sync()
        t0 = time.perf_counter()
        rep = gpq.des_run(m, mode="graph", batch=args.batch)
        dt = time.perf_counter() - t0
        h = m.trace_hash()
        hashes.add(h)
        m.model_finish()
        q.queue_finish()
        timed = i >= args.warmup
        log("step %d%s: %d events, %d windows, %.4fs (%.3e events/s), hash %016x" %
            (i, "" if timed else " (warmup)", rep["events"], rep["windows"], dt, rep["events"] / dt, h))
        if timed:
            total_events += rep["events"]
            total_s += dt
    value = total_events / total_s
    parity = "no-golden"
    if golden is not None:
        parity = "match" if hashes == {int(golden["trace_hash"], 16)} else "MISMATCH"

    line = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * total_s / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "queue": args.queue, "num_lps": wl["num_lps"],
                   "init_per_lp": wl["init_per_lp"], "end_time": wl["end_time"], "lookahead": 0.1,
                   "drive": "hipgraph x%d" % args.batch, "events_per_step": total_events // args.steps,
                   "window": ("merge and commit in one kernel per step, windows one lookahead apart"
                              if DRAIN == "fused" else
                              "drain each LP's safe window" if DRAIN else "one event per LP per window"),
                   "laq_far_per_lp": FAR,
                   "laq_calendar": ({"windows": CAL, "log_records_per_lp": CAL_SLOTS} if CAL and DRAIN == "fused"
                                    else None),
                   "trace_hash_vs_cpu": parity},
    }
    def roofline_line():
        dom, kernels, counters = roofline(gpq, args.queue, wl, log)
        k = kernels[dom]
        kname = dom
        if args.queue == "laq" and dom == "queue_post":
            kname = "laq_post_far_kernel" if FAR else "laq_post_kernel"
        if dom == "merge_commit_step":
            kname = "phold_calendar_kernel" if CAL else "phold_fused_kernel"
        out = {"bound": "hbm", "kernel": dom, "kernel_symbol": kname, "achieved": k["gbps"], "peak": HBM_PEAK_GBPS,
               "unit": "GB/s", "frac": k["gbps"] / HBM_PEAK_GBPS,
               "traffic": traffic_from_profiles(kname, args.workload),
               "alg_bytes_per_launch": k["alg_bytes"], "mean_launch_us": k["us"], "launches": k["launches"],
               "alg_bytes_model": ("88 B per committed event + 4 B per LP and step" if dom != "queue_post" else
                                   "sorted insert 16*(2k+2d) + 4 B per scanned LP head and per block partial"),
               "other_kernels": {n: {"achieved": v["gbps"], "mean_launch_us": v["us"]}
                                 for n, v in kernels.items() if n != dom}}
        if "overhead_bytes" in k:
            out["overhead_bytes_per_launch"] = k["overhead_bytes"]
        return out

    def section(key, fn, *a, **kw):
        # a failure of one part is reported in the line under its key, with the error text,
        # instead of ending the run before the line is printed
        try:
            line[key] = fn(*a, **kw)
        except Exception as e:   # noqa: BLE001
            line[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            log("%s failed: %s" % (key, line[key]["error"]))

    if not args.no_roofline and args.queue in ("laq", "lhq", "lsq"):
        section("roofline", roofline_line)
    if not args.no_cpu:
        section("cpu_baseline", cpu_baseline, wl, log)
    if not args.no_autotune and args.queue == "laq":
        section("phold_autotune", phold_autotune_bench, gpq, wl, log)
    if not args.no_queues:
        if not args.no_autotune:
            section("sort_autotune", sort_autotune_bench, gpq, wl, log)
        section("per_queue", per_queue_bench, gpq, wl, log, drive=args.per_queue_drive)
        if "error" not in line["per_queue"]:
            hashes_q = {v["trace_hash"] for v in line["per_queue"].values()}
            line["per_queue_hash_equal"] = len(hashes_q) == 1 and (golden is None or
                                                                    hashes_q == {golden["trace_hash"]})
    if not args.no_batch:
        section("phold_batch", phold_batch_bench, gpq, log)
    if not args.no_astar:
        section("astar_batch", astar_bench, gpq, log, cpu=not args.no_cpu)
        section("astar_single", astar_single_bench, gpq, log, cpu=not args.no_cpu, drive=args.astar_drive)
    if not args.no_serving:
        section("serving", serving_bench, gpq, log, rates=[float(r) for r in args.serve_rates.split(",")])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
