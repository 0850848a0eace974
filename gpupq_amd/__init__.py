"""gpupq_amd -- MI355X-native data-parallel priority queues and the PHOLD / A*
models that drive them.

The names mirror the reference's host interfaces:
  queue.h  (src/queue.h:86-115): Queue.queue_init / queue_post_init / queue_pre /
           queue_post / queue_get_min_ts / queue_finish / queue_check_phold
  model.h  (src/model.h:9-19):  Phold.model_init / model_handle_next /
           model_get_events / model_finish
  des.h    (src/des.h:3-5):     des_init / des_run / des_finish
All work runs in hand-written HIP kernels in libgpupq.so (gfx950); importing
this package without the built library raises ImportError.
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import DesReport, Event, GpqError, PholdConfig, QueueConfig, check

__all__ = ["Queue", "Phold", "PholdBatch", "Astar", "AstarBatch", "autotune_batch", "query_synthetic", "serve_trace", "serve_run", "des_init",
           "des_run", "des_finish",
           "GpqError", "lib", "event_dtype"]

lib = _lib.load()

# numpy view of gpq_event
event_dtype = np.dtype([("ts", "<u4"), ("dst", "<u4"), ("src", "<u4"), ("seq", "<u4")])


def f32_bits(x):
    return int(np.float32(x).view(np.uint32))


def bits_f32(u):
    return float(np.uint32(u).view(np.float32))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Queue:
    """One queue handle.  Reference: queue_init() (src/queue.h:87) with the
    compile-time choices of src/config.h as constructor arguments."""

    def __init__(self, kind="laq", num_lps=1024, capacity_per_lp=64, stage_per_lp=16, node_size=0,
                 heap_nodes=0, total_capacity=0, lookahead=0.1, horizon=float("inf"), block_threads=256,
                 collect_stats=False, far_per_lp=0, far_step=0.0, update_ids=0, update_per_lp=0, heaps=0,
                 calendar_windows=0, sort_tile=0):
        self.kind = kind
        self.num_lps = num_lps
        cfg = QueueConfig(_lib.QUEUE_KINDS[kind], num_lps, capacity_per_lp, stage_per_lp, node_size,
                          heap_nodes, total_capacity, lookahead, horizon, block_threads, 1 if collect_stats else 0,
                          far_per_lp, far_step, update_ids, update_per_lp, heaps, calendar_windows, sort_tile)
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(lib.gpq_queue_create(ctypes.byref(cfg), ctypes.byref(h)), "queue_init")
        self.handle = h
        self._models = weakref.WeakSet()     # models on this queue: finished before it

    # -- queue.h
    def queue_post_init(self):
        check(lib.gpq_queue_post_init(self.handle), "queue_post_init")

    def queue_pre(self):
        check(lib.gpq_queue_pre(self.handle), "queue_pre")

    def queue_post(self):
        check(lib.gpq_queue_post(self.handle), "queue_post")

    def queue_get_min_ts(self):
        out = ctypes.c_float()
        check(lib.gpq_queue_get_min_ts(self.handle, ctypes.byref(out)), "queue_get_min_ts")
        return out.value

    def queue_check_phold(self):
        """Pending item count (conservation check, src/heap_queue.cu:1908-1945)."""
        out = ctypes.c_uint64()
        check(lib.gpq_queue_pending(self.handle, ctypes.byref(out)), "queue_check_phold")
        return out.value

    def error_flags(self):
        out = ctypes.c_uint32()
        check(lib.gpq_queue_error_flags(self.handle, ctypes.byref(out)), "error_flags")
        return out.value

    def stats(self):
        """Work counters of the post kernels (needs collect_stats=True for all but post_launches)."""
        st = _lib.QueueStats()
        check(lib.gpq_queue_stats_get(self.handle, ctypes.byref(st)), "stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def sync(self):
        check(lib.gpq_queue_sync(self.handle), "sync")

    def queue_finish(self):
        if self.handle:
            for m in list(self._models):         # model_finish() (src/des.cu:220) before queue_finish() (:91)
                m.model_finish()
            check(lib.gpq_queue_destroy(self.handle), "queue_finish")
            self.handle = None

    close = queue_finish

    # -- device verbs driven from the host (src/queue.cuh:46-70)
    def insert(self, events):
        ev = np.ascontiguousarray(events, dtype=event_dtype)
        check(lib.gpq_queue_insert_host(self.handle, _ptr(ev), len(ev)), "queue_insert")

    def insert_or_update(self, events):
        """queue_insert_or_update (src/local_array_queue.cu:830-846): in decrease-key mode
        (update_ids > 0) an id already in the queue gets the smaller key at the next post."""
        ev = np.ascontiguousarray(events, dtype=event_dtype)
        check(lib.gpq_queue_insert_or_update_host(self.handle, _ptr(ev), len(ev)), "queue_insert_or_update")

    def peek(self, set_done=False, max_items=None):
        max_items = max_items or max(self.num_lps, 1) * 4
        out = np.zeros(max_items, dtype=event_dtype)
        idx = np.zeros(max_items, dtype=np.int32)
        found = ctypes.c_uint32()
        check(lib.gpq_queue_peek_host(self.handle, 1 if set_done else 0, _ptr(out), _ptr(idx), max_items,
                                      ctypes.byref(found)), "queue_peek")
        n = min(found.value, max_items)
        order = np.lexsort((out["seq"][:n], out["src"][:n], out["ts"][:n]))
        return out[:n][order], idx[:n][order]

    def lp_contents(self, lp, cap=4096):
        out = np.zeros(cap, dtype=event_dtype)
        n = ctypes.c_uint32()
        check(lib.gpq_queue_lp_contents(self.handle, lp, _ptr(out), cap, ctypes.byref(n)), "lp_contents")
        return out[:min(n.value, cap)]

    def heap_node(self, node, cap=1 << 20):
        """Records of one parallel-heap node (GPQ_HQ only), in key order."""
        out = np.zeros(cap, dtype=event_dtype)
        n = ctypes.c_uint32()
        check(lib.gpq_heap_node_contents(self.handle, node, _ptr(out), cap, ctypes.byref(n)), "heap_node")
        return out[:min(n.value, cap)]

    def queue_length(self, lp):
        n = ctypes.c_uint32()
        check(lib.gpq_queue_length(self.handle, lp, ctypes.byref(n)), "queue_length")
        return n.value

    def queue_clear(self, lp):
        check(lib.gpq_queue_clear_lp(self.handle, lp), "queue_clear")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.queue_finish()

    def __del__(self):
        try:
            self.queue_finish()
        except Exception:
            pass


def phold_autotune(num_lps, init_per_lp=32, seed=42, end_time=50.0, lookahead=0.1, p_self=0.5, sample_end_time=0.0,
                   block_threads=256):
    """The PHOLD layout autotuner (gpq_phold_autotune): times a fixed list of LAQ layouts on the
    workload cut to end time sample_end_time (0 = 5) and returns (queue keyword arguments for
    Queue("laq", ...), the Phold drain argument, [every candidate with its status and rate])."""
    qc = QueueConfig(_lib.QUEUE_KINDS["laq"], num_lps, 0, 0, 0, 0, 0, lookahead, end_time, block_threads, 0, 0, 0.0)
    pc = PholdConfig(num_lps, init_per_lp, seed, lookahead, end_time, min(int(round(p_self * 2 ** 32)), 0xFFFFFFFF),
                     0, 0)
    res = (_lib.PholdTuning * 16)()
    n = ctypes.c_uint32()
    check(lib.gpq_phold_autotune(ctypes.byref(pc), ctypes.byref(qc), sample_end_time, res, 16, ctypes.byref(n)),
          "phold_autotune")
    kw = dict(capacity_per_lp=qc.capacity_per_lp, stage_per_lp=qc.stage_per_lp, far_per_lp=qc.far_per_lp,
              far_step=qc.far_step, calendar_windows=qc.calendar_windows)
    drain = {0: False, 1: True, 2: "fused"}[pc.drain]
    cands = [dict(capacity_per_lp=r.capacity_per_lp, stage_per_lp=r.stage_per_lp, far_per_lp=r.far_per_lp,
                  far_step=r.far_step, calendar_windows=r.calendar_windows, drain={0: False, 1: True, 2: "fused"}[r.drain],
                  status={0: "ran", 1: "does not fit", 2: "trace differs"}[r.status],
                  events_per_s=r.events_per_second) for r in res[:n.value]]
    return kw, drain, cands


def autotune_sort(kind, num_lps, init_per_lp=32, seed=42, end_time=50.0, lookahead=0.1, p_self=0.5,
                  sample_end_time=0.0, **queue_kwargs):
    """The sort-tile autotuner of the global queues (gpq_queue_autotune_sort) for kind "gaq" or
    "hq" and the other Queue arguments given: returns (the picked sort_tile, [every candidate
    with its status and rate])."""
    cfg_args = dict(capacity_per_lp=64, stage_per_lp=16, node_size=0, heap_nodes=0, total_capacity=0,
                    block_threads=256, heaps=0)
    cfg_args.update(queue_kwargs)
    qc = QueueConfig(_lib.QUEUE_KINDS[kind], num_lps, cfg_args["capacity_per_lp"], cfg_args["stage_per_lp"],
                     cfg_args["node_size"], cfg_args["heap_nodes"], cfg_args["total_capacity"], lookahead, end_time,
                     cfg_args["block_threads"], 0, 0, 0.0, 0, 0, cfg_args["heaps"], 0, 0)
    pc = PholdConfig(num_lps, init_per_lp, seed, lookahead, end_time, min(int(round(p_self * 2 ** 32)), 0xFFFFFFFF),
                     0, 0)
    res = (_lib.SortTuning * 8)()
    n = ctypes.c_uint32()
    check(lib.gpq_queue_autotune_sort(ctypes.byref(pc), ctypes.byref(qc), sample_end_time, res, 8, ctypes.byref(n)),
          "queue_autotune_sort")
    cands = [dict(sort_tile=r.sort_tile, status={0: "ran", 1: "could not run", 2: "trace differs"}[r.status],
                  events_per_s=r.events_per_second) for r in res[:n.value]]
    return qc.sort_tile, cands


class Phold:
    """PHOLD model on a queue (src/phold.cu, src/model.h:9-19)."""

    def __init__(self, queue, init_per_lp=16, seed=42, end_time=None, lookahead=None, p_self=0.5, checks=True,
                 drain=False):
        """drain: False (one event per LP per step), True (each LP commits its whole safe window)
        or "fused" (the same, with the LAQ merge inside the step kernel; two-tier LAQ only)."""
        self.queue = queue
        lookahead = queue.cfg.lookahead if lookahead is None else lookahead
        end_time = queue.cfg.horizon if end_time is None else end_time
        self.cfg = PholdConfig(queue.num_lps, init_per_lp, seed, lookahead, end_time,
                               min(int(round(p_self * 2 ** 32)), 0xFFFFFFFF), 1 if checks else 0,
                               2 if drain == "fused" else (1 if drain else 0))
        h = ctypes.c_void_p()
        check(lib.gpq_phold_create(queue.handle, ctypes.byref(self.cfg), ctypes.byref(h)), "phold_create")
        self.handle = h
        queue._models.add(self)

    def model_init(self):
        check(lib.gpq_phold_init(self.handle), "model_init")

    def model_handle_next(self):
        check(lib.gpq_phold_handle_next(self.handle), "model_handle_next")

    def model_get_events(self):
        out = ctypes.c_uint64()
        check(lib.gpq_phold_get_events(self.handle, ctypes.byref(out)), "model_get_events")
        return out.value

    def model_finish(self):
        if self.handle:
            check(lib.gpq_phold_finish(self.handle), "model_finish")
            check(lib.gpq_phold_destroy(self.handle), "model_finish")
            self.handle = None

    def lp_state(self):
        n = self.queue.num_lps
        cnt = np.zeros(n, dtype=np.uint32)
        hsh = np.zeros(n, dtype=np.uint64)
        check(lib.gpq_phold_lp_state(self.handle, _ptr(cnt), _ptr(hsh)), "lp_state")
        return cnt, hsh

    def trace_hash(self):
        out = ctypes.c_uint64()
        check(lib.gpq_phold_trace_hash(self.handle, ctypes.byref(out)), "trace_hash")
        return out.value

    def profile(self, batch=64):
        """Runs the step loop with HIP events around every kernel; mean durations in us."""
        p = _lib.KernelProfile()
        check(lib.gpq_des_profile(self.handle, batch, ctypes.byref(p)), "des_profile")
        return {k: getattr(p, k) for k, _ in p._fields_}

    def __del__(self):
        try:
            self.model_finish()
        except Exception:
            pass


def des_init(kind="laq", **queue_kwargs):
    """des_init() (src/des.cu:74-87): creates the configured queue."""
    return Queue(kind=kind, **queue_kwargs)


def des_run(model, mode="graph", batch=32):
    """des_run() (src/des.cu:134-221): steps until no pending event is below end_time.
    Returns the report as a dict (events/s as in src/des.cu:95-96)."""
    rep = DesReport()
    check(lib.gpq_des_run(model.handle, _lib.DRIVE_MODES[mode], batch, ctypes.byref(rep)), "des_run")
    return {k: getattr(rep, k) for k, _ in DesReport._fields_}


def des_finish(model):
    """des_finish() (src/des.cu:89-132): releases the model and its queue."""
    q = model.queue
    model.model_finish()
    q.queue_finish()


# ---------------------------------------------------------------------------------------
# A* (src/a_star.cu)
# ---------------------------------------------------------------------------------------
class Astar:
    """One grid query whose open lists are the LPs of a per-LP queue (src/a_star.cu:269-593,
    model_init :698-842).  The queue needs lookahead = horizon = +inf.  Costs are exact
    integers: 4-connected unit steps (Manhattan h) or 8-connected 10/14 (octile h)."""

    def __init__(self, queue, width, height, connectivity=4, source=(0, 0), target=None, seed=7, threshold=0.2,
                 stream_tag=0, query=0, window=0, tile=0):
        target = (width - 1, height - 1) if target is None else target
        thr = threshold if isinstance(threshold, int) else min(int(threshold * 2 ** 32), 0xFFFFFFFF)
        self.queue = queue
        self.width, self.height = width, height
        self.cfg = _lib.AstarConfig(width, height, connectivity, source[0], source[1], target[0], target[1], seed,
                                    stream_tag, query, thr, min(window, 0xFFFFFFFF), tile)
        h = ctypes.c_void_p()
        check(lib.gpq_astar_create(queue.handle, ctypes.byref(self.cfg), ctypes.byref(h)), "astar_create")
        self.handle = h
        queue._models.add(self)

    def model_init(self):
        check(lib.gpq_astar_init(self.handle), "astar model_init")

    def set_grid(self, blocked):
        b = np.ascontiguousarray(blocked, dtype=np.uint8).reshape(-1)
        assert b.size == self.width * self.height
        check(lib.gpq_astar_set_grid(self.handle, _ptr(b)), "astar set_grid")

    def grid(self):
        out = np.zeros(self.width * self.height, dtype=np.uint8)
        check(lib.gpq_astar_grid(self.handle, _ptr(out)), "astar grid")
        return out

    def model_handle_next(self):
        check(lib.gpq_astar_handle_next(self.handle), "astar model_handle_next")

    def path(self):
        """An optimal path of the finished search as (x, y) pairs, source first; [] if unreachable."""
        n = ctypes.c_uint32()
        check(lib.gpq_astar_path(self.handle, None, 0, ctypes.byref(n)), "astar path")
        cells = np.zeros(max(n.value, 1), dtype=np.uint32)
        check(lib.gpq_astar_path(self.handle, _ptr(cells), n.value, ctypes.byref(n)), "astar path")
        return [(int(c) % self.width, int(c) // self.width) for c in cells[:n.value]]

    def run(self, mode="graph", batch=16):
        rep = _lib.AstarReport()
        check(lib.gpq_astar_run(self.handle, _lib.DRIVE_MODES[mode], batch, ctypes.byref(rep)), "astar run")
        return {k: getattr(rep, k) for k, _ in _lib.AstarReport._fields_}

    def model_finish(self):
        if self.handle:
            check(lib.gpq_astar_destroy(self.handle), "astar model_finish")
            self.handle = None

    def __del__(self):
        try:
            self.model_finish()
        except Exception:
            pass


def query_synthetic(seed, q):
    """Query q of the synthetic batched workload (side = round(256 * 2^(3u)), density
    U[0.1, 0.3], 8-connected, free source and target)."""
    s = _lib.QuerySpec()
    check(lib.gpq_query_synthetic(seed, q, ctypes.byref(s)), "query_synthetic")
    return s


def autotune_batch(specs, sample=64, heap_capacity=0, max_resident=0):
    """Fastest (threads_per_query, stage_capacity) for these queries, timed on the first
    `sample`; returns (kwargs for AstarBatch, seconds of the best candidate)."""
    arr = (_lib.QuerySpec * len(specs))(*specs)
    t = _lib.BatchTuning(0, heap_capacity, 0, max_resident)
    sec = ctypes.c_double()
    check(lib.gpq_astar_batch_autotune(arr, len(specs), sample, ctypes.byref(t), ctypes.byref(sec)), "autotune")
    return dict(threads_per_query=t.threads_per_query, heap_capacity=t.heap_capacity,
                stage_capacity=t.stage_capacity, max_resident=t.max_resident), sec.value


class AstarBatch:
    """Independent grid queries, one workgroup each, pulled from a device work counter by one
    persistent launch (src/a_star.cu:149-164); mode "plain" searches them one at a time with a
    host synchronization per expansion step."""

    def __init__(self, specs, threads_per_query=0, heap_capacity=0, stage_capacity=0, max_resident=0):
        self.specs = (_lib.QuerySpec * len(specs))(*specs)
        self.n = len(specs)
        self.tuning = _lib.BatchTuning(threads_per_query, heap_capacity, stage_capacity, max_resident)
        h = ctypes.c_void_p()
        check(lib.gpq_astar_batch_create(self.specs, self.n, ctypes.byref(self.tuning), ctypes.byref(h)),
              "astar_batch_create")
        self.handle = h

    def run(self, mode="graph"):
        sec = ctypes.c_double()
        check(lib.gpq_astar_batch_run(self.handle, _lib.DRIVE_MODES[mode], ctypes.byref(sec)), "astar_batch_run")
        return sec.value

    def results(self):
        costs = np.zeros(self.n, dtype=np.int64)
        exp = np.zeros(self.n, dtype=np.uint64)
        check(lib.gpq_astar_batch_results(self.handle, _ptr(costs), _ptr(exp)), "astar_batch_results")
        return costs, exp

    def timing(self):
        """Per-query (pickup, completion) device times of the last run, microseconds."""
        st = np.zeros(self.n, dtype=np.float64)
        fi = np.zeros(self.n, dtype=np.float64)
        check(lib.gpq_astar_batch_timing(self.handle, _ptr(st), _ptr(fi)), "astar_batch_timing")
        return st, fi

    def grid(self, q):
        side = self.specs[q].side
        out = np.zeros(side * side, dtype=np.uint8)
        check(lib.gpq_astar_batch_grid(self.handle, q, _ptr(out)), "astar_batch_grid")
        return out.reshape(side, side)

    def path(self, q):
        """An optimal path of query q as (x, y) cells, source first ([] when unreachable); q is
        searched again alone (src/a_star.cu:417-438 backtracks the closed list the same way)."""
        side = self.specs[q].side
        n = ctypes.c_uint32()
        check(lib.gpq_astar_batch_path(self.handle, q, None, 0, ctypes.byref(n)), "astar_batch_path")
        out = np.zeros(max(n.value, 1), dtype=np.uint32)
        check(lib.gpq_astar_batch_path(self.handle, q, _ptr(out), len(out), ctypes.byref(n)), "astar_batch_path")
        return [(int(c) % side, int(c) // side) for c in out[:n.value]]

    def close(self):
        if self.handle:
            check(lib.gpq_astar_batch_destroy(self.handle), "astar_batch_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# batched PHOLD
# ---------------------------------------------------------------------------------------
class PholdBatch:
    """Many PHOLD instances of uneven size advanced together in one local array queue;
    instances = [(num_lps, init_per_lp, seed, end_time[, lookahead]), ...] (lookahead by default
    `lookahead`, p_self 0.5).  calendar_windows > 0: calendar-mode LAQs for both drives, the
    graph drive with all instances in one calendar on one window sequence (capacity_per_lp is
    then a window's log records per LP, far_per_lp the list capacity per LP)."""

    def __init__(self, instances, capacity_per_lp=128, stage_per_lp=32, lookahead=0.1, p_self=0.5, far_per_lp=0,
                 calendar_windows=0):
        ps = min(int(round(p_self * 2 ** 32)), 0xFFFFFFFF)
        self.n = len(instances)
        rows = [tuple(i) + (lookahead,) if len(i) == 4 else tuple(i) for i in instances]
        self.inst = (_lib.PholdInstance * self.n)(*[_lib.PholdInstance(l, k, s, t, la, ps)
                                                     for (l, k, s, t, la) in rows])
        h = ctypes.c_void_p()
        check(lib.gpq_phold_batch_create(self.inst, self.n, capacity_per_lp, stage_per_lp, far_per_lp,
                                         ctypes.byref(h)), "phold_batch_create")
        self.handle = h
        if calendar_windows:
            check(lib.gpq_phold_batch_set_calendar(self.handle, calendar_windows), "phold_batch_set_calendar")

    def run(self, mode="graph", batch=16):
        sec = ctypes.c_double()
        check(lib.gpq_phold_batch_run(self.handle, _lib.DRIVE_MODES[mode], batch, ctypes.byref(sec)),
              "phold_batch_run")
        return sec.value

    def results(self):
        h = np.zeros(self.n, dtype=np.uint64)
        e = np.zeros(self.n, dtype=np.uint64)
        check(lib.gpq_phold_batch_results(self.handle, _ptr(h), _ptr(e)), "phold_batch_results")
        return h, e

    def close(self):
        if self.handle:
            check(lib.gpq_phold_batch_destroy(self.handle), "phold_batch_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------------------
def _serve_config(requests, seed=1, rate=0.0, phold_fraction=0.5, hbm_budget=0, phold_log2=(12, 18),
                  phold_init_per_lp=16, phold_end_time=20.0, phold_batch=16, astar_lanes=4, preallocated=False):
    return _lib.ServeConfig(requests, seed, rate, phold_fraction, hbm_budget, phold_log2[0], phold_log2[1],
                            phold_init_per_lp, phold_end_time, phold_batch, astar_lanes, 1 if preallocated else 0)


def _requests_dict(arr):
    return [{k: getattr(r, k) for k, _ in _lib.Request._fields_} for r in arr]


def serve_trace(requests, **kw):
    """The synthetic request trace: kind (0 PHOLD, 1 A*), size and arrival time of each request."""
    cfg = _serve_config(requests, **kw)
    out = (_lib.Request * requests)()
    check(lib.gpq_serve_trace(ctypes.byref(cfg), out), "serve_trace")
    return _requests_dict(out)


def serve_run(requests, mode="graph", **kw):
    """Serves the trace; mode "graph" = continuous batching, "plain" = one request at a time
    (preallocated=True: the plain drive's PHOLD queues and models are created before the clock).
    Returns (report dict, list of per-request dicts)."""
    cfg = _serve_config(requests, **kw)
    rep = _lib.ServeReport()
    out = (_lib.Request * requests)()
    check(lib.gpq_serve_run(ctypes.byref(cfg), _lib.DRIVE_MODES[mode], ctypes.byref(rep), out), "serve_run")
    return {k: getattr(rep, k) for k, _ in _lib.ServeReport._fields_}, _requests_dict(out)


def write_ppm(filename, blocked, width, height, path=()):
    """Binary PPM (P6) of a grid: obstacles black, free cells white, path cells red (the
    reference's image dump, src/a_star.cu:596-690)."""
    g = np.asarray(blocked, dtype=np.uint8).reshape(height, width)
    img = np.full((height, width, 3), 255, dtype=np.uint8)
    img[g != 0] = 0
    for (x, y) in path:
        img[y, x] = (255, 0, 0)
    with open(filename, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (width, height))
        f.write(img.tobytes())
