"""ctypes declarations of include/gpupq.h.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GPUPQ_LIB") or os.path.join(HERE, "libgpupq.so")   # GPUPQ_LIB: another build

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
c_void_p = ctypes.c_void_p

GPQ_OK, GPQ_ERR_INVALID, GPQ_ERR_CAPACITY, GPQ_ERR_HIP, GPQ_ERR_NOMEM, GPQ_ERR_STATE, GPQ_ERR_UNSUPPORTED = range(7)
QUEUE_KINDS = {"laq": 0, "lsq": 1, "lhq": 2, "hq": 3, "gaq": 4}
DRIVE_MODES = {"plain": 0, "graph": 1}


class Event(ctypes.Structure):
    _fields_ = [("ts", c_u32), ("dst", c_u32), ("src", c_u32), ("seq", c_u32)]


class QueueConfig(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("num_lps", c_u32), ("capacity_per_lp", c_u32),
                ("stage_per_lp", c_u32), ("node_size", c_u32), ("heap_nodes", c_u32),
                ("total_capacity", c_u64), ("lookahead", c_f32), ("horizon", c_f32),
                ("block_threads", c_u32), ("collect_stats", c_u32), ("far_per_lp", c_u32), ("far_step", c_f32),
                ("update_ids", c_u32), ("update_per_lp", c_u32), ("heaps", c_u32), ("calendar_windows", c_u32),
                ("sort_tile", c_u32)]


class PholdConfig(ctypes.Structure):
    _fields_ = [("num_lps", c_u32), ("init_per_lp", c_u32), ("seed", c_u32), ("lookahead", c_f32),
                ("end_time", c_f32), ("p_self_u32", c_u32), ("checks", c_u32), ("drain", c_u32)]


class SortTuning(ctypes.Structure):
    _fields_ = [("sort_tile", c_u32), ("status", c_u32), ("events_per_second", ctypes.c_double)]


class PholdTuning(ctypes.Structure):
    _fields_ = [("capacity_per_lp", c_u32), ("stage_per_lp", c_u32), ("far_per_lp", c_u32), ("far_step", c_f32),
                ("drain", c_u32), ("status", c_u32), ("events_per_second", ctypes.c_double),
                ("calendar_windows", c_u32)]


class DesReport(ctypes.Structure):
    _fields_ = [("steps", c_u64), ("windows", c_u64), ("events", c_u64), ("sim_seconds", ctypes.c_double),
                ("events_per_second", ctypes.c_double), ("final_min_ts", c_f32), ("error_flags", c_u32)]


class QueueStats(ctypes.Structure):
    _fields_ = [("post_launches", c_u64), ("inserted", c_u64), ("displaced", c_u64), ("merged_lps", c_u64),
                ("scanned_lps", c_u64), ("aux_bytes", c_u64), ("sorted", c_u64), ("merged", c_u64),
                ("selected", c_u64)]


class KernelProfile(ctypes.Structure):
    _fields_ = [("steps", c_u64), ("events", c_u64), ("step_kernel_us", ctypes.c_double),
                ("post_kernel_us", ctypes.c_double), ("window_kernel_us", ctypes.c_double),
                ("seconds", ctypes.c_double)]


class AstarConfig(ctypes.Structure):
    _fields_ = [("width", c_u32), ("height", c_u32), ("connectivity", c_u32), ("sx", c_u32), ("sy", c_u32),
                ("tx", c_u32), ("ty", c_u32), ("seed", c_u32), ("stream_tag", c_u32), ("query", c_u32),
                ("threshold", c_u32), ("window", c_u32), ("tile", c_u32)]


class AstarReport(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_int64), ("expanded", c_u64), ("closed", c_u64), ("steps", c_u64),
                ("seconds", ctypes.c_double), ("error_flags", c_u32)]


class QuerySpec(ctypes.Structure):
    _fields_ = [("side", c_u32), ("threshold", c_u32), ("sx", c_u32), ("sy", c_u32), ("tx", c_u32), ("ty", c_u32),
                ("connectivity", c_u32), ("seed", c_u32), ("stream_tag", c_u32), ("query", c_u32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("threads_per_query", c_u32), ("heap_capacity", c_u32), ("stage_capacity", c_u32),
                ("max_resident", c_u32)]


class PholdInstance(ctypes.Structure):
    _fields_ = [("num_lps", c_u32), ("init_per_lp", c_u32), ("seed", c_u32), ("end_time", c_f32),
                ("lookahead", c_f32), ("p_self_u32", c_u32)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("requests", c_u32), ("seed", c_u32), ("rate", ctypes.c_double), ("phold_fraction", ctypes.c_double),
                ("hbm_budget", c_u64), ("phold_min_log2", c_u32), ("phold_max_log2", c_u32),
                ("phold_init_per_lp", c_u32), ("phold_end_time", c_f32), ("phold_batch", c_u32),
                ("astar_lanes", c_u32), ("plain_preallocated", c_u32)]


class Request(ctypes.Structure):
    _fields_ = [("kind", c_u32), ("size", c_u32), ("arrival_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("finish_ms", ctypes.c_double), ("result", c_u64), ("work", c_u64), ("bytes", c_u64)]


class ServeReport(ctypes.Structure):
    _fields_ = [("completed", c_u32), ("peak_inflight", c_u32), ("seconds", ctypes.c_double),
                ("requests_per_second", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("phold_p99_ms", ctypes.c_double), ("astar_p99_ms", ctypes.c_double),
                ("peak_bytes", c_u64), ("phold_peak_bytes", c_u64), ("pool_bytes", c_u64)]


# (name, restype, argtypes) for every entry point of include/gpupq.h
SIGNATURES = [
    ("gpq_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("gpq_last_error", ctypes.c_char_p, []),
    ("gpq_version", ctypes.c_char_p, []),
    ("gpq_queue_create", ctypes.c_int, [ctypes.POINTER(QueueConfig), ctypes.POINTER(c_void_p)]),
    ("gpq_queue_destroy", ctypes.c_int, [c_void_p]),
    ("gpq_queue_post_init", ctypes.c_int, [c_void_p]),
    ("gpq_queue_pre", ctypes.c_int, [c_void_p]),
    ("gpq_queue_post", ctypes.c_int, [c_void_p]),
    ("gpq_queue_get_min_ts", ctypes.c_int, [c_void_p, ctypes.POINTER(c_f32)]),
    ("gpq_queue_pending", ctypes.c_int, [c_void_p, ctypes.POINTER(c_u64)]),
    ("gpq_queue_error_flags", ctypes.c_int, [c_void_p, ctypes.POINTER(c_u32)]),
    ("gpq_queue_sync", ctypes.c_int, [c_void_p]),
    ("gpq_queue_insert_host", ctypes.c_int, [c_void_p, c_void_p, c_u32]),
    ("gpq_queue_insert_or_update_host", ctypes.c_int, [c_void_p, c_void_p, c_u32]),
    ("gpq_queue_peek_host", ctypes.c_int, [c_void_p, ctypes.c_int, c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_queue_lp_contents", ctypes.c_int, [c_void_p, c_u32, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_heap_node_contents", ctypes.c_int, [c_void_p, c_u32, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_queue_length", ctypes.c_int, [c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_queue_clear_lp", ctypes.c_int, [c_void_p, c_u32]),
    ("gpq_phold_create", ctypes.c_int, [c_void_p, ctypes.POINTER(PholdConfig), ctypes.POINTER(c_void_p)]),
    ("gpq_phold_autotune", ctypes.c_int, [ctypes.POINTER(PholdConfig), ctypes.POINTER(QueueConfig), c_f32,
                                          ctypes.POINTER(PholdTuning), c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_phold_destroy", ctypes.c_int, [c_void_p]),
    ("gpq_phold_init", ctypes.c_int, [c_void_p]),
    ("gpq_phold_handle_next", ctypes.c_int, [c_void_p]),
    ("gpq_phold_get_events", ctypes.c_int, [c_void_p, ctypes.POINTER(c_u64)]),
    ("gpq_phold_finish", ctypes.c_int, [c_void_p]),
    ("gpq_phold_lp_state", ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    ("gpq_phold_trace_hash", ctypes.c_int, [c_void_p, ctypes.POINTER(c_u64)]),
    ("gpq_des_run", ctypes.c_int, [c_void_p, ctypes.c_int, c_u32, ctypes.POINTER(DesReport)]),
    ("gpq_queue_stats_get", ctypes.c_int, [c_void_p, ctypes.POINTER(QueueStats)]),
    ("gpq_des_profile", ctypes.c_int, [c_void_p, c_u32, ctypes.POINTER(KernelProfile)]),
    ("gpq_astar_create", ctypes.c_int, [c_void_p, ctypes.POINTER(AstarConfig), ctypes.POINTER(c_void_p)]),
    ("gpq_astar_destroy", ctypes.c_int, [c_void_p]),
    ("gpq_astar_init", ctypes.c_int, [c_void_p]),
    ("gpq_astar_set_grid", ctypes.c_int, [c_void_p, c_void_p]),
    ("gpq_astar_handle_next", ctypes.c_int, [c_void_p]),
    ("gpq_astar_run", ctypes.c_int, [c_void_p, ctypes.c_int, c_u32, ctypes.POINTER(AstarReport)]),
    ("gpq_astar_grid", ctypes.c_int, [c_void_p, c_void_p]),
    ("gpq_astar_path", ctypes.c_int, [c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_query_synthetic", ctypes.c_int, [c_u32, c_u32, ctypes.POINTER(QuerySpec)]),
    ("gpq_astar_batch_create", ctypes.c_int, [c_void_p, c_u32, ctypes.POINTER(BatchTuning), ctypes.POINTER(c_void_p)]),
    ("gpq_astar_batch_destroy", ctypes.c_int, [c_void_p]),
    ("gpq_astar_batch_autotune", ctypes.c_int, [c_void_p, c_u32, c_u32, ctypes.POINTER(BatchTuning),
                                                ctypes.POINTER(ctypes.c_double)]),
    ("gpq_astar_batch_run", ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    ("gpq_astar_batch_results", ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    ("gpq_astar_batch_timing", ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    ("gpq_astar_batch_path", ctypes.c_int, [c_void_p, c_u32, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_astar_batch_grid", ctypes.c_int, [c_void_p, c_u32, c_void_p]),
    ("gpq_phold_batch_create", ctypes.c_int, [c_void_p, c_u32, c_u32, c_u32, c_u32, ctypes.POINTER(c_void_p)]),
    ("gpq_phold_batch_set_calendar", ctypes.c_int, [c_void_p, c_u32]),
    ("gpq_phold_batch_destroy", ctypes.c_int, [c_void_p]),
    ("gpq_phold_batch_run", ctypes.c_int, [c_void_p, ctypes.c_int, c_u32, ctypes.POINTER(ctypes.c_double)]),
    ("gpq_phold_batch_results", ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    ("gpq_serve_trace", ctypes.c_int, [ctypes.POINTER(ServeConfig), c_void_p]),
    ("gpq_serve_run", ctypes.c_int, [ctypes.POINTER(ServeConfig), ctypes.c_int, ctypes.POINTER(ServeReport),
                                     c_void_p]),
    ("gpq_selftest_exp", ctypes.c_int, [c_u32, c_u32, c_void_p]),
    ("gpq_selftest_sort", ctypes.c_int, [c_void_p, c_u32, c_void_p]),
    ("gpq_selftest_merge", ctypes.c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p]),
    ("gpq_selftest_select", ctypes.c_int, [c_void_p, c_u32, c_u32, c_f32, c_void_p, c_void_p, ctypes.POINTER(c_u32),
                                           c_void_p, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    ("gpq_selftest_philox", ctypes.c_int, [c_u32, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32, c_void_p]),
    ("gpq_queue_autotune_sort", ctypes.c_int, [c_void_p, c_void_p, c_f32, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    ("gpq_queue_device_view", ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int), c_void_p, c_u64]),
    ("gpq_queue_stream", ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    ("gpq_sort_events_noalloc", ctypes.c_int, [c_void_p, c_u32, c_void_p, c_void_p]),
]

_lib = None


def load(path=LIB_PATH):
    """Loads libgpupq.so and declares every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("gpupq_amd: %s is not built (run `python gpupq_amd/build.py`); "
                          "there is no fallback implementation" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GpqError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__("%s: %s (%s)" % (where, message, _status_name(status)))
        self.status = status


def _status_name(s):
    return {0: "ok", 1: "invalid", 2: "capacity", 3: "hip", 4: "nomem", 5: "state", 6: "unsupported"}.get(s, str(s))


def check(status, where):
    if status != GPQ_OK:
        msg = _lib.gpq_last_error().decode() if _lib is not None else ""
        raise GpqError(status, where, msg)
