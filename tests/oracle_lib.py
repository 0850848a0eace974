"""ctypes binding of the CPU oracle (oracle/liboracle.so).  Test infrastructure:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)


class PholdCfg(ctypes.Structure):
    _fields_ = [("num_lps", ctypes.c_uint32), ("init_per_lp", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("end_time", ctypes.c_float),
                ("lookahead", ctypes.c_float), ("p_self_u32", ctypes.c_uint32)]


class PholdRes(ctypes.Structure):
    _fields_ = [("events", ctypes.c_uint64), ("trace_hash", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("seconds", ctypes.c_double),
                ("max_queue_len", ctypes.c_uint32), ("in_flight_end", ctypes.c_uint32),
                ("monotone_ok", ctypes.c_uint32)]


class QuerySpec(ctypes.Structure):
    _fields_ = [("side", ctypes.c_int32), ("threshold", ctypes.c_uint32),
                ("sx", ctypes.c_int32), ("sy", ctypes.c_int32),
                ("tx", ctypes.c_int32), ("ty", ctypes.c_int32)]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError("oracle not built: run `make -C oracle`")
        self.lib = lib = ctypes.CDLL(path)
        lib.orc_exp_from_u32.restype = ctypes.c_float
        lib.orc_exp_from_u32.argtypes = [ctypes.c_uint32]
        lib.orc_phold_combine.restype = ctypes.c_uint64
        lib.orc_phold_combine.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        lib.orc_trace_step.restype = ctypes.c_uint64
        lib.orc_trace_step.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        lib.orc_ring_search.restype = ctypes.c_int
        lib.orc_ring_search.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64]
        lib.orc_ring_insert_sorted.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64]
        lib.orc_ring_quicksort.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_safe_select_i64.restype = ctypes.c_int
        lib.orc_safe_select_i64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int64, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
        for fn in ("orc_astar_grid",):
            getattr(lib, fn).restype = ctypes.c_int64
            getattr(lib, fn).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        lib.orc_dijkstra_grid.restype = ctypes.c_int64
        lib.orc_dijkstra_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_bfs_grid4.restype = ctypes.c_int64
        lib.orc_bfs_grid4.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_grid_bernoulli.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        lib.orc_batch_query_spec.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(QuerySpec), ctypes.c_void_p]
        lib.orc_batch_query_shape.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_uint32)]

    # -- random numbers
    def philox(self, ctr, key):
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        self.lib.orc_philox4x32_10(c, k, o)
        return list(o)

    def exp_table(self, first=0, count=1 << 24):
        out = np.empty(count, dtype=np.float32)
        self.lib.orc_exp_table(ctypes.c_uint32(first), ctypes.c_uint32(count), _ptr(out))
        return out

    # -- PHOLD
    def phold(self, variant, num_lps, init_per_lp, end_time, seed=42, lookahead=0.1, p_self_u32=0x80000000,
              threads=1):
        """variant: "seq", "brute", "windowed" or "parallel" (windowed drain on `threads` host threads)."""
        cfg = PholdCfg(num_lps, init_per_lp, seed, np.float32(end_time), np.float32(lookahead), p_self_u32)
        res = PholdRes()
        cnt = np.zeros(num_lps, dtype=np.uint32)
        hsh = np.zeros(num_lps, dtype=np.uint64)
        if variant == "parallel":
            rc = self.lib.orc_phold_parallel(ctypes.byref(cfg), ctypes.c_uint32(threads), _ptr(cnt), _ptr(hsh),
                                             ctypes.byref(res))
            if rc != 0:
                raise RuntimeError("oracle phold failed: %d" % rc)
            return res, cnt, hsh
        fn = {"seq": self.lib.orc_phold_seq, "brute": self.lib.orc_phold_brute,
              "windowed": self.lib.orc_phold_windowed}[variant]
        rc = fn(ctypes.byref(cfg), _ptr(cnt), _ptr(hsh), ctypes.byref(res))
        if rc != 0:
            raise RuntimeError("oracle phold failed: %d" % rc)
        return res, cnt, hsh

    def combine(self, cnt, hsh):
        cnt = np.ascontiguousarray(cnt, dtype=np.uint32)
        hsh = np.ascontiguousarray(hsh, dtype=np.uint64)
        return int(self.lib.orc_phold_combine(_ptr(cnt), _ptr(hsh), len(cnt)))

    def trace_step(self, h, ts_bits, src, seq):
        return int(self.lib.orc_trace_step(h, ts_bits, src, seq))

    # -- LAQ ring primitives (int64 keys)
    def ring_search(self, ring, read, write, key):
        r = np.ascontiguousarray(ring, dtype=np.int64)
        return self.lib.orc_ring_search(_ptr(r), len(r), read, write, key)

    def ring_insert_sorted(self, ring, read, write, key):
        r = np.array(ring, dtype=np.int64)
        self.lib.orc_ring_insert_sorted(_ptr(r), len(r), read, write, key)
        return r

    def ring_quicksort(self, ring, left, right):
        r = np.array(ring, dtype=np.int64)
        self.lib.orc_ring_quicksort(_ptr(r), len(r), left, right)
        return r

    def safe_select(self, ts, lp, num_lps, lookahead, bound=None):
        ts = np.ascontiguousarray(ts, dtype=np.int64)
        lp = np.ascontiguousarray(lp, dtype=np.int32)
        marks = np.zeros(len(ts), dtype=np.uint8)
        n = self.lib.orc_safe_select_i64(_ptr(ts), _ptr(lp), len(ts), num_lps, lookahead,
                                         0 if bound is None else 1, 0 if bound is None else bound, _ptr(marks))
        return n, marks

    # -- grids
    def astar(self, blocked, w, h, conn, s, t):
        b = np.ascontiguousarray(blocked, dtype=np.uint8)
        exp = ctypes.c_uint64()
        c = self.lib.orc_astar_grid(_ptr(b), w, h, conn, s[0], s[1], t[0], t[1], ctypes.byref(exp))
        return int(c), int(exp.value)

    def dijkstra(self, blocked, w, h, conn, s, t):
        b = np.ascontiguousarray(blocked, dtype=np.uint8)
        return int(self.lib.orc_dijkstra_grid(_ptr(b), w, h, conn, s[0], s[1], t[0], t[1]))

    def bfs4(self, blocked, w, h, s, t):
        b = np.ascontiguousarray(blocked, dtype=np.uint8)
        return int(self.lib.orc_bfs_grid4(_ptr(b), w, h, s[0], s[1], t[0], t[1]))

    def grid(self, w, h, seed, stream_tag, query, threshold):
        b = np.zeros(w * h, dtype=np.uint8)
        self.lib.orc_grid_bernoulli(_ptr(b), w, h, seed, stream_tag, query, threshold)
        return b

    def query_shape(self, seed, q):
        side = ctypes.c_int32()
        thr = ctypes.c_uint32()
        self.lib.orc_batch_query_shape(seed, q, ctypes.byref(side), ctypes.byref(thr))
        return side.value, thr.value

    def query_spec(self, seed, q, blocked):
        b = np.ascontiguousarray(blocked, dtype=np.uint8)
        spec = QuerySpec()
        self.lib.orc_batch_query_spec(seed, q, ctypes.byref(spec), _ptr(b))
        return spec
