"""The device operator API (src/queue.cuh:46-70) used from outside the library: a PHOLD model
compiled on its own against include/gpupq_device.h (tests/ext/phold_ext.hip, nothing of
libgpupq linked) gets each queue's view from gpq_queue_device_view and is driven through the
host queue API in the order of des_run (src/des.cu:154-214: get_min_ts, queue_pre,
model_handle_next, queue_post).  Its per-LP counts and trace hashes must equal the oracle's,
for every queue kind."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXT_PATH = os.path.join(ROOT, "tests", "ext", "libphold_ext.so")

KINDS = {
    "laq": dict(capacity_per_lp=64, stage_per_lp=16),
    "laq2": dict(capacity_per_lp=32, stage_per_lp=16, far_per_lp=64),     # two-tier mode
    "lhq": dict(capacity_per_lp=64, stage_per_lp=16),
    "lsq": dict(capacity_per_lp=64, stage_per_lp=16),
    "gaq": dict(capacity_per_lp=64, stage_per_lp=2),
    "hq": dict(capacity_per_lp=40, stage_per_lp=1),
    "hq4": dict(capacity_per_lp=40, stage_per_lp=1, heaps=4),
}


@pytest.fixture(scope="module")
def ext():
    if not os.path.exists(EXT_PATH):
        pytest.fail("tests/ext/libphold_ext.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(EXT_PATH)
    lib.ext_view_bytes.restype = ctypes.c_uint64
    lib.ext_view_bytes.argtypes = [ctypes.c_int]
    lib.ext_phold_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                     ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
    lib.ext_phold_destroy.argtypes = [ctypes.c_void_p]
    lib.ext_phold_destroy.restype = None
    lib.ext_phold_populate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                                       ctypes.c_void_p]
    lib.ext_phold_handle_next.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.ext_phold_state.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ext_lp_verbs.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return lib


def _queue(gpq, name, L, T):
    args = dict(KINDS[name])
    kind = {"laq2": "laq", "hq4": "hq"}.get(name, name)
    if kind == "hq":
        heaps = args.get("heaps", 1)
        args["node_size"] = max(-(-L // heaps), 32)
    return gpq.Queue(kind, num_lps=L, lookahead=0.1, horizon=T, **args)


def _view(gpq, ext, q):
    code = gpq._lib.QUEUE_KINDS[q.kind]
    size = ext.ext_view_bytes(code)
    buf = ctypes.create_string_buffer(size)
    kind = ctypes.c_int(-1)
    gpq._lib.check(gpq.lib.gpq_queue_device_view(q.handle, ctypes.byref(kind), buf, size), "device_view")
    assert kind.value == code
    stream = ctypes.c_void_p()
    gpq._lib.check(gpq.lib.gpq_queue_stream(q.handle, ctypes.byref(stream)), "queue_stream")
    return code, buf, stream


def _run_external(gpq, ext, q, L, k, T, seed):
    code, view, stream = _view(gpq, ext, q)
    model = ctypes.c_void_p()
    assert ext.ext_phold_create(L, k, seed, 0.1, 0x80000000, ctypes.byref(model)) == 0
    try:
        for j in range(k):           # model_init (src/phold.cu:124-163): one event per LP, then post_init
            assert ext.ext_phold_populate(model, code, view, j, stream) == 0
            q.queue_post_init()
        assert q.queue_check_phold() == L * k
        steps = 0
        while q.queue_get_min_ts() < T:          # des_run (src/des.cu:154-214)
            q.queue_pre()
            assert ext.ext_phold_handle_next(model, code, view, stream) == 0
            q.queue_post()
            steps += 1
            assert steps < 100000
        assert q.error_flags() == 0
        assert q.queue_check_phold() == L * k         # conservation (src/phold.cu:113)
        cnt = np.zeros(L, dtype=np.uint32)
        hsh = np.zeros(L, dtype=np.uint64)
        assert ext.ext_phold_state(model, cnt.ctypes.data_as(ctypes.c_void_p), hsh.ctypes.data_as(ctypes.c_void_p)) == 0
        return cnt, hsh, steps
    finally:
        ext.ext_phold_destroy(model)


@pytest.mark.parametrize("name", sorted(KINDS))
def test_external_phold_model_matches_oracle(gpq, oracle, ext, name):
    L, k, T, seed = 1000, 4, 6.0, 23
    q = _queue(gpq, name, L, T)
    try:
        cnt, hsh, steps = _run_external(gpq, ext, q, L, k, T, seed)
    finally:
        q.queue_finish()
    res, ocnt, ohsh = oracle.phold("seq", L, k, T, seed=seed)
    assert steps > 10
    assert np.array_equal(cnt, ocnt) and np.array_equal(hsh, ohsh)
    assert oracle.combine(cnt, hsh) == res.trace_hash and int(cnt.sum()) == res.events


def test_external_model_equals_library_model(gpq, ext):
    # the same workload through the library's own PHOLD model on the same queue kind
    L, k, T, seed = 4096, 4, 5.0, 5
    q = _queue(gpq, "laq", L, T)
    try:
        cnt, hsh, _ = _run_external(gpq, ext, q, L, k, T, seed)
    finally:
        q.queue_finish()
    q2 = _queue(gpq, "laq", L, T)
    m = gpq.Phold(q2, init_per_lp=k, seed=seed, end_time=T)
    m.model_init()
    gpq.des_run(m, mode="plain")
    c2, h2 = m.lp_state()
    m.model_finish()
    q2.queue_finish()
    assert np.array_equal(cnt, c2) and np.array_equal(hsh, h2)


@pytest.mark.parametrize("name", ["laq", "laq2", "lhq", "lsq"])
def test_external_length_is_empty_clear(gpq, ext, name):
    # queue_length / queue_is_empty / queue_clear from an external kernel against the host API
    L = 8
    q = _queue(gpq, name, L, 1e9)
    try:
        ev = np.zeros(3 + 2 + 1, dtype=gpq.event_dtype)
        for i, (ts, dst) in enumerate([(1.0, 0), (2.0, 0), (3.0, 0), (4.0, 1), (5.0, 1), (6.0, 5)]):
            ev[i] = (gpq.f32_bits(ts), dst, dst, i)
        q.insert(ev)
        q.queue_post()
        code, view, stream = _view(gpq, ext, q)
        lps = np.array([0, 1, 2, 5], dtype=np.uint32)
        clear = np.array([0, 1, 0, 0], dtype=np.uint32)
        length = np.zeros(4, dtype=np.int32)
        empty = np.zeros(4, dtype=np.uint32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)    # noqa: E731
        assert ext.ext_lp_verbs(code, view, p(lps), 4, p(clear), p(length), p(empty), stream) == 0
        assert list(length) == [3, 2, 0, 1]
        assert list(empty) == [0, 0, 1, 0]
        assert q.queue_length(1) == 0 and q.queue_length(0) == 3     # LP 1 was cleared
        assert [gpq.bits_f32(e["ts"]) for e in q.lp_contents(0)] == [1.0, 2.0, 3.0]
    finally:
        q.queue_finish()


def test_device_view_checks(gpq, ext):
    q = _queue(gpq, "lhq", 64, 5.0)
    try:
        buf = ctypes.create_string_buffer(4096)
        st = gpq.lib.gpq_queue_device_view(q.handle, None, buf, 3)            # wrong size
        assert st == gpq._lib.GPQ_ERR_INVALID
    finally:
        q.queue_finish()
    # a LAQ bound to a PHOLD model in merge-and-commit mode has alternating staging buffers
    q = gpq.Queue("laq", num_lps=256, capacity_per_lp=64, stage_per_lp=32, far_per_lp=128, lookahead=0.1,
                  horizon=5.0)
    m = gpq.Phold(q, init_per_lp=4, seed=1, end_time=5.0, drain="fused")
    m.model_init()
    buf = ctypes.create_string_buffer(int(ext.ext_view_bytes(0)))
    assert gpq.lib.gpq_queue_device_view(q.handle, None, buf, len(buf)) == gpq._lib.GPQ_ERR_STATE
    m.model_finish()
    q.queue_finish()
