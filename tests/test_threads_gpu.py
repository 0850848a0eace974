"""The threading contract of the C ABI: a handle is used by one host thread at a time, and
different handles may be driven from different host threads at once.  Two PHOLD simulations
(each with its own queue, model and stream) run concurrently from two Python threads; ctypes
releases the GIL inside every library call, so the calls overlap.  Each result must equal the
oracle's, and equal the same simulation run alone."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (queue kind, queue args, LPs, events per LP, end time, seed, drain)
    ("laq", dict(capacity_per_lp=64, stage_per_lp=32, far_per_lp=128), 4096, 16, 10.0, 42, "fused"),
    ("lhq", dict(capacity_per_lp=128, stage_per_lp=32), 2048, 16, 10.0, 7, True),
    ("hq", dict(capacity_per_lp=40, stage_per_lp=1, node_size=1024), 1024, 8, 8.0, 3, False),
    ("gaq", dict(capacity_per_lp=64, stage_per_lp=2), 3000, 8, 8.0, 11, False),
]


def _run(gpq, case, mode, out, key, barrier=None):
    kind, args, L, k, T, seed, drain = case
    try:
        q = gpq.Queue(kind, num_lps=L, lookahead=0.1, horizon=T, **args)
        m = gpq.Phold(q, init_per_lp=k, seed=seed, end_time=T, drain=drain)
        m.model_init()
        if barrier is not None:
            barrier.wait()
        rep = gpq.des_run(m, mode=mode, batch=8)
        cnt, hsh = m.lp_state()
        out[key] = (rep["events"], m.trace_hash(), cnt, hsh)
        m.model_finish()
        q.queue_finish()
    except Exception as e:     # noqa: BLE001 -- reported by the main thread
        out[key] = e


@pytest.mark.parametrize("mode", ["graph", "plain"])
def test_two_handles_from_two_threads(gpq, oracle, mode):
    for a, b in [(0, 1), (2, 3), (0, 3)]:
        out = {}
        barrier = threading.Barrier(2)
        ts = [threading.Thread(target=_run, args=(gpq, CASES[i], mode, out, i, barrier)) for i in (a, b)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=300)
        for i in (a, b):
            r = out.get(i)
            assert not isinstance(r, Exception) and r is not None, "case %d: %r" % (i, r)
            kind, _, L, k, T, seed, _ = CASES[i]
            res, ocnt, ohsh = oracle.phold("seq", L, k, T, seed=seed)
            assert r[0] == res.events and r[1] == res.trace_hash, "case %d (%s)" % (i, kind)
            assert np.array_equal(r[2], ocnt) and np.array_equal(r[3], ohsh)


def test_many_threads_each_with_its_own_handle(gpq, oracle):
    # eight host threads, each creating, running and destroying its own LAQ simulation
    out = {}
    case = CASES[0]
    ts = [threading.Thread(target=_run, args=(gpq, case[:5] + (100 + j, case[6]), "graph", out, j))
          for j in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    for j in range(8):
        r = out.get(j)
        assert not isinstance(r, Exception) and r is not None, "thread %d: %r" % (j, r)
        res, _, _ = oracle.phold("seq", case[2], case[3], case[4], seed=100 + j)
        assert r[0] == res.events and r[1] == res.trace_hash, "thread %d" % j
