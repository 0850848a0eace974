"""The global queues' record primitives on the GPU: tile sort + merge passes, the
multi-segment merge-path merge, and the one-safe-event-per-LP selection, against numpy
sorting and the oracle's selection (pinned by the reference's known-answer tests)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def events(gpq, ts, dst=None, src=None, seq=None):
    n = len(ts)
    a = np.zeros(n, dtype=gpq.event_dtype)
    a["ts"] = np.asarray(ts, dtype=np.float32).view(np.uint32)
    a["dst"] = 0 if dst is None else dst
    a["src"] = 0 if src is None else src
    a["seq"] = np.arange(n) if seq is None else seq
    return a


def keysort(a):
    return a[np.lexsort((a["seq"], a["src"], a["ts"]))]


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dev_sort(gpq, a):
    out = np.zeros_like(a)
    gpq._lib.check(gpq.lib.gpq_selftest_sort(P(a), len(a), P(out)), "sort")
    return out


def test_sort_enqueue_buffer_fixture(gpq):
    # src/test_hq.cu:991-1008: {10, 5, 20, 1} -> 1, 5, 10, 20
    out = dev_sort(gpq, events(gpq, [10, 5, 20, 1]))
    assert [gpq.bits_f32(t) for t in out["ts"]] == [1.0, 5.0, 10.0, 20.0]


@pytest.mark.parametrize("n", [1, 7, 2047, 2048, 2049, 5000, 65536 + 17, 300001, (1 << 20) + 3])
def test_sort_matches_numpy(gpq, n):
    rng = np.random.default_rng(n)
    a = events(gpq, rng.integers(0, 1000, n) * 0.5, dst=rng.integers(0, 7, n), src=rng.integers(0, 50, n),
               seq=rng.integers(0, 1 << 31, n))
    out = dev_sort(gpq, a)
    assert np.array_equal(out, keysort(a))


@pytest.mark.parametrize("na,nb", [(0, 5), (5, 0), (1, 1), (3000, 10), (10, 3000), (70001, 50001), (1 << 18, 1 << 17),
                                   ((1 << 20) + 5, 300), (300, (1 << 20) + 5)])
def test_merge_is_stable_and_sorted(gpq, na, nb):
    rng = np.random.default_rng(na * 7 + nb)
    a = keysort(events(gpq, rng.integers(0, 500, na) * 1.0, src=rng.integers(0, 3, na), seq=rng.integers(0, 4, na)))
    b = keysort(events(gpq, rng.integers(0, 500, nb) * 1.0, src=rng.integers(0, 3, nb), seq=rng.integers(0, 4, nb)))
    a["dst"] = 0          # tag the source array so ties can be checked
    b["dst"] = 1
    out = np.zeros(na + nb, dtype=gpq.event_dtype)
    gpq._lib.check(gpq.lib.gpq_selftest_merge(P(a) if na else None, na, P(b) if nb else None, nb, P(out)), "merge")
    key = lambda e: (int(e["ts"]), int(e["src"]), int(e["seq"]))
    expect = sorted([(key(e), 0, i) for i, e in enumerate(a)] + [(key(e), 1, i) for i, e in enumerate(b)])
    assert [key(e) for e in out] == [k for k, _, _ in expect]
    assert list(out["dst"]) == [w for _, w, _ in expect]          # ties: first input first


def dev_select(gpq, s, num_lps, wend, bound=None):
    n = len(s)
    deq = np.zeros(n, dtype=gpq.event_dtype)
    rest = np.zeros(n, dtype=gpq.event_dtype)
    nd, nr, pre = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    b = None
    if bound is not None:
        b = np.zeros(1, dtype=gpq.event_dtype)
        b[0] = bound
    gpq._lib.check(gpq.lib.gpq_selftest_select(P(s), n, num_lps, wend, P(b) if b is not None else None, P(deq),
                                               ctypes.byref(nd), P(rest), ctypes.byref(nr), ctypes.byref(pre)), "select")
    return deq[:nd.value], rest[:nr.value], pre.value


def test_select_fixtures(gpq):
    # src/test_gaq.cu:108-151 and src/test_hq.cu:1011-1034: [1(lp0), 2(lp0), 3(lp1), 4(lp1)] -> 1 and 3
    s = events(gpq, [1, 2, 3, 4], dst=[0, 0, 1, 1])
    deq, rest, pre = dev_select(gpq, s, 2, 1 + 10.0)
    assert [gpq.bits_f32(t) for t in deq["ts"]] == [1.0, 3.0]
    assert [gpq.bits_f32(t) for t in rest["ts"]] == [2.0, 4.0] and pre == 4
    # src/test_hq.cu:959-989: merged [5,12,15,17] of LP 0 with root maximum 15 -> only 5
    s = events(gpq, [5, 12, 15, 17], dst=[0, 0, 0, 0], seq=[0, 1, 2, 3])
    deq, rest, pre = dev_select(gpq, s, 2, 15.0, bound=s[2])
    assert [gpq.bits_f32(t) for t in deq["ts"]] == [5.0] and pre == 2


@pytest.mark.parametrize("n,lps", [(1000, 7), (100000, 4096), (300000, 97)])
def test_select_matches_oracle(gpq, oracle, n, lps):
    rng = np.random.default_rng(n)
    ts = np.sort(rng.choice(np.arange(n * 4), n, replace=False)).astype(np.float32)   # distinct
    s = events(gpq, ts, dst=rng.integers(0, lps, n))
    ts_bits = s["ts"].astype(np.int64)
    wend = float(ts[0] + np.float32(n / 3))
    wend_bits = int(np.float32(wend).view(np.uint32))
    bound = s[n // 2]
    k, marks = oracle.safe_select(ts_bits, s["dst"].astype(np.int32), lps, wend_bits - int(ts_bits[0]),
                                  bound=int(bound["ts"]))
    deq, rest, pre = dev_select(gpq, s, lps, wend, bound=bound)
    assert len(deq) == k
    assert np.array_equal(deq, s[marks.astype(bool)])
    scanned = pre
    assert np.array_equal(rest, s[:scanned][~marks[:scanned].astype(bool)])


def test_select_large_prefix(gpq, oracle):
    # 9 M records all inside the window: the per-block counts (one per 1024 records) take more
    # than one sweep of the block scan (8 x 1024 counts per sweep)
    n, lps = 9_000_000, 1 << 20
    rng = np.random.default_rng(11)
    ts = np.arange(n, dtype=np.float32)          # distinct, exact in float32
    s = events(gpq, ts, dst=rng.integers(0, lps, n))
    ts_bits = s["ts"].astype(np.int64)
    wend = float(n + 1)
    wend_bits = int(np.float32(wend).view(np.uint32))
    k, marks = oracle.safe_select(ts_bits, s["dst"].astype(np.int32), lps, wend_bits - int(ts_bits[0]))
    deq, rest, pre = dev_select(gpq, s, lps, wend)
    assert pre == n and len(deq) == k
    m = marks.astype(bool)
    assert np.array_equal(deq, s[m])
    assert np.array_equal(rest, s[~m])
