"""Batched PHOLD on the GPU: instances of uneven size share one local array queue and
advance together; each instance's trace hash and event count equal the sequential oracle's
for that instance alone, and the plain one-instance-per-queue path gives the same."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INSTANCES = [
    (300, 8, 101, 6.0),
    (1024, 16, 42, 10.0),
    (4096, 16, 7, 5.0),
    (17, 4, 9, 30.0),
    (2000, 32, 11, 3.0),
    (256, 1, 12, 0.05),      # ends before most events: few or no commits
    (513, 16, 13, 20.0),
]


@pytest.mark.parametrize("tier", ["one", "two"])
def test_batch_matches_oracle_and_plain(gpq, oracle, tier):
    kw = {} if tier == "one" else dict(capacity_per_lp=64, far_per_lp=128)
    with gpq.PholdBatch(INSTANCES, **kw) as b:
        b.run("graph", batch=8)
        hg, eg = b.results()
        b.run("plain")
        hp, ep = b.results()
    assert np.array_equal(hg, hp) and np.array_equal(eg, ep)
    for i, (L, k, seed, T) in enumerate(INSTANCES):
        res, _, _ = oracle.phold("seq", L, k, T, seed=seed)
        assert int(eg[i]) == res.events, i
        assert int(hg[i]) == res.trace_hash, i


def test_batch_equals_single_instance_path(gpq):
    L, k, seed, T = 3000, 16, 77, 8.0
    with gpq.PholdBatch([(L, k, seed, T), (700, 8, 5, 12.0)]) as b:
        b.run("graph", batch=16)
        h, e = b.results()
    q = gpq.Queue("laq", num_lps=L, capacity_per_lp=128, stage_per_lp=32, lookahead=0.1, horizon=T)
    m = gpq.Phold(q, init_per_lp=k, seed=seed, end_time=T, drain=True)
    m.model_init()
    rep = gpq.des_run(m, mode="graph", batch=16)
    assert rep["events"] == int(e[0]) and m.trace_hash() == int(h[0])
    gpq.des_finish(m)


CAL = dict(calendar_windows=64, capacity_per_lp=8, stage_per_lp=32, far_per_lp=64)


@pytest.mark.parametrize("windows", [64, 16])
def test_calendar_batch_matches_oracle_and_plain(gpq, oracle, windows):
    # all instances in one calendar on one window sequence (graph) against each alone in its own
    # calendar queue (plain) and the oracle; end times from 0.05 to 30 in one calendar
    kw = dict(CAL, calendar_windows=windows)
    with gpq.PholdBatch(INSTANCES, **kw) as b:
        b.run("graph", batch=8)
        hg, eg = b.results()
        b.run("plain")
        hp, ep = b.results()
    assert np.array_equal(hg, hp) and np.array_equal(eg, ep)
    for i, (L, k, seed, T) in enumerate(INSTANCES):
        res, _, _ = oracle.phold("seq", L, k, T, seed=seed)
        assert int(eg[i]) == res.events, i
        assert int(hg[i]) == res.trace_hash, i


def test_calendar_batch_mixed_lookaheads(gpq, oracle):
    # the shared windows are as wide as the smallest lookahead; every instance keeps its own
    inst = [(300, 8, 101, 6.0, 0.1), (1024, 16, 42, 10.0, 0.25), (700, 8, 5, 4.0, 0.5), (64, 4, 3, 12.0, 1.5)]
    with gpq.PholdBatch(inst, **CAL) as b:
        b.run("graph", batch=16)
        hg, eg = b.results()
    for i, (L, k, seed, T, la) in enumerate(inst):
        res, _, _ = oracle.phold("seq", L, k, T, seed=seed, lookahead=la)
        assert int(eg[i]) == res.events and int(hg[i]) == res.trace_hash, i


def test_calendar_batch_dense_instance(gpq, oracle):
    # 64 events per LP: about 6 per LP and window, so many LPs take the slow path, and a block's log
    # of a window holds about 1,500 records, past the ones whose keys the step keeps in LDS
    inst = [(256, 64, 21, 2.0), (300, 8, 101, 6.0), (600, 64, 22, 1.5)]
    with gpq.PholdBatch(inst, **CAL) as b:
        b.run("graph", batch=8)
        hg, eg = b.results()
        b.run("plain")
        hp, ep = b.results()
    assert np.array_equal(hg, hp) and np.array_equal(eg, ep)
    for i, (L, k, seed, T) in enumerate(inst):
        res, _, _ = oracle.phold("seq", L, k, T, seed=seed)
        assert int(eg[i]) == res.events and int(hg[i]) == res.trace_hash, i


def test_calendar_batch_runs_twice_and_rejects_bad_settings(gpq, oracle):
    with gpq.PholdBatch(INSTANCES[:3], **CAL) as b:
        b.run("graph", batch=16)
        h1, e1 = b.results()
        b.run("graph", batch=4)
        h2, e2 = b.results()
    assert np.array_equal(h1, h2) and np.array_equal(e1, e2)
    with pytest.raises(gpq.GpqError):
        gpq.PholdBatch(INSTANCES[:2], calendar_windows=48)
    with gpq.PholdBatch([(256, 32, 1, 5.0)], **dict(CAL, far_per_lp=16)) as b:
        with pytest.raises(gpq.GpqError):       # the lists cannot park an LP's 32 initial events
            b.run("graph")
