"""The calendar mode of the local array queue (one append log per upcoming window, grouped by LP
when its window comes) under the PHOLD merge-and-commit step, against the CPU oracle: per-LP
counts and trace hashes at small sizes, the committed goldens up to 2^20 LPs, plain == graph, and
the rarer paths: a short calendar (events beyond it wait in the beyond-the-calendar lists and
move into their logs step by step), and LPs with more than eight events in one window (64
events per LP: committed by a whole wave)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BENCH = dict(capacity_per_lp=16, stage_per_lp=32, far_per_lp=128, calendar_windows=64)


def run_cal(gpq, L, k, T, mode="graph", batch=16, checks=True, seed=42, **over):
    args = dict(BENCH)
    args.update(over)
    q = gpq.Queue("laq", num_lps=L, lookahead=0.1, horizon=T, **args)
    m = gpq.Phold(q, init_per_lp=k, seed=seed, end_time=T, checks=checks, drain="fused")
    m.model_init()
    assert q.queue_check_phold() == L * k
    rep = gpq.des_run(m, mode=mode, batch=batch)
    cnt, hsh = m.lp_state()
    out = dict(rep=rep, cnt=cnt, hsh=hsh, hash=m.trace_hash(), pending=q.queue_check_phold(),
               flags=q.error_flags(), q=q, m=m)
    return out


def close(g):
    g["m"].model_finish()
    g["q"].queue_finish()


@pytest.mark.parametrize("over", [{}, dict(capacity_per_lp=4), dict(calendar_windows=8),
                                  dict(capacity_per_lp=4, calendar_windows=16)])
@pytest.mark.parametrize("L,k,T", [(1, 1, 20.0), (4, 2, 5.0), (64, 8, 20.0), (97, 3, 7.5), (4096, 16, 20.0),
                                   (64, 64, 5.0)])
def test_calendar_matches_oracle_per_lp(gpq, oracle, L, k, T, over):
    if over.get("capacity_per_lp", 16) * 11 < 2 * k:
        pytest.skip("a window's log (capacity_per_lp records per LP) is too short for this population")
    res, cnt, hsh = oracle.phold("seq", L, k, T)
    g = run_cal(gpq, L, k, T, **over)
    try:
        assert g["flags"] == 0
        assert g["rep"]["events"] == res.events
        assert np.array_equal(g["cnt"], cnt) and np.array_equal(g["hsh"], hsh)
        assert g["hash"] == res.trace_hash
        assert g["pending"] == L * k
    finally:
        close(g)


def test_calendar_plain_equals_graph_and_golden(gpq, golden):
    gd = golden["phold_1024_k16_t100"]
    a = run_cal(gpq, 1024, 16, 100.0, mode="plain")
    b = run_cal(gpq, 1024, 16, 100.0, mode="graph", batch=7)
    try:
        for g in (a, b):
            assert g["flags"] == 0 and g["pending"] == 1024 * 16
            assert "%016x" % g["hash"] == gd["trace_hash"]
            assert [int(x) for x in g["cnt"]] == gd["lp_count"]
            assert ["%016x" % int(x) for x in g["hsh"]] == gd["lp_hash"]
        # windows one lookahead apart: about 100 / 0.1 steps after the initial windows
        assert a["rep"]["windows"] == b["rep"]["windows"]
        assert abs((a["rep"]["windows"] - 16) - 100.0 / 0.1) <= 3
    finally:
        close(a)
        close(b)


def test_calendar_contents_between_steps(gpq, oracle):
    """Between host-driven steps the queue reports every pending event, in key order."""
    L, k, T = 64, 8, 3.0
    q = gpq.Queue("laq", num_lps=L, lookahead=0.1, horizon=T, **dict(BENCH, capacity_per_lp=4, calendar_windows=8))
    m = gpq.Phold(q, init_per_lp=k, seed=42, end_time=T, drain="fused")
    m.model_init()
    try:
        for _ in range(6):
            m.model_handle_next()
            q.queue_post()
            assert q.queue_check_phold() == L * k
            total = 0
            for lp in range(L):
                c = q.lp_contents(lp)
                keys = [(int(e["ts"]), int(e["src"]), int(e["seq"])) for e in c]
                assert keys == sorted(keys) and all(int(e["dst"]) == lp for e in c)
                assert q.queue_length(lp) == len(c)
                total += len(c)
            assert total == L * k
        rep = gpq.des_run(m, mode="plain")
        res, _, _ = oracle.phold("seq", L, k, T)
        assert rep["events"] == res.events and m.trace_hash() == res.trace_hash
    finally:
        m.model_finish()
        q.queue_finish()


@pytest.mark.parametrize("name", ["mid_l4096_k16_t20", "mid_l65536_k32_t10", "phold_2p20_k32_t50"])
def test_calendar_matches_goldens(gpq, golden, name):
    gd = golden[name]
    g = run_cal(gpq, gd["num_lps"], gd["init_per_lp"], gd["end_time"], checks=False, batch=64)
    try:
        assert g["flags"] == 0
        assert g["rep"]["events"] == gd["events"]
        assert "%016x" % g["hash"] == gd["trace_hash"]
        assert g["pending"] == gd["num_lps"] * gd["init_per_lp"]
    finally:
        close(g)


def test_calendar_rejects_other_uses(gpq):
    import gpupq_amd
    q = gpq.Queue("laq", num_lps=256, lookahead=0.1, horizon=5.0, **BENCH)
    try:
        with pytest.raises(gpupq_amd.GpqError):
            gpq.Phold(q, init_per_lp=4, seed=1, end_time=5.0, drain=True)      # the calendar serves drain 2 only
        with pytest.raises(gpupq_amd.GpqError):
            q.queue_clear(3)
    finally:
        q.queue_finish()
    for bad in (dict(calendar_windows=12), dict(calendar_windows=512), dict(far_per_lp=0), dict(capacity_per_lp=65)):
        with pytest.raises(gpupq_amd.GpqError):
            gpq.Queue("laq", num_lps=256, lookahead=0.1, horizon=5.0, **dict(BENCH, **bad))
