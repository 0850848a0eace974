"""Serving on the GPU: the served (continuous batching) and plain drives give every request
the same result, and those results equal the oracle's (PHOLD trace hash and per-LP counters
of the sequential oracle; A* optimal cost)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(seed=3, phold_log2=(8, 11), phold_end_time=4.0, phold_batch=8, astar_lanes=3)


def test_trace_is_deterministic_and_mixed(gpq):
    a = gpq.serve_trace(40, rate=500.0, **SMALL)
    b = gpq.serve_trace(40, rate=500.0, **SMALL)
    assert a == b
    kinds = [r["kind"] for r in a]
    assert 0 < sum(kinds) < len(kinds)
    arr = [r["arrival_ms"] for r in a]
    assert arr == sorted(arr) and arr[-1] > 0
    assert all(256 <= r["size"] <= 2048 for r in a)


def test_served_equals_plain_equals_oracle(gpq, oracle):
    n = 12
    rep_s, served = gpq.serve_run(n, mode="graph", rate=0.0, **SMALL)
    rep_p, plain = gpq.serve_run(n, mode="plain", rate=0.0, **SMALL)
    assert rep_s["completed"] == rep_p["completed"] == n
    for s, p in zip(served, plain):
        assert (s["kind"], s["size"], s["result"]) == (p["kind"], p["size"], p["result"])
        if s["kind"] == 0:
            assert s["work"] == p["work"]          # committed events (A* expansions may differ)
        assert s["finish_ms"] >= s["arrival_ms"]
    # PHOLD requests against the sequential oracle; A* against the oracle's A*
    for i, r in enumerate(served):
        if r["kind"] == 0:
            word3 = oracle.philox((i, 0, 6, 0), (SMALL["seed"], 0))[3]
            res, _, _ = oracle.phold("seq", r["size"], 16, SMALL["phold_end_time"], seed=int(word3))
            assert r["result"] == res.trace_hash and r["work"] == res.events
        else:
            s = gpq.query_synthetic(SMALL["seed"], i)
            grid = oracle.grid(s.side, s.side, SMALL["seed"], 1, i, s.threshold)
            c, _ = oracle.astar(grid, s.side, s.side, 8, (s.sx, s.sy), (s.tx, s.ty))
            assert r["result"] == (c if c >= 0 else 2 ** 64 - 1)


def test_served_under_arrivals_and_tight_budget(gpq):
    n = 16
    rep, reqs = gpq.serve_run(n, mode="graph", rate=400.0, hbm_budget=64 << 20, **SMALL)
    assert rep["completed"] == n
    assert rep["p99_ms"] >= rep["p50_ms"] > 0
    # the shared PHOLD pool and the PHOLD requests admitted at once stay within the budget
    assert 0 < rep["pool_bytes"] <= (64 << 20)
    assert 0 < rep["phold_peak_bytes"] <= (64 << 20)
    _, alone = gpq.serve_run(n, mode="plain", rate=0.0, **SMALL)
    assert [r["result"] for r in reqs] == [r["result"] for r in alone]
