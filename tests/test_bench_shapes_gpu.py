"""Parity at the benchmarked shapes: the 512-query batched A* workload at the autotuned launch
shape, and a served trace whose PHOLD requests span 2^12..2^18 LPs (16 events per LP, end
time 20), each result against the CPU oracle."""
import os

import numpy as np
import pytest

import gpupq_amd

pytestmark = pytest.mark.gpu

ASTAR_SEED = 2024
# the shape the batched A* autotuner picks for the 512 synthetic queries on one MI355X
TUNED = dict(threads_per_query=512, heap_capacity=2048, stage_capacity=8)


def _host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def _oracle_costs(oracle, specs, seed):
    out = []
    for q, s in enumerate(specs):
        grid = oracle.grid(s.side, s.side, seed, 1, q, s.threshold)
        c, _ = oracle.astar(grid, s.side, s.side, 8, (s.sx, s.sy), (s.tx, s.ty))
        out.append(c)
    return np.array(out, dtype=np.int64)


@pytest.fixture(scope="module")
def astar512(gpq, oracle):
    specs = [gpq.query_synthetic(ASTAR_SEED, q) for q in range(512)]
    return specs, _oracle_costs(oracle, specs, ASTAR_SEED)


def test_astar_batch_512_at_tuned_shape_matches_oracle(gpq, astar512):
    specs, ref = astar512
    with gpq.AstarBatch(specs, **TUNED) as b:
        b.run("graph")
        costs, expanded = b.results()
        start, finish = b.timing()
    assert np.array_equal(costs, ref)
    assert (expanded > 0).all()
    assert (finish >= start).all()


def test_astar_batch_512_at_autotuned_pick_matches_oracle(gpq, astar512):
    specs, ref = astar512
    tuning, best = gpq.autotune_batch(specs, sample=64)
    assert best > 0 and tuning["threads_per_query"] % 64 == 0
    with gpq.AstarBatch(specs, **tuning) as b:
        b.run("graph")
        costs, _ = b.results()
    assert np.array_equal(costs, ref)


# 16 requests of the serving distribution (half PHOLD with 2^U[12,18] LPs, 16 events per LP,
# end time 20; half A* of the batched distribution) under seed 133: five PHOLD requests from
# 4366 to 261149 LPs (2^12.1 .. 2^17.99)
SERVE = dict(seed=133, phold_fraction=0.5, phold_log2=(12, 18), phold_init_per_lp=16, phold_end_time=20.0,
             hbm_budget=8 << 30)


def test_serving_at_benchmark_sizes_served_equals_plain_equals_oracle(gpq, oracle):
    n = 16
    trace = gpq.serve_trace(n, rate=0.0, **SERVE)
    sizes = [r["size"] for r in trace if r["kind"] == 0]
    assert min(sizes) < 1 << 13 and max(sizes) > 250000
    rep_s, served = gpq.serve_run(n, mode="graph", rate=0.0, **SERVE)
    rep_p, plain = gpq.serve_run(n, mode="plain", rate=0.0, **SERVE)
    rep_a, prealloc = gpq.serve_run(n, mode="plain", rate=0.0, preallocated=True, **SERVE)
    assert rep_s["completed"] == rep_p["completed"] == rep_a["completed"] == n
    assert rep_s["pool_bytes"] <= SERVE["hbm_budget"]
    assert rep_s["phold_peak_bytes"] <= SERVE["hbm_budget"]
    threads = _host_threads()
    for i, (s, p, a) in enumerate(zip(served, plain, prealloc)):
        assert s["result"] == p["result"] == a["result"], "request %d" % i
        if s["kind"] == 0:
            assert s["work"] == p["work"] == a["work"]
            word3 = oracle.philox((i, 0, 6, 0), (SERVE["seed"], 0))[3]
            res, _, _ = oracle.phold("parallel", s["size"], 16, 20.0, seed=int(word3), threads=threads)
            assert s["result"] == res.trace_hash and s["work"] == res.events, "PHOLD request %d" % i
        else:
            q = gpq.query_synthetic(SERVE["seed"], i)
            grid = oracle.grid(q.side, q.side, SERVE["seed"], 1, i, q.threshold)
            c, _ = oracle.astar(grid, q.side, q.side, 8, (q.sx, q.sy), (q.tx, q.ty))
            assert s["result"] == (c if c >= 0 else 2 ** 64 - 1), "A* request %d" % i


def test_serving_rejects_a_request_above_the_budget(gpq):
    # the largest PHOLD request (261149 LPs at about 3.6 KB per LP) cannot fit 256 MiB
    tight = dict(SERVE, hbm_budget=256 << 20)
    for mode in ("graph", "plain"):
        with pytest.raises(gpupq_amd.GpqError) as e:
            gpq.serve_run(16, mode=mode, rate=0.0, **tight)
        assert e.value.status == 2 and "budget" in str(e.value)        # GPQ_ERR_CAPACITY
