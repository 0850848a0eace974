"""Pins the CPU oracle (oracle/liboracle.so) before anything is compared with it.

Each check is chosen so that a plausible mistake in the oracle fails it: published
known-answer vectors and an independent implementation for Philox, exhaustive
comparison with glibc for the Exp sampler, three structurally different PHOLD
simulations that must agree bit-for-bit, the reference's own known-answer tests
for the ring / safe-selection primitives, and textbook shortest-path routines and
closed forms for A*."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Random123 known-answer vectors for philox4x32-10 (ctr, key -> out).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_known_answers(oracle):
    for ctr, key, out in PHILOX_KAT:
        assert oracle.philox(ctr, key) == out


def test_philox_matches_rocrand_engine(oracle):
    exe = os.path.join(ROOT, "oracle", "philox_rocrand_xcheck")
    if not os.path.exists(exe):
        pytest.skip("rocRAND host cross-check not built")
    rng = np.random.default_rng(5)
    cases = []
    for _ in range(200):
        # the engine reaches counter {c0, c1, c2, c3} via offset 4*(c1:c0), which must fit in 64 bits
        c = [int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2 ** 29))] + \
            [int(x) for x in rng.integers(0, 2 ** 32, 2)]
        k = [int(x) for x in rng.integers(0, 2 ** 32, 2)]
        cases.append((c, k))
    inp = "".join("%x %x %x %x %x %x\n" % (*c, *k) for c, k in cases)
    res = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
    lines = res.stdout.strip().splitlines()
    assert len(lines) == len(cases)
    for (c, k), line in zip(cases, lines):
        assert [int(x, 16) for x in line.split()] == oracle.philox(c, k)


def test_exp_sampler_exhaustive_vs_glibc(oracle):
    out = oracle.exp_table(0, 1 << 24)
    n = np.arange(1, (1 << 24) + 1, dtype=np.float64)
    ref = (-np.log(n * 2.0 ** -24)).astype(np.float32)
    assert np.all(out >= 0)
    ulp = np.abs(out.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    ulp[(out == 0) & (ref == 0)] = 0          # +0 vs -0 at u = 1
    assert ulp.max() <= 1
    assert (ulp > 0).sum() < 64                # correctly rounded except near midpoints
    assert abs(out.astype(np.float64).mean() - 1.0) < 1e-3


@pytest.mark.parametrize("L,k,T", [(1, 1, 20.0), (4, 2, 5.0), (16, 4, 10.0), (64, 8, 20.0), (97, 3, 7.5)])
def test_phold_three_simulations_agree(oracle, L, k, T):
    rs = {v: oracle.phold(v, L, k, T) for v in ("seq", "brute", "windowed")}
    base_res, base_cnt, base_hash = rs["seq"]
    assert base_res.events > 0
    for v, (res, cnt, hsh) in rs.items():
        assert res.events == base_res.events, v
        assert res.trace_hash == base_res.trace_hash, v
        assert np.array_equal(cnt, base_cnt), v
        assert np.array_equal(hsh, base_hash), v
        assert res.in_flight_end == L * k, v          # conservation
        assert res.monotone_ok == 1, v
    assert oracle.combine(base_cnt, base_hash) == base_res.trace_hash


def test_phold_windowed_matches_at_scale(oracle):
    a = oracle.phold("seq", 4096, 16, 20.0)
    b = oracle.phold("windowed", 4096, 16, 20.0)
    assert a[0].trace_hash == b[0].trace_hash and a[0].events == b[0].events
    assert b[0].steps > 100


@pytest.mark.parametrize("threads", [1, 2, 3, 8])
def test_phold_parallel_cpu_matches_sequential(oracle, threads):
    # the all-cores CPU baseline (windowed drain, LPs split over host threads) commits the
    # sequential trace, including with more threads than a small LP count divides evenly
    for L, k, T in [(5, 2, 6.0), (1000, 8, 10.0)]:
        a = oracle.phold("seq", L, k, T)
        b = oracle.phold("parallel", L, k, T, threads=threads)
        assert b[0].events == a[0].events and b[0].trace_hash == a[0].trace_hash
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        assert b[0].in_flight_end == L * k and b[0].monotone_ok == 1


def test_phold_rate_matches_model(oracle):
    # steady state: L*k events in flight, each lives lookahead + Exp(1) = 1.1 on average
    L, k, T = 1024, 16, 100.0
    res, cnt, _ = oracle.phold("seq", L, k, T)
    expect = L * k / 1.1 * T
    assert abs(res.events - expect) / expect < 0.02
    assert cnt.min() > 0


def test_phold_hash_sensitivity(oracle):
    base = oracle.phold("seq", 64, 8, 10.0)[0].trace_hash
    assert oracle.phold("seq", 64, 8, 10.0, seed=43)[0].trace_hash != base
    assert oracle.phold("seq", 64, 8, 10.0, p_self_u32=0x40000000)[0].trace_hash != base
    assert oracle.phold("seq", 64, 8, 10.0, lookahead=0.2)[0].trace_hash != base
    h = oracle.trace_step(0, 1, 2, 3)
    assert h != oracle.trace_step(0, 1, 3, 2) and h != oracle.trace_step(0, 2, 2, 3)
    cnt = np.array([1, 2], np.uint32)
    hs = np.array([5, 7], np.uint64)
    assert oracle.combine(cnt, hs) != oracle.combine(cnt[::-1], hs[::-1])


def test_golden_small_records_reproduce(oracle, golden):
    for name in ("tiny_l4_k2_t5", "small_l64_k8_t20", "phold_1024_k16_t100"):
        g = golden[name]
        res, cnt, hsh = oracle.phold("seq", g["num_lps"], g["init_per_lp"], g["end_time"], seed=g["seed"])
        assert res.events == g["events"]
        assert "%016x" % res.trace_hash == g["trace_hash"]
        assert [int(x) for x in cnt] == g["lp_count"]


# --- the reference's known-answer tests for the LAQ ring primitives (src/test_laq.cu) ---

def test_ring_quicksort_fixture(oracle):
    # src/test_laq.cu:39-52: {6,10,13,5,8,3,2} -> 2,3,5,6,8,10,13 (FEL_SIZE 8)
    ring = [6, 10, 13, 5, 8, 3, 2, 0]
    out = oracle.ring_quicksort(ring, 0, 6)
    assert list(out[:7]) == [2, 3, 5, 6, 8, 10, 13]


def test_ring_quicksort_wrapping(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        cap = 16
        left = int(rng.integers(0, cap))
        n = int(rng.integers(1, cap))
        ring = rng.integers(0, 100, cap)
        idx = [(left + i) % cap for i in range(n)]
        out = oracle.ring_quicksort(ring, left, (left + n - 1) % cap)
        assert list(out[idx]) == sorted(ring[idx])


def test_ring_search_fixture(oracle):
    # src/test_laq.cu:147-166
    ring = np.zeros(8, np.int64)
    ring[6], ring[7], ring[0], ring[1] = 2, 4, 6, 8
    assert oracle.ring_search(ring, 6, 2, 1) == 5
    assert oracle.ring_search(ring, 6, 2, 5) == 7
    assert oracle.ring_search(ring, 6, 2, 9) == 2


def test_ring_insert_sorted_fixtures(oracle):
    # src/test_laq.cu:168-189
    ring = np.zeros(8, np.int64)
    ring[2:6] = [2, 4, 6, 8]
    out = oracle.ring_insert_sorted(ring, 2, 6, 5)
    assert list(out[2:7]) == [2, 4, 5, 6, 8]
    # src/test_laq.cu:191-213 (wrapping)
    ring = np.zeros(8, np.int64)
    ring[5], ring[6], ring[7], ring[0], ring[1] = 2, 4, 6, 8, 10
    out = oracle.ring_insert_sorted(ring, 5, 2, 5)
    assert [out[i] for i in (5, 6, 7, 0, 1, 2)] == [2, 4, 5, 6, 8, 10]
    # src/test_laq.cu:215-236 (new minimum, wrapping)
    out = oracle.ring_insert_sorted(ring, 5, 2, 1)
    assert [out[i] for i in (5, 6, 7, 0, 1, 2)] == [1, 2, 4, 6, 8, 10]


def test_ring_insert_sorted_matches_sort(oracle):
    rng = np.random.default_rng(2)
    cap = 32
    for _ in range(100):
        read = int(rng.integers(0, cap))
        n = int(rng.integers(0, cap - 1))
        vals = np.sort(rng.integers(0, 50, n))
        ring = np.full(cap, -7, np.int64)
        for i, v in enumerate(vals):
            ring[(read + i) % cap] = v
        key = int(rng.integers(0, 50))
        out = oracle.ring_insert_sorted(ring, read, (read + n) % cap, key)
        got = [out[(read + i) % cap] for i in range(n + 1)]
        assert got == sorted(list(vals) + [key])


# --- one safe event per LP (src/test_gaq.cu, src/test_hq.cu) ---

def test_safe_select_gaq_fixture(oracle):
    # src/test_gaq.cu:108-151: [1(n0),2(n0),3(n1),4(n1)] -> dequeue ts 1 and ts 3
    n, marks = oracle.safe_select([1, 2, 3, 4], [0, 0, 1, 1], 2, 10)
    assert n == 2 and list(marks) == [1, 0, 1, 0]


def test_safe_select_hq_fixtures(oracle):
    # src/test_hq.cu:1011-1034: [0(n0),1(n0),2(n1),3(n1)] -> marks T,F,T,F
    n, marks = oracle.safe_select([0, 1, 2, 3], [0, 0, 1, 1], 2, 10)
    assert list(marks) == [1, 0, 1, 0]
    # src/test_hq.cu:959-989 (not safe): merged [5,12,15,17] of LP 0, root max 15 -> only ts 5
    n, marks = oracle.safe_select([5, 12, 15, 17], [0, 0, 0, 0], 2, 10, bound=15)
    assert n == 1 and list(marks) == [1, 0, 0, 0]
    # src/test_hq.cu:882-917 (no insert): root [18(n0),23(n1)] -> both dequeued
    n, marks = oracle.safe_select([18, 23], [0, 1], 2, 10, bound=23)
    assert n == 2
    # src/test_hq.cu:689-722 (multi insert): root [10(n0),11(n1)] + 20..36 -> dequeue 10, 11
    ts = [10, 11, 20, 21, 25, 26, 30, 31, 35, 36]
    lp = [0, 1, 0, 0, 0, 0, 0, 0, 0, 0]
    n, marks = oracle.safe_select(ts, lp, 2, 10, bound=11)
    assert n == 2 and list(np.nonzero(marks)[0]) == [0, 1]


# --- A* / shortest paths ---

def _rand_grid(rng, w, h, p):
    return (rng.random(w * h) < p).astype(np.uint8)


def test_astar_equals_dijkstra_and_bfs_4conn(oracle):
    rng = np.random.default_rng(3)
    for _ in range(40):
        w, h = int(rng.integers(2, 60)), int(rng.integers(2, 60))
        g = _rand_grid(rng, w, h, float(rng.uniform(0, 0.4)))
        s = (int(rng.integers(0, w)), int(rng.integers(0, h)))
        t = (int(rng.integers(0, w)), int(rng.integers(0, h)))
        g[s[1] * w + s[0]] = 0
        g[t[1] * w + t[0]] = 0
        a, _ = oracle.astar(g, w, h, 4, s, t)
        assert a == oracle.dijkstra(g, w, h, 4, s, t) == oracle.bfs4(g, w, h, s, t)


def test_astar_equals_dijkstra_8conn(oracle):
    rng = np.random.default_rng(4)
    for _ in range(40):
        w, h = int(rng.integers(2, 60)), int(rng.integers(2, 60))
        g = _rand_grid(rng, w, h, float(rng.uniform(0, 0.4)))
        s = (int(rng.integers(0, w)), int(rng.integers(0, h)))
        t = (int(rng.integers(0, w)), int(rng.integers(0, h)))
        g[s[1] * w + s[0]] = 0
        g[t[1] * w + t[0]] = 0
        a, _ = oracle.astar(g, w, h, 8, s, t)
        assert a == oracle.dijkstra(g, w, h, 8, s, t)


def test_astar_closed_forms(oracle):
    w, h = 37, 23
    g = np.zeros(w * h, np.uint8)
    for s, t in [((0, 0), (36, 22)), ((5, 17), (30, 2)), ((3, 3), (3, 3))]:
        dx, dy = abs(s[0] - t[0]), abs(s[1] - t[1])
        assert oracle.astar(g, w, h, 4, s, t)[0] == dx + dy
        assert oracle.astar(g, w, h, 8, s, t)[0] == 10 * max(dx, dy) + 4 * min(dx, dy)
    # enclosed target is unreachable
    g2 = np.zeros(w * h, np.uint8)
    for x, y in [(9, 10), (11, 10), (10, 9), (10, 11), (9, 9), (11, 11), (9, 11), (11, 9)]:
        g2[y * w + x] = 1
    assert oracle.astar(g2, w, h, 4, (0, 0), (10, 10))[0] == -1
    assert oracle.astar(g2, w, h, 8, (0, 0), (10, 10))[0] == -1
    # a wall with one gap forces a detour of known length (4-connected)
    g3 = np.zeros(w * h, np.uint8)
    for y in range(h):
        if y != 20:
            g3[y * w + 18] = 1
    assert oracle.astar(g3, w, h, 4, (0, 0), (36, 0))[0] == 36 + 2 * 20


def test_grid_generator_statistics(oracle):
    thr = int(0.2 * 2 ** 32)
    g = oracle.grid(512, 512, 7, 2, 0, thr)
    assert abs(g.mean() - 0.2) < 0.005
    g2 = oracle.grid(512, 512, 8, 2, 0, thr)
    assert not np.array_equal(g, g2)
    sides = []
    for q in range(256):
        side, t = oracle.query_shape(11, q)
        assert 256 <= side <= 2048
        assert 0.1 * 2 ** 32 - 1 <= t <= 0.3 * 2 ** 32
        sides.append(side)
    lg = np.log2(sides)
    assert abs(lg.mean() - 9.5) < 0.2           # log-uniform on [8, 11]


def test_query_spec_picks_free_cells(oracle):
    for q in range(8):
        side, thr = oracle.query_shape(11, q)
        g = oracle.grid(side, side, 11, 2, q, thr)
        spec = oracle.query_spec(11, q, g)
        assert spec.side == side
        assert g[spec.sy * side + spec.sx] == 0 and g[spec.ty * side + spec.tx] == 0
