"""Global queues on the GPU: the reference's GAQ and HQ known-answer tests restated at
the queue API level (src/test_gaq.cu, src/test_hq.cu), and randomized rounds checked
against a sorted model (GAQ: exact dequeue set; HQ: every dequeued record is its LP's
minimum inside the window, the heap invariant holds node by node, nothing is lost)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ev(gpq, items):
    a = np.zeros(len(items), dtype=gpq.event_dtype)
    for i, (ts, dst, src, seq) in enumerate(items):
        a[i] = (gpq.f32_bits(ts), dst, src, seq)
    return a


def tsl(gpq, arr):
    return [gpq.bits_f32(t) for t in arr["ts"]]


def heap(gpq, r=2, lps=2, nodes=8, stage=4):
    return gpq.Queue("hq", num_lps=lps, node_size=r, heap_nodes=nodes, stage_per_lp=stage, lookahead=10.0,
                     horizon=1e9)


# ---------------- GAQ (src/test_gaq.cu) ----------------

def test_gaq_insert_init_and_cont(gpq):
    with gpq.Queue("gaq", num_lps=2, total_capacity=16, stage_per_lp=4, lookahead=10.0, horizon=1e9) as q:
        # src/test_gaq.cu:55-77: threads insert (n0,1),(n1,2) and (n0,3),(n1,4) -> 1,2,3,4
        q.insert(ev(gpq, [(1, 0, 0, 0), (2, 1, 0, 1), (3, 0, 1, 0), (4, 1, 1, 1)]))
        q.queue_post()
        assert tsl(gpq, q.lp_contents(0)) == [1.0, 3.0] and tsl(gpq, q.lp_contents(1)) == [2.0, 4.0]
        assert q.queue_get_min_ts() == 1.0 and q.queue_check_phold() == 4
        # src/test_gaq.cu:79-106: then (n0,3),(n1,5) -> 1,2,3,3,4,5
        q.insert(ev(gpq, [(3, 0, 0, 2), (5, 1, 1, 2)]))
        q.queue_post()
        assert tsl(gpq, q.lp_contents(0)) == [1.0, 3.0, 3.0] and tsl(gpq, q.lp_contents(1)) == [2.0, 4.0, 5.0]
        assert q.queue_check_phold() == 6


def test_gaq_dequeue_one_per_lp(gpq):
    # src/test_gaq.cu:108-151: [1(n0), 2(n0), 3(n1), 4(n1)] -> slot 0 gets ts 1, slot 1 gets ts 3
    with gpq.Queue("gaq", num_lps=2, total_capacity=16, stage_per_lp=4, lookahead=10.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(1, 0, 0, 0), (2, 0, 0, 1), (3, 1, 1, 0), (4, 1, 1, 1)]))
        q.queue_post()
        q.queue_pre()
        items, idx = q.peek()
        assert tsl(gpq, items) == [1.0, 3.0] and sorted(idx) == [0, 1]
        q.queue_post()
        assert q.queue_check_phold() == 2 and q.queue_get_min_ts() == 2.0


# ---------------- HQ (src/test_hq.cu) ----------------

def test_hq_post_init_fixture(gpq):
    # src/test_hq.cu:724-783: 18,23 -> root; 31,32 -> node 1; 25,27 -> node 2
    with heap(gpq) as q:
        for pair in [(18, 23), (31, 32), (25, 27)]:
            q.insert(ev(gpq, [(pair[0], 0, 0, pair[0]), (pair[1], 0, 0, pair[1])]))
            q.queue_post_init()
        assert [tsl(gpq, q.heap_node(v)) for v in range(3)] == [[18.0, 23.0], [31.0, 32.0], [25.0, 27.0]]
        assert q.queue_get_min_ts() == 18.0          # src/test_hq.cu:919-932


def build(gpq, q, nodes):
    for vals, dsts in nodes:
        q.insert(ev(gpq, [(t, d, d, int(t)) for t, d in zip(vals, dsts)]))
        q.queue_post_init()


def test_hq_delete_update_no_insert(gpq):
    # src/test_hq.cu:882-917: root [18(n0), 23(n1)], [27,29], [30,31] -> dequeue 18, 23;
    # the last node refills the root, delete-update swaps it with node 1
    with heap(gpq) as q:
        build(gpq, q, [((18, 23), (0, 1)), ((27, 29), (0, 0)), ((30, 31), (0, 0))])
        q.queue_post()
        items, _ = q.peek()
        assert tsl(gpq, items) == [18.0, 23.0]
        assert [tsl(gpq, q.heap_node(v)) for v in range(3)] == [[27.0, 29.0], [30.0, 31.0], []]


def test_hq_delete_update_not_safe(gpq):
    # src/test_hq.cu:959-989: root [5,15], node 1 [18,23], inserts 12, 17 (all LP 0):
    # only 5 is dequeued; nodes [12,15], [18,23], [17]
    with heap(gpq) as q:
        build(gpq, q, [((5, 15), (0, 0)), ((18, 23), (0, 0))])
        q.insert(ev(gpq, [(12, 0, 0, 12), (17, 0, 0, 17)]))
        q.queue_post()
        items, _ = q.peek()
        assert tsl(gpq, items) == [5.0]
        assert [tsl(gpq, q.heap_node(v)) for v in range(3)] == [[12.0, 15.0], [18.0, 23.0], [17.0]]


def test_hq_multi_insert(gpq):
    # src/test_hq.cu:689-722: root [10(n0), 11(n1)], node 1 [15,16], 8 inserts 20..36 ->
    # dequeue 10, 11; nodes hold 15,16,20,21,25,26,30,31 and the last chunk 35,36
    with heap(gpq, nodes=8, stage=4) as q:
        build(gpq, q, [((10, 11), (0, 1)), ((15, 16), (0, 0))])
        q.insert(ev(gpq, [(t, 0, 0, t) for t in (20, 21, 25, 26, 30, 31, 35, 36)]))
        q.queue_post()
        items, _ = q.peek()
        assert tsl(gpq, items) == [10.0, 11.0]
        got = [tsl(gpq, q.heap_node(v)) for v in range(5)]
        assert sorted(x for n in got[:4] for x in n) == [15.0, 16.0, 20.0, 21.0, 25.0, 26.0, 30.0, 31.0]
        assert got[4] == [35.0, 36.0]
        check_heap(gpq, q, 8)


def test_hq_insert_update_fixture(gpq):
    # src/test_hq.cu:605-640 (insert_update): nodes [18,23], [27,29], [30,31] and two new records
    # 28, 31: the root keeps 18, 23; the chunk 28, 31 travels to the next free node (3) through
    # node 1, where 28 stays and 29 moves on: [18,23], [27,28], [30,31], [29,31]
    with heap(gpq) as q:
        build(gpq, q, [((18, 23), (0, 1)), ((27, 29), (0, 0)), ((30, 31), (0, 0))])
        q.insert(ev(gpq, [(28, 0, 1, 28), (31, 0, 1, 131)]))
        q.queue_post_init()
        assert [tsl(gpq, q.heap_node(v)) for v in range(4)] == [[18.0, 23.0], [27.0, 28.0], [30.0, 31.0],
                                                                 [29.0, 31.0]]
        check_heap(gpq, q, 8)


def test_hq_insert_update_partial_node_fixture(gpq):
    # src/test_hq.cu:642-680 (insert_update2): nodes [18,23], [27,29], [30] and two new records
    # 28, 31: the reference fills node 2's free slot first ([28,30]) and starts node 3 with 31.
    # Here the same seven records end in the same heap order, root [18,23] and node 1 [27,29]
    # unchanged
    with heap(gpq) as q:
        build(gpq, q, [((18, 23), (0, 1)), ((27, 29), (0, 0)), ((30,), (0,))])
        q.insert(ev(gpq, [(28, 0, 1, 28), (31, 0, 1, 131)]))
        q.queue_post_init()
        nodes = [tsl(gpq, q.heap_node(v)) for v in range(4)]
        assert nodes[0] == [18.0, 23.0] and nodes[1] == [27.0, 29.0]
        assert sorted(x for n in nodes[2:] for x in n) == [28.0, 30.0, 31.0]
        assert check_heap(gpq, q, 8) == 7


def test_hq_copy_last_nonempty_root_fixture(gpq):
    # src/test_hq.cu:850-879 (copy_last_nonempty_root): after the step's dequeue the root holds
    # only 25; the last node [30,31] refills it (root [25,31], node 2 [30]) and delete-update
    # restores the order: the root keeps the two smallest of root and children, 25 and 27
    with heap(gpq) as q:
        build(gpq, q, [((18, 25), (0, 0)), ((27, 29), (1, 1)), ((30, 31), (1, 1))])
        q.queue_post()
        items, _ = q.peek()
        assert tsl(gpq, items) == [18.0]           # one safe event per LP: 25 waits behind 18
        nodes = [tsl(gpq, q.heap_node(v)) for v in range(3)]
        assert nodes[0] == [25.0, 27.0]
        assert sorted(nodes[1] + nodes[2]) == [29.0, 30.0, 31.0]
        assert check_heap(gpq, q, 8) == 5


def check_heap(gpq, q, nodes, heaps=1, lps_per_heap=None):
    """Every node sorted, parent maximum <= child minimum, counts add up; with several heaps,
    node v of heap h is h * nodes + v and holds only records of heap h's LPs."""
    key = lambda e: (int(e["ts"]), int(e["src"]), int(e["seq"]))
    total = 0
    for h in range(heaps):
        content = [q.heap_node(h * nodes + v) for v in range(nodes)]
        for v, c in enumerate(content):
            ks = [key(e) for e in c]
            assert ks == sorted(ks), "heap %d node %d unsorted" % (h, v)
            total += len(c)
            if heaps > 1:
                assert all(int(e["dst"]) // lps_per_heap == h for e in c), "record in the wrong heap"
            for ch in (2 * v + 1, 2 * v + 2):
                if ch < nodes and len(content[ch]) and len(c):
                    assert ks[-1] <= key(content[ch][0]), "heap %d order violated at %d -> %d" % (h, v, ch)
    assert total == q.queue_check_phold()
    return total


@pytest.mark.parametrize("kind,r,heaps", [("gaq", 0, 1), ("hq", 4, 1), ("hq", 16, 1), ("hq", 64, 1), ("hq", 8, 3),
                                          ("hq", 16, 13)])
def test_random_rounds_against_model(gpq, kind, r, heaps):
    rng = np.random.default_rng(r + 11 + heaps)
    L, la = 13, 0.5
    nodes = 128
    per_heap = -(-L // heaps)
    stage = max(1, min(8, 7 * r // L)) if kind == "hq" else 8
    if kind == "gaq":
        q = gpq.Queue("gaq", num_lps=L, total_capacity=4096, stage_per_lp=8, lookahead=la, horizon=1e9)
    else:
        q = gpq.Queue("hq", num_lps=L, node_size=r, heap_nodes=nodes, stage_per_lp=stage, lookahead=la, horizon=1e9,
                      heaps=heaps)
    enq_cap = L * stage                                   # inserts per step (each heap takes them all)
    model = {}
    seq = 0
    now = 0.0
    with q:
        for rnd in range(40):
            batch = []
            for _ in range(int(rng.integers(0, min(enq_cap, L) + 1))):
                t = float(np.float32(now + rng.integers(0, 12) * 0.25))
                batch.append((t, int(rng.integers(0, L)), int(rng.integers(0, 3)), seq))
                seq += 1
            q.insert(ev(gpq, batch))
            for (t, d, s_, k) in batch:
                model.setdefault(d, []).append((gpq.f32_bits(t), s_, k))
            q.queue_post()
            q.queue_pre()
            items, _ = q.peek()
            pend = sorted(x for v in model.values() for x in v)
            if pend:
                gmin = pend[0][0]
                assert q.queue_get_min_ts() == gpq.bits_f32(gmin)
                wend = gpq.f32_bits(np.float32(gpq.bits_f32(gmin)) + np.float32(la))
                got = [(int(e["ts"]), int(e["src"]), int(e["seq"]), int(e["dst"])) for e in items]
                assert len({d for *_, d in got}) == len(got)                     # at most one per LP
                for (t, s_, k, d) in got:
                    assert (t, s_, k) == min(model[d]) and t < wend             # the LP's minimum, safe
                if kind == "gaq":
                    expect = sorted((min(v)[0], min(v)[1], min(v)[2], d) for d, v in model.items()
                                    if v and min(v)[0] < wend)
                    assert sorted(got) == expect
                else:
                    assert len(got) > 0 or all(not v for v in model.values())
                for (t, s_, k, d) in got:
                    model[d].remove((t, s_, k))
            if kind == "hq":
                check_heap(gpq, q, nodes, heaps, per_heap)
            now += 0.5
        assert q.queue_check_phold() == sum(len(v) for v in model.values())
        assert q.error_flags() == 0


# ---------------- sort tiles of the global queues ----------------

@pytest.mark.parametrize("tile", [1024, 2048, 4096])
@pytest.mark.parametrize("kind", ["gaq", "hq"])
def test_sort_tile_keeps_the_trace(gpq, oracle, golden, kind, tile):
    # every tile of the record sort and merge kernels commits the oracle's trace (per LP), and
    # the 65536-LP golden
    def run(L, k, T):
        kw = dict(capacity_per_lp=64, stage_per_lp=2) if kind == "gaq" else \
            dict(capacity_per_lp=40, stage_per_lp=1, node_size=max(L, 32))
        q = gpq.Queue(kind, num_lps=L, lookahead=0.1, horizon=T, sort_tile=tile, **kw)
        m = gpq.Phold(q, init_per_lp=k, seed=42, end_time=T)
        m.model_init()
        rep = gpq.des_run(m, mode="graph", batch=16)
        c, h = m.lp_state()
        out = (rep["events"], m.trace_hash(), c, h, q.error_flags())
        m.model_finish()
        q.queue_finish()
        return out
    res, cnt, hsh = oracle.phold("seq", 3000, 8, 6.0)
    ev_, th, c, h, fl = run(3000, 8, 6.0)
    assert fl == 0 and ev_ == res.events and th == res.trace_hash
    assert np.array_equal(c, cnt) and np.array_equal(h, hsh)
    gd = golden["mid_l65536_k32_t10"]
    ev_, th, _, _, fl = run(65536, 32, 10.0)
    assert fl == 0 and ev_ == gd["events"] and "%016x" % th == gd["trace_hash"]


@pytest.mark.parametrize("kind", ["gaq", "hq"])
def test_sort_tile_autotune(gpq, oracle, kind):
    L, k, T = 8192, 16, 4.0
    kw = dict(capacity_per_lp=64, stage_per_lp=2) if kind == "gaq" else \
        dict(capacity_per_lp=40, stage_per_lp=1, node_size=L)
    tile, cands = gpq.autotune_sort(kind, L, init_per_lp=k, seed=42, end_time=T, sample_end_time=1.0, **kw)
    assert [c["sort_tile"] for c in cands] == [2048, 1024, 4096]
    assert all(c["status"] == "ran" and c["events_per_s"] > 0 for c in cands)
    picked = [c for c in cands if c["sort_tile"] == tile]
    assert len(picked) == 1 and picked[0]["events_per_s"] * 1.02 >= max(c["events_per_s"] for c in cands)
    q = gpq.Queue(kind, num_lps=L, lookahead=0.1, horizon=T, sort_tile=tile, **kw)
    m = gpq.Phold(q, init_per_lp=k, seed=42, end_time=T)
    m.model_init()
    gpq.des_run(m, mode="graph", batch=16)
    res, _, _ = oracle.phold("seq", L, k, T)
    assert m.trace_hash() == res.trace_hash
    m.model_finish()
    q.queue_finish()


def test_sort_tile_rejected(gpq):
    with pytest.raises(gpq.GpqError):
        gpq.Queue("gaq", num_lps=64, capacity_per_lp=8, stage_per_lp=2, lookahead=0.1, horizon=1.0, sort_tile=3000)


@pytest.mark.parametrize("kind", ["gaq", "hq"])
def test_work_counters(gpq, oracle, kind):
    # the counters behind the primitives' algorithmic bytes: every record inserted is sorted
    # once (the L*k initial events and one successor per committed event), every merge sees at
    # least the sorted inserts, and every committed event came from a scanned prefix
    L, k, T = 2048, 8, 5.0
    kw = dict(capacity_per_lp=64, stage_per_lp=2) if kind == "gaq" else \
        dict(capacity_per_lp=40, stage_per_lp=1, node_size=L)
    q = gpq.Queue(kind, num_lps=L, lookahead=0.1, horizon=T, collect_stats=True, **kw)
    m = gpq.Phold(q, init_per_lp=k, seed=42, end_time=T)
    m.model_init()
    rep = gpq.des_run(m, mode="graph", batch=16)
    st = q.stats()
    th = m.trace_hash()
    m.model_finish()
    q.queue_finish()
    res, _, _ = oracle.phold("seq", L, k, T)
    assert rep["events"] == res.events and th == res.trace_hash
    assert st["sorted"] == L * k + rep["events"]
    assert st["merged"] >= st["sorted"]
    assert st["selected"] >= rep["events"]
    # without collect_stats the counters stay zero
    with gpq.Queue(kind, num_lps=L, lookahead=0.1, horizon=T, **kw) as q2:
        s2 = q2.stats()
        assert s2["sorted"] == s2["merged"] == s2["selected"] == 0
