"""A* on the GPU against the oracle (exact integer costs).

Single query (open lists = LPs of a per-LP queue) and batched queries (one workgroup per
query).  The oracle's A*, Dijkstra and BFS are themselves pinned in test_oracle.py; here
the checks are: the device obstacle generator equals the oracle's bit for bit; costs equal
the oracle's on random grids; closed forms on obstacle-free grids; source == target gives
0; an enclosed target gives -1; plain and persistent drives give the same costs."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = float("inf")


def open_lists(gpq, kind, lists=64, cap=64, stage=32):
    if kind == "laq2":        # two-tier LAQ: a short ring per list backed by an unsorted far store
        return gpq.Queue("laq", num_lps=lists, capacity_per_lp=32, stage_per_lp=16, far_per_lp=1024, lookahead=INF,
                         horizon=INF)
    return gpq.Queue(kind, num_lps=lists, capacity_per_lp=cap, stage_per_lp=stage, lookahead=INF, horizon=INF)


def thr(p):
    return min(int(p * 2 ** 32), 0xFFFFFFFF)


def oracle_grid(oracle, w, h, seed, tag, q, t, forced=()):
    g = oracle.grid(w, h, seed, tag, q, t)
    for (x, y) in forced:
        g[y * w + x] = 0
    return g


@pytest.mark.parametrize("kind", ["laq", "laq2", "lhq", "lsq"])
@pytest.mark.parametrize("conn", [4, 8])
def test_single_query_matches_oracle(gpq, oracle, kind, conn):
    w, h = 97, 61
    for seed in (1, 2, 3):
        s, t = (3, 5), (90, 57)
        with open_lists(gpq, kind) as q:
            a = gpq.Astar(q, w, h, connectivity=conn, source=s, target=t, seed=seed, threshold=0.25)
            a.model_init()
            grid = a.grid()
            ref_grid = oracle_grid(oracle, w, h, seed, 0, 0, thr(0.25), forced=(s, t))
            assert np.array_equal(grid, ref_grid)
            rep = a.run("graph", 8)
            cost, _ = oracle.astar(ref_grid, w, h, conn, s, t)
            assert rep["cost"] == cost == oracle.dijkstra(ref_grid, w, h, conn, s, t)
            assert rep["error_flags"] == 0 and rep["expanded"] >= 1
            if cost >= 0:
                check_path(a.path(), ref_grid, w, conn, s, t, cost)
            a.model_finish()


@pytest.mark.parametrize("conn,ring,stage", [(4, 32, 16), (4, 64, 32), (8, 64, 32)])
def test_single_query_two_tier_open_lists(gpq, oracle, conn, ring, stage):
    """The two-tier LAQ as open lists (the heap-backed LAQ's A* exchange, src/local_array_queue.cu:
    150-180,612-627): each list keeps its best entries in a short sorted ring, the rest in an
    unsorted far store; rings refill from the far stores as they run short.  Costs exact, plain ==
    graph, refills counted, no overflow.  Eight lists: their open sets outgrow the rings, with
    many entries of one f on the obstacle-free grid.  A list's staging takes the inserts of one
    step (up to lists x connectivity in all)."""
    cases = [(97, 61, (3, 5), (90, 57), 0.25, 1), (300, 200, (0, 0), (299, 199), 0.3, 4),
             (256, 256, (128, 0), (128, 255), 0.0, 2), (1000, 700, (5, 690), (990, 3), 0.35, 11)]
    refilled = 0
    for (w, h, s, t, p, seed) in cases:
        with gpq.Queue("laq", num_lps=8, capacity_per_lp=ring, stage_per_lp=stage, far_per_lp=4096, lookahead=INF,
                       horizon=INF, collect_stats=True) as q:
            a = gpq.Astar(q, w, h, connectivity=conn, source=s, target=t, seed=seed, threshold=p)
            a.model_init()
            ref_grid = oracle_grid(oracle, w, h, seed, 0, 0, thr(p), forced=(s, t))
            rep = a.run("graph", 16)
            cost = oracle.dijkstra(ref_grid, w, h, conn, s, t)
            assert rep["cost"] == cost
            assert rep["error_flags"] == 0
            if cost >= 0:
                check_path(a.path(), ref_grid, w, conn, s, t, cost)
            refilled += q.stats()["aux_bytes"]
            a.model_init()
            assert a.run("plain")["cost"] == cost
            a.model_finish()
    assert refilled > 0                               # the far stores were used and refilled the rings


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("window", [0, 25, 0xFFFFFFFF])
def test_single_query_tile_search_matches_oracle(gpq, oracle, conn, tile, window):
    """The tile search (one workgroup per tile, LDS frontier) keeps the exact cost."""
    cases = [(97, 61, (3, 5), (90, 57), 0.25), (200, 150, (199, 0), (0, 149), 0.3), (64, 64, (0, 0), (63, 63), 0.2)]
    for (w, h, s, t, p) in cases:
        for seed in (1, 2, 3):
            with open_lists(gpq, "lhq") as q:
                a = gpq.Astar(q, w, h, connectivity=conn, source=s, target=t, seed=seed, threshold=p, tile=tile,
                              window=window)
                a.model_init()
                ref_grid = oracle_grid(oracle, w, h, seed, 0, 0, thr(p), forced=(s, t))
                rep = a.run("graph", 8)
                cost, _ = oracle.astar(ref_grid, w, h, conn, s, t)
                assert rep["cost"] == cost == oracle.dijkstra(ref_grid, w, h, conn, s, t)
                assert rep["error_flags"] == 0 and rep["closed"] <= rep["expanded"]
                if cost >= 0:
                    check_path(a.path(), ref_grid, w, conn, s, t, cost)
                a.model_finish()


@pytest.mark.parametrize("conn", [4, 8])
def test_single_query_tile_search_large(gpq, oracle, conn):
    """Larger random grids near the percolation density: long detours, many f levels."""
    w, h = 1000, 700
    for seed, p in ((11, 0.35), (12, 0.2)):
        s, t = (5, 690), (990, 3)
        with open_lists(gpq, "lhq") as q:
            a = gpq.Astar(q, w, h, connectivity=conn, source=s, target=t, seed=seed, threshold=p, tile=32)
            a.model_init()
            ref_grid = oracle_grid(oracle, w, h, seed, 0, 0, thr(p), forced=(s, t))
            rep = a.run("graph", 16)
            assert rep["cost"] == oracle.dijkstra(ref_grid, w, h, conn, s, t)
            assert rep["error_flags"] == 0
            if rep["cost"] >= 0:
                check_path(a.path(), ref_grid, w, conn, s, t, rep["cost"])
            a.model_finish()


def test_single_query_tile_search_edges(gpq, oracle):
    with open_lists(gpq, "lhq") as q:
        a = gpq.Astar(q, 1, 1, source=(0, 0), target=(0, 0), threshold=0, tile=16)
        a.model_init()
        assert a.run("plain")["cost"] == 0
        a.model_finish()
        a = gpq.Astar(q, 40, 40, source=(17, 17), target=(17, 17), threshold=0.3, tile=16)
        a.model_init()
        assert a.run("graph", 4)["cost"] == 0
        a.model_finish()
        # an enclosed target: the search ends with nothing pending
        w, h = 48, 40
        grid = np.zeros(w * h, dtype=np.uint8)
        for (x, y) in ((29, 20), (31, 20), (30, 19), (30, 21)):
            grid[y * w + x] = 1
        a = gpq.Astar(q, w, h, source=(1, 1), target=(30, 20), threshold=0, tile=16)
        a.set_grid(grid)
        rep = a.run("graph", 4)
        assert rep["cost"] == -1 and a.path() == []
        a.model_finish()
        # plain == graph
        a = gpq.Astar(q, 120, 90, connectivity=8, source=(2, 3), target=(110, 80), seed=5, threshold=0.25, tile=32)
        a.model_init()
        r1 = a.run("plain")
        a.model_init()
        r2 = a.run("graph", 5)
        ref = oracle_grid(oracle, 120, 90, 5, 0, 0, thr(0.25), forced=((2, 3), (110, 80)))
        assert r1["cost"] == r2["cost"] == oracle.dijkstra(ref, 120, 90, 8, (2, 3), (110, 80))
        a.model_finish()


def test_single_query_tile_must_be_16_or_32(gpq):
    with open_lists(gpq, "lhq") as q:
        with pytest.raises(Exception):
            gpq.Astar(q, 32, 32, tile=8)


def check_path(path, grid, w, conn, s, t, cost):
    """A valid path: source to target, free cells, unit steps, edge costs summing to the cost."""
    assert path[0] == tuple(s) and path[-1] == tuple(t)
    total = 0
    for (x0, y0), (x1, y1) in zip(path, path[1:]):
        assert grid[y1 * w + x1] == 0
        dx, dy = abs(x1 - x0), abs(y1 - y0)
        assert max(dx, dy) == 1 and (conn == 8 or dx + dy == 1)
        total += 1 if conn == 4 else (14 if dx and dy else 10)
    assert total == cost


@pytest.mark.parametrize("conn", [4, 8])
def test_single_query_closed_forms(gpq, conn):
    w, h = 200, 130
    with open_lists(gpq, "lhq", lists=128) as q:
        for (s, t) in [((0, 0), (199, 129)), ((150, 3), (7, 100)), ((5, 5), (5, 5))]:
            a = gpq.Astar(q, w, h, connectivity=conn, source=s, target=t, threshold=0)
            a.model_init()
            rep = a.run("plain")
            dx, dy = abs(s[0] - t[0]), abs(s[1] - t[1])
            expect = dx + dy if conn == 4 else 10 * max(dx, dy) + 4 * min(dx, dy)
            assert rep["cost"] == expect
            a.model_finish()


def test_single_query_enclosed_target(gpq, oracle):
    w, h = 40, 30
    grid = np.zeros((h, w), dtype=np.uint8)
    grid[9:14, 19] = grid[9:14, 23] = 1
    grid[9, 19:24] = grid[13, 19:24] = 1        # a closed box around (21, 11)
    with open_lists(gpq, "laq") as q:
        a = gpq.Astar(q, w, h, connectivity=8, source=(0, 0), target=(21, 11), threshold=0)
        a.model_init()
        a.set_grid(grid)
        assert a.run("graph", 4)["cost"] == -1 == oracle.dijkstra(grid.reshape(-1), w, h, 8, (0, 0), (21, 11))
        assert a.path() == []
        a.model_finish()


def test_single_query_plain_equals_graph(gpq, oracle):
    w, h = 128, 128
    with open_lists(gpq, "laq", lists=256) as q:
        a = gpq.Astar(q, w, h, connectivity=4, seed=9, threshold=0.2)
        a.model_init()
        r1 = a.run("plain")
        a.model_init()
        r2 = a.run("graph", 16)
        assert r1["cost"] == r2["cost"]
        ref = oracle_grid(oracle, w, h, 9, 0, 0, thr(0.2), forced=((0, 0), (w - 1, h - 1)))
        assert r1["cost"] == oracle.bfs4(ref, w, h, (0, 0), (w - 1, h - 1))
        a.model_finish()


@pytest.mark.parametrize("tile", [0, 16, 32, "laq2"])
def test_single_query_4096_grid(gpq, oracle, tile):
    """The 4096 x 4096, 4-connected, Bernoulli(0.2) seed 7 query, (0,0) -> (4095,4095): open
    lists in a local heap queue (tile 0), in two-tier local array queues (the benchmark's: a
    32-record ring and a 256-record far store per list) and the tile search."""
    w = h = 4096
    if tile == "laq2":
        tile = 0
        q = gpq.Queue("laq", num_lps=1 << 16, capacity_per_lp=32, stage_per_lp=16, far_per_lp=256, lookahead=INF,
                      horizon=INF)
    else:
        q = gpq.Queue("lhq", num_lps=64 if tile else 1 << 16, capacity_per_lp=64, stage_per_lp=16, lookahead=INF,
                      horizon=INF)
    with q:
        a = gpq.Astar(q, w, h, connectivity=4, seed=7, threshold=0.2, tile=tile)
        a.model_init()
        rep = a.run("graph", 32)
        grid = a.grid()
        ref = oracle_grid(oracle, w, h, 7, 0, 0, thr(0.2), forced=((0, 0), (w - 1, h - 1)))
        assert np.array_equal(grid, ref)
        assert rep["cost"] == oracle.bfs4(ref, w, h, (0, 0), (w - 1, h - 1))
        assert rep["error_flags"] == 0
        check_path(a.path(), ref, w, 4, (0, 0), (w - 1, h - 1), rep["cost"])
        a.model_finish()


# ---------------------------------------------------------------------------------------
# batched
# ---------------------------------------------------------------------------------------

def synthetic(gpq, seed, n):
    return [gpq.query_synthetic(seed, q) for q in range(n)]


def test_synthetic_specs_match_oracle(gpq, oracle):
    for q in range(12):
        s = gpq.query_synthetic(11, q)
        side, t = oracle.query_shape(11, q)
        assert (s.side, s.threshold, s.connectivity) == (side, t, 8)
        ref = oracle.query_spec(11, q, oracle.grid(side, side, 11, 1, q, t))
        assert (s.sx, s.sy, s.tx, s.ty) == (ref.sx, ref.sy, ref.tx, ref.ty)


@pytest.mark.parametrize("k", [64, 256])
def test_batch_matches_oracle(gpq, oracle, k):
    specs = synthetic(gpq, 11, 24)
    with gpq.AstarBatch(specs, threads_per_query=k) as b:
        b.run("graph")
        costs, expanded = b.results()
        for q, s in enumerate(specs):
            grid = oracle.grid(s.side, s.side, 11, 1, q, s.threshold)
            if q < 3:
                assert np.array_equal(b.grid(q).reshape(-1), grid)
            c, _ = oracle.astar(grid, s.side, s.side, 8, (s.sx, s.sy), (s.tx, s.ty))
            assert costs[q] == c, "query %d" % q
            assert expanded[q] >= 1 or s.sx == s.tx and s.sy == s.ty


def test_batch_plain_equals_graph_and_reruns(gpq):
    specs = synthetic(gpq, 5, 6)
    with gpq.AstarBatch(specs, threads_per_query=64, max_resident=2) as b:
        b.run("plain")
        c1, _ = b.results()
        for _ in range(3):                       # generation tags: reruns in the same slots
            b.run("graph")
            c2, _ = b.results()
            assert np.array_equal(c1, c2)


def test_batch_edge_cases(gpq, oracle):
    S = gpq._lib.QuerySpec
    specs = [
        S(64, 0, 10, 10, 10, 10, 8, 1, 1, 0),            # source == target
        S(300, 0, 0, 0, 299, 299, 8, 1, 1, 1),           # obstacle-free: closed form
        S(300, 0, 7, 250, 280, 3, 4, 1, 1, 2),           # obstacle-free 4-connected
        S(256, thr(0.35), 0, 0, 255, 255, 4, 3, 1, 3),   # dense 4-connected
        S(1, 0, 0, 0, 0, 0, 8, 1, 1, 4),                 # 1 x 1 grid
        S(2048, thr(0.1), 0, 0, 2047, 2047, 8, 3, 1, 5), # the largest side
    ]
    with gpq.AstarBatch(specs, threads_per_query=128) as b:
        b.run("graph")
        costs, _ = b.results()
    assert costs[0] == 0 and costs[4] == 0
    assert costs[1] == 14 * 299
    assert costs[2] == 273 + 247
    for i in (3, 5):
        s = specs[i]
        grid = oracle_grid(oracle, s.side, s.side, s.seed, 1, s.query, s.threshold,
                           forced=((s.sx, s.sy), (s.tx, s.ty)))
        assert costs[i] == oracle.dijkstra(grid, s.side, s.side, s.connectivity, (s.sx, s.sy), (s.tx, s.ty))


def test_batch_rejects_bad_specs(gpq):
    S = gpq._lib.QuerySpec
    with pytest.raises(gpq.GpqError):
        gpq.AstarBatch([S(4096, 0, 0, 0, 1, 1, 8, 1, 1, 0)])
    with pytest.raises(gpq.GpqError):
        gpq.AstarBatch([S(64, 0, 0, 0, 64, 1, 8, 1, 1, 0)])
    with pytest.raises(gpq.GpqError):
        gpq.AstarBatch([S(64, 0, 0, 0, 1, 1, 8, 1, 1, 0)], threads_per_query=96)


def test_batch_autotune_keeps_results(gpq):
    specs = synthetic(gpq, 13, 16)
    kw, sec = gpq.autotune_batch(specs, sample=8)
    assert kw["threads_per_query"] in (64, 128, 256, 512, 1024) and kw["stage_capacity"] in (8, 16, 32)
    assert sec > 0
    with gpq.AstarBatch(specs, **kw) as b:
        b.run("graph")
        c1, _ = b.results()
    with gpq.AstarBatch(specs, threads_per_query=64) as b:
        b.run("graph")
        c2, _ = b.results()
    assert np.array_equal(c1, c2)


def test_path_image(gpq, tmp_path):
    w, h = 48, 32
    with open_lists(gpq, "lhq") as q:
        a = gpq.Astar(q, w, h, connectivity=8, source=(1, 1), target=(46, 30), seed=4, threshold=0.2)
        a.model_init()
        a.run("graph", 8)
        path = a.path()
        f = tmp_path / "grid.ppm"
        gpq.write_ppm(str(f), a.grid(), w, h, path)
        data = f.read_bytes()
        assert data.startswith(b"P6\n48 32\n255\n") and len(data) == len(b"P6\n48 32\n255\n") + 3 * w * h
        a.model_finish()


@pytest.mark.parametrize("conn", [4, 8])
def test_single_query_laq_decrease_key(gpq, oracle, conn):
    # open lists on a LAQ in decrease-key mode: one entry per cell, cheaper paths update it;
    # exact costs, the same as the duplicate-insert lists, and a second search on the same queue
    w, h = 97, 61
    for seed in (1, 2, 3):
        s, t = (3, 5), (90, 57)
        with gpq.Queue("laq", num_lps=64, capacity_per_lp=128, stage_per_lp=32, lookahead=INF, horizon=INF,
                       update_ids=w * h, collect_stats=True) as q:
            a = gpq.Astar(q, w, h, connectivity=conn, source=s, target=t, seed=seed, threshold=0.25)
            a.model_init()
            ref_grid = oracle_grid(oracle, w, h, seed, 0, 0, thr(0.25), forced=(s, t))
            cost, _ = oracle.astar(ref_grid, w, h, conn, s, t)
            rep = a.run("graph", 8)
            assert rep["cost"] == cost and rep["error_flags"] == 0
            if cost >= 0:
                check_path(a.path(), ref_grid, w, conn, s, t, cost)
            a.model_init()                                   # the lists start empty again
            rep2 = a.run("plain", 1)
            assert rep2["cost"] == cost and rep2["error_flags"] == 0
            a.model_finish()


def test_single_query_decrease_key_needs_enough_ids(gpq):
    with gpq.Queue("laq", num_lps=64, capacity_per_lp=128, stage_per_lp=32, lookahead=INF, horizon=INF,
                   update_ids=100) as q:
        with pytest.raises(gpq.GpqError):
            gpq.Astar(q, 97, 61)


def test_batch_slot_reuse_past_tag_exhaustion(gpq, oracle):
    # one slot serves 66 queries: the 64th exhausts the slot's generation tags and clears it;
    # it is smaller than the queries before and after it, so the clear must cover the whole
    # slot and not only its own cells (the 65th and 66th read cells the 64th never touched)
    S = gpq._lib.QuerySpec
    specs = []
    for q in range(66):
        side = 8 if q == 63 else 128
        specs.append(S(side, thr(0.2), 0, 0, side - 1, side - 1, 8, 9, 1, q))
    with gpq.AstarBatch(specs, threads_per_query=64, max_resident=1) as b:
        b.run("graph")
        costs, _ = b.results()
        for q in (0, 62, 63, 64, 65):
            s = specs[q]
            grid = b.grid(q).reshape(-1)
            c, _ = oracle.astar(grid, s.side, s.side, 8, (s.sx, s.sy), (s.tx, s.ty))
            assert costs[q] == c, "query %d" % q


def test_batch_paths(gpq, oracle):
    # an optimal path of every query, read back from the query searched again alone: valid steps
    # over free cells whose costs sum to the batch's (and the oracle's) cost
    specs = synthetic(gpq, 11, 10)
    S = gpq._lib.QuerySpec
    specs.append(S(64, thr(0.0), 3, 3, 3, 3, 8, 1, 1, 10))             # source == target
    specs.append(S(96, thr(0.25), 0, 0, 95, 95, 4, 5, 1, 11))          # 4-connected
    with gpq.AstarBatch(specs, threads_per_query=128) as b:
        b.run("graph")
        costs, _ = b.results()
        for q, s in enumerate(specs):
            grid = b.grid(q).reshape(-1)
            path = b.path(q)
            if costs[q] < 0:
                assert path == []
                continue
            check_path(path, grid, s.side, s.connectivity, (s.sx, s.sy), (s.tx, s.ty), costs[q])
            if q < 10:
                c, _ = oracle.astar(grid, s.side, s.side, 8, (s.sx, s.sy), (s.tx, s.ty))
                assert costs[q] == c
        # the batch's results stay as they were
        assert np.array_equal(b.results()[0], costs)
