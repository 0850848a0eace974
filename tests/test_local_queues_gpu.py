"""Per-LP queues (local array, local heap, local splay) on the GPU: the reference's
LAQ known-answer tests restated at the queue API level, random insert/dequeue
rounds against a sorted-list model, and the overflow error paths (forced from the
host through tiny capacities)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ev(gpq, items):
    """items: (ts, dst, src, seq) tuples; ts given as a float."""
    a = np.zeros(len(items), dtype=gpq.event_dtype)
    for i, (ts, dst, src, seq) in enumerate(items):
        a[i] = (gpq.f32_bits(ts), dst, src, seq)
    return a


def key_list(arr):
    return [(int(e["ts"]), int(e["src"]), int(e["seq"])) for e in arr]


def test_insert_ringbuffer_fixture(gpq):
    # src/test_laq.cu:115-145: inserts (n0,1),(n0,2),(n1,3),(n1,4) -> LP0 holds 1,2 and LP1 holds 3,4
    with gpq.Queue("laq", num_lps=2, capacity_per_lp=8, stage_per_lp=4, lookahead=10.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(1, 0, 0, 0), (2, 0, 0, 1), (3, 1, 1, 0), (4, 1, 1, 1)]))
        q.queue_post()
        assert [gpq.bits_f32(e["ts"]) for e in q.lp_contents(0)] == [1.0, 2.0]
        assert [gpq.bits_f32(e["ts"]) for e in q.lp_contents(1)] == [3.0, 4.0]
        assert q.queue_length(0) == 2 and q.queue_length(1) == 2


def test_reduce_fixture(gpq):
    # src/test_laq.cu:93-113: heads 6 (LP0: 6,10) and 5 (LP1: 5,13) -> global min 5
    with gpq.Queue("laq", num_lps=2, capacity_per_lp=8, stage_per_lp=4, lookahead=10.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(10, 0, 0, 0), (6, 0, 0, 1), (13, 1, 1, 0), (5, 1, 1, 1)]))
        q.queue_post()
        assert q.queue_get_min_ts() == 5.0
        # insertion order does not matter: contents are sorted (src/test_laq.cu:54-91)
        assert [gpq.bits_f32(e["ts"]) for e in q.lp_contents(0)] == [6.0, 10.0]
        assert [gpq.bits_f32(e["ts"]) for e in q.lp_contents(1)] == [5.0, 13.0]


def test_peek_is_one_safe_head_per_lp(gpq):
    # window [5, 5 + 10): LP0 head 6 and LP1 head 5 are safe, LP2 head 20 is not
    with gpq.Queue("laq", num_lps=3, capacity_per_lp=8, stage_per_lp=4, lookahead=10.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(6, 0, 0, 0), (7, 0, 0, 1), (5, 1, 1, 0), (20, 2, 2, 0)]))
        q.queue_post()
        items, idx = q.peek(set_done=True)
        assert sorted(gpq.bits_f32(t) for t in items["ts"]) == [5.0, 6.0]
        q.queue_post()
        assert q.queue_get_min_ts() == 7.0
        assert q.queue_check_phold() == 2


def test_ties_keep_total_order(gpq):
    # equal timestamps are ordered by (src, seq)
    with gpq.Queue("laq", num_lps=1, capacity_per_lp=16, stage_per_lp=8, lookahead=1.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(3, 0, 5, 1), (3, 0, 2, 9), (3, 0, 5, 0), (1, 0, 9, 9)]))
        q.queue_post()
        q.insert(ev(gpq, [(3, 0, 2, 8), (3, 0, 4, 0)]))
        q.queue_post()
        got = [(gpq.bits_f32(e["ts"]), int(e["src"]), int(e["seq"])) for e in q.lp_contents(0)]
        assert got == [(1.0, 9, 9), (3.0, 2, 8), (3.0, 2, 9), (3.0, 4, 0), (3.0, 5, 0), (3.0, 5, 1)]


@pytest.mark.parametrize("kind", ["laq", "lhq", "lsq"])
@pytest.mark.parametrize("cap,stage", [(8, 8), (64, 16), (128, 64)])
def test_random_rounds_against_sorted_model(gpq, kind, cap, stage):
    """Random insert rounds (ties in ts, several sources) and one-safe-head-per-LP dequeue
    rounds, compared after every round with a per-LP sorted-list model."""
    rng = np.random.default_rng(cap)
    L = 37
    model = [[] for _ in range(L)]
    seqs = np.zeros(L, np.int64)
    with gpq.Queue(kind, num_lps=L, capacity_per_lp=cap, stage_per_lp=stage, lookahead=0.25,
                   horizon=1e9) as q:
        for rnd in range(25):
            batch = []
            for lp in range(L):
                room = cap - len(model[lp])
                k = int(rng.integers(0, min(stage, room) + 1))
                for _ in range(k):
                    ts = float(np.float32(rng.integers(0, 40) * 0.125 + rnd))
                    src = int(rng.integers(0, 5))
                    batch.append((ts, lp, src, int(seqs[src])))
                    seqs[src] += 1
            order = rng.permutation(len(batch))
            q.insert(ev(gpq, [batch[i] for i in order]))
            for (ts, lp, src, seq) in batch:
                model[lp].append((gpq.f32_bits(ts), src, seq))
            q.queue_post()
            for lp in range(L):
                model[lp].sort()
            # one dequeue round: heads inside [gmin, gmin + 0.25)
            heads = [m[0] for m in model if m]
            if heads:
                gmin = min(h[0] for h in heads)
                assert q.queue_get_min_ts() == gpq.bits_f32(gmin)
                wend = gpq.f32_bits(np.float32(gpq.bits_f32(gmin)) + np.float32(0.25))
                items, _ = q.peek(set_done=True)
                expect = sorted(m[0] for m in model if m and m[0][0] < wend)
                assert sorted(key_list(items)) == expect
                for lp in range(L):
                    if model[lp] and model[lp][0][0] < wend:
                        model[lp].pop(0)
            for lp in (0, 7, L - 1):
                assert key_list(q.lp_contents(lp)) == model[lp]
        assert q.queue_check_phold() == sum(len(m) for m in model)
        assert q.error_flags() == 0


@pytest.mark.parametrize("kind", ["laq", "lhq", "lsq"])
def test_stage_overflow_is_reported(gpq, kind):
    with gpq.Queue(kind, num_lps=2, capacity_per_lp=16, stage_per_lp=2, lookahead=1.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(1, 0, 0, 0), (2, 0, 0, 1), (3, 0, 0, 2)]))
        assert q.error_flags() & 1
        q.queue_post()
        assert len(q.lp_contents(0)) == 2


@pytest.mark.parametrize("kind", ["laq", "lhq", "lsq"])
def test_queue_overflow_is_reported(gpq, kind):
    with gpq.Queue(kind, num_lps=2, capacity_per_lp=4, stage_per_lp=8, lookahead=1.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(float(i), 1, 0, i) for i in range(5)]))
        q.queue_post()
        assert q.error_flags() & 2
        assert q.queue_length(1) <= 4
        assert q.queue_length(0) == 0


@pytest.mark.parametrize("kind", ["laq", "lhq", "lsq"])
def test_phold_overflow_surfaces_as_capacity_error(gpq, kind):
    q = gpq.Queue(kind, num_lps=64, capacity_per_lp=16, stage_per_lp=4, lookahead=0.1, horizon=50.0)
    m = gpq.Phold(q, init_per_lp=12, seed=42, end_time=50.0)
    m.model_init()
    with pytest.raises(gpq.GpqError) as ei:
        gpq.des_run(m, mode="graph", batch=8)
    assert ei.value.status == 2
    m.model_finish()
    q.queue_finish()


@pytest.mark.parametrize("kind", ["laq", "lhq", "lsq"])
def test_clear_and_length(gpq, kind):
    with gpq.Queue(kind, num_lps=2, capacity_per_lp=8, stage_per_lp=4, lookahead=1.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(1, 0, 0, 0), (2, 0, 0, 1), (3, 1, 1, 0)]))
        q.queue_post()
        q.queue_clear(0)
        q.sync()
        assert q.queue_length(0) == 0 and q.queue_length(1) == 1
        q.queue_post()
        assert q.queue_get_min_ts() == 3.0


def test_laq_decrease_key(gpq):
    # LOCAL_ARRAY_QUEUE_UPDATE_ELEMENTS (src/local_array_queue.cu:20-32, 350-416, 830-846): an id
    # (seq) already in the queue keeps one entry, which takes the smaller key at the next post
    inf = float("inf")
    with gpq.Queue("laq", num_lps=2, capacity_per_lp=8, stage_per_lp=4, lookahead=inf, horizon=inf,
                   update_ids=64) as q:
        def lp0():
            return [(gpq.bits_f32(e["ts"]), int(e["seq"])) for e in q.lp_contents(0)]
        q.insert_or_update(ev(gpq, [(10, 0, 0, 5), (20, 0, 0, 6), (4, 1, 0, 9)]))
        q.queue_post()
        assert lp0() == [(10.0, 5), (20.0, 6)]
        q.insert_or_update(ev(gpq, [(8, 0, 0, 6)]))           # cheaper: id 6 moves to the front
        q.queue_post()
        assert lp0() == [(8.0, 6), (10.0, 5)] and q.queue_get_min_ts() == 4.0
        q.insert_or_update(ev(gpq, [(12, 0, 0, 5)]))          # dearer: dropped
        q.queue_post()
        assert lp0() == [(8.0, 6), (10.0, 5)]
        q.insert_or_update(ev(gpq, [(9, 0, 0, 7)]))           # new id: inserted
        q.queue_post()
        assert lp0() == [(8.0, 6), (9.0, 7), (10.0, 5)]
        items, _ = q.peek(set_done=True)                      # pops id 6 (LP0) and id 9 (LP1)
        assert sorted(int(s) for s in items["seq"]) == [6, 9]
        q.queue_post()
        q.insert_or_update(ev(gpq, [(30, 0, 0, 6)]))          # id 6 left the queue: a plain insert
        q.queue_post()
        assert lp0() == [(9.0, 7), (10.0, 5), (30.0, 6)]
        q.insert_or_update(ev(gpq, [(7, 0, 0, 5), (6, 0, 0, 5)]))   # two updates of one id in one step
        q.queue_post()
        assert lp0() == [(6.0, 5), (9.0, 7), (30.0, 6)]
        q.insert_or_update(ev(gpq, [(2, 0, 0, 5)]))           # buffered, then its entry is popped
        q.peek(set_done=True)
        q.queue_post()                                        # so it is inserted anew
        assert lp0() == [(2.0, 5), (9.0, 7), (30.0, 6)]
        assert q.queue_check_phold() == 3 and q.error_flags() == 0


def int_keys(gpq, items):
    """items: (key, dst, src, seq) with integer keys (the open-list mode's, A*'s f values)."""
    a = np.zeros(len(items), dtype=gpq.event_dtype)
    for i, it in enumerate(items):
        a[i] = it
    return a


def test_equal_keys_in_one_post_are_all_kept(gpq):
    # a model may stage one key twice in a step: both copies stay, side by side
    with gpq.Queue("laq", num_lps=1, capacity_per_lp=16, stage_per_lp=8, lookahead=1.0, horizon=1e9) as q:
        q.insert(ev(gpq, [(3, 0, 1, 1), (3, 0, 1, 1), (2, 0, 0, 0), (3, 0, 1, 1)]))
        q.queue_post()
        q.insert(ev(gpq, [(3, 0, 1, 1), (1, 0, 0, 0)]))
        q.queue_post()
        got = [(gpq.bits_f32(e["ts"]), int(e["src"]), int(e["seq"])) for e in q.lp_contents(0)]
        assert got == [(1.0, 0, 0), (2.0, 0, 0)] + [(3.0, 1, 1)] * 4
        assert q.error_flags() == 0


@pytest.mark.parametrize("cap,stage,far,dups", [(16, 8, 256, False), (64, 16, 512, False),
                                                (128, 32, 1024, False), (16, 8, 256, True)])
def test_laq_open_lists_two_tier_random_rounds(gpq, cap, stage, far, dups):
    """Two-tier LAQ with lookahead = +inf (open lists, a horizon per LP): random insert rounds
    with many equal timestamps (and, with dups, equal full keys), a pop of every LP's head on
    some rounds only (so rings overflow into the far stores), then pops until every LP is empty.
    Every pop must be the LP's smallest entry of a per-LP sorted-list model, and the contents
    must equal the model's."""
    rng = np.random.default_rng(cap + far + dups)
    L = 37
    model = [[] for _ in range(L)]
    seqs = np.zeros(L, np.int64)
    inf = float("inf")

    def pop_round():
        items, _ = q.peek(set_done=True)
        expect = sorted(m[0] for m in model if m)
        assert sorted(key_list(items)) == expect
        for m in model:
            if m:
                m.pop(0)

    with gpq.Queue("laq", num_lps=L, capacity_per_lp=cap, stage_per_lp=stage, far_per_lp=far, lookahead=inf,
                   horizon=inf, collect_stats=True) as q:
        for rnd in range(20):
            batch = []
            for lp in range(L):
                for _ in range(int(rng.integers(0, stage + 1))):
                    src = int(rng.integers(0, 5))
                    seq = int(rng.integers(0, 3)) if dups else int(seqs[src])
                    batch.append((int(rng.integers(rnd, rnd + (8 if dups else 40))), lp, src, seq))
                    seqs[src] += 1
            order = rng.permutation(len(batch))
            q.insert(int_keys(gpq, [batch[i] for i in order]))
            for (key, lp, src, seq) in batch:
                model[lp].append((key, src, seq))
            q.queue_post()
            for m in model:
                m.sort()
            for lp in (0, 11, L - 1):
                assert key_list(q.lp_contents(lp)) == model[lp]
            if rnd % 3 == 2:
                pop_round()
                q.queue_post()
        assert q.queue_check_phold() == sum(len(m) for m in model)
        while any(model):
            pop_round()
            q.queue_post()
        assert q.queue_check_phold() == 0
        assert q.error_flags() == 0
        assert q.stats()["aux_bytes"] > 0          # refills from the far stores happened


def test_laq_open_lists_equal_keys_pass_through_a_short_ring(gpq):
    """48 entries with one ts through a ring of 32: the horizon is a full key, so the far store
    takes part of them and refills the ring in (ts, src, seq) order."""
    inf = float("inf")
    with gpq.Queue("laq", num_lps=1, capacity_per_lp=32, stage_per_lp=16, far_per_lp=64, lookahead=inf,
                   horizon=inf, collect_stats=True) as q:
        want = []
        for r in range(3):
            items = [(5, 0, (7 * i) % 3, 16 * r + i) for i in range(16)]
            want += [(5, src, seq) for (_, _, src, seq) in items]
            q.insert(int_keys(gpq, items))
            q.queue_post()
        assert key_list(q.lp_contents(0)) == sorted(want)
        got = []
        while len(got) < 48:
            items, _ = q.peek(set_done=True)
            assert len(items) == 1
            got += key_list(items)
            q.queue_post()
        assert got == sorted(want)
        assert q.error_flags() == 0 and q.queue_check_phold() == 0 and q.stats()["aux_bytes"] > 0


def test_laq_open_lists_far_overflow_is_reported(gpq):
    inf = float("inf")
    with gpq.Queue("laq", num_lps=1, capacity_per_lp=32, stage_per_lp=16, far_per_lp=16, lookahead=inf,
                   horizon=inf) as q:
        for r in range(2):                         # 16 + 16 fill the ring
            q.insert(int_keys(gpq, [(5, 0, 0, 16 * r + i) for i in range(16)]))
            q.queue_post()
        assert q.queue_check_phold() == 32 and q.error_flags() == 0
        q.insert(int_keys(gpq, [(5, 0, 0, 32 + i) for i in range(16)]))
        q.queue_post()                             # 48 > 32: more than 16 go to the far store
        assert q.error_flags() & 2


def test_laq_open_lists_needs_room_for_two_stages(gpq):
    with pytest.raises(gpq.GpqError):
        gpq.Queue("laq", num_lps=4, capacity_per_lp=16, stage_per_lp=16, far_per_lp=64, lookahead=float("inf"),
                  horizon=float("inf"))


def test_laq_decrease_key_needs_one_tier(gpq):
    with pytest.raises(gpq.GpqError):
        gpq.Queue("laq", num_lps=4, capacity_per_lp=64, stage_per_lp=4, lookahead=0.1, horizon=1.0, far_per_lp=64,
                  update_ids=16)
