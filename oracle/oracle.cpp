// oracle.cpp -- CPU restatement of the GPUPQ hot path.  TEST INFRASTRUCTURE ONLY:
// loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker; never linked into or called by the product library.
//
// Each function cites the reference file:line it follows (paths relative to the
// reference root).  The benchmark workloads redefine the reference's numerics
// (float32 timestamps, Philox streams, total-order tie-break, k events per LP);
// the algorithmic structure below follows the reference.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <atomic>
#include <queue>
#include <thread>
#include <vector>

// ------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Round: two 32x32->64 products of counter words 0 and 2
// with the multipliers, the high halves xor'ed with the other counter words and
// the key; the key is bumped by the Weyl constants between rounds.
// ------------------------------------------------------------------------------
extern "C" void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t pa = (uint64_t)0xD2511F53u * c0;
    const uint64_t pb = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(pb >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)pb;
    const uint32_t n2 = (uint32_t)(pa >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)pa;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ------------------------------------------------------------------------------
// Exp(1) from one draw.  -ln(u), u = n * 2^-24, n = (x>>8)+1 in [1, 2^24].
// n = m * 2^k, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
// series to s^13 (|s| <= 0.1716 so the truncation is < 1e-12 relative).  Every
// operation is an IEEE double add/mul/div with a fixed order and no contraction,
// so any IEEE-754 machine gets the same bits.  The reference draws
// logf(curand_uniform) (src/phold.cu:29-34); that generator is replaced here.
// ------------------------------------------------------------------------------
extern "C" float orc_exp_from_u32(uint32_t x) {
  const uint32_t n = (x >> 8) + 1u;
  int k = 31 - __builtin_clz(n);
  uint64_t scale_bits = (uint64_t)(1023 - k) << 52;   // 2^-k, exact
  double scale;
  std::memcpy(&scale, &scale_bits, sizeof(scale));
  double m = (double)n * scale;                         // exact, in [1, 2)
  if (m >= 0x1.6a09e667f3bcdp+0) { m = m * 0.5; k += 1; }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 0x1.3b13b13b13b14p-3;                      // 2/13
  p = 0x1.745d1745d1746p-3 + z * p;                     // 2/11
  p = 0x1.c71c71c71c71cp-3 + z * p;                     // 2/9
  p = 0x1.2492492492492p-2 + z * p;                     // 2/7
  p = 0x1.999999999999ap-2 + z * p;                     // 2/5
  p = 0x1.5555555555555p-1 + z * p;                     // 2/3
  p = 2.0 + z * p;
  const double ln_m = s * p;
  const double r = (double)(24 - k) * 0x1.62e42fefa39efp-1 - ln_m;
  return (float)r;
}

extern "C" void orc_exp_table(uint32_t first, uint32_t count, float* out) {
  for (uint32_t i = 0; i < count; ++i) out[i] = orc_exp_from_u32((first + i) << 8);
}

// ------------------------------------------------------------------------------
// PHOLD
// ------------------------------------------------------------------------------
namespace {

struct Ev {
  uint32_t ts;   // float bits; ts >= 0 so uint order == float order
  uint32_t dst;
  uint32_t src;
  uint32_t seq;
};

// Total order (ts, src, seq).  The reference orders on ts only (src/queue_item.h:22-28)
// and lets ties fall to whichever structure holds them; (src, seq) makes every key unique.
inline bool key_less(const Ev& a, const Ev& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}
struct KeyGreater {
  bool operator()(const Ev& a, const Ev& b) const { return key_less(b, a); }
};

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline uint64_t mix64(uint64_t x) {  // splitmix64 finalizer
  x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27; x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// Draw block for the n-th event of LP lp on stream `stream`: counter {n, lp, stream, 0}.
inline void draw(uint32_t seed, uint32_t n, uint32_t lp, uint32_t stream, uint32_t out[4]) {
  const uint32_t ctr[4] = {n, lp, stream, 0u};
  const uint32_t key[2] = {seed, 0u};
  orc_philox4x32_10(ctr, key, out);
}

// Initial events: k per LP, owned and sourced by the LP itself, ts = Exp(1) from
// stream 1 (replaces the uniform-destination population of src/phold.cu:63-79).
void populate(const orc_phold_cfg* c, std::vector<Ev>& out) {
  out.clear();
  out.reserve((size_t)c->num_lps * c->init_per_lp);
  for (uint32_t lp = 0; lp < c->num_lps; ++lp)
    for (uint32_t j = 0; j < c->init_per_lp; ++j) {
      uint32_t r[4];
      draw(c->seed, j, lp, 1u, r);
      out.push_back(Ev{fbits(orc_exp_from_u32(r[0])), lp, lp, j});
    }
}

// One committed event at LP lp (src/phold.cu:93-116): count it, extend the LP's
// trace hash, draw the destination and ts' = (ts + lookahead) + Exp(1).
struct LpState {
  std::vector<uint32_t> count;
  std::vector<uint64_t> hash;
  std::vector<uint32_t> last_ts;
  bool monotone = true;
};

inline Ev handle(const orc_phold_cfg* c, LpState& st, const Ev& e) {
  const uint32_t lp = e.dst;
  const uint32_t n = st.count[lp]++;
  if (e.ts < st.last_ts[lp]) st.monotone = false;   // src/des.cu:52-64
  st.last_ts[lp] = e.ts;
  st.hash[lp] = orc_trace_step(st.hash[lp], e.ts, e.src, e.seq);
  uint32_t r[4];
  draw(c->seed, n, lp, 0u, r);
  const float ex = orc_exp_from_u32(r[0]);
  const uint32_t dst = (r[1] < c->p_self_u32) ? lp
                       : (uint32_t)(((uint64_t)r[2] * c->num_lps) >> 32);
  const float t_new = (bitsf(e.ts) + c->lookahead) + ex;
  return Ev{fbits(t_new), dst, lp, c->init_per_lp + n};
}

void init_state(const orc_phold_cfg* c, LpState& st) {
  st.count.assign(c->num_lps, 0u);
  st.hash.assign(c->num_lps, 0ull);
  st.last_ts.assign(c->num_lps, 0u);
  st.monotone = true;
}

void finish(const orc_phold_cfg* c, const LpState& st, uint32_t* lp_count, uint64_t* lp_hash,
            orc_phold_result* res, double secs, size_t pending) {
  uint64_t total = 0;
  for (uint32_t i = 0; i < c->num_lps; ++i) total += st.count[i];
  if (lp_count) std::memcpy(lp_count, st.count.data(), sizeof(uint32_t) * c->num_lps);
  if (lp_hash) std::memcpy(lp_hash, st.hash.data(), sizeof(uint64_t) * c->num_lps);
  res->events = total;
  res->trace_hash = orc_phold_combine(st.count.data(), st.hash.data(), c->num_lps);
  res->seconds = secs;
  res->in_flight_end = (uint32_t)pending;
  res->monotone_ok = st.monotone ? 1u : 0u;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" uint64_t orc_trace_step(uint64_t h, uint32_t ts_bits, uint32_t src, uint32_t seq) {
  h = mix64(h ^ (((uint64_t)ts_bits << 32) | src));
  h = mix64(h ^ ((uint64_t)seq + 0x9E3779B97F4A7C15ull));
  return h;
}

extern "C" uint64_t orc_phold_combine(const uint32_t* lp_count, const uint64_t* lp_hash, uint32_t num_lps) {
  uint64_t h = 0x6A09E667F3BCC908ull;
  for (uint32_t i = 0; i < num_lps; ++i) {
    h = mix64(h ^ lp_hash[i]);
    h = mix64(h ^ (((uint64_t)lp_count[i] << 32) | i));
  }
  return h;
}

// Sequential simulation: the global minimum is popped while its ts < T
// (src/des.cu:154-214 with one event per iteration).
extern "C" int orc_phold_seq(const orc_phold_cfg* c, uint32_t* lp_count, uint64_t* lp_hash,
                             orc_phold_result* res) {
  if (!c || !res || c->num_lps == 0) return -1;
  std::memset(res, 0, sizeof(*res));
  std::vector<Ev> init;
  populate(c, init);
  LpState st;
  init_state(c, st);
  std::priority_queue<Ev, std::vector<Ev>, KeyGreater> pq(KeyGreater(), std::move(init));
  const uint32_t t_end = fbits(c->end_time);
  const double t0 = now_s();
  while (!pq.empty() && pq.top().ts < t_end) {
    const Ev e = pq.top();
    pq.pop();
    pq.push(handle(c, st, e));
  }
  finish(c, st, lp_count, lp_hash, res, now_s() - t0, pq.size());
  return 0;
}

// Brute force: every pending event in one flat vector, the minimum found by a full
// linear scan each time (no heap, no per-LP structure).
extern "C" int orc_phold_brute(const orc_phold_cfg* c, uint32_t* lp_count, uint64_t* lp_hash,
                               orc_phold_result* res) {
  if (!c || !res || c->num_lps == 0) return -1;
  std::memset(res, 0, sizeof(*res));
  std::vector<Ev> pending;
  populate(c, pending);
  LpState st;
  init_state(c, st);
  const uint32_t t_end = fbits(c->end_time);
  const double t0 = now_s();
  for (;;) {
    size_t best = pending.size();
    for (size_t i = 0; i < pending.size(); ++i)
      if (best == pending.size() || key_less(pending[i], pending[best])) best = i;
    if (best == pending.size() || pending[best].ts >= t_end) break;
    const Ev e = pending[best];
    pending[best] = handle(c, st, e);
  }
  finish(c, st, lp_count, lp_hash, res, now_s() - t0, pending.size());
  return 0;
}

// Windowed (conservative lookahead) simulation: per step gmin = min head ts over
// LPs, wend = min(gmin + lookahead, T); every LP whose head ts < wend commits
// exactly that head (src/local_array_queue.cu:760-779); new events are merged
// after the step (src/local_array_queue.cu:577-600).
extern "C" int orc_phold_windowed(const orc_phold_cfg* c, uint32_t* lp_count, uint64_t* lp_hash,
                                  orc_phold_result* res) {
  if (!c || !res || c->num_lps == 0) return -1;
  std::memset(res, 0, sizeof(*res));
  std::vector<Ev> init;
  populate(c, init);
  const uint32_t L = c->num_lps;
  std::vector<std::priority_queue<Ev, std::vector<Ev>, KeyGreater>> q(L);
  for (const Ev& e : init) q[e.dst].push(e);
  LpState st;
  init_state(c, st);
  const float t_end = c->end_time;
  std::vector<Ev> staged;
  uint64_t steps = 0;
  uint32_t max_len = 0;
  const double t0 = now_s();
  for (;;) {
    uint32_t gmin = 0xFFFFFFFFu;
    for (uint32_t lp = 0; lp < L; ++lp) {
      if (!q[lp].empty()) gmin = std::min(gmin, q[lp].top().ts);
      max_len = std::max<uint32_t>(max_len, (uint32_t)q[lp].size());
    }
    if (gmin == 0xFFFFFFFFu || !(bitsf(gmin) < t_end)) break;
    float wend = bitsf(gmin) + c->lookahead;
    if (t_end < wend) wend = t_end;
    staged.clear();
    for (uint32_t lp = 0; lp < L; ++lp) {
      if (q[lp].empty() || !(bitsf(q[lp].top().ts) < wend)) continue;
      const Ev e = q[lp].top();
      q[lp].pop();
      staged.push_back(handle(c, st, e));
    }
    for (const Ev& e : staged) q[e.dst].push(e);
    ++steps;
  }
  size_t pending = 0;
  for (uint32_t lp = 0; lp < L; ++lp) pending += q[lp].size();
  finish(c, st, lp_count, lp_hash, res, now_s() - t0, pending);
  res->steps = steps;
  res->max_queue_len = max_len;
  return 0;
}

// Windowed drain on `threads` host threads: every LP commits all of its events below
// wend = min(gmin + lookahead, T) in key order, then the events created in the window are
// delivered.  An event created in a window lies at or beyond its end ((ts + lookahead) + Exp
// >= fl(gmin + lookahead) for ts >= gmin), so each LP commits the sequential trace.  Thread t
// owns the LPs [t*L/threads, (t+1)*L/threads) and hands events for other threads' LPs over
// through one outbox per (source thread, destination thread) pair.  This is the all-cores CPU
// baseline; the trace equals orc_phold_seq's.
namespace {
struct SpinBarrier {
  std::atomic<uint32_t> arrived{0}, phase{0};
  uint32_t n;
  explicit SpinBarrier(uint32_t n_) : n(n_) {}
  void wait() {
    const uint32_t p = phase.load(std::memory_order_acquire);
    if (arrived.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
      arrived.store(0, std::memory_order_relaxed);
      phase.store(p + 1, std::memory_order_release);
    } else {
      while (phase.load(std::memory_order_acquire) == p) std::this_thread::yield();
    }
  }
};
}  // namespace

extern "C" int orc_phold_parallel(const orc_phold_cfg* c, uint32_t threads, uint32_t* lp_count, uint64_t* lp_hash,
                                  orc_phold_result* res) {
  if (!c || !res || c->num_lps == 0 || threads == 0) return -1;
  std::memset(res, 0, sizeof(*res));
  const uint32_t L = c->num_lps;
  if (threads > L) threads = L;
  std::vector<Ev> init;
  populate(c, init);
  using PQ = std::priority_queue<Ev, std::vector<Ev>, KeyGreater>;
  std::vector<PQ> q(L);
  for (const Ev& e : init) q[e.dst].push(e);
  init.clear();
  init.shrink_to_fit();
  LpState st;
  init_state(c, st);
  auto lo = [&](uint32_t t) { return (uint32_t)((uint64_t)L * t / threads); };
  auto owner = [&](uint32_t lp) {   // largest t with lo(t) <= lp
    uint32_t t = (uint32_t)((uint64_t)lp * threads / L);
    while (t + 1 < threads && lo(t + 1) <= lp) ++t;
    while (lo(t) > lp) --t;
    return t;
  };
  std::vector<std::vector<Ev>> box((size_t)threads * threads);
  std::vector<uint32_t> tmin(threads, 0xFFFFFFFFu);
  std::vector<uint8_t> mono(threads, 1);
  std::vector<uint64_t> tsteps(threads, 0);
  SpinBarrier bar(threads);
  const uint32_t t_end = fbits(c->end_time);
  auto work = [&](uint32_t t) {
    const uint32_t a = lo(t), b = lo(t + 1);
    uint32_t m = 0xFFFFFFFFu;
    for (uint32_t lp = a; lp < b; ++lp)
      if (!q[lp].empty()) m = std::min(m, q[lp].top().ts);
    tmin[t] = m;
    bar.wait();
    uint64_t steps = 0;
    for (;;) {
      uint32_t gmin = 0xFFFFFFFFu;
      for (uint32_t u = 0; u < threads; ++u) gmin = std::min(gmin, tmin[u]);
      if (gmin == 0xFFFFFFFFu || !(gmin < t_end)) break;
      float wend = bitsf(gmin) + c->lookahead;
      if (c->end_time < wend) wend = c->end_time;
      const uint32_t wbits = fbits(wend);
      ++steps;
      for (uint32_t lp = a; lp < b; ++lp) {
        PQ& h = q[lp];
        while (!h.empty() && h.top().ts < wbits) {
          const Ev e = h.top();
          h.pop();
          const uint32_t n = st.count[lp]++;
          if (e.ts < st.last_ts[lp]) mono[t] = 0;
          st.last_ts[lp] = e.ts;
          st.hash[lp] = orc_trace_step(st.hash[lp], e.ts, e.src, e.seq);
          uint32_t r[4];
          draw(c->seed, n, lp, 0u, r);
          const float ex = orc_exp_from_u32(r[0]);
          const uint32_t dst = (r[1] < c->p_self_u32) ? lp : (uint32_t)(((uint64_t)r[2] * L) >> 32);
          const Ev ne{fbits((bitsf(e.ts) + c->lookahead) + ex), dst, lp, c->init_per_lp + n};
          if (dst >= a && dst < b) box[(size_t)t * threads + t].push_back(ne);
          else box[(size_t)t * threads + owner(dst)].push_back(ne);
        }
      }
      bar.wait();                                  // every outbox of this window is complete
      for (uint32_t u = 0; u < threads; ++u) {
        std::vector<Ev>& in = box[(size_t)u * threads + t];
        for (const Ev& e : in) q[e.dst].push(e);
        in.clear();
      }
      m = 0xFFFFFFFFu;
      for (uint32_t lp = a; lp < b; ++lp)
        if (!q[lp].empty()) m = std::min(m, q[lp].top().ts);
      bar.wait();                                  // every thread read the old minima
      tmin[t] = m;
      bar.wait();                                  // the new minima are published
    }
    tsteps[t] = steps;
  };
  const double t0 = now_s();
  std::vector<std::thread> pool;
  for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  const double secs = now_s() - t0;
  size_t pending = 0;
  for (uint32_t lp = 0; lp < L; ++lp) pending += q[lp].size();
  for (uint32_t t = 0; t < threads; ++t) st.monotone = st.monotone && mono[t];
  finish(c, st, lp_count, lp_hash, res, secs, pending);
  res->steps = tsteps[0];
  return 0;
}

// ------------------------------------------------------------------------------
// LAQ ring primitives (int64 keys, as the reference's `long ts`).
// ------------------------------------------------------------------------------

// i mod cap with a non-negative result (src/sort.cu:17-26).
extern "C" int orc_ring_wrap(int i, int cap) {
  int r = i % cap;
  return r < 0 ? r + cap : r;
}

// Position search on the sorted run [read, write) (src/sort.cu:48-66): walks from
// the slot before `read` and returns the slot after which `key` belongs; the
// reference's check `i + 1 == write` is on the unwrapped index.
extern "C" int orc_ring_search(const int64_t* ring, int cap, int read, int write, int64_t key) {
  int i = orc_ring_wrap(read - 1, cap);
  const int stop = orc_ring_wrap(write + 1, cap);
  while (i != stop) {
    if (i + 1 == write) return i + 1;
    if (ring[orc_ring_wrap(i + 1, cap)] > key) return i;
    i = orc_ring_wrap(i + 1, cap);
  }
  return i;
}

// Backward insertion into the sorted run [read, write) whose slot `write` is free
// (src/sort.cu:329-356): larger keys move up one slot; equal keys stay in front.
extern "C" void orc_ring_insert_sorted(int64_t* ring, int cap, int read, int write, int64_t key) {
  int pos = write;
  while (pos != read) {
    const int prev = orc_ring_wrap(pos - 1, cap);
    if (ring[prev] <= key) break;
    ring[pos] = ring[prev];
    pos = prev;
  }
  ring[pos] = key;
}

// In-place sort of the ring range [left, right] (inclusive, wrapping when
// right < left), as quicksort_seq does for a modified LP (src/sort.cu:105-131).
extern "C" void orc_ring_quicksort(int64_t* ring, int cap, int left, int right) {
  if (left == right) return;
  if (right < left) right += cap;
  std::vector<std::pair<int, int>> stack{{left, right}};
  auto at = [&](int i) -> int64_t& { return ring[orc_ring_wrap(i, cap)]; };
  while (!stack.empty()) {
    auto [lo, hi] = stack.back();
    stack.pop_back();
    if (lo >= hi) continue;
    std::swap(at(lo + (hi - lo) / 2), at(lo));
    const int64_t pivot = at(lo);
    int store = lo;
    for (int i = lo + 1; i <= hi; ++i)
      if (at(i) < pivot) std::swap(at(++store), at(i));
    std::swap(at(lo), at(store));
    stack.push_back({lo, store - 1});
    stack.push_back({store + 1, hi});
  }
}

// One safe event per LP over a sorted array: src/global_array_queue.cu:147-179 and
// src/heap_queue.cu:1560-1604 (the latter also stops past the root's maximum).
extern "C" int orc_safe_select_i64(const int64_t* ts, const int32_t* lp, int n, int num_lps,
                                   int64_t lookahead, int bound_valid, int64_t bound,
                                   uint8_t* marks) {
  std::vector<uint8_t> seen(num_lps, 0);
  int marked = 0;
  for (int i = 0; i < n; ++i) marks[i] = 0;
  for (int i = 0; i < n; ++i) {
    if (ts[i] >= ts[0] + lookahead) break;
    if (bound_valid && ts[i] > bound) break;
    if (marked == num_lps) break;
    if (!seen[lp[i]]) { seen[lp[i]] = 1; marks[i] = 1; ++marked; }
  }
  return marked;
}

// ------------------------------------------------------------------------------
// Grids and shortest paths.
// ------------------------------------------------------------------------------

// Cell c is an obstacle iff word (c & 3) of Philox({c >> 2, query, stream_tag, 2},
// {seed, 0}) is below `threshold` (= density * 2^32).
extern "C" void orc_grid_bernoulli(uint8_t* blocked, int w, int h, uint32_t seed, uint32_t stream_tag,
                                   uint32_t query, uint32_t threshold) {
  const uint64_t cells = (uint64_t)w * h;
  const uint32_t key[2] = {seed, 0u};
  for (uint64_t base = 0; base < cells; base += 4) {
    const uint32_t ctr[4] = {(uint32_t)(base >> 2), query, stream_tag, 2u};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    for (int j = 0; j < 4 && base + j < cells; ++j) blocked[base + j] = r[j] < threshold ? 1 : 0;
  }
}

extern "C" void orc_batch_query_shape(uint32_t seed, uint32_t q, int32_t* side, uint32_t* threshold) {
  const uint32_t key[2] = {seed, 0u};
  const uint32_t ctr[4] = {q, 0u, 3u, 0u};
  uint32_t r[4];
  orc_philox4x32_10(ctr, key, r);
  const double u0 = (double)r[0] * 0x1p-32, u1 = (double)r[1] * 0x1p-32;
  *side = (int32_t)std::floor(256.0 * std::exp2(3.0 * u0) + 0.5);
  const double density = 0.1 + 0.2 * u1;
  *threshold = (uint32_t)(density * 4294967296.0);
}

// Source and target: first attempt a = 0, 1, ... of Philox({a, q, 4 | 5, 0}) whose
// cell ((r0*side)>>32, (r1*side)>>32) is free.
extern "C" void orc_batch_query_spec(uint32_t seed, uint32_t q, orc_query_spec* spec, const uint8_t* blocked) {
  orc_batch_query_shape(seed, q, &spec->side, &spec->threshold);
  const uint32_t key[2] = {seed, 0u};
  const int side = spec->side;
  for (int which = 0; which < 2; ++which) {
    for (uint32_t a = 0;; ++a) {
      const uint32_t ctr[4] = {a, q, 4u + (uint32_t)which, 0u};
      uint32_t r[4];
      orc_philox4x32_10(ctr, key, r);
      const int x = (int)(((uint64_t)r[0] * side) >> 32), y = (int)(((uint64_t)r[1] * side) >> 32);
      if (!blocked[(size_t)y * side + x]) {
        if (which == 0) { spec->sx = x; spec->sy = y; } else { spec->tx = x; spec->ty = y; }
        break;
      }
    }
  }
}

namespace {
struct Nbr { int dx, dy; uint32_t cost; };
const Nbr kN4[4] = {{1, 0, 1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, 1}};
const Nbr kN8[8] = {{1, 0, 10}, {-1, 0, 10}, {0, 1, 10}, {0, -1, 10},
                    {1, 1, 14}, {-1, 1, 14}, {1, -1, 14}, {-1, -1, 14}};

inline uint32_t heur(int conn, int x, int y, int tx, int ty) {
  const uint32_t dx = (uint32_t)std::abs(x - tx), dy = (uint32_t)std::abs(y - ty);
  if (conn == 4) return dx + dy;
  return 10u * std::max(dx, dy) + 4u * std::min(dx, dy);
}

int64_t best_first(const uint8_t* blocked, int w, int h, int conn, int sx, int sy, int tx, int ty,
                   bool use_h, uint64_t* expanded) {
  if (expanded) *expanded = 0;
  if (sx < 0 || sy < 0 || tx < 0 || ty < 0 || sx >= w || tx >= w || sy >= h || ty >= h) return -1;
  if (blocked[(size_t)sy * w + sx] || blocked[(size_t)ty * w + tx]) return -1;
  const size_t cells = (size_t)w * h;
  std::vector<uint32_t> g(cells, 0xFFFFFFFFu);
  // key: f (high 32) | h (low 32) -- ties toward the goal; cell carried alongside.
  typedef std::pair<uint64_t, uint32_t> Item;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> open;
  const uint32_t s = (uint32_t)(sy * w + sx), t = (uint32_t)(ty * w + tx);
  g[s] = 0;
  const uint32_t hs = use_h ? heur(conn, sx, sy, tx, ty) : 0u;
  open.push({((uint64_t)hs << 32) | hs, s});
  const Nbr* nb = conn == 4 ? kN4 : kN8;
  const int nn = conn == 4 ? 4 : 8;
  uint64_t exp = 0;
  while (!open.empty()) {
    const Item it = open.top();
    open.pop();
    const uint32_t u = it.second;
    const uint32_t f = (uint32_t)(it.first >> 32), hu = (uint32_t)it.first;
    if (f - hu != g[u]) continue;          // stale entry
    if (u == t) { if (expanded) *expanded = exp; return g[u]; }
    ++exp;
    const int ux = (int)(u % w), uy = (int)(u / w);
    for (int i = 0; i < nn; ++i) {
      const int vx = ux + nb[i].dx, vy = uy + nb[i].dy;
      if (vx < 0 || vy < 0 || vx >= w || vy >= h) continue;
      const uint32_t v = (uint32_t)(vy * w + vx);
      if (blocked[v]) continue;
      const uint32_t gv = g[u] + nb[i].cost;
      if (gv < g[v]) {
        g[v] = gv;
        const uint32_t hv = use_h ? heur(conn, vx, vy, tx, ty) : 0u;
        open.push({((uint64_t)(gv + hv) << 32) | hv, v});
      }
    }
  }
  if (expanded) *expanded = exp;
  return -1;
}
}  // namespace

// Exact integer A* (the reference's expansion, src/a_star.cu:269-593, keeps f in fp16 and
// pops up to 8 entries per step, so it is not guaranteed optimal; this restatement returns the optimum).
extern "C" int64_t orc_astar_grid(const uint8_t* blocked, int w, int h, int conn, int sx, int sy,
                                  int tx, int ty, uint64_t* expanded) {
  return best_first(blocked, w, h, conn, sx, sy, tx, ty, true, expanded);
}

extern "C" int64_t orc_dijkstra_grid(const uint8_t* blocked, int w, int h, int conn, int sx, int sy,
                                     int tx, int ty) {
  return best_first(blocked, w, h, conn, sx, sy, tx, ty, false, nullptr);
}

extern "C" int64_t orc_bfs_grid4(const uint8_t* blocked, int w, int h, int sx, int sy, int tx, int ty) {
  if (blocked[(size_t)sy * w + sx] || blocked[(size_t)ty * w + tx]) return -1;
  std::vector<int32_t> dist((size_t)w * h, -1);
  std::vector<uint32_t> frontier{(uint32_t)(sy * w + sx)};
  dist[(size_t)sy * w + sx] = 0;
  const uint32_t t = (uint32_t)(ty * w + tx);
  size_t head = 0;
  while (head < frontier.size()) {
    const uint32_t u = frontier[head++];
    if (u == t) return dist[u];
    const int ux = (int)(u % w), uy = (int)(u / w);
    for (const Nbr& n : kN4) {
      const int vx = ux + n.dx, vy = uy + n.dy;
      if (vx < 0 || vy < 0 || vx >= w || vy >= h) continue;
      const uint32_t v = (uint32_t)(vy * w + vx);
      if (blocked[v] || dist[v] >= 0) continue;
      dist[v] = dist[u] + 1;
      frontier.push_back(v);
    }
  }
  return -1;
}
