// gpupq_device/core.cuh -- device-side vocabulary shared by every queue view and by model
// kernels written against them: the event record and its total order, the counter-based
// random stream and Exp(1) sampler of the PHOLD workloads, the trace hash, the wave64
// helpers, the window state a queue publishes, and the staging insert of the per-LP queues.
// Header-only HIP for gfx950.
//
// Reference counterparts: src/queue_item.h:11-32 (event, operator<),
// src/phold.cu:29-34 (rand_exponential), src/util.cuh:13-16 (get_lp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpq {

constexpr int kWave = 64;
constexpr uint32_t kInfBits = 0x7F800000u;   // +inf as float bits: "no event"

// 16-byte event record.  ts holds float32 bits; every timestamp is >= 0, so the
// unsigned order of the bits is the float order.  dst names the LP that owns the
// event (the reference's `node`, one node per LP); (src, seq) is unique per event.
struct __attribute__((aligned(16))) Event {
  uint32_t ts;
  uint32_t dst;
  uint32_t src;
  uint32_t seq;
};

// Total order (ts, src, seq); the reference compares ts only (src/queue_item.h:22-28).
__device__ __forceinline__ bool key_less(const Event& a, const Event& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}

__device__ __forceinline__ Event load_event(const Event* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return Event{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void store_event(Event* p, const Event& e) {
  *reinterpret_cast<uint4*>(p) = make_uint4(e.ts, e.dst, e.src, e.seq);
}

__device__ __forceinline__ float bits_to_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f_to_bits(float f) { return __float_as_uint(f); }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64 products,
// Weyl key bumps between rounds.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    if (round) { k.x += 0x9E3779B9u; k.y += 0xBB67AE85u; }
    const uint64_t pa = (uint64_t)0xD2511F53u * c.x;
    const uint64_t pb = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo_a = (uint32_t)pa, hi_a = (uint32_t)(pa >> 32);
    const uint32_t lo_b = (uint32_t)pb, hi_b = (uint32_t)(pb >> 32);
    c = make_uint4(hi_b ^ c.y ^ k.x, lo_b, hi_a ^ c.w ^ k.y, lo_a);
  }
  return c;
}

// Exp(1) from one 32-bit draw: u = ((x>>8)+1) * 2^-24 in (0, 1], returns (float)(-ln u).
// ln of the integer n = m * 2^k, m in [sqrt(1/2), sqrt(2)), by the atanh series
// 2(s + s^3/3 + ... + s^13/13), s = (m-1)/(m+1), in IEEE double with a fixed
// operation order (the library is compiled with -ffp-contract=off), so the value
// is bit-identical on every IEEE-754 machine.
__device__ __forceinline__ float exp_from_u32(uint32_t x) {
  const uint32_t n = (x >> 8) + 1u;
  int k = 31 - __clz(n);
  const double scale = __longlong_as_double((long long)((uint64_t)(1023 - k) << 52));
  double m = (double)n * scale;
  if (m >= 0x1.6a09e667f3bcdp+0) { m = m * 0.5; k += 1; }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 0x1.3b13b13b13b14p-3;
  p = 0x1.745d1745d1746p-3 + z * p;
  p = 0x1.c71c71c71c71cp-3 + z * p;
  p = 0x1.2492492492492p-2 + z * p;
  p = 0x1.999999999999ap-2 + z * p;
  p = 0x1.5555555555555p-1 + z * p;
  p = 2.0 + z * p;
  const double ln_m = s * p;
  const double r = (double)(24 - k) * 0x1.62e42fefa39efp-1 - ln_m;
  return (float)r;
}

// ---------------------------------------------------------------------------
// Trace hash: per-LP chain over committed (ts bits, src, seq), splitmix64 mixing.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27; x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}
__device__ __forceinline__ uint64_t trace_step(uint64_t h, uint32_t ts, uint32_t src, uint32_t seq) {
  h = mix64(h ^ (((uint64_t)ts << 32) | src));
  h = mix64(h ^ ((uint64_t)seq + 0x9E3779B97F4A7C15ull));
  return h;
}

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, kWave);
  return v;
}

// Adds v (summed over the wave) to a work counter; every lane of the wave must call it.
__device__ __forceinline__ void stats_add(unsigned long long* stats, int slot, uint32_t v) {
  const uint32_t s = wave_sum_u32(v);
  if (lane_id() == 0 && s) atomicAdd(&stats[slot], (unsigned long long)s);
}

__device__ __forceinline__ Event shfl_event(const Event& e, int src_lane) {
  return Event{(uint32_t)__shfl((int)e.ts, src_lane, kWave), (uint32_t)__shfl((int)e.dst, src_lane, kWave),
               (uint32_t)__shfl((int)e.src, src_lane, kWave), (uint32_t)__shfl((int)e.seq, src_lane, kWave)};
}

// Device-resident window state shared by a queue and the model driving it.
// gmin: minimum pending timestamp (bits); wend: exclusive end of the safe window,
// min(gmin + lookahead, horizon); done: nonzero once nothing is left below the horizon.
struct WindowState {
  uint32_t gmin;
  uint32_t wend;
  uint32_t done;
  uint32_t near_end;         // LAQ two-tier: ts bits of the near horizon (events below it are in the rings)
  uint64_t windows;          // number of windows computed
  uint32_t error;            // GPQ_FLAG_* bits
  uint32_t near_next;        // LAQ two-tier: nonzero = the next post moves the near horizon here
  uint32_t sel;              // PHOLD merge-and-commit step: staging buffer read by the next step (0 main, 1 alt)
};

enum : uint32_t {
  kErrStageOverflow = 1u,    // more inserts for one LP in one step than the staging area holds
  kErrQueueOverflow = 2u,    // an LP's queue exceeded its capacity
  kErrHeapOverflow = 4u,     // the parallel heap ran out of nodes
  kErrBadState = 8u,
};

// queue_insert for the per-LP queues: the event is staged for its owning LP and folded
// in by queue_post; false (plus a device flag) when the LP's staging area is full.
__device__ __forceinline__ bool stage_insert(Event* stage, uint32_t* ins, uint32_t stage_cap, WindowState* ws,
                                             const Event& e) {
  const uint32_t pos = atomicAdd(&ins[e.dst], 1u);
  if (pos >= stage_cap) {
    atomicOr(&ws->error, kErrStageOverflow);
    return false;
  }
  store_event(&stage[(size_t)e.dst * stage_cap + pos], e);
  return true;
}

// Reserves one slot of a shared append counter per calling lane with one atomic per wave
// (every active lane of the wave that reaches the call takes part).
__device__ __forceinline__ uint32_t wave_append_slot(uint32_t* counter) {
  const uint64_t mask = __ballot(1);
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, leader, kWave);
  return base + rank;
}

// Striped insert area of a global queue (GAQ, one HQ heap): kInsertStripes segments of seg_cap
// records, then a spill segment of spill_cap records; counter k at seg_m[k * kStripeWords],
// k = 0 .. kInsertStripes, 64 bytes apart (kStripeCounterWords words in all): packed into two
// lines they took every wave's append atomic (GAQ 6.85e8 against 7.03e8 events/s at 2^20 LPs).  A
// workgroup appends to stripe blockIdx.x mod kInsertStripes (one counter atomic per wave), so
// that waves of different workgroups do not all contend for one counter; a lane whose stripe is
// full appends to the spill segment.  queue_post reads the stripes and then the spill segment,
// in that order, as one run of inserts.  All active lanes of the wave must take part.
constexpr uint32_t kInsertStripes = 32;
constexpr uint32_t kStripeWords = 16;
constexpr uint32_t kStripeCounterWords = (kInsertStripes + 1u) * kStripeWords;

__device__ __forceinline__ bool striped_append(Event* area, uint32_t* seg_m, uint32_t seg_cap, uint32_t spill_cap,
                                               const Event& e) {
  const uint32_t s = blockIdx.x & (kInsertStripes - 1u);
  const uint32_t pos = wave_append_slot(&seg_m[s * kStripeWords]);
  if (pos < seg_cap) {
    store_event(area + (size_t)s * seg_cap + pos, e);
    return true;
  }
  const uint32_t sp = wave_append_slot(&seg_m[kInsertStripes * kStripeWords]);   // the lanes whose stripe is full
  if (sp < spill_cap) {
    store_event(area + (size_t)kInsertStripes * seg_cap + sp, e);
    return true;
  }
  return false;
}

}  // namespace gpq
