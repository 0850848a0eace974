// phold_fused.cuh -- the PHOLD merge-and-commit step on the two-tier local array queue,
// shared by the single-instance model (phold.hip) and the pool of instances (phold_batch.hip).
#pragma once

#include "laq_merge.cuh"

namespace gpq {

// ---------------------------------------------------------------------------------
// Merge-and-commit step on the two-tier local array queue (drain mode 2).
//
// One launch per step does queue_post and the model step together.  A group of G lanes per
// LP merges the events staged for it in the previous step (and, when the near horizon moves,
// its far events below the new horizon) into the sorted ring and takes the records below the
// window end off the front of the run; after a block barrier one thread per LP commits them
// in key order from the ring slots just written (in L2), and the events they schedule are
// staged in the other staging buffer.  Window k ends at w_k = min(w_{k-1} + lookahead, end
// time) in float arithmetic, which needs no minimum reduction: every pending event is
// >= w_{k-1}, so every event scheduled in step k is >= fl(w_{k-1} + lookahead) >= w_k.  Each
// LP commits its events in key order, so the trace is the one of the reference's step
// (src/phold.cu:81-117, src/des.cu:134-221) and of the other drive modes.
// ---------------------------------------------------------------------------------
constexpr int kFusedBlock = 256;   // LPs per block (the LAQ post block)

// The PHOLD instance a block of LPs belongs to: its window, its per-LP state (indexed by the
// queue's LP), and its parameters; the instance's LP i is the queue's LP base + i.
struct FusedModel {
  const WindowState* win;        // sel, wend, near_end, near_next, done of the instance
  uint32_t* count;
  uint64_t* hash;
  uint32_t* last_ts;             // nullptr: checks off
  uint32_t base, lps, init_per_lp, seed, p_self;
  float lookahead;
};

// The step for the 256 LPs of this block (all of one instance).  Kernels wrap it with the
// launch bounds: four blocks per CU (at five the compiler spills the merge state, measured
// slower).
template <int G, int SL>
__device__ __forceinline__ void fused_block(const LaqView& q, const FusedModel& m, uint32_t settle) {
  __shared__ uint32_t s_meta[kFusedBlock];     // by work slot: k | n << 7 | far count << 14, then d | first << 8
  __shared__ uint32_t s_rt[kFusedBlock];       // by work slot: rd | thread << 8 | head below the window end << 16
  __shared__ uint32_t s_cnt[kFusedBlock], s_last[kFusedBlock];
  __shared__ uint64_t s_hash[kFusedBlock];
  __shared__ uint32_t s_hist[kFusedBlock];
  __shared__ laq::Key3 s_near[kFusedBlock / G][laq::kFarList];   // merge input of each group
  __shared__ uint32_t s_count, s_inserted, s_moved;
  __shared__ unsigned long long s_aux;
  const WindowState* ws = m.win;
  if (!settle && ws->done) return;
  // settle: fold the last step's staged events in, commit nothing, keep the horizon
  const uint32_t sel = ws->sel;
  const uint32_t wend = settle ? 0u : ws->wend;
  const uint32_t moving = settle ? 0u : ws->near_next;
  const uint32_t N = moving ? moving : ws->near_end;
  Event* const stg_in = sel ? q.stage_alt : q.stage;
  uint32_t* const ins_in = sel ? q.ins_alt : q.ins;
  Event* const stg_out = sel ? q.stage : q.stage_alt;
  uint32_t* const ins_out = sel ? q.ins : q.ins_alt;
  const uint32_t lp = blockIdx.x * kFusedBlock + threadIdx.x;
  if (threadIdx.x == 0) {
    s_count = 0;
    s_inserted = 0;
    s_moved = 0;
    s_aux = 0;
  }
  __syncthreads();
  if (lp < q.num_lps && lp - m.base < m.lps) {
    // every word in one round trip (almost every LP is listed in a PHOLD step)
    const uint32_t staged = ins_in[lp];
    const uint32_t f = q.nfar[lp];
    const uint32_t drainable = q.head[lp] < wend ? 1u : 0u;
    const uint32_t len = q.len[lp], rd = q.rd[lp], cnt = m.count[lp];
    const uint64_t hash = m.hash[lp];
    const uint32_t last = m.last_ts ? m.last_ts[lp] : 0u;
    if (staged || (moving && f) || drainable) {
      const uint32_t slot = atomicAdd(&s_count, 1u);
      const uint32_t k = staged < q.stage_cap ? staged : q.stage_cap;
      s_meta[slot] = k | (len << 7) | (f << 14);
      s_rt[slot] = rd | (threadIdx.x << 8) | (drainable << 16);
      s_cnt[slot] = cnt;
      s_hash[slot] = hash;
      s_last[slot] = last;
      if (q.stats) {
        atomicAdd(&s_inserted, staged);
        if (moving && f) atomicAdd(&s_aux, 32ull * f);   // far records read and written once
      }
    }
  }
  __syncthreads();
  const int lane = lane_id();
  const int g = lane % G;
  const int gbase = lane - g;
  const int group = (int)threadIdx.x / G;
  const uint64_t gmask = (G == kWave ? ~0ull : ((1ull << G) - 1ull)) << gbase;
  laq::Key3* const near = &s_near[group][0];
  uint32_t* const hist = s_hist + group * G;
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  uint32_t moved = 0;
  for (uint32_t i = (uint32_t)group; i < s_count; i += kFusedBlock / G) {
    const uint32_t meta = s_meta[i], rt = s_rt[i];
    const uint32_t k = meta & 127u, f = meta >> 14;
    uint32_t n = (meta >> 7) & 127u, r = rt & 255u;
    const uint32_t lpt = blockIdx.x * kFusedBlock + ((rt >> 8) & 255u);
    uint32_t nn = 0, d = 0;
    laq::Run<SL> old;                             // the run's loads overlap the candidates' loads
    laq::load_run<G, SL>(q, lpt, n, r, g, old.k);
    if (k || (moving && f)) {
      uint32_t n_far, fmin;
      const uint32_t n_near = laq::split_candidates<G>(q, lpt, stg_in + (size_t)lpt * q.stage_cap, k, f, moving != 0u,
                                                       N, near, g, gmask, &n_far, &fmin);
      laq::wave_lds_sync();
      if (g == 0) {
        ins_in[lpt] = 0;
        q.nfar[lpt] = n_far < q.far_cap ? n_far : q.far_cap;
        q.far_min[lpt] = fmin;
      }
      nn = n_near < laq::kFarList ? n_near : laq::kFarList;
    }
    if (n + nn > (1u << q.cap_log2)) {
      if (g == 0) atomicOr(&q.ws->error, kErrQueueOverflow);
    } else if (nn || (rt >> 16)) {               // something to merge or to commit
      uint32_t done = 0;
      while (nn - done > (uint32_t)G - 1u) {
        if (done) {                               // the previous pass's ring writes, then the new run
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
          laq::load_run<G, SL>(q, lpt, n, r, g, old.k);
        }
        moved += laq::laq_merge_pass_inl<G, SL, false, laq::Key3, true>(q, lpt, near + done, (uint32_t)G - 1u, n, r, g,
                                                                    gbase, hist, 0u, nullptr, &old);
        done += (uint32_t)G - 1u;
      }
      if (done) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        laq::load_run<G, SL>(q, lpt, n, r, g, old.k);
      }
      moved += laq::laq_merge_pass_inl<G, SL, true, laq::Key3, true>(q, lpt, near + done, nn - done, n, r, g, gbase, hist,
                                                                 wend, &d, &old);
      if (g == 0) {
        q.rd[lpt] = r;
        q.len[lpt] = n;
      }
    }
    if (g == 0) s_meta[i] = d | (((r - d) & mask) << 8);
    laq::wave_lds_sync();                         // near[] and hist are free for the group's next LP
  }
  __syncthreads();                                // every merged run is in the ring
  // commits: one thread per listed LP, its d records from ring slot `first` on, in key order,
  // kCommitBatch at a time: the destinations depend only on the LP's event counter, so the
  // batch's staging-slot atomics are issued together and their latencies overlap
  constexpr uint32_t kCommitBatch = 4;
  if (threadIdx.x < s_count) {
    const uint32_t i = threadIdx.x;
    const uint32_t dm = s_meta[i], d = dm & 255u;
    if (d) {
      const uint32_t lpt = blockIdx.x * kFusedBlock + ((s_rt[i] >> 8) & 255u);
      const uint32_t local = lpt - m.base;        // the LP's index inside its instance
      const Event* ring = q.ring + ((size_t)lpt << q.cap_log2);
      const uint32_t first = dm >> 8;
      const uint32_t cnt = s_cnt[i];
      uint64_t h = s_hash[i];
      uint32_t last = s_last[i];
      for (uint32_t j0 = 0; j0 < d; j0 += kCommitBatch) {
        Event rec[kCommitBatch];
        uint32_t draw[kCommitBatch], dst[kCommitBatch], slot[kCommitBatch];
#pragma unroll
        for (uint32_t b = 0; b < kCommitBatch; ++b)
          if (j0 + b < d) rec[b] = load_event(&ring[(first + j0 + b) & mask]);
#pragma unroll
        for (uint32_t b = 0; b < kCommitBatch; ++b)
          if (j0 + b < d) {
            const uint4 rr = philox4x32_10(make_uint4(cnt + j0 + b, local, 0u, 0u), make_uint2(m.seed, 0u));
            draw[b] = rr.x;
            dst[b] = rr.y < m.p_self ? lpt : m.base + (uint32_t)(((uint64_t)rr.z * m.lps) >> 32);
            slot[b] = atomicAdd(&ins_out[dst[b]], 1u);
          }
#pragma unroll
        for (uint32_t b = 0; b < kCommitBatch; ++b)
          if (j0 + b < d) {
            const Event& cur = rec[b];
            h = trace_step(h, cur.ts, cur.src, cur.seq);
            if (m.last_ts) {
              if (cur.ts < last) atomicOr(&q.ws->error, kErrBadState);
              last = cur.ts;
            }
            const float t_new = (bits_to_f(cur.ts) + m.lookahead) + exp_from_u32(draw[b]);
            if (slot[b] < q.stage_cap)
              store_event(&stg_out[(size_t)dst[b] * q.stage_cap + slot[b]],
                          Event{f_to_bits(t_new), dst[b], local, m.init_per_lp + cnt + j0 + b});
            else
              atomicOr(&q.ws->error, kErrStageOverflow);
          }
      }
      m.count[lpt] = cnt + d;
      m.hash[lpt] = h;
      if (m.last_ts) m.last_ts[lpt] = last;
    }
  }
  if (q.stats) {
    if (g == 0 && moved) atomicAdd(&s_moved, moved);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t lps_here = min((uint32_t)kFusedBlock, q.num_lps - blockIdx.x * kFusedBlock);
      atomicAdd(&q.stats[0], (unsigned long long)s_inserted);
      atomicAdd(&q.stats[1], (unsigned long long)s_moved);
      atomicAdd(&q.stats[2], (unsigned long long)s_count);
      atomicAdd(&q.stats[3], (unsigned long long)lps_here);
      if (s_aux) atomicAdd(&q.stats[4], s_aux);
    }
  }
}

// queue_post after a merge-and-commit step, for one instance's window: the window advances
// by the lookahead (capped at the end time); a window that reached the end time ends the run;
// a window beyond the near horizon schedules the horizon's move (by `step`) for the next
// step; the staging buffers swap.
__device__ __forceinline__ void advance_window(WindowState* ws, float lookahead, float end_time, float step) {
  WindowState w = *ws;
  if (w.done) return;
  if (w.near_next) {
    w.near_end = w.near_next;
    w.near_next = 0;
  }
  const float we = bits_to_f(w.wend);
  w.windows += 1;
  w.sel ^= 1u;
  w.gmin = w.wend;                                // a lower bound of the pending minimum
  if (!(we < end_time)) {
    w.done = 1;
  } else {
    float wn = we + lookahead;
    if (end_time < wn) wn = end_time;
    const float nf = bits_to_f(w.near_end);
    if (wn > nf) {
      float nx = nf + step;
      if (!(nx >= wn)) nx = wn + step;
      w.near_next = f_to_bits(nx);
    }
    w.wend = f_to_bits(wn);
  }
  *ws = w;
}

}  // namespace gpq
