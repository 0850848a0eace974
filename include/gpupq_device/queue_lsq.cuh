// queue_lsq.cuh -- device API of the local splay queue (LSQ): one top-down splay tree
// per LP (Sleator and Tarjan, "Self-adjusting binary search trees", JACM 32(3), 1985,
// simple top-down splay), owned by one thread.  Reference: src/local_splay_queue.cu
// (splay :108-156, insert :158-202, delete_min :233-273, peek :337-360, set_done
// :362-369, pool malloc_/free_ :72-104).
//
// Nodes live in a per-LP pool of `cap` slots addressed by 32-bit indices (the
// reference uses 8-byte pointers and finds a free slot by a linear scan marked by
// ts == -1): item[lp*cap + i] (16 B) and link[lp*cap + i] = {left, right} (8 B);
// free slots form a list threaded through `left`, head in freelist[lp].
#pragma once

#include "core.cuh"

namespace gpq {

constexpr uint32_t kNil = 0xFFFFFFFFu;

struct LsqView {
  Event* item;
  uint2* link;
  Event* stage;
  uint32_t* root;
  uint32_t* freelist;
  uint32_t* len;
  uint32_t* ins;
  uint32_t* head;
  WindowState* ws;
  unsigned long long* stats;
  uint32_t num_lps;
  uint32_t cap;
  uint32_t stage_cap;

  __device__ __forceinline__ uint32_t slots() const { return num_lps; }
  __device__ __forceinline__ bool insert(const Event& e) const { return stage_insert(stage, ins, stage_cap, ws, e); }

  // Per-LP tree operations; node indices are LP-local.
  struct Tree {
    Event* it;
    uint2* ln;
    uint32_t visits;
    __device__ __forceinline__ uint32_t L(uint32_t t) const { return ln[t].x; }
    __device__ __forceinline__ uint32_t R(uint32_t t) const { return ln[t].y; }
    __device__ __forceinline__ void setL(uint32_t t, uint32_t v) { ln[t].x = v; }
    __device__ __forceinline__ void setR(uint32_t t, uint32_t v) { ln[t].y = v; }
    __device__ __forceinline__ Event key(uint32_t t) const { return load_event(&it[t]); }

    // Simple top-down splay of key k in the tree rooted at t; ties descend left, as the
    // reference does (`i <= t->item`).  The header node is kept in registers.
    __device__ uint32_t splay(const Event& k, uint32_t t) {
      if (t == kNil) return t;
      uint32_t hl = kNil, hr = kNil;   // header's left / right
      uint32_t l = kNil, r = kNil;     // kNil here means "the header"
      for (;;) {
        ++visits;
        if (!key_less(key(t), k)) {                  // k <= t
          uint32_t tl = L(t);
          if (tl == kNil) break;
          if (key_less(k, key(tl))) {                 // rotate right
            setL(t, R(tl));
            setR(tl, t);
            t = tl;
            tl = L(t);
            if (tl == kNil) break;
          }
          if (r == kNil) hl = t; else setL(r, t);     // link right
          r = t;
          t = tl;
        } else {
          uint32_t tr = R(t);
          if (tr == kNil) break;
          if (key_less(key(tr), k)) {                 // rotate left
            setR(t, L(tr));
            setL(tr, t);
            t = tr;
            tr = R(t);
            if (tr == kNil) break;
          }
          if (l == kNil) hr = t; else setR(l, t);     // link left
          l = t;
          t = tr;
        }
      }
      if (l == kNil) hr = L(t); else setR(l, L(t));   // assemble
      if (r == kNil) hl = R(t); else setL(r, R(t));
      setL(t, hr);
      setR(t, hl);
      return t;
    }

    __device__ uint32_t leftmost(uint32_t t) {
      if (t == kNil) return t;
      while (L(t) != kNil) { t = L(t); ++visits; }
      return t;
    }
  };

  __device__ __forceinline__ Tree tree(uint32_t lp) const {
    return Tree{item + (size_t)lp * cap, link + (size_t)lp * cap, 0u};
  }

  // queue_peek: splays LP lp's minimum to the root (src/local_splay_queue.cu:337-360).
  __device__ __forceinline__ int peek(Event** out, int lp) const {
    if (!(head[lp] < ws->wend)) return -1;
    Tree t = tree((uint32_t)lp);
    const uint32_t rt = root[lp];
    const uint32_t m = t.leftmost(rt);
    const uint32_t nr = t.splay(t.key(m), rt);
    root[lp] = nr;
    *out = item + (size_t)lp * cap + nr;
    return (int)((uint32_t)lp * cap + nr);
  }

  // queue_set_done: deletes the root placed there by peek (delete_min, src/local_splay_queue.cu:233-273).
  __device__ __forceinline__ void set_done(int index) const {
    const uint32_t lp = (uint32_t)index / cap;
    const uint32_t node = (uint32_t)index % cap;
    Tree t = tree(lp);
    uint32_t x;
    if (t.L(node) == kNil) {
      x = t.R(node);
    } else {                                           // equal keys to its left
      x = t.splay(t.key(node), t.L(node));
      t.setR(x, t.R(node));
    }
    t.setL(node, freelist[lp]);                        // free_ (src/local_splay_queue.cu:96-104)
    freelist[lp] = node;
    root[lp] = x;
    len[lp] -= 1u;
    const uint32_t m = t.leftmost(x);
    head[lp] = m == kNil ? kInfBits : t.key(m).ts;
  }

  // queue_clear: the whole pool of the LP becomes free again (src/local_splay_queue.cu:469-494).
  __device__ __forceinline__ void clear(int lp) const {
    uint2* ln = link + (size_t)lp * cap;
    for (uint32_t i = 0; i < cap; ++i) ln[i] = make_uint2(i + 1u < cap ? i + 1u : kNil, kNil);
    root[lp] = kNil;
    freelist[lp] = 0;
    len[lp] = 0;
    head[lp] = kInfBits;
  }
  __device__ __forceinline__ bool is_empty(int lp) const { return len[lp] == 0 && ins[lp] == 0; }
  __device__ __forceinline__ int length(int lp) const { return (int)len[lp]; }
  __device__ __forceinline__ void insert_or_update(const Event& e, int) const { insert(e); }
};

}  // namespace gpq
