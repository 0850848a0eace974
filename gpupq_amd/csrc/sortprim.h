// sortprim.h -- record sort / merge / select primitives for the global queues (GAQ, HQ),
// replacing the reference's moderngpu mergesort (lib/moderngpu/src/moderngpu/
// kernel_mergesort.hxx:27-148) and cub DeviceReduce::Min.  All sizes that depend on the
// simulation live in device memory, so the launch sequence is fixed and graph-capturable;
// grids are sized from capacities and surplus blocks exit.
#pragma once

#include "gpq_host.h"

namespace gpq {

// Tiles of the sort and merge kernels: 1024, 2048 (the default) or 4096 records per block
// (128, 256 or 512 threads x 8 records), chosen per queue (gpq_queue_config.sort_tile).
constexpr uint32_t kMergeTile = 2048;   // default outputs per merge block
constexpr uint32_t kSortTile = 2048;    // default records per tile-sort block
bool valid_sort_tile(uint32_t tile);
// Merge passes of a sort of up to cap records with tiles of `tile` records.
uint32_t sort_passes(uint32_t cap, uint32_t tile);

// Stable merge of A = a1[0:n_a1) ++ a2[0:n_a2) (both parts sorted, a1 before a2) with
// sorted B (ties take A first).  Output rank o goes to out[s][o - start_s], where segment
// s covers ranks [end[s-1], end[s]) (end[-1] = 0, unused segments have end = total).
struct MergeTask {
  const Event* a1;
  const Event* a2;
  const Event* b;
  Event* out[3];
  uint32_t n_a1;
  uint32_t n_a2;
  uint32_t n_b;
  uint32_t end[3];
};

// Sorts *d_n (<= cap) records of `in` (not modified) into buf0/buf1; returns the buffer
// holding the sorted run.  Passes: one tile sort, then ceil(log2(cap / tile)) merge passes.
// `acct` (optional): a device counter that receives the records sorted (*d_n), for the
// algorithmic-byte accounting of queues created with collect_stats.
// `seg_pre` (optional): `in` is a striped insert area (core.cuh) with seg_cap-record stripes,
// its segments' records laid out by gather_inserts (*d_n = their total).
gpq_status sort_records(const Event* in, Event* buf0, Event* buf1, const uint32_t* d_n, uint32_t cap,
                        hipStream_t stream, Event** sorted, uint32_t tile = kSortTile,
                        unsigned long long* acct = nullptr, const uint32_t* seg_pre = nullptr, uint32_t seg_cap = 0);

// Records per stripe of an insert area for `enq_cap` inserts per step: twice an even share, at
// least 1024 (a stripe that fills up spills into the shared segment of enq_cap records).
inline uint32_t insert_stripe_cap(uint64_t enq_cap) {
  const uint64_t c = 2 * ((enq_cap + kInsertStripes - 1) / kInsertStripes);
  return (uint32_t)(c < 1024 ? 1024 : c);
}

// The step's inserts in a striped insert area (counters seg_m[k * kStripeWords], k <=
// kInsertStripes; stripes of seg_cap records, spill segment of spill_cap): seg_pre[k] = records
// in the segments before k (kInsertStripes + 2 words), *d_m = their total, capped at max_total
// (beyond it the queue's error flag is set and the last segments are cut short); the counters are
// reset to zero.
gpq_status gather_inserts(uint32_t* seg_m, uint32_t seg_cap, uint32_t spill_cap, uint32_t max_total, uint32_t* seg_pre,
                          uint32_t* d_m, WindowState* ws, hipStream_t stream);

// Runs `ntasks` merge tasks (device array) in one launch; max_total bounds every task's
// n_a1 + n_a2 + n_b.  `acct` (optional) receives the records merged (the tasks' totals).
// `splits` (optional, ntasks * stride u32 with stride >= merge_split_words(max_total, tile)):
// the merge-path splits of every tile boundary are computed by a partition launch first (one
// wave per boundary), instead of by each merge block before it can load its input.
// Merges whose max_total is below partition_min search their splits in the merge blocks even
// with `splits` given.
constexpr uint64_t kMergePartitionMin = 1ull << 20;
gpq_status merge_tasks(const MergeTask* d_tasks, uint32_t ntasks, uint64_t max_total, hipStream_t stream,
                       uint32_t tile = kMergeTile, unsigned long long* acct = nullptr, uint32_t* splits = nullptr,
                       uint32_t stride = 0, uint64_t partition_min = kMergePartitionMin);
uint32_t merge_split_words(uint64_t max_total, uint32_t tile);

// One safe event per LP over the sorted array s[0:*d_n), s = *d_s: the prefix of records with
// ts < *d_wend and key <= *d_bound (when *d_bound_valid) is scanned, the first record of
// each LP in it is moved to `deq` (in order, count -> *d_deq_n) and the others to `rest`
// (in order, count -> *d_rest_n); *d_prefix receives the prefix length.  `first` (one
// u32 per LP, all 0xFFFFFFFF between calls) and `scratch` (>= 2 * ceil(cap / 1024) + 2
// u32) are workspace; `acct` (optional) receives the prefix length.  Reference:
// src/global_array_queue.cu:147-179 (serial scan) and src/heap_queue.cu:478-547 (atomicMin
// per LP).
struct SelectBufs {
  uint32_t* first;
  uint32_t* scratch;
  uint32_t* d_prefix;
  uint32_t* d_deq_n;
  uint32_t* d_rest_n;
};
gpq_status select_safe(const Event* const* d_s, const uint32_t* d_n, uint32_t cap, const uint32_t* d_wend,
                       const Event* d_bound, const uint32_t* d_bound_valid, const SelectBufs& b, Event* deq,
                       Event* rest, hipStream_t stream, unsigned long long* acct = nullptr);

// Copies records: dst[i] = src[i] for i < *d_n (bounded by cap).
gpq_status copy_records(const Event* src, Event* dst, const uint32_t* d_n, uint32_t cap, hipStream_t stream);

}  // namespace gpq
