/*
 * oracle.h -- CPU restatement of the GPUPQ priority-queue hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpupq_amd/, include/)
 * includes, links or calls this code; only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it, and only
 * as the checker.
 *
 * What it restates (reference file:line, relative to the reference root):
 *   - PHOLD event handling           src/phold.cu:81-117   (one event -> one new event)
 *   - population                     src/phold.cu:63-79,124-163 (k initial events per LP here)
 *   - the DES loop + termination     src/des.cu:134-221    (commit exactly the events with ts < T)
 *   - safe-window rule               src/local_array_queue.cu:745-779,
 *                                    src/global_array_queue.cu:147-179,
 *                                    src/heap_queue.cu:478-547,1560-1604
 *   - LAQ ring primitives            src/sort.cu:17-26,48-66,75-131,329-380
 *   - A* expansion structure         src/a_star.cu:269-593 (restated as exact integer A*)
 *
 * Pinning (see tests/test_oracle.py):
 *   - Philox4x32-10: Random123 known-answer vectors and rocRAND's host-compiled
 *     engine (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h).
 *   - orc_exp_from_u32: exhaustive comparison with glibc log() over all 2^24 inputs.
 *   - orc_phold_seq: equals an independent brute-force linear-scan simulation and the
 *     windowed (one safe event per LP per step) simulation; conservation and
 *     per-LP monotone timestamps; event rate matches L*k/1.1 per time unit.
 *   - ring primitives / safe selection: the reference's known-answer tests
 *     (src/test_laq.cu, src/test_gaq.cu, src/test_hq.cu) restated as fixtures.
 *   - A*: equals Dijkstra and BFS; closed forms on obstacle-free grids.
 */
#ifndef GPUPQ_ORACLE_H
#define GPUPQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers -------------------------------------------------- */
void  orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Exp(1) sample from one 32-bit draw: u = ((x>>8)+1) * 2^-24 in (0,1], returns (float)(-ln u). */
float orc_exp_from_u32(uint32_t x);
/* out[i] = orc_exp_from_u32((first + i) << 8) for i < count: one value per distinct u. */
void  orc_exp_table(uint32_t first, uint32_t count, float* out);

/* ---- PHOLD ------------------------------------------------------------ */
typedef struct {
  uint32_t num_lps;          /* L */
  uint32_t init_per_lp;      /* k initial events per LP */
  uint32_t seed;             /* Philox key = {seed, 0} */
  float    end_time;         /* commit exactly the events with ts < end_time */
  float    lookahead;        /* 0.1f */
  uint32_t p_self_u32;       /* self-destination iff draw[1] < p_self_u32 (0x80000000 = 0.5) */
} orc_phold_cfg;

typedef struct {
  uint64_t events;           /* committed events */
  uint64_t trace_hash;       /* orc_phold_combine over the per-LP arrays */
  uint64_t steps;            /* windowed variant only: number of windows */
  double   seconds;          /* wall clock of the event loop only */
  uint32_t max_queue_len;    /* windowed variant only: max per-LP pending count */
  uint32_t in_flight_end;    /* events pending at the end (conservation: == L*k) */
  uint32_t monotone_ok;      /* 1 if every LP saw non-decreasing timestamps */
} orc_phold_result;

/* Sequential std::priority_queue simulation over the total order (ts, src, seq). */
int orc_phold_seq(const orc_phold_cfg* cfg, uint32_t* lp_count, uint64_t* lp_hash, orc_phold_result* res);
/* Independent brute force: per-LP unsorted bags + global linear scan for the minimum. Tiny inputs only. */
int orc_phold_brute(const orc_phold_cfg* cfg, uint32_t* lp_count, uint64_t* lp_hash, orc_phold_result* res);
/* Windowed: per step, every LP whose head has ts < min(gmin + lookahead, T) processes exactly its head. */
int orc_phold_windowed(const orc_phold_cfg* cfg, uint32_t* lp_count, uint64_t* lp_hash, orc_phold_result* res);
/* Windowed drain on `threads` host threads (LPs split in contiguous ranges): the all-cores CPU
 * baseline; commits the same trace as orc_phold_seq. */
int orc_phold_parallel(const orc_phold_cfg* cfg, uint32_t threads, uint32_t* lp_count, uint64_t* lp_hash,
                       orc_phold_result* res);
uint64_t orc_phold_combine(const uint32_t* lp_count, const uint64_t* lp_hash, uint32_t num_lps);
/* Per-LP chained trace hash step for one committed event (ts bits, src LP, seq). */
uint64_t orc_trace_step(uint64_t h, uint32_t ts_bits, uint32_t src, uint32_t seq);

/* ---- LAQ ring primitives (keys are int64 as in the reference's `long ts`) --------- */
int  orc_ring_wrap(int i, int cap);                                            /* sort.cu:17-26 */
int  orc_ring_search(const int64_t* ring, int cap, int read, int write, int64_t key); /* sort.cu:48-66 */
void orc_ring_insert_sorted(int64_t* ring, int cap, int read, int write, int64_t key); /* sort.cu:329-356 */
void orc_ring_quicksort(int64_t* ring, int cap, int left, int right);          /* sort.cu:75-131 */

/* One safe event per LP over a key-sorted array (GAQ mark_safe_events / HQ partition).
 * ts[] sorted ascending, lp[] owner of each item; marks[i]=1 for the first item of each LP
 * with ts < ts[0] + lookahead and (bound_valid ? ts <= bound : 1).  Returns #marked. */
int orc_safe_select_i64(const int64_t* ts, const int32_t* lp, int n, int num_lps,
                        int64_t lookahead, int bound_valid, int64_t bound, uint8_t* marks);

/* ---- A* / shortest paths on grids ---------------------------------------------- */
/* blocked[y*w+x] != 0 is an obstacle.  conn 4: unit costs, Manhattan h.
 * conn 8: straight 10, diagonal 14 (corner cutting allowed, as in src/a_star.cu:393-459),
 * octile h = 10*max + 4*min.  Returns the optimal cost, or -1 if unreachable. */
int64_t orc_astar_grid(const uint8_t* blocked, int w, int h, int conn,
                       int sx, int sy, int tx, int ty, uint64_t* expanded);
int64_t orc_dijkstra_grid(const uint8_t* blocked, int w, int h, int conn,
                          int sx, int sy, int tx, int ty);
int64_t orc_bfs_grid4(const uint8_t* blocked, int w, int h, int sx, int sy, int tx, int ty);

/* Synthetic grids (the same definition the product's generator implements). */
void orc_grid_bernoulli(uint8_t* blocked, int w, int h, uint32_t seed, uint32_t stream_tag,
                        uint32_t query, uint32_t threshold);
typedef struct { int32_t side; uint32_t threshold; int32_t sx, sy, tx, ty; } orc_query_spec;
/* Batched-A* query q: side = round(256 * 2^(3u)), density U[0.1,0.3], source/target free. */
void orc_batch_query_spec(uint32_t seed, uint32_t q, orc_query_spec* spec, const uint8_t* blocked);
void orc_batch_query_shape(uint32_t seed, uint32_t q, int32_t* side, uint32_t* threshold);

#ifdef __cplusplus
}
#endif
#endif
