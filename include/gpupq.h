/*
 * gpupq.h -- C ABI of the MI355X (gfx950) priority-queue library.
 *
 * Plain C types only.  Every entry point returns a gpq_status; nothing exits or
 * aborts.  Handles own all of their device memory (allocated once at create,
 * released at destroy); no entry point allocates inside a simulation step.
 * A handle is not re-entrant; different handles may be used from different
 * host threads.  Device work of a handle is ordered on that handle's HIP stream.
 *
 * The reference binds one queue and one model at compile time through macro
 * renaming (src/queue.h:33-84, src/queue.cuh:7-40, src/config.h:8-21).  Here the
 * queue kind and every size are runtime fields of gpq_queue_config, and each
 * reference entry point maps to one function below (cited per function).
 * The device-side verbs of src/queue.cuh:46-70 (insert / peek / set_done /
 * is_empty / length / clear / insert_or_update) are the header-only device API
 * of include/gpupq_device.h, compiled into the caller's model kernels; a queue's
 * view for them comes from gpq_queue_device_view.
 */
#ifndef GPUPQ_H
#define GPUPQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GPQ_OK = 0,
  GPQ_ERR_INVALID = 1,     /* bad argument or configuration */
  GPQ_ERR_CAPACITY = 2,    /* a queue, staging area or heap overflowed (device flag) */
  GPQ_ERR_HIP = 3,         /* a HIP runtime call failed */
  GPQ_ERR_NOMEM = 4,       /* device allocation failed or exceeds the HBM budget */
  GPQ_ERR_STATE = 5,       /* call out of order, or a runtime invariant was violated */
  GPQ_ERR_UNSUPPORTED = 6  /* combination not provided (e.g. queue kind x model) */
} gpq_status;

typedef enum {
  GPQ_LAQ = 0,   /* per-LP sorted ring ("local array queue"), src/local_array_queue.cu */
  GPQ_LSQ = 1,   /* per-LP splay tree,                        src/local_splay_queue.cu */
  GPQ_LHQ = 2,   /* per-LP binary heap,                       src/local_heap_queue.cu */
  GPQ_HQ  = 3,   /* global parallel heap with r-item nodes,   src/heap_queue.cu */
  GPQ_GAQ = 4    /* global sorted array,                      src/global_array_queue.cu */
} gpq_queue_kind;

/* 16-byte PHOLD event: ts holds the float32 bits of a timestamp >= 0.
 * Order: (ts, src, seq).  Reference: struct event {long ts; int node;} (src/queue_item.h:13-16). */
typedef struct {
  uint32_t ts;
  uint32_t dst;   /* owning LP (the reference's `node`, one node per LP) */
  uint32_t src;   /* LP that generated the event */
  uint32_t seq;   /* per-source sequence number */
} gpq_event;

typedef struct {
  gpq_queue_kind kind;
  uint32_t num_lps;          /* NUM_LPS (src/config.h:68) */
  uint32_t capacity_per_lp;  /* LAQ/LSQ/LHQ: items per LP (LAQ: power of two); FEL_SIZE analogue */
  uint32_t stage_per_lp;     /* inserts one LP may receive per step (<= 64) */
  uint32_t node_size;        /* HQ: r = HQ_NODE_SIZE (src/config.h:88); must be >= num_lps */
  uint32_t heap_nodes;       /* HQ: node count (HQ_HEAP_SIZE); 0 = derive from capacity_per_lp */
  uint64_t total_capacity;   /* GAQ: items in the global array; 0 = num_lps * capacity_per_lp */
  float lookahead;           /* LOOKAHEAD (src/config.h:65); 0.1 in the float32 workloads */
  float horizon;             /* events with ts >= horizon are never dequeued (MAX_TS analogue) */
  uint32_t block_threads;    /* launch width of the per-LP kernels (multiple of 64, <= 1024; 0 = 256) */
  uint32_t collect_stats;    /* 1: post kernels count algorithmic work (gpq_queue_stats_get); costs atomics */
  uint32_t far_per_lp;       /* LAQ two-tier mode (0 = off): unsorted far store per LP for events at or
                                beyond the near horizon; the sorted ring keeps only events below it.
                                With lookahead = +inf (open lists, as A*'s) each LP has a horizon of its
                                own (a full key): a ring holding fewer than 8 records refills from its far
                                store, a full ring moves its largest records there (the heap-backed LAQ's
                                exchange, src/local_array_queue.cu:150-180,515-681); needs
                                capacity_per_lp >= 2 * stage_per_lp */
  float far_step;            /* LAQ two-tier with a finite lookahead: the near horizon moves to
                                (minimum + lookahead + far_step) when a window reaches it
                                (0 = 4 * lookahead) */
  uint32_t update_ids;       /* LAQ decrease-key mode (0 = off; one-tier LAQ only): item ids are
                                gpq_event.seq < update_ids (the A* cell); queue_insert_or_update of an
                                id already in the queue replaces its entry when the new key is smaller
                                (LOCAL_ARRAY_QUEUE_UPDATE_ELEMENTS, src/local_array_queue.cu:20-32,
                                350-416, 830-846) */
  uint32_t update_per_lp;    /* decrease-key mode: updates one LP may receive per step (0 = stage_per_lp) */
  uint32_t heaps;            /* HQ: independent heaps, HQ_NUM_QUEUES (src/config.h:87; 0 = 1).  Heap h
                                holds the LPs [h * ceil(num_lps / heaps), (h + 1) * ceil(num_lps / heaps))
                                (src/heap_queue.cu:1811); node_size and heap_nodes apply to each heap, and
                                each heap can take num_lps * stage_per_lp inserts per step; the window is
                                global (the minimum over all heaps) */
  uint32_t calendar_windows;  /* LAQ calendar mode (0 = off; a power of two in [8, 256]; needs far_per_lp):
                                for the PHOLD merge-and-commit step (drain 2) only, whose windows end one
                                lookahead apart: one unsorted append log for each of the next
                                calendar_windows windows and each block of 256 LPs, holding
                                capacity_per_lp (<= 64) records per LP of the block; a step reads only
                                the logs of the window it commits.  Events beyond the calendar wait in
                                two lists, and events at or past the horizon in a parked list, of
                                num_lps * far_per_lp records each.  Same committed trace. */
  uint32_t sort_tile;         /* GAQ / HQ: records per block of the record sort and merge kernels:
                                1024, 2048 or 4096 (0 = 2048); gpq_queue_autotune_sort picks one */
} gpq_queue_config;

typedef struct gpq_queue gpq_queue;
typedef struct gpq_phold gpq_phold;

/* ---- status ---------------------------------------------------------------------------- */
const char* gpq_status_string(gpq_status s);
/* Text of the most recent error on the calling thread ("" if none). */
const char* gpq_last_error(void);
/* Library version string. */
const char* gpq_version(void);

/* ---- host queue API (src/queue.h:86-115) ------------------------------------------------ */
/* queue_init() (src/queue.h:87): allocates and clears the queue. */
gpq_status gpq_queue_create(const gpq_queue_config* cfg, gpq_queue** out);
/* queue_finish() (src/queue.h:103) plus release of all device memory.  Destroy every model
 * created on the queue (gpq_phold_destroy, gpq_astar_destroy) first, as the reference calls
 * model_finish() (src/des.cu:220) before queue_finish() (src/des.cu:91). */
gpq_status gpq_queue_destroy(gpq_queue* q);
/* queue_post_init() (src/queue.h:92): bulk maintenance after a chunk of initial inserts. */
gpq_status gpq_queue_post_init(gpq_queue* q);
/* queue_pre() (src/queue.h:95): maintenance before a model step (selects the dequeue set for HQ/GAQ). */
gpq_status gpq_queue_pre(gpq_queue* q);
/* queue_post() (src/queue.h:97): folds the step's inserts in and publishes the next window. */
gpq_status gpq_queue_post(gpq_queue* q);
/* queue_get_min_ts() (src/queue.h:100): smallest pending timestamp (+inf if empty); synchronizes. */
gpq_status gpq_queue_get_min_ts(gpq_queue* q, float* out);
/* queue_check_phold() (src/queue.h:114, src/heap_queue.cu:1908-1945): pending item count.
 * PHOLD conserves events, so pending == num_lps * init_per_lp after every post; this replaces
 * the per-step get_enqueue_count() == get_dequeue_count() check (src/des.cu:31-43). */
gpq_status gpq_queue_pending(gpq_queue* q, uint64_t* pending);
/* Device error flags raised so far (bit 0 staging overflow, 1 queue overflow, 2 heap overflow). */
gpq_status gpq_queue_error_flags(gpq_queue* q, uint32_t* flags);
/* Waits for all queued device work of the handle. */
gpq_status gpq_queue_sync(gpq_queue* q);

/* Work counters of the post kernels, accumulated while collect_stats is set: the inputs
 * of the algorithmic-byte model (sorted insert of k records displacing d: 16*(2k+2d) bytes). */
typedef struct {
  uint64_t post_launches;   /* post / post_init launches */
  uint64_t inserted;        /* staged records folded in (k summed over LPs) */
  uint64_t displaced;       /* existing records moved to make room (d summed over LPs) */
  uint64_t merged_lps;      /* LP merges performed */
  uint64_t scanned_lps;     /* LP heads folded into the window reductions */
  uint64_t aux_bytes;       /* far-store moves of the two-tier LAQ (bytes) */
  uint64_t sorted;          /* GAQ / HQ: records sorted (the step's inserts) */
  uint64_t merged;          /* GAQ / HQ: records merged (a + b summed over the merges) */
  uint64_t selected;        /* GAQ / HQ: records scanned by the safe-event selection */
} gpq_queue_stats;
gpq_status gpq_queue_stats_get(gpq_queue* q, gpq_queue_stats* out);

/* Host-driven use of the device verbs (the reference's test kernels do the same,
 * src/test_laq.cu:115-145, src/test_gaq.cu:41-61): */
/* queue_insert() for each of n host events, one device thread per event. */
gpq_status gpq_queue_insert_host(gpq_queue* q, const gpq_event* items, uint32_t n);
/* queue_insert_or_update(item, item.dst) (src/queue.cuh:66, src/local_array_queue.cu:830-846) for
 * each of n host events; a plain insert unless the queue is a LAQ in decrease-key mode. */
gpq_status gpq_queue_insert_or_update_host(gpq_queue* q, const gpq_event* items, uint32_t n);
/* queue_peek() for every peek slot (LP for local queues, dequeue position for HQ/GAQ);
 * writes up to `max` found items and their indices, *found = count.  With set_done != 0
 * also calls queue_set_done() on each found item. */
gpq_status gpq_queue_peek_host(gpq_queue* q, int set_done, gpq_event* items, int32_t* index,
                               uint32_t max, uint32_t* found);
/* Contents owned by one LP in key order (local queues; HQ/GAQ: all pending items of that LP). */
gpq_status gpq_queue_lp_contents(gpq_queue* q, uint32_t lp, gpq_event* out, uint32_t cap, uint32_t* n);
/* Records of one node of the parallel heap (GPQ_HQ), in key order; node 0 is the root
 * (the reference's tests read heap_queue_data.fel per node, src/test_hq.cu:580-917).  With
 * several heaps, node h * nodes_per_heap + v is node v of heap h. */
gpq_status gpq_heap_node_contents(gpq_queue* q, uint32_t node, gpq_event* out, uint32_t cap, uint32_t* n);
/* The device operator API for model kernels compiled outside the library (src/queue.cuh:46-70:
 * queue_insert / queue_peek / queue_set_done / queue_is_empty / queue_length / queue_clear /
 * queue_insert_or_update, which the reference links into its model kernels with nvcc -rdc,
 * src/Makefile:21).  Copies q's device view -- the struct of its kind declared in
 * include/gpupq_device.h (gpq::LaqView, LhqView, LsqView, GaqView or HqView), whose device
 * methods are those verbs -- to *view; view_bytes must equal the size of that struct.  *kind (may
 * be NULL) receives the kind.  The caller passes the struct by value to its kernels and launches
 * them on the queue's stream between gpq_queue_pre and gpq_queue_post, as des_run orders
 * queue_pre, model_handle_next and queue_post (src/des.cu:154-214).  GPQ_ERR_STATE for a LAQ bound
 * to a PHOLD model in merge-and-commit mode (drain 2), whose staging buffers alternate per step. */
gpq_status gpq_queue_device_view(gpq_queue* q, gpq_queue_kind* kind, void* view, uint64_t view_bytes);
/* The HIP stream (hipStream_t) on which the queue orders its device work. */
gpq_status gpq_queue_stream(gpq_queue* q, void** stream);
/* queue_length(lp) / queue_is_empty(lp) / queue_clear(lp) (src/queue.cuh:58-66). */
gpq_status gpq_queue_length(gpq_queue* q, uint32_t lp, uint32_t* n);
gpq_status gpq_queue_clear_lp(gpq_queue* q, uint32_t lp);

/* ---- PHOLD model (src/model.h:9-19, src/phold.cu) -------------------------------------- */
typedef struct {
  uint32_t num_lps;        /* must equal the queue's num_lps */
  uint32_t init_per_lp;    /* k initial events per LP */
  uint32_t seed;           /* Philox key {seed, 0} */
  float lookahead;         /* ts' = (ts + lookahead) + Exp(1) */
  float end_time;          /* commit exactly the events with ts < end_time */
  uint32_t p_self_u32;     /* P(dst == self) * 2^32; 0x80000000 = 0.5 */
  uint32_t checks;         /* 1: per-LP monotone-timestamp check (src/des.cu:52-64) */
  uint32_t drain;          /* 0: one safe event per LP per step, as src/phold.cu:81-117 does;
                              1: each LP commits all of its events below the window end in key
                              order in one step (per-LP queues only).  Events created in a step
                              lie at or beyond the window end, so the committed trace is the same;
                              2: as 1, with the queue's merge done inside the model step (one kernel
                              per step) and each window ending one lookahead after the previous one;
                              needs the two-tier LAQ with capacity_per_lp <= 64.  queue_post then
                              only moves the window; events inserted from outside the model step
                              are not supported in this mode.  Same committed trace. */
} gpq_phold_config;

gpq_status gpq_phold_create(gpq_queue* q, const gpq_phold_config* cfg, gpq_phold** out);

/* One layout the PHOLD autotuner timed, and its rate (0 unless status is 0). */
typedef struct {
  uint32_t capacity_per_lp;
  uint32_t stage_per_lp;
  uint32_t far_per_lp;
  float far_step;               /* time units; 0 = 4 * lookahead */
  uint32_t drain;
  uint32_t status;              /* 0 ran; 1 could not hold the workload; 2 its trace hash differed
                                   from the first layout's (never picked) */
  double events_per_second;
  uint32_t calendar_windows;    /* LAQ calendar mode (0 = off); capacity_per_lp is then a window's
                                   log records per LP */
} gpq_phold_tuning;

/* Autotuner (no reference counterpart; the reference fixes FEL_SIZE at compile time,
 * src/local_array_queue.cuh:20): times a fixed list of LAQ layouts -- ring capacity per LP,
 * staging slots, far store, near-horizon step and step kind -- on the model's own workload cut
 * to end time sample_end_time (0 = 5, at most model->end_time), each in a fresh queue driven
 * by GPQ_DRIVE_GRAPH, and writes the fastest (a later candidate must beat the pick so far by
 * more than 2%) into *queue (kind = GPQ_LAQ, capacity_per_lp, stage_per_lp, far_per_lp, far_step,
 * calendar_windows;
 * update_ids and update_per_lp cleared; the other fields are kept) and model->drain.  results
 * (may be NULL) receives up to max_results candidates in the order timed, *n_results (may be
 * NULL) their count.  GPQ_ERR_CAPACITY when no layout holds the workload. */
gpq_status gpq_phold_autotune(gpq_phold_config* model, gpq_queue_config* queue, float sample_end_time,
                              gpq_phold_tuning* results, uint32_t max_results, uint32_t* n_results);
/* One sort tile the sort autotuner timed, and its rate (0 unless status is 0). */
typedef struct {
  uint32_t sort_tile;
  uint32_t status;              /* 0 ran; 1 could not run; 2 its trace hash differed from the first's */
  double events_per_second;
} gpq_sort_tuning;

/* Sort-tile autotuner for the global queues (no reference counterpart; the reference's
 * moderngpu mergesort fixes its tile at compile time, lib/moderngpu/src/moderngpu/
 * kernel_mergesort.hxx:27-148): times the tiles 2048, 1024 and 4096 of the record sort and
 * merge kernels on the model's own workload cut to end time sample_end_time (0 = 2, at most
 * model->end_time) through queue->kind (GPQ_GAQ or GPQ_HQ), each in a fresh queue driven by
 * GPQ_DRIVE_GRAPH, checks that every tile commits the same trace, and writes the fastest (a later
 * tile must beat the pick so far by more than 2%) into queue->sort_tile.  results (may be NULL)
 * receives up to max_results candidates in the order timed, *n_results (may be NULL) their count. */
gpq_status gpq_queue_autotune_sort(const gpq_phold_config* model, gpq_queue_config* queue, float sample_end_time,
                                   gpq_sort_tuning* results, uint32_t max_results, uint32_t* n_results);
gpq_status gpq_phold_destroy(gpq_phold* m);
/* model_init() (src/phold.cu:124-163): k initial events per LP, each chunk followed by queue_post_init(). */
gpq_status gpq_phold_init(gpq_phold* m);
/* model_handle_next() (src/phold.cu:185-192): one event step over all peek slots (asynchronous). */
gpq_status gpq_phold_handle_next(gpq_phold* m);
/* model_get_events() (src/phold.cu:169-182): committed events so far. */
gpq_status gpq_phold_get_events(gpq_phold* m, uint64_t* events);
/* model_finish() (src/model.h:19). */
gpq_status gpq_phold_finish(gpq_phold* m);
/* Per-LP committed counts and chained trace hashes (arrays of num_lps). */
gpq_status gpq_phold_lp_state(gpq_phold* m, uint32_t* counts, uint64_t* hashes);
/* Combined trace hash over the per-LP arrays, LPs in order. */
gpq_status gpq_phold_trace_hash(gpq_phold* m, uint64_t* hash);

/* ---- DES driver (src/des.h:3-5, src/des.cu:134-221) ------------------------------------ */
typedef enum {
  GPQ_DRIVE_PLAIN = 0,   /* per step: get_min_ts (host sync + D2H), pre, model step, post */
  GPQ_DRIVE_GRAPH = 1    /* `batch` steps captured once in a hipGraph, replayed; done flag polled per replay */
} gpq_drive_mode;

typedef struct {
  uint64_t steps;           /* model steps launched (graph mode: rounded up to whole batches) */
  uint64_t windows;         /* windows published by the queue (non-empty steps + trailing no-ops) */
  uint64_t events;          /* committed events */
  double sim_seconds;       /* wall clock of the step loop (src/des.cu:148-215) */
  double events_per_second; /* events / sim_seconds (src/des.cu:95-96) */
  float final_min_ts;       /* smallest pending timestamp at the end */
  uint32_t error_flags;
} gpq_des_report;

/* des_run(): steps the model until no pending event is below end_time. batch = steps per graph. */
gpq_status gpq_des_run(gpq_phold* m, gpq_drive_mode mode, uint32_t batch, gpq_des_report* report);

/* Per-kernel timing of a des_run: steps are launched eagerly, `batch` at a time, with HIP
 * events recorded on the queue's stream around the model step kernel and around the
 * queue's post kernels; durations are averaged over all profiled steps. */
typedef struct {
  uint64_t steps;
  uint64_t events;
  double step_kernel_us;    /* mean duration of the model step (handle_next) kernel */
  double post_kernel_us;    /* mean duration of queue_pre + queue_post kernels, without the window
                               publication of the per-LP queues (their post is one kernel) */
  double window_kernel_us;  /* mean duration of the window publication kernel (per-LP queues; 0 otherwise) */
  double seconds;           /* wall clock of the profiled loop */
} gpq_kernel_profile;
gpq_status gpq_des_profile(gpq_phold* m, uint32_t batch, gpq_kernel_profile* out);

/* ---- A* on grids (src/a_star.cu, src/model.h:9-19) ----------------------------------------
 * Grid cell c = y * width + x is an obstacle iff word (c & 3) of Philox4x32-10 of counter
 * {c >> 2, query, stream_tag, 2} under key {seed, 0} is below `threshold` (density * 2^32);
 * source and target are forced free.  connectivity 4: unit costs, Manhattan heuristic;
 * 8: costs 10 / 14, octile heuristic 10 * max + 4 * min (diagonal moves may cut corners, as
 * src/a_star.cu:393-459 allows).  Costs and f are exact uint32. */
typedef struct {
  uint32_t width, height;
  uint32_t connectivity;     /* 4 or 8 */
  uint32_t sx, sy, tx, ty;
  uint32_t seed, stream_tag, query;
  uint32_t threshold;        /* obstacle density * 2^32 */
  uint32_t window;           /* a list pops its best entry only if f <= (smallest f of all lists) + window;
                                0xFFFFFFFF: every non-empty list pops each step */
  uint32_t tile;             /* 0: the open lists are the queue's LPs, a cell's list a hash of the cell,
                                each list popping its best entry per step (src/a_star.cu:325-374);
                                16 or 32: the grid is cut into tile x tile squares, each with its own
                                open list (an inbox in HBM), searched by one workgroup per step that
                                expands its tile's entries inside the window layer after layer with
                                the tile's g values in LDS; entries for other tiles go to their inboxes.
                                The queue then only provides the stream and the error flags. */
} gpq_astar_config;

typedef struct {
  int64_t cost;              /* optimal path cost, -1 if the target is unreachable */
  uint64_t expanded;         /* expansions (a cell may be expanded again after a cheaper path) */
  uint64_t closed;           /* distinct cells in the closed set */
  uint64_t steps;            /* model steps launched */
  double seconds;            /* wall clock of the step loop */
  uint32_t error_flags;
} gpq_astar_report;

typedef struct gpq_astar gpq_astar;

/* One query searched by many open lists at once: the LPs of `open_lists` (a per-LP queue:
 * GPQ_LAQ, GPQ_LHQ or GPQ_LSQ; its lookahead and horizon must be +inf) are the open lists.
 * model_init (src/a_star.cu:698-842) generates the grid and pushes the source. */
gpq_status gpq_astar_create(gpq_queue* open_lists, const gpq_astar_config* cfg, gpq_astar** out);
gpq_status gpq_astar_destroy(gpq_astar* a);
gpq_status gpq_astar_init(gpq_astar* a);
/* Replaces the generated obstacle map (w*h bytes, nonzero = obstacle) and restarts the search. */
gpq_status gpq_astar_set_grid(gpq_astar* a, const uint8_t* blocked);
/* model_handle_next (src/a_star.cu:854-872): every open list pops its best entry, expands it
 * and sends improved neighbours to their lists; queue_post folds them in. */
gpq_status gpq_astar_handle_next(gpq_astar* a);
/* des_run for A*: steps until the best goal cost is <= the smallest f of any open list. */
gpq_status gpq_astar_run(gpq_astar* a, gpq_drive_mode mode, uint32_t batch, gpq_astar_report* rep);
/* The grid the search ran on (w*h bytes). */
gpq_status gpq_astar_grid(gpq_astar* a, uint8_t* blocked);
/* An optimal path of the finished search, source to target, as cell indices y * width + x
 * (the closed-list backtrack of src/a_star.cu:90-94,417-438).  *len = path length (0 when the
 * target is unreachable); at most cap cells are written. */
gpq_status gpq_astar_path(gpq_astar* a, uint32_t* cells, uint32_t cap, uint32_t* len);

/* Batched A*: many independent queries, each searched by one workgroup with per-thread
 * local heaps; one persistent launch pulls queries from a device work counter (the reference's
 * get_new_agent, src/a_star.cu:149-164).  Query q of a batch uses the shape and endpoints of
 * the synthetic distribution (side = round(256 * 2^(3u)), density U[0.1, 0.3], 8-connected)
 * unless explicit specs are given. */
typedef struct {
  uint32_t side;             /* square grid side */
  uint32_t threshold;        /* obstacle density * 2^32 */
  uint32_t sx, sy, tx, ty;
  uint32_t connectivity;     /* 4 or 8 */
  uint32_t seed, stream_tag, query;   /* obstacle generator (see gpq_astar_config) */
} gpq_query_spec;

typedef struct {
  uint32_t threads_per_query;   /* workgroup size = open lists per query (64..1024, multiple of 64) */
  uint32_t heap_capacity;       /* entries per open list */
  uint32_t stage_capacity;      /* inserts one open list may receive per iteration */
  uint32_t max_resident;        /* queries searched at once (0 = fill the device) */
} gpq_batch_tuning;

typedef struct gpq_astar_batch gpq_astar_batch;

/* Synthetic query q of the batched workload (seed = master seed). */
gpq_status gpq_query_synthetic(uint32_t seed, uint32_t q, gpq_query_spec* out);
gpq_status gpq_astar_batch_create(const gpq_query_spec* specs, uint32_t count, const gpq_batch_tuning* tuning,
                                  gpq_astar_batch** out);
gpq_status gpq_astar_batch_destroy(gpq_astar_batch* b);
/* Autotuner: times threads_per_query in {64..1024} x stage_capacity in {8, 16, 32} on the
 * first `sample` queries (0 = all) and writes the fastest into *tuning (its heap_capacity
 * and max_resident are kept as given; 0 = defaults).  best_seconds (may be NULL) receives
 * that candidate's time. */
gpq_status gpq_astar_batch_autotune(const gpq_query_spec* specs, uint32_t count, uint32_t sample,
                                    gpq_batch_tuning* tuning, double* best_seconds);
/* Runs every query: GPQ_DRIVE_GRAPH = one persistent launch (captured once in a hipGraph);
 * GPQ_DRIVE_PLAIN = the one-query-per-launch path: each query alone, one launch per expansion
 * step with a host synchronization per step (the reference's des_run pattern). */
gpq_status gpq_astar_batch_run(gpq_astar_batch* b, gpq_drive_mode mode, double* seconds);
/* Per-query optimal cost (-1 unreachable) and expansions. */
gpq_status gpq_astar_batch_results(gpq_astar_batch* b, int64_t* costs, uint64_t* expanded);
/* Device wall-clock times of the last run, per query, in microseconds from the earliest
 * pickup: when a workgroup picked the query up and when its search completed. */
gpq_status gpq_astar_batch_timing(gpq_astar_batch* b, double* start_us, double* finish_us);
/* An optimal path of query q, source to target, as cell indices y * side + x (the closed-list
 * backtrack of src/a_star.cu:90-94,417-438): q is searched again alone in workgroup slot 0 and
 * the path is read back from that slot's g values.  *len = path length (0 when the target is
 * unreachable); at most cap cells are written.  Overwrites q's timing of the last run. */
gpq_status gpq_astar_batch_path(gpq_astar_batch* b, uint32_t q, uint32_t* cells, uint32_t cap, uint32_t* len);
/* Obstacle map of query q (side*side bytes). */
gpq_status gpq_astar_batch_grid(gpq_astar_batch* b, uint32_t q, uint8_t* blocked);

/* ---- Batched PHOLD: many instances of uneven size in one queue ---------------------------
 * The instances share one local array queue, each owning a range of whole 256-LP blocks and
 * its own window; one hipGraph replay of (model step, queue merge, per-instance window
 * publication) advances all of them.  Each instance's trace hash and event count equal
 * those of the same instance run alone (gpq_phold_* with drain on). */
typedef struct {
  uint32_t num_lps;
  uint32_t init_per_lp;
  uint32_t seed;
  float end_time;
  float lookahead;
  uint32_t p_self_u32;
} gpq_phold_instance;

typedef struct gpq_phold_batch gpq_phold_batch;

/* capacity_per_lp / stage_per_lp: the LAQ sizes (0: 128 / 32); far_per_lp > 0: two-tier LAQ
 * (every instance with its own near horizon; capacity_per_lp <= 128). */
gpq_status gpq_phold_batch_create(const gpq_phold_instance* inst, uint32_t count, uint32_t capacity_per_lp,
                                  uint32_t stage_per_lp, uint32_t far_per_lp, gpq_phold_batch** out);
/* Calendar mode (0: off; else a power of two in [8, 256] windows): both drives use calendar-mode
 * LAQs and the merge-and-commit step; capacity_per_lp is then a window's log records per LP and
 * far_per_lp (>= every init_per_lp) the list capacity per LP.  GPQ_DRIVE_GRAPH puts all
 * instances in one calendar LAQ on one window sequence (width: the smallest lookahead; last end:
 * the latest end time), one kernel per step for all of them; each instance's committed events
 * are still exactly its events below its own end time, so its results equal its run alone. */
gpq_status gpq_phold_batch_set_calendar(gpq_phold_batch* b, uint32_t calendar_windows);
gpq_status gpq_phold_batch_destroy(gpq_phold_batch* b);
/* GPQ_DRIVE_GRAPH: all instances in one pool, `batch` steps per graph replay.
 * GPQ_DRIVE_PLAIN: each instance alone in its own queue, a host synchronization per step.
 * seconds: wall clock of the step loops (population excluded). */
gpq_status gpq_phold_batch_run(gpq_phold_batch* b, gpq_drive_mode mode, uint32_t batch, double* seconds);
gpq_status gpq_phold_batch_results(gpq_phold_batch* b, uint64_t* trace_hashes, uint64_t* events);

/* ---- Serving: a stream of mixed requests (PHOLD simulations and A* queries) --------------
 * Request i of the synthetic trace: Philox4x32-10 of counter {i, 0, 6, 0} under key {seed, 0}
 * gives the gap to the previous arrival (Exp with mean 1/rate; rate 0: all arrive at t = 0),
 * the kind (word 1 < phold_fraction * 2^32: PHOLD), the PHOLD size (LPs = round(2^U[min, max]),
 * init_per_lp events per LP, end time, lookahead 0.1, p_self 0.5) and the PHOLD seed (word 3).
 * An A* request i is gpq_query_synthetic(seed, i). */
typedef struct {
  uint32_t requests;
  uint32_t seed;
  double rate;                       /* arrivals per second */
  double phold_fraction;             /* share of PHOLD requests */
  uint64_t hbm_budget;               /* bytes of the PHOLD pool admitted requests share (0: 4 GiB) */
  uint32_t phold_min_log2, phold_max_log2;   /* 0: 12 and 18 */
  uint32_t phold_init_per_lp;        /* 0: 16 */
  float phold_end_time;              /* 0: 20 */
  uint32_t phold_batch;              /* steps per graph replay of a served PHOLD request (0: 16) */
  uint32_t astar_lanes;              /* concurrent A* launches when served (0: 4) */
  uint32_t plain_preallocated;       /* GPQ_DRIVE_PLAIN only: 1 = every PHOLD request's queue and model
                                        are created before the clock starts and populated inside it,
                                        so neither drive allocates inside the clock; 0 = each request
                                        creates and destroys its own inside the clock (process per run) */
} gpq_serve_config;

typedef struct {
  uint32_t kind;                     /* 0 PHOLD, 1 A* */
  uint32_t size;                     /* PHOLD: LPs; A*: grid side */
  double arrival_ms, start_ms, finish_ms;   /* from the start of the run */
  uint64_t result;                   /* PHOLD: trace hash; A*: optimal cost (UINT64_MAX: unreachable) */
  uint64_t work;                     /* PHOLD: committed events; A*: expansions */
  uint64_t bytes;                    /* HBM held while admitted */
} gpq_request;

typedef struct {
  uint32_t completed;
  uint32_t peak_inflight;
  double seconds;                    /* first arrival to last completion */
  double requests_per_second;
  double p50_ms, p99_ms, mean_ms;    /* latency = completion - arrival */
  double phold_p99_ms, astar_p99_ms;
  uint64_t peak_bytes;               /* most HBM held by admitted requests at once (PHOLD + A*) */
  uint64_t phold_peak_bytes;         /* most HBM held by admitted PHOLD requests at once (<= hbm_budget) */
  uint64_t pool_bytes;               /* served: HBM of the shared PHOLD pool (<= hbm_budget) */
} gpq_serve_report;

/* The trace alone (kind, size, arrival_ms of each request). */
gpq_status gpq_serve_trace(const gpq_serve_config* cfg, gpq_request* out);
/* Serves the trace.  GPQ_DRIVE_GRAPH: continuous batching -- admission in arrival order under
 * the HBM budget into one shared PHOLD pool advanced by graph replays, A* queries gathered
 * into persistent-kernel launches on astar_lanes streams.  GPQ_DRIVE_PLAIN: one request at a
 * time, a host synchronization per step.  A PHOLD request whose queue alone exceeds hbm_budget
 * is rejected with GPQ_ERR_CAPACITY before anything runs.  out (may be NULL) receives every
 * request. */
gpq_status gpq_serve_run(const gpq_serve_config* cfg, gpq_drive_mode mode, gpq_serve_report* rep, gpq_request* out);

/* ---- self-test hooks: the device random-number path evaluated in bulk ------------------- */
/* out[i] = device Exp(1) sample of draw ((first + i) << 8), i < count (count <= 2^24). */
gpq_status gpq_selftest_exp(uint32_t first, uint32_t count, float* out);
/* out[4*i .. 4*i+3] = device Philox4x32-10 of counter {ctr0 + i, ctr1, ctr2, ctr3}, key {key0, key1}. */
gpq_status gpq_selftest_philox(uint32_t ctr0, uint32_t ctr1, uint32_t ctr2, uint32_t ctr3, uint32_t key0,
                               uint32_t key1, uint32_t count, uint32_t* out);
/* Record primitives of the global queues on host arrays: sort by (ts, src, seq); stable
 * merge of two sorted arrays; one-safe-event-per-LP selection over a sorted array (prefix
 * ts < wend and key <= *bound when bound != NULL; first record per LP -> deq, others of the
 * prefix -> rest, in order). */
gpq_status gpq_selftest_sort(const gpq_event* in, uint32_t n, gpq_event* out);
gpq_status gpq_selftest_merge(const gpq_event* a, uint32_t na, const gpq_event* b, uint32_t nb, gpq_event* out);
gpq_status gpq_selftest_select(const gpq_event* sorted, uint32_t n, uint32_t num_lps, float wend,
                               const gpq_event* bound, gpq_event* deq, uint32_t* ndeq, gpq_event* rest,
                               uint32_t* nrest, uint32_t* prefix);

/* ---- allocation-free record sort ---------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): sorts n device
 * records d_in by (ts, src, seq) into d_out (device; d_out == d_in sorts in place) on `stream`
 * (a hipStream_t; NULL = the default stream), with the kernels of gpq_selftest_sort.  Its
 * partner buffer and count word live in a library-owned device array of 16 MiB + 256 B that
 * every GPU it runs on holds until the process ends; it calls no hipMalloc or hipFree.
 * n <= 2^20 (GPQ_ERR_CAPACITY above).  Calls that overlap on one GPU share that array and race:
 * the caller serialises them.  Returns once the work is queued (asynchronous). */
gpq_status gpq_sort_events_noalloc(const gpq_event* d_in, uint32_t n, gpq_event* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
