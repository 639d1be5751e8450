/*
 * albd.h -- the thin C ABI between libalbba's C host code and its HIP
 * kernels.  Plain C: device memory appears as typed pointers, the stream and
 * events stay inside albd_ctx.  Every launch function is asynchronous on the
 * context's stream unless its comment says otherwise, and returns 0 or a
 * negative ALBBA_E* code.
 */
#ifndef ALBD_H
#define ALBD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALBD_HIDDEN __attribute__((visibility("hidden")))
/* Counter slots: blocks spread their per-level atomics over this many 64-byte
 * slots (same-address device atomics serialise at roughly 100 ns each), and
 * the select kernel sums them at the end of the level. */
#define ALBD_NSLOT 256

typedef struct albd_ctx albd_ctx;

/* Per-level counters, one copy per slot. */
typedef struct albd_slot {
    unsigned long long next_count;   /* vertices discovered this level        */
    unsigned long long next_edges;   /* sum of out-degree of those vertices   */
    unsigned long long next_words;   /* nonzero next-frontier words           */
    unsigned long long scanned;      /* pull: column entries gathered         */
    unsigned long long words;        /* pull: position words scanned          */
    unsigned long long bypassed;     /* pull: words skipped, all visited      */
    unsigned long long col_bypassed; /* pull + column bypass: gathers skipped */
    unsigned long long pad_;         /* one slot per 64 bytes                 */
} albd_slot;

typedef struct albd_counters {
    unsigned long long queue_len;    /* current frontier queue length         */
    unsigned long long qnext_len;    /* queue being built                     */
    unsigned long long active_len;   /* length of the current unvisited-word list */
    unsigned long long push_edges;   /* sum of degrees over the queue         */
    /* written by the select kernel at the end of every level */
    unsigned long long frontier;       /* vertices found by the last level     */
    unsigned long long frontier_edges; /* their out-edges                      */
    unsigned long long frontier_words; /* nonzero words of their bitmap        */
    unsigned long long reached;        /* vertices with a level so far         */
    unsigned long long reached_edges;  /* their out-edges                      */
    unsigned long long scanned;        /* pull: column entries gathered        */
    unsigned long long words;          /* pull: position words scanned         */
    unsigned long long bypassed;       /* pull: words bypassed, all visited    */
    unsigned long long col_bypassed;   /* pull + column bypass: column entries
                                          whose frontier word was empty        */
    int sel_choice;                    /* policy's choice for the next level   */
    int heur_choice;                   /* the reference heuristic's choice     */
    float features[12];                /* selector features of the next level  */
    int persist_levels;                /* levels run by the small-frontier kernel */
    int persist_in_queue;              /* 1: its last frontier is in st.queue   */
    int persist_error;                 /* 1: its grid barrier gave up           */
    int pad_;
    unsigned int seq;                  /* host mailbox: sequence of the copy    */
    unsigned int pad2_;
    /* Unvisited-word list lengths, slot level % 3: built by that level's
     * preparation, read by its pull, handed over (to active_len) by the next
     * select kernel, which also zeroes slot (level + 2) % 3 for the level
     * after.  Three slots let the select kernel of one level and the
     * preparation of the next run in one launch.  Outside the block the
     * select kernel writes back. */
    unsigned long long list_cnt[3];
} albd_counters;

/* levels one small-frontier call may run (st.level_log holds 4 per level) */
#define ALBD_PERSIST_LEVELS 4096

/* grid state of the small-frontier kernel, zeroed before each launch */
typedef struct albd_persist {
    unsigned long long barrier;      /* arrivals at the grid barrier        */
    unsigned int error;              /* a workgroup gave up waiting         */
    unsigned int pad_;
} albd_persist;

/* Policy inputs of the select kernel. */
typedef struct albd_select {
    const float *params;   /* MLP parameters on the device, or NULL (heuristic) */
    int32_t level;         /* index of the level being decided                  */
    int32_t prev_choice;   /* choice committed for the previous level           */
    int32_t valid_mask;    /* bit c set: choice c may run                        */
    int32_t first;         /* 1 right after init: the root is the frontier       */
    int32_t forced;        /* >= 0: the policy's choice for this level is this one
                              (a scripted level); -1: the MLP's, or the heuristic's */
} albd_select_args;

/* Read-only graph arrays on the device. */
typedef struct albd_graph {
    int32_t n, C, pad, nslices, nwords;
    int32_t sparse_lanes;       /* pull: at most this many live lanes -> row phase */
    int32_t win;                /* positions [k*win, (k+1)*win) hold the vertices of
                                 * the same id range (the sigma windows) */
    int64_t npos, nnz, padded_nnz, nvops;
    const int64_t *slice_ptr;   /* [nslices+1] */
    const int32_t *slice_width; /* [nslices]   */
    const int64_t *vop_ptr;     /* [nslices+1] */
    const uint8_t *vop_len;     /* [nvops]     */
    const int32_t *colidx;      /* [padded_nnz]*/
    const int32_t *row_order;   /* [npos] position -> vertex id (-1 past n) */
    const int32_t *row_len;     /* [npos]      */
    const int64_t *push_ptr;    /* [npos+1]    */
    const int32_t *push_col;    /* [nnz]       */
    const int64_t *pull_ptr;    /* [npos+1] row-major rows of the layout (may alias push) */
    const int32_t *pull_col;    /* [nnz]       */
    const uint64_t *haslen;     /* [nwords] bit set where row_len > 0 */
    const int32_t *pos_of;      /* [n] vertex id -> position */
    const uint16_t *row_off;    /* [n] row_order[p] - start of p's window, when
                                 * windows are at most 65 536 rows; else null */
    /* win: the sigma window (rows [k*win, (k+1)*win) of the layout hold the
     * vertices of that id range), or 0 when the rows were ordered across
     * windows first (the first-column order): then no window holds its own
     * id range */
} albd_graph;

/* Mutable BFS state on the device. */
typedef struct albd_state {
    uint64_t *visited;          /* [nwords]  */
    uint64_t *front;            /* [nwords]  current frontier bitmap */
    uint64_t *next;             /* [nwords]  next frontier bitmap    */
    uint64_t *summary;          /* [nwords/64+1] nonzero-word bits of front */
    int32_t *queue;             /* [npos]    current frontier queue (positions) */
    int32_t *queue_next;        /* [npos]    */
    int64_t *prefix;            /* [npos+1]  degree prefix of the queue */
    int64_t *block_sums;        /* [ALBD_SCAN_BLOCKS+1] */
    uint32_t *pmin;             /* [npos]    push: min parent id per position */
    unsigned long long *lp;     /* [npos]    level << 32 | parent vertex id, by position */
    int32_t *level;             /* [n]       vertex-id space, written at the end */
    int32_t *parent;            /* [n]       */
    int32_t *active;            /* [nwords]  unvisited-word list */
    int32_t *active_next;       /* [nwords]  */
    albd_counters *ctr;         /* device counters */
    albd_slot *slots;           /* [ALBD_NSLOT] per-level counter slots */
    int64_t *level_log;         /* [4*ALBD_PERSIST_LEVELS] found, edges, queue per level */
    unsigned long long *pkey;   /* [npos] small-frontier (level, parent) keys, all ones at rest */
    albd_persist *persist;      /* small-frontier grid state */
} albd_state;

#define ALBD_SCAN_BLOCKS 1024

/* ---- context / memory (synchronous unless noted) ---- */
ALBD_HIDDEN int albd_device_count(int *count);
ALBD_HIDDEN int albd_ctx_create(int device, albd_ctx **out);
ALBD_HIDDEN void albd_ctx_destroy(albd_ctx *c);
ALBD_HIDDEN int albd_ctx_device(const albd_ctx *c);
ALBD_HIDDEN int albd_num_cus(const albd_ctx *c);
ALBD_HIDDEN int albd_malloc(albd_ctx *c, size_t bytes, void **p);
ALBD_HIDDEN void albd_free(albd_ctx *c, void *p);
ALBD_HIDDEN int albd_h2d(albd_ctx *c, void *dst, const void *src, size_t bytes);
ALBD_HIDDEN int albd_d2h(albd_ctx *c, void *dst, const void *src, size_t bytes);
ALBD_HIDDEN int albd_d2d(albd_ctx *c, void *dst, const void *src, size_t bytes); /* async */
ALBD_HIDDEN int albd_sync(albd_ctx *c);
ALBD_HIDDEN const char *albd_error(void);
/* pinned host mirror of the counters, filled by albd_read_counters */
ALBD_HIDDEN albd_counters *albd_host_counters(albd_ctx *c);
ALBD_HIDDEN int albd_read_counters(albd_ctx *c, const albd_state *st); /* sync */

/* ---- events (timing on the context stream) ---- */
#define ALBD_NEVENTS 10
ALBD_HIDDEN int albd_event_record(albd_ctx *c, int idx);
ALBD_HIDDEN int albd_event_elapsed(albd_ctx *c, int a, int b, float *ms); /* syncs b */
ALBD_HIDDEN int albd_event_wait(albd_ctx *c, int idx); /* host waits for the event */

/* ---- pinned host staging ---- */
ALBD_HIDDEN int albd_pinned_alloc(albd_ctx *c, size_t bytes, void **p);
ALBD_HIDDEN void albd_pinned_free(albd_ctx *c, void *p);
/* host -> device copy on the context's stream, not waited for */
ALBD_HIDDEN int albd_h2d_async(albd_ctx *c, void *dst, const void *src, size_t bytes);

/* ---- fills (kernels, safe under graph capture) ---- */
ALBD_HIDDEN int albd_fill_u32(albd_ctx *c, void *p, uint32_t v, size_t count);
ALBD_HIDDEN int albd_fill_u64(albd_ctx *c, void *p, uint64_t v, size_t count);
/* dst[i] = src[i] for count 64-bit words (16-byte aligned pointers) */
ALBD_HIDDEN int albd_copy_u64(albd_ctx *c, void *dst, const void *src, size_t count);

/* ---- SELL-C-sigma build on the device ---- */
typedef struct albd_build_input {
    const int64_t *row_ptr; /* CSR (validated by the caller: row_ptr[0] = 0,
                               non-decreasing, row_ptr[n] = nnz), host memory
                               unless on_device                               */
    const int32_t *col;
    int64_t nnz;
    int32_t n, C, sigma, pad;
    int32_t symmetric;      /* 1: the push CSR is the layout's own rows      */
    int32_t on_device;      /* 1: row_ptr/col are device pointers (no upload) */
    int32_t sort_first;     /* 1: first-column order before the window sort
                               (ALBBA_SELL_SORT_FIRST_COL)                   */
} albd_build_input;
typedef struct albd_build_report {
    float upload_ms;        /* CSR host -> device                            */
    float build_ms;         /* every build step on the device                */
    int32_t rows_sorted;    /* 1: every row was already in ascending order   */
    int32_t pad_;
} albd_build_report;
/* device allocation for the graph's own arrays (the caller owns them) */
typedef void *(*albd_build_alloc)(void *arg, size_t bytes);
/* Fill every layout array of *g (row order, positions, slices, columns, the
 * push and pull CSRs) from the CSR; synchronous.  Temporaries are freed. */
ALBD_HIDDEN int albd_build_graph(albd_ctx *c, const albd_build_input *in, albd_graph *g,
                                 albd_build_alloc alloc, void *alloc_arg, albd_build_report *rep);

/* W = A^T of a device CSR with int32 row pointers (the reference's working
 * matrix, bfs/src/main.c:117-126): checks row_ptr (0 .. nnz, non-decreasing)
 * and the column range (*bad = 1 and no output when either fails), then
 * allocates *wp (int64, n+1) and *wc (nnz) -- freed by the caller with
 * albd_free -- with every row ascending.  Synchronous. */
ALBD_HIDDEN int albd_transpose_csr(albd_ctx *c, const int32_t *rp32, const int32_t *col, int32_t n,
                                   int64_t nnz, int64_t **wp, int32_t **wc, int *bad);

/* Generated graph on the device (dev/gen.hip): the edge rule of `kind`
 * (ALBBA_GEN_*), ids mapped through perm (host array, n entries), both
 * directions, self loops and duplicates dropped, rows ascending.  Allocates
 * *rp (int64, n+1) and *col (m) -- freed by the caller with albd_free.
 * Synchronous. */
typedef struct albd_gen_spec {
    int32_t kind, scale;
    int32_t n, bits;        /* bits: id field width of the sort keys      */
    int64_t ne;             /* edges drawn                                 */
    double a, ab, abc;      /* RMAT thresholds                             */
    uint64_t seed;
    int32_t rows, cols;
    int64_t H;              /* grid: horizontal edges                      */
    double p_remove;
} albd_gen_spec;
ALBD_HIDDEN int albd_gen_csr(albd_ctx *c, const albd_gen_spec *spec, const int32_t *perm_host,
                             int64_t **rp, int32_t **col, int64_t *m);

/* ---- BFS pieces ---- */
/* haslen[w] from row_len */
ALBD_HIDDEN int albd_build_haslen(albd_ctx *c, const albd_graph *g, uint64_t *haslen);
/* row_off from row_order (g->win at most 65 536) */
ALBD_HIDDEN int albd_build_row_off(albd_ctx *c, const albd_graph *g, uint16_t *row_off);
/* reset every state array and seed the root (position rp, vertex id ro) */
ALBD_HIDDEN int albd_bfs_init(albd_ctx *c, const albd_graph *g, const albd_state *st,
                              int32_t rp, int32_t ro);
/* level/parent from position order (st->lp) to vertex order */
/* if_done: queued before the host knew the BFS had ended; the kernel runs only
 * if the last select kernel found the frontier empty */
ALBD_HIDDEN int albd_to_vertex_order(albd_ctx *c, const albd_graph *g, const albd_state *st, int if_done);
/* zero st->next (and, flag 1, the per-level slot counters) */
ALBD_HIDDEN int albd_level_begin(albd_ctx *c, const albd_graph *g, const albd_state *st);

/* pull variants */
enum { ALBD_PULL_ALL = 0, ALBD_PULL_LIST = 1, ALBD_PULL_LIST_CB = 2 };
/* compact the unvisited-word list of `level`: from all words (from_list = 0)
 * or from the previous list st->active of length ctr->active_len (from_list =
 * 1; prev_len bounds it on the host); result in st->active_next, length in
 * ctr->list_cnt[level % 3] */
ALBD_HIDDEN int albd_build_active(albd_ctx *c, const albd_graph *g, albd_state *st,
                                  int from_list, int64_t prev_len, int32_t level);
ALBD_HIDDEN int albd_build_summary(albd_ctx *c, const albd_graph *g, const albd_state *st);
/* out[nwords] (device) = stored entries of every position word */
ALBD_HIDDEN int albd_word_entries(albd_ctx *c, const albd_graph *g, int64_t *out);
/* As albd_pull, queued before the level's direction is known: the kernel
 * returns at once unless the select kernel chose this variant (spec = 1: its
 * learned choice, 0: the heuristic's) and the frontier is not empty; spec =
 * -1 is albd_pull. */
ALBD_HIDDEN int albd_pull_spec(albd_ctx *c, const albd_graph *g, const albd_state *st, int variant,
                               int32_t level, int64_t items_bound, int spec);
/* one launch: zero st->next, compact the unvisited-word list (as
 * albd_build_active), and with `summary` build the frontier summary */
ALBD_HIDDEN int albd_prep(albd_ctx *c, const albd_graph *g, const albd_state *st, int from_list,
                          int summary, int32_t level);
/* albd_select_post and albd_prep of level sel->level in one launch (the
 * select kernel in one block, the preparation in the others) */
ALBD_HIDDEN int albd_select_prep_post(albd_ctx *c, const albd_graph *g, const albd_state *st,
                                      const albd_select_args *sel, unsigned int *seq, int from_list,
                                      int summary);
ALBD_HIDDEN int albd_pull(albd_ctx *c, const albd_graph *g, const albd_state *st, int variant,
                          int32_t level, int64_t items_bound);
ALBD_HIDDEN int albd_cb_supported(const albd_graph *g);

/* the push kernels of one level (scan, expansion, finalize), queued before
 * the level's direction is known: they return at once unless the select
 * kernel chose the push (spec = 1: its learned choice, 0: the heuristic's)
 * for a nonempty frontier of at most ALBD_SPEC_PUSH_MAX vertices, the bound
 * below which albd_push_scan takes its one-block scan.  The queue must hold
 * the frontier (the last level pushed); the finalize lists the next queue as
 * albd_push_finalize does with build_max. */
#define ALBD_SPEC_PUSH_MAX 8192
ALBD_HIDDEN int albd_push_spec(albd_ctx *c, const albd_graph *g, const albd_state *st, int32_t level, int spec,
                               int64_t build_max);
/* push: queue from the front bitmap (when the last level was a pull) */
ALBD_HIDDEN int albd_front_to_queue(albd_ctx *c, const albd_graph *g, const albd_state *st);
/* push: prefix of degrees over the queue (qlen known on the host) */
ALBD_HIDDEN int albd_push_scan(albd_ctx *c, const albd_graph *g, const albd_state *st,
                               int64_t qlen);
ALBD_HIDDEN int albd_push_expand(albd_ctx *c, const albd_graph *g, const albd_state *st);
/* push: merge next into visited, write level/parent from pmin, and list
 * queue_next (ctr->qnext_len) while the level's frontier edges
 * (ctr->frontier_edges) are at most build_max; else it stays empty */
ALBD_HIDDEN int albd_push_finalize(albd_ctx *c, const albd_graph *g, const albd_state *st,
                                   int32_t level, int64_t build_max);
/* Small-frontier push levels in one persistent kernel of `blocks`
 * workgroups (0: the default; capped at what can be co-resident): starting from the
 * queue (st.queue, ctr->queue_len), run push levels level0, level0+1, ...
 * while the new frontier has at most edge_cap out-edges and vertex_cap
 * vertices, at most max_levels levels.
 * On return the state is that of one finished push level: st.next holds the
 * last frontier's bitmap, its queue is in st.queue_next (ctr->persist_in_queue
 * = 0) or st.queue (= 1), ctr->qnext_len its length, the slots its counts;
 * ctr->persist_levels levels ran, st.level_log holds their sizes.  A grid
 * barrier that waits more than spin_limit polls sets ctr->persist_error and
 * the kernel ends early (the keys in st.pkey must then be refilled). */
ALBD_HIDDEN int albd_push_small(albd_ctx *c, const albd_graph *g, const albd_state *st,
                                int32_t level0, int64_t edge_cap, int64_t vertex_cap,
                                int32_t max_levels, int32_t blocks, uint32_t spin_limit);
/* end of a level: sum the slot counters into ctr->frontier*, update the
 * running totals, compute the 12 features and the next level's choice
 * (MLP argmin when sel->params is set, else the reference heuristic) */
ALBD_HIDDEN int albd_select(albd_ctx *c, const albd_graph *g, const albd_state *st,
                            const albd_select_args *sel);
/* albd_select, then the counters on the host (albd_host_counters): the
 * kernel posts them into a pinned host-coherent mailbox and the host polls
 * it, instead of a copy and a stream synchronisation (synchronous) */
ALBD_HIDDEN int albd_select_read(albd_ctx *c, const albd_graph *g, const albd_state *st,
                                 const albd_select_args *sel);
/* the launch half of albd_select_read (asynchronous); albd_mail_wait(seq)
 * completes it */
ALBD_HIDDEN int albd_select_post(albd_ctx *c, const albd_graph *g, const albd_state *st,
                                 const albd_select_args *sel, unsigned int *seq);
/* mailbox plumbing for albd_select_read: the next sequence number and the
 * mailbox's device address; the wait fills albd_host_counters */
ALBD_HIDDEN albd_counters *albd_mail_next(albd_ctx *c, unsigned int *seq);
ALBD_HIDDEN int albd_mail_wait(albd_ctx *c, const albd_state *st, unsigned int seq);
ALBD_HIDDEN void albd_mail_stats(const albd_ctx *c, unsigned long *hits, unsigned long *misses);

/* ---- batched BFS (batch.hip) ---- */
/* One instance: a device graph, its own scratch, a root.  The control block
 * ctl (32 words) must be zero before the launch: [0] team barrier,
 * [1] found-list length, [2] error flag, [4..19] ring of four level counters
 * (pulled vertices, out-edges, -, nonzero words), [24..31] results. */
typedef struct albd_batch_item {
    albd_graph g;
    uint64_t *visited;           /* [nwords] */
    uint64_t *bm[3];             /* [nwords] frontier bitmaps, rotating */
    int32_t *found;              /* [npos] frontier queues, level after level */
    unsigned long long *key;     /* [npos] (level, parent id), all ones at rest */
    int32_t *level, *parent;     /* [n] results in vertex order */
    unsigned long long *ctl;     /* [32] see above */
    int32_t root;                /* vertex id */
    int32_t mapping;             /* pull: ALBD_MAP_LANE_ROW or ALBD_MAP_WAVE_ROW */
} albd_batch_item;

enum { ALBD_MAP_LANE_ROW = 0, ALBD_MAP_WAVE_ROW = 1 };
/* results in ctl[24..31] */
enum { ALBD_RES_LEVELS = 24, ALBD_RES_REACHED, ALBD_RES_EDGES, ALBD_RES_T0, ALBD_RES_T1,
       ALBD_RES_PUSHES, ALBD_RES_PULLS, ALBD_RES_ERROR };

typedef struct albd_batch_args {
    const albd_batch_item *items; /* [count] device */
    const int32_t *wg_team;      /* [workgroups] team of each workgroup */
    const int32_t *team_first;   /* [teams] first workgroup of the team */
    const int32_t *team_size;    /* [teams] workgroups in the team */
    const int32_t *team_ptr;     /* [teams+1] range of team_items per team */
    const int32_t *team_items;   /* [count] instances, run in this order per team */
    const float *params;         /* MLP parameters (learned policy) or NULL (heuristic) */
    uint32_t spin_limit;         /* polls before a team barrier gives up */
} albd_batch_args;

/* workgroups that can be resident at once for albd_batch_launch */
ALBD_HIDDEN int albd_batch_capacity(albd_ctx *c, int *workgroups);
/* one launch of `workgroups` co-resident workgroups (at most
 * albd_batch_capacity) over all teams */
ALBD_HIDDEN int albd_batch_launch(albd_ctx *c, const albd_batch_args *a, int32_t workgroups);
/* The batched BFS's fixed device scratch (albba_batch_run_noalloc): a
 * file-scope device array of ALBD_BATCH_SCRATCH_BYTES per device.  _init
 * looks up its address on `device` (current) once; albd_ctx_create calls it.
 * albd_batch_scratch returns that address for the context's device. */
#define ALBD_BATCH_SCRATCH_BYTES ((size_t)512 << 10)
ALBD_HIDDEN int albd_batch_scratch_init(int device);
ALBD_HIDDEN int albd_batch_scratch(albd_ctx *c, void **ptr, size_t *bytes);
/* ticks per millisecond of the device clock read by the batch kernel */
ALBD_HIDDEN double albd_clock_ticks_per_ms(albd_ctx *c);

/* ---- selector MLP (mlp.hip); every pointer is device memory ---- */
#define ALBD_MLP_SCRATCH_PER_SAMPLE (32 + 32 + 4 + 32 + 32 + 1)
/* samples per partial sum of the gradient */
#define ALBD_MLP_CHUNK 64
/* floats of scratch albd_mlp_loss_grad needs for ns samples: the per-sample
 * activations and errors, then one partial gradient (+ loss) per chunk */
#define ALBD_MLP_SCRATCH_FLOATS(ns)                                                                \
    ((size_t)(ns) * ALBD_MLP_SCRATCH_PER_SAMPLE +                                                  \
     (((size_t)(ns) + ALBD_MLP_CHUNK - 1) / ALBD_MLP_CHUNK) * (ALBBA_MLP_NPARAM + 1))
/* out[ns*4] = MLP(x[ns*12]) */
ALBD_HIDDEN int albd_mlp_forward(albd_ctx *c, const float *params, const float *x, int32_t ns,
                                 float *out);
/* loss[0] = mean squared error, grad[1604] = d loss / d params;
 * scratch holds ALBD_MLP_SCRATCH_FLOATS(ns) floats */
ALBD_HIDDEN int albd_mlp_loss_grad(albd_ctx *c, const float *params, const float *x,
                                   const float *t, int32_t ns, float *scratch, float *grad,
                                   float *loss);
ALBD_HIDDEN int albd_mlp_adam(albd_ctx *c, float *params, const float *grad, float *m, float *v,
                              int32_t np, float lr, float b1, float b2, float eps, int32_t step);

#ifdef __cplusplus
}
#endif
#endif
