/*
 * albba.h -- public C ABI of the MI355X-native ALBBA linear-algebra BFS.
 *
 * Two groups of entry points live here:
 *
 *  1. The reference's own C entry points for the graph-load -> SELL-C-sigma
 *     build -> BFS path, with the reference's names, argument meaning and
 *     (void / exit-free) error behaviour, so that a caller of the reference
 *     relinks against libalbba.so without source changes:
 *       read_Dmatrix            bfs/src/biio.h:127
 *       read_Dmatrix_convert    bfs/src/biio.h:98
 *       binary_read_csr         bfs/src/biio.h:21
 *       binary_write_csr        bfs/src/biio.h:73
 *       mmio_info / mmio_data   bfs/src/mmio_highlevel.h:11,140
 *       convert_csr_to_csc      bfs/src/sparse_matrix.c:230
 *       sellcs_preprocess       bfs/src/bfs_SELLCSv2.c:34
 *       run_custom_bfs_test     bfs/src/bfs.h:13 (bfs_SELLCSv2.c:126)
 *       free_sellcs_struct      bfs/src/sparse_matrix.c:943
 *
 *  2. The albba_* API: status-returning functions (0 = ok, negative =
 *     ALBBA_E*), opaque handles that own device memory, no globals touched.
 *     A context is bound to one HIP device and one HIP stream and must not be
 *     shared between host threads.
 *
 * All integers crossing the boundary are plain C types; no HIP or torch types
 * appear in any signature.
 */
#ifndef ALBBA_H
#define ALBBA_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALBBA_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Reference-compatible types (bfs/src/sparse_matrix.h:18-98)               */
/* ------------------------------------------------------------------------ */
typedef double elem_t;

typedef struct SparseMatrixCSR_STRUCT {
    char *name;
    elem_t *values;
    int *column_indices;
    int *row_pointers;
    int nrows;
    int ncolumns;
    int nnz;
    int isSymmetric;
} SparseMatrixCSR;

typedef struct SparseMatrixCSC_STRUCT {
    char *name;
    elem_t *values;
    int *row_indices;
    int *column_pointers;
    int nrows;
    int ncolumns;
    int nnz;
} SparseMatrixCSC;

/* Same field list as bfs/src/sparse_matrix.h:43-63.  sellcs_preprocess fills
 * it exactly as the reference does (DFC layout, no padding, column indices in
 * vertex-id space, vop_lengths = active lanes - 1) for any C in [1,256]. */
typedef struct SparseMatrixSELLCS_STRUCT {
    char *name;
    elem_t *values;
    int *column_indices;
    int nrows;
    int ncolumns;
    int nnz;
    int C;
    int sigma;
    int nslices;
    int *slice_widths;
    int *slice_pointers;
    int *row_order;
    uint8_t *vop_lengths;
    int *vop_pointers;
    int *slice_widths_pointers;
    int *max_index;
    int *min_index;
} SparseMatrixSELLCS;

/* Reference globals read by the retained entry points
 * (bfs/src/bfs_SELLCSv2.c:29-31, bfs/src/main.c:14-22). */
extern ALBBA_API int sigma_window;   /* sigma used by run_custom_bfs_test, default 16384 */
extern ALBBA_API int perf_analysis;  /* 1: fill iter_times[] (always filled here)        */
extern ALBBA_API int verify;         /* 1: fill frontier[]   (always filled here)        */

/* bfs/src/biio.h:21 -- 0 on success, -1 on a missing file or short read. */
ALBBA_API int binary_read_csr(int *row, int *col, int *nnz, int **row_ptr,
                              int **col_idx, double **val, char *filename);
/* bfs/src/biio.h:73 */
ALBBA_API void binary_write_csr(int row, int col, int nnz, int *row_ptr,
                                int *col_idx, double *val, char *filename);
/* bfs/src/biio.h:98 -- loads the .cbd sibling, else parses .mtx and writes it. */
ALBBA_API void read_Dmatrix_convert(int *row, int *col, int *nnz, int **row_ptr,
                                    int **col_idx, double **val, int *isSymmetric,
                                    char *filename);
/* bfs/src/biio.h:127 -- dispatch on ".cbd" / ".mtx"; anything else prints
 * "File type unsupport" and leaves the outputs untouched. */
ALBBA_API void read_Dmatrix(int *row, int *col, int *nnz, int **row_ptr,
                            int **col_idx, double **val, int *isSymmetric,
                            char *filename);
/* bfs/src/mmio_highlevel.h:11 -- 0 ok, -1 open failure, -2 bad banner, -4 bad size line. */
ALBBA_API int mmio_info(int *m, int *n, int *nnz, int *isSymmetric, char *filename);
/* bfs/src/mmio_highlevel.h:140 */
ALBBA_API int mmio_data(int *csrRowPtr, int *csrColIdx, double *csrVal, char *filename);
/* bfs/src/sparse_matrix.c:230 -- rows of the CSC come out ascending. */
ALBBA_API void convert_csr_to_csc(const SparseMatrixCSR *csr, SparseMatrixCSC *csc);
/* bfs/src/bfs_SELLCSv2.c:34 -- caller presets out->C and out->sigma. */
ALBBA_API void sellcs_preprocess(const SparseMatrixCSR *csr, SparseMatrixSELLCS *out,
                                 bool sort_flag);
/* bfs/src/sparse_matrix.c:943 */
ALBBA_API void free_sellcs_struct(SparseMatrixSELLCS *sellcs);
/* bfs/src/bfs.h:13 -- BFS from vertex 0 on the GPU.  Fills frontier[level]
 * (frontier size per level), *num_iter, *num_vex, *bfs_time (ms, level loop
 * only), *conv_time (ms, SELL build) and iter_times[level] (ms); sets *nnzx
 * to 0 at the end as the reference loop does.  bfs_x is not touched. */
ALBBA_API void run_custom_bfs_test(const SparseMatrixCSR *matrix, elem_t *bfs_x,
                                   int *colidx_x, int *nnzx, int *frontier,
                                   int *num_vex, double *conv_time,
                                   double *iter_times, double *bfs_time,
                                   int *num_iter, bool bit_flag, bool sort_flag);

/* ------------------------------------------------------------------------ */
/* albba_* API                                                               */
/* ------------------------------------------------------------------------ */
enum {
    ALBBA_OK = 0,
    ALBBA_EINVAL = -1,     /* bad argument (root out of range, bad C/sigma ...) */
    ALBBA_ENOMEM = -2,     /* host allocation failed                           */
    ALBBA_EIO = -3,        /* file missing, short read, parse error            */
    ALBBA_EHIP = -4,       /* a HIP call failed (message: albba_last_error)    */
    ALBBA_ENODEV = -5,     /* no HIP device / library built without kernels    */
    ALBBA_ERANGE = -6,     /* n or nnz beyond the int32 limits of the format   */
    ALBBA_EDEVMEM = -7     /* device allocation failed or over the byte cap    */
};

/* Host CSR with 64-bit row pointers (the working matrix of the BFS). */
typedef struct albba_csr {
    int32_t n;            /* rows == columns                                 */
    int64_t nnz;          /* stored directed entries                         */
    int64_t *row_ptr;     /* [n+1]                                           */
    int32_t *col;         /* [nnz]                                           */
    int32_t symmetric;    /* 1 if A == A^T (generated graphs are)            */
} albba_csr;

/* Device-ready SELL-C-sigma layout (host copy).  Row p of the layout is
 * vertex row_order[p]; column indices are stored as POSITIONS (pos_of[]
 * applied) and each row lists its neighbours in ascending original id, so
 * the first frontier hit in a row scan is the min-id parent.  Column v of
 * slice s holds act = vop_lengths[vop]+1 entries (lanes 0..act-1), padded to
 * a multiple of `pad` entries with -1. */
typedef struct albba_sellcs {
    int32_t n, C, sigma, pad, nslices;
    int64_t nnz;              /* unpadded entries                           */
    int64_t padded_nnz;       /* storage length of column_indices           */
    int64_t nvops;            /* sum of slice widths                        */
    int32_t *row_order;       /* [n]  position -> vertex id                 */
    int32_t *pos_of;          /* [n]  vertex id -> position                 */
    int32_t *slice_widths;    /* [nslices]                                  */
    int64_t *slice_pointers;  /* [nslices+1] padded element offsets         */
    int64_t *vop_pointers;    /* [nslices+1]                                */
    uint8_t *vop_lengths;     /* [nvops] active lanes - 1                   */
    int32_t *column_indices;  /* [padded_nnz]                               */
    int32_t *row_lengths;     /* [n] in-degree per position                 */
    /* push (out-edge) CSR over positions: transpose of the layout's matrix */
    int64_t *push_ptr;        /* [n+1]                                      */
    int32_t *push_col;        /* [nnz]                                      */
    /* row-major copy of the layout's rows over positions (in-neighbours in
     * ascending id), read by the pull for long rows; NULL when the matrix is
     * symmetric, where the push CSR holds the same rows */
    int64_t *pull_ptr;        /* [n+1] or NULL                              */
    int32_t *pull_col;        /* [nnz] or NULL                              */
    int32_t order_flags;      /* ALBBA_SELL_* row-order flags it was built with */
} albba_sellcs;

typedef struct albba_ctx albba_ctx;       /* one device + one stream           */
typedef struct albba_graph albba_graph;   /* device-resident graph + BFS state */
typedef struct albba_selector albba_selector; /* learned direction selector    */

/* Kernel choices of one BFS level (the selector's classes). */
enum {
    ALBBA_CHOICE_PUSH = 0,       /* push SpMSpV from the frontier queue           */
    ALBBA_CHOICE_PULL = 1,       /* pull SpMV over every slice                     */
    ALBBA_CHOICE_PULL_SB = 2,    /* pull over the compacted unvisited-slice list   */
    ALBBA_CHOICE_PULL_CB = 3,    /* pull with frontier-word column bypass (LDS)    */
    ALBBA_NUM_CHOICES = 4
};

/* Direction switch policies. */
enum {
    ALBBA_SWITCH_HEURISTIC = 0,  /* reference rule: pull iff |F| > (int)(0.04 n)   */
    ALBBA_SWITCH_LEARNED = 1,    /* MLP selector, heuristic as fallback            */
    ALBBA_SWITCH_FIXED = 2,      /* always `fixed_choice`                          */
    ALBBA_SWITCH_SCRIPT = 3,     /* per-level choices from `script`                */
    ALBBA_SWITCH_BEST = 4        /* time every choice per level, commit the fastest */
};

#define ALBBA_NUM_FEATURES 12

typedef struct albba_bfs_opts {
    int32_t switch_policy;       /* ALBBA_SWITCH_*                                 */
    int32_t fixed_choice;        /* for ALBBA_SWITCH_FIXED                         */
    const int32_t *script;       /* for ALBBA_SWITCH_SCRIPT, one choice per level  */
    int32_t script_len;          /* levels past the end use the heuristic          */
    const albba_selector *selector; /* for ALBBA_SWITCH_LEARNED                    */
    int32_t time_levels;         /* 1: record per-level kernel times (events)      */
    int32_t timing_reps;         /* ALBBA_SWITCH_BEST / sampling: reps per choice  */
    /* Heuristic and learned policies: when a level pushes and its frontier has
     * at most this many out-edges, after two such levels in a row (a slowly
     * growing frontier), the following levels run inside one persistent push
     * kernel (one grid barrier per level, no host round trip) until the
     * frontier outgrows the cap or the heuristic's push threshold (0
     * disables; default 65536).  Results are identical either way. */
    int64_t small_frontier_edges;
    int32_t small_frontier_blocks; /* workgroups of that kernel (0: default)    */
    int32_t reserved_;
} albba_bfs_opts;

/* Per-level record; arrays are caller-owned with `capacity` entries. */
typedef struct albba_bfs_stats {
    int32_t num_levels;          /* levels run (iterations of the level loop)     */
    int32_t capacity;            /* entries available in the arrays below         */
    int64_t reached;             /* vertices with level >= 0                      */
    int64_t edges_in_component;  /* sum of out-degree over reached vertices       */
    double bfs_ms;               /* level loop incl. init, HIP events             */
    int32_t *frontier;           /* [capacity] |F| entering each level (or NULL)  */
    int32_t *choice;             /* [capacity] kernel choice per level (or NULL)  */
    float *level_ms;             /* [capacity] kernel time per level (or NULL)    */
    float *features;             /* [capacity*12] selector features (or NULL)     */
    float *choice_ms;            /* [capacity*4] all-choices timings (BEST only)  */
    int64_t *entries;            /* [capacity] pull: column entries gathered;
                                    push: frontier edges expanded (or NULL)       */
    int64_t *words;              /* [capacity] pull: position words scanned;
                                    push: queue length (or NULL)                  */
    int64_t *bypassed;           /* [capacity] pull: 64-row words not scanned
                                    because every row in them is visited or
                                    empty (the reference's slice bypass,
                                    bfs/src/bfs_par_kernels.c:152-153); push: 0
                                    (or NULL)                                     */
    float *kernel_ms;            /* [capacity] main kernel only (pull or push
                                    expansion), needs time_levels (or NULL)       */
    int64_t *col_bypassed;       /* [capacity] pull with column bypass: column
                                    entries whose frontier word was empty, so
                                    their gather was skipped (the reference's
                                    width bypass, bfs_par_kernels.c:613,628-631,
                                    counts skipped columns); otherwise 0 (or NULL) */
    int64_t *swept;              /* [capacity] pull: stored entries of the words
                                    it scanned, the column entries a full scan
                                    of those slices reads (the reference's
                                    SpMV, bfs/src/bfs_par_kernels.c:128-198,
                                    scans every column of a slice that is not
                                    bypassed); needs time_levels and a list
                                    pull (the slice- or column-bypass choice),
                                    else -1 (or NULL)                             */
} albba_bfs_stats;

/* Last error message of the calling thread (static storage). */
ALBBA_API const char *albba_last_error(void);
ALBBA_API const char *albba_version(void);

/* ---- host graph construction ------------------------------------------- */
ALBBA_API void albba_csr_free(albba_csr *g);
/* .cbd / .mtx through read_Dmatrix semantics, then the working matrix of the
 * reference (transpose, rows ascending; bfs/src/main.c:117-126). */
ALBBA_API int albba_read_graph(const char *path, albba_csr *out);
/* Write a CSR as .cbd (bfs/src/biio.h:73 layout, values = 1.0). */
ALBBA_API int albba_write_cbd(const char *path, const albba_csr *g);
/* Seeded generators.  All permute vertex ids with perm_seed, symmetrize,
 * drop self loops and duplicates, and return rows sorted ascending. */
ALBBA_API int albba_gen_rmat(int scale, int edgefactor, double a, double b, double c,
                             uint64_t seed, uint64_t perm_seed, albba_csr *out);
ALBBA_API int albba_gen_er(int32_t n, double mean_degree, uint64_t seed,
                           uint64_t perm_seed, albba_csr *out);
ALBBA_API int albba_gen_grid(int32_t rows, int32_t cols, double p_remove,
                             uint64_t seed, uint64_t perm_seed, albba_csr *out);
/* Roots: `count` distinct vertices with degree > 0, drawn with root_seed.
 * If largest_cc != 0 the roots are drawn from the largest connected
 * component (of the symmetric closure).  Returns the number written. */
ALBBA_API int albba_pick_roots(const albba_csr *g, int count, uint64_t root_seed,
                               int largest_cc, int32_t *roots);
/* The same graphs generated on the device (bit for bit the CSR of the
 * albba_gen_* function of that kind), the layout built there as by
 * albba_graph_build.  host_copy, if not NULL, receives the CSR (free with
 * albba_csr_free).  The id permutation is drawn on the host (Fisher-Yates is
 * sequential); edges, permutation, symmetrization, self-loop removal, sort
 * and dedup run on the device.  No reference counterpart: the reference only
 * loads files (bfs/src/main.c:88-99). */
enum { ALBBA_GEN_RMAT = 0, ALBBA_GEN_ER = 1, ALBBA_GEN_GRID = 2 };
typedef struct albba_gen_params {
    int32_t kind;                 /* ALBBA_GEN_*                                */
    int32_t scale, edgefactor;    /* RMAT                                       */
    double a, b, c;               /* RMAT quadrant probabilities                */
    int32_t n;                    /* ER vertices                                */
    int32_t rows, cols;           /* grid                                       */
    double mean_degree;           /* ER                                         */
    double p_remove;              /* grid                                       */
    uint64_t seed, perm_seed;
} albba_gen_params;
typedef struct albba_gen_report {
    float perm_ms;   /* the permutation on the host                         */
    float gen_ms;    /* edges to CSR on the device                          */
    float build_ms;  /* the layout on the device                            */
    float total_ms;
} albba_gen_report;
ALBBA_API int albba_graph_generate(albba_ctx *ctx, const albba_gen_params *p, int C, int sigma, int pad,
                                   albba_graph **out, albba_csr *host_copy, albba_gen_report *report);
/* Counter-based RNG used by every generator: splitmix64 finaliser of
 * seed + (idx+1) * 0x9E3779B97F4A7C15. */
ALBBA_API uint64_t albba_rng_u64(uint64_t seed, uint64_t idx);

/* ---- SELL-C-sigma ------------------------------------------------------- */
/* C in {16,32,64,128} (rows per slice), sigma a multiple of C (or >= n),
 * pad in {1,16,32} entries per column.  Builds the push CSR as well. */
ALBBA_API int albba_sellcs_build(const albba_csr *g, int C, int sigma, int pad,
                                 albba_sellcs **out);
/* Row-order flags of albba_sellcs_build_ex / albba_graph_build_ex.
 * ALBBA_SELL_SORT_FIRST_COL: before the sigma-window sort, order the rows
 * by their first column index, descending and stable (the reference's
 * `-s 1`, sellcs_preprocess(..., sort_flag = true): bfs/src/bfs_SELLCSv2.c:
 * 40-45, sort_by_1st_colidx bfs/src/sorting_utils.c:255-281, then the window
 * sort of get_order_by_row_size_radix_sorted :282-297).  A row's first column
 * is its smallest in-neighbour id; an empty row takes the next stored entry
 * (the next non-empty row's first column), 0 past the last entry. */
enum { ALBBA_SELL_SORT_FIRST_COL = 1 };
/* albba_sellcs_build with row-order flags (ALBBA_SELL_*); flags = 0 is
 * albba_sellcs_build. */
ALBBA_API int albba_sellcs_build_ex(const albba_csr *g, int C, int sigma, int pad, int flags,
                                    albba_sellcs **out);
ALBBA_API void albba_sellcs_free(albba_sellcs *s);

/* ---- device ------------------------------------------------------------- */
ALBBA_API int albba_device_count(int *count);
ALBBA_API int albba_ctx_create(int device, albba_ctx **out);
ALBBA_API void albba_ctx_destroy(albba_ctx *ctx);
ALBBA_API int albba_ctx_synchronize(albba_ctx *ctx);
/* Upload a SELL layout (+ push CSR) and allocate the BFS state on the
 * context's device. */
ALBBA_API int albba_graph_upload(albba_ctx *ctx, const albba_sellcs *s, albba_graph **out);
/* Build the layout on the device from a host CSR -- row lengths, the
 * sigma-window sort, slices, active lanes, padded columns, the push and pull
 * CSRs -- and allocate the BFS state: the same device arrays, bit for bit, as
 * albba_sellcs_build followed by albba_graph_upload (same arguments and
 * errors; rows out of ascending order are sorted on the device).  The
 * reference builds on the host, timed as conv_time (bfs/src/bfs_SELLCSv2.c:
 * 34-59 sellcs_preprocess, called at :161-168).  report may be NULL. */
typedef struct albba_build_report {
    float upload_ms;     /* CSR host -> device copy                          */
    float build_ms;      /* every build step on the device                   */
    int32_t rows_sorted; /* 1: the rows were already in ascending id order  */
    int32_t reserved_;
} albba_build_report;
ALBBA_API int albba_graph_build(albba_ctx *ctx, const albba_csr *g, int C, int sigma, int pad,
                                albba_graph **out, albba_build_report *report);
/* albba_graph_build with row-order flags (ALBBA_SELL_*): the same device
 * arrays as albba_sellcs_build_ex with the same flags + albba_graph_upload. */
ALBBA_API int albba_graph_build_ex(albba_ctx *ctx, const albba_csr *g, int C, int sigma, int pad,
                                   int flags, albba_graph **out, albba_build_report *report);
/* Load a .cbd file (bfs/src/biio.h:21-71 layout) straight into device
 * memory -- the row pointers and columns stream through pinned buffers, file
 * reads overlapping the copies; the values are not read -- then form the
 * reference's working matrix W = A^T with rows ascending on the device
 * (bfs/src/main.c:117-126, the DIRECTED build; the .cbd path sets no symmetry
 * flag) and build the layout there (as albba_graph_build).  The same device
 * graph as albba_read_graph + albba_sellcs_build + albba_graph_upload.
 * .mtx files are text: load them with albba_read_graph.  report may be NULL. */
typedef struct albba_load_report {
    float read_ms;   /* file -> device, overlapped                          */
    float prep_ms;   /* checks and the transpose on the device              */
    float build_ms;  /* the layout on the device                            */
    float total_ms;  /* the whole call                                      */
} albba_load_report;
ALBBA_API int albba_graph_load(albba_ctx *ctx, const char *path, int C, int sigma, int pad,
                               albba_graph **out, albba_load_report *report);
/* Copy a device graph's layout to the host as an albba_sellcs (free with
 * albba_sellcs_free); row_lengths and push_ptr cover nslices * C positions. */
ALBBA_API int albba_graph_layout(albba_ctx *ctx, const albba_graph *g, albba_sellcs **out);
ALBBA_API void albba_graph_free(albba_graph *g);
ALBBA_API int64_t albba_graph_device_bytes(const albba_graph *g);
/* Vertices and stored entries of a device graph. */
ALBBA_API int albba_graph_size(const albba_graph *g, int32_t *n, int64_t *nnz);

/* Device copy bandwidth: copies `bytes` (rounded down to 16) between two
 * fresh device buffers `reps` times on the context's stream and returns the
 * mean time of one copy in *ms.  No counterpart in the reference: it gives
 * the achievable HBM bandwidth next to the BFS kernels' figures and is the
 * known-size kernel that memory traffic counters are scaled by. */
ALBBA_API int albba_bench_copy(albba_ctx *ctx, int64_t bytes, int reps, float *ms);

ALBBA_API void albba_bfs_opts_default(albba_bfs_opts *o);
/* BFS from vertex `root` (original id).  Results stay on the device until
 * albba_bfs_results copies them out.  stats may be NULL. */
ALBBA_API int albba_bfs(albba_ctx *ctx, albba_graph *g, int32_t root,
                        const albba_bfs_opts *opts, albba_bfs_stats *stats);
/* BFS from roots[0..nroots) in turn on one graph (Graph500's sequence of
 * roots; the reference runs one root per process, bfs/src/main.c:196-213).
 * Each BFS's first kernels are queued while the previous one's last kernels
 * still run: one synchronisation at the end instead of one per root.  opts:
 * one set for every root (nopts = 1) or one per root (nopts = nroots); stats:
 * nroots entries, or NULL.  The graph's results are the last root's. */
ALBBA_API int albba_bfs_multi(albba_ctx *ctx, albba_graph *g, const int32_t *roots, int32_t nroots,
                              const albba_bfs_opts *opts, int32_t nopts, albba_bfs_stats *stats);
/* level[v]: BFS distance or -1; parent[v]: min-id parent, root for the root,
 * -1 if unreached.  Either pointer may be NULL. */
ALBBA_API int albba_bfs_results(albba_ctx *ctx, albba_graph *g, int32_t *level,
                                int32_t *parent);

/* ---- batched BFS ---------------------------------------------------------
 * Many BFS instances (graphs and roots of uneven size) in one persistent
 * launch: the resident workgroups are split into teams, each team runs the
 * whole level loop of its instances with the direction of every level chosen
 * on the device.  No counterpart in the reference, which runs one BFS per
 * process; results follow the same level/min-id-parent contract as
 * albba_bfs. */
typedef struct albba_batch albba_batch;

enum { ALBBA_MAP_LANE_ROW = 0, ALBBA_MAP_WAVE_ROW = 1 }; /* pull: a lane or a wave per row */

typedef struct albba_batch_result {
    int32_t num_levels;
    int32_t error;               /* nonzero: the instance did not finish       */
    int64_t reached;             /* vertices with a level                      */
    int64_t edges_in_component;  /* out-edges summed over the reached vertices */
    double ms;                   /* the instance's run time on the device clock */
    double done_ms;              /* from the first instance's start to its end */
    int32_t push_levels;
    int32_t pull_levels;
    int32_t team_size;           /* workgroups that ran it                     */
    int32_t team;                /* team index                                 */
} albba_batch_result;

/* Device scratch for one BFS on each of graphs[0..count) (a graph may appear
 * more than once); the graphs must outlive the batch. */
ALBBA_API int albba_batch_create(albba_ctx *ctx, albba_graph *const *graphs, int32_t count,
                                 albba_batch **out);
ALBBA_API void albba_batch_free(albba_batch *b);
/* BFS from roots[i] (vertex id) on instance i, all in one launch.  opts: the
 * heuristic (default) or learned policy; mapping[i] (NULL: lane per row)
 * picks instance i's pull mapping; workgroups 0 uses every resident
 * workgroup.  results[count] and total_ms (launch time) may be NULL.
 * Team sizes: the first run splits the workgroups by each instance's share of
 * the stored entries; every later run of the batch sizes them from that
 * first run's level count and run time of each instance (a high-diameter
 * instance, whose levels are latency-bound, gets the team that brings it
 * under the shortest feasible makespan); albba_batch_forget drops that
 * history. */
ALBBA_API int albba_batch_run(albba_batch *b, const int32_t *roots, const albba_bfs_opts *opts,
                              const int32_t *mapping, int32_t workgroups,
                              albba_batch_result *results, float *total_ms);
/* albba_batch_run without any allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors as albba_batch_run, except that the run's device tables go to a
 * fixed device array of the library (512 KiB per device, set up by
 * albba_ctx_create) instead of memory allocated per run, and batches of more
 * than ALBBA_BATCH_NOALLOC_MAX instances are rejected (ALBBA_ERANGE).  Every
 * call on a device shares that array: calls that overlap on one device race
 * (one stream per context; do not run two at once). */
#define ALBBA_BATCH_NOALLOC_MAX 1024
ALBBA_API int albba_batch_run_noalloc(albba_batch *b, const int32_t *roots, const albba_bfs_opts *opts,
                                      const int32_t *mapping, int32_t workgroups,
                                      albba_batch_result *results, float *total_ms);
/* Forget the previous run's per-instance level counts and times: the next
 * run splits the workgroups by the instances' shares of the stored entries. */
ALBBA_API void albba_batch_forget(albba_batch *b);
/* level/parent of instance i after albba_batch_run (either may be NULL) */
ALBBA_API int albba_batch_results(albba_batch *b, int32_t i, int32_t *level, int32_t *parent);

/* ---- autotuner -------------------------------------------------------------
 * Picks the layout of one graph by measurement: every (C, sigma) candidate is
 * built, uploaded and timed on the sample roots through albba_bfs with the
 * given options (mean BFS time after one warm-up run), then both pull
 * mappings of the batched BFS are timed on the fastest layout (one instance
 * per sample root).  The reference fixes C = 256 and takes sigma from -w
 * (bfs/src/bfs_SELLCSv2.c:161-163); it has no tuner. */
typedef struct albba_tune_config {
    int32_t C;          /* slice height                                  */
    int32_t sigma;      /* sorting window (>= n: one global window)      */
    int32_t mapping;    /* batched pull: ALBBA_MAP_*                     */
    int32_t status;     /* 0, or the error code of this candidate       */
    double ms;          /* mean BFS time over the sample roots           */
    double batch_ms;    /* batched launch time with this mapping         */
} albba_tune_config;

/* candidates == NULL: C in {32, 64, 128} x sigma in {C, 256, 4096, n}
 * (ncand ignored, 12 results).  results[ncand or 12] may be NULL. */
ALBBA_API int albba_autotune(albba_ctx *ctx, const albba_csr *g, const int32_t *roots,
                             int32_t nroots, const albba_bfs_opts *opts,
                             const albba_tune_config *candidates, int32_t ncand,
                             albba_tune_config *results, albba_tune_config *best);

/* ---- learned direction selector ----------------------------------------- */
#define ALBBA_MLP_IN 12
#define ALBBA_MLP_H1 32
#define ALBBA_MLP_H2 32
#define ALBBA_MLP_OUT 4
#define ALBBA_MLP_NPARAM (ALBBA_MLP_IN * ALBBA_MLP_H1 + ALBBA_MLP_H1 + \
                          ALBBA_MLP_H1 * ALBBA_MLP_H2 + ALBBA_MLP_H2 + \
                          ALBBA_MLP_H2 * ALBBA_MLP_OUT + ALBBA_MLP_OUT)

ALBBA_API int albba_selector_create(albba_ctx *ctx, uint64_t init_seed, albba_selector **out);
ALBBA_API void albba_selector_free(albba_selector *s);
ALBBA_API int albba_selector_get_params(const albba_selector *s, float *params);
ALBBA_API int albba_selector_set_params(albba_selector *s, const float *params);
ALBBA_API int albba_selector_load(albba_selector *s, const char *path);
ALBBA_API int albba_selector_save(const albba_selector *s, const char *path);
/* The learned policy lists the next level's frontier queue after a push
 * level only while that level's frontier out-edges are at most ratio * n
 * (the next frontier is at most that large; a pull next needs no queue, a
 * push after an unlisted level lists the bitmap).  The ratio is part of the
 * trained selector (chosen on validation graphs from the levels where the
 * selector pushed twice in a row) and travels in its file; 0 (unset, and
 * every version-1 file) takes the heuristic's push threshold (int)(0.04 n). */
ALBBA_API int albba_selector_set_queue_ratio(albba_selector *s, float ratio);
ALBBA_API float albba_selector_queue_ratio(const albba_selector *s);
/* Train on the GPU: x[nsamples*12] features, t[nsamples*4] targets
 * (relative log kernel times).  Adam, full batch.  Writes the final loss. */
ALBBA_API int albba_selector_train(albba_selector *s, const float *x, const float *t,
                                   int32_t nsamples, int32_t steps, float lr,
                                   float *final_loss);
/* One forward/backward pass on the GPU: loss and d(loss)/d(params). */
ALBBA_API int albba_selector_loss_grad(albba_selector *s, const float *x, const float *t,
                                       int32_t nsamples, float *loss, float *grad);
/* Batched inference on the GPU: out[nsamples*4] predicted scores. */
ALBBA_API int albba_selector_predict(albba_selector *s, const float *x, int32_t nsamples,
                                     float *out);

#ifdef __cplusplus
}
#endif
#endif /* ALBBA_H */
