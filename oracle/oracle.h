/*
 * CPU oracle for the ALBBA BFS path.  TEST INFRASTRUCTURE ONLY: used by
 * tests/, by the smoke check and by the CPU-baseline leg of bench.py as the
 * checker; the product (libalbba.so) never links or calls it.
 *
 * Graphs are CSR "working matrices" W (bfs/src/main.c:117-126): row v lists
 * the in-neighbours of v, so a BFS discovers v from u when u is in row v.
 * Every function here is a plain restatement written for clarity, not speed,
 * except where a config size needs linear time.
 */
#ifndef ALBBA_ORACLE_H
#define ALBBA_ORACLE_H
#include <stdint.h>

/* SELL-C-sigma exactly as bfs/src/sparse_matrix.c:523-623 +
 * bfs/src/sorting_utils.c:173-253 + bfs/src/utils.c:198-241 build it, for any
 * C in [1,256]: rows stably sorted by length (descending) in sigma windows,
 * column-major packing without padding, vop_lengths = active lanes - 1,
 * per-column min/max column index.  Arrays are malloc'ed by the callee. */
typedef struct orc_sell {
    int32_t n, C, sigma, nslices, nvops;
    int64_t nnz;
    int32_t *row_order;      /* [n]          */
    int32_t *slice_widths;   /* [nslices]    */
    int32_t *vop_pointers;   /* [nslices+1]  */
    uint8_t *vop_lengths;    /* [nvops]      */
    int32_t *slice_pointers; /* [nslices+1]  */
    int32_t *column_indices; /* [nnz]        */
    int32_t *max_index;      /* [nvops]      */
    int32_t *min_index;      /* [nvops]      */
} orc_sell;

int orc_sell_build(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t C, int32_t sigma,
                   orc_sell *out);
void orc_sell_free(orc_sell *s);

/* Textbook queue BFS over out-edges (the transpose of W), then the min-id
 * parent of every reached v: min{u in row v of W : level[u] = level[v]-1}.
 * level[] = -1 and parent[] = -1 for unreached vertices; parent[root] = root.
 * Returns the number of levels (max level + 1). */
int orc_bfs_queue(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                  int32_t *level, int32_t *parent);

/* bfs/src/bfs_ref.h:26-67 + bfs/src/utils.c:305-334: level-synchronous BFS as
 * a dense SpMV y = W x followed by the visited-mask check, from `root`.
 * frontier[iter] receives the frontier size entering iteration iter (the
 * reference's num_vex_ref); level[] is the iteration a vertex became new.
 * Returns iter_ref (number of iterations).  O(L * m): small graphs only. */
int orc_bfs_dense_ref(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                      int32_t *frontier, int32_t *level);

/* Scalar + OpenMP restatement of ALBBA's level loop
 * (bfs/src/bfs_SELLCSv2.c:126-765 with bfs/src/bfs_par_kernels.c:128-198,
 * 566-656, 696-797) over an orc_sell layout: pull bitmap SpMV with visited
 * slice bypass when |F| > (int)(0.04 n), otherwise the sparse-frontier
 * SpMSpV with slice-range and column-range bypass; frontier format
 * conversions between a dense int mask and a sparse id list, with the
 * dense2sparse offsets taken from the scanned counts of the same id ranges
 * (the reference's offsets come from permuted slice counts, which can drop
 * vertices).  Extended with the root argument, level[] and the min-id
 * parent[].  `tasks` = number of slice ranges balanced by stored entries
 * (the reference fixes 8).  Per-level outputs (may be NULL): frontier
 * sizes, mode (0 = SpMSpV, 1 = SpMV), slice bypass and column bypass
 * counts; capacity = their length.  Returns the number of levels or -1. */
int orc_albba_bfs(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                  int32_t *frontier, int32_t *mode, int64_t *bypass_slice, int64_t *bypass_col,
                  int32_t capacity);

/* orc_albba_bfs stopped after at most max_levels levels (a bounded sample
 * of a long BFS, for timing); returns the levels run. */
int orc_albba_bfs_levels(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                         int32_t *frontier, int32_t *mode, int64_t *bypass_slice, int64_t *bypass_col,
                         int32_t capacity, int32_t max_levels);

/* orc_albba_bfs with the reference's dense2sparse offsets (the F9 defect)
 * in place of the corrected ones, to show where it loses frontier vertices;
 * meaningful at C = 256, the slice width the reference's `<< 8` assumes. */
int orc_albba_bfs_f9(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                     int32_t *frontier, int32_t *mode, int32_t capacity);

/* Graph500-style BFS tree check on W: 0 if valid, else a negative code
 *  -1 root wrong, -2 parent of a reached vertex unreached or level gap != 1,
 *  -3 tree edge missing from the graph, -4 an edge skips a level,
 *  -5 reached vertex count inconsistent, -6 parent not the min-id choice. */
int orc_validate(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                 const int32_t *level, const int32_t *parent);

#endif
