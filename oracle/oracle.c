/*
 * CPU oracle for the ALBBA BFS path -- TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Nothing in libalbba.so links, loads or calls this file.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* SELL-C-sigma (bfs/src/sparse_matrix.c:523-623, sorting_utils.c:173-253,
 * utils.c:198-241)                                                        */

typedef struct { int32_t len, idx; } kv_t;

static int cmp_len_desc(const void *a, const void *b)
{
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->len != y->len) return x->len > y->len ? -1 : 1; /* longer first */
    return (x->idx > y->idx) - (x->idx < y->idx);           /* stable      */
}

void orc_sell_free(orc_sell *s)
{
    free(s->row_order); free(s->slice_widths); free(s->vop_pointers); free(s->vop_lengths);
    free(s->slice_pointers); free(s->column_indices); free(s->max_index); free(s->min_index);
    memset(s, 0, sizeof(*s));
}

int orc_sell_build(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t C, int32_t sigma,
                   orc_sell *s)
{
    memset(s, 0, sizeof(*s));
    if (n < 1 || C < 1 || C > 256 || sigma < 1) return -1;
    s->n = n; s->C = C; s->sigma = sigma; s->nnz = row_ptr[n];
    kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (size_t)n);
    s->row_order = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!kv || !s->row_order) return -2;
    for (int32_t i = 0; i < n; ++i) { kv[i].len = (int32_t)(row_ptr[i + 1] - row_ptr[i]); kv[i].idx = i; }
    for (int64_t lo = 0; lo < n; lo += sigma) {
        int64_t hi = lo + sigma < n ? lo + sigma : n;
        qsort(kv + lo, (size_t)(hi - lo), sizeof(kv_t), cmp_len_desc);
    }
    for (int32_t i = 0; i < n; ++i) s->row_order[i] = kv[i].idx;
    const int32_t ns = (n + C - 1) / C;
    s->nslices = ns;
    s->slice_widths = (int32_t *)malloc(sizeof(int32_t) * (size_t)ns);
    s->vop_pointers = (int32_t *)malloc(sizeof(int32_t) * ((size_t)ns + 1));
    s->slice_pointers = (int32_t *)malloc(sizeof(int32_t) * ((size_t)ns + 1));
    int64_t nv = 0;
    for (int32_t sl = 0; sl < ns; ++sl) {
        int32_t w = 0;
        for (int64_t p = (int64_t)sl * C; p < n && p < (int64_t)(sl + 1) * C; ++p) w = kv[p].len > w ? kv[p].len : w;
        s->slice_widths[sl] = w;
        s->vop_pointers[sl] = (int32_t)nv;
        nv += w;
    }
    s->vop_pointers[ns] = (int32_t)nv;
    s->nvops = (int32_t)nv;
    s->vop_lengths = (uint8_t *)malloc((size_t)(nv > 0 ? nv : 1));
    s->max_index = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    s->min_index = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    s->column_indices = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->nnz > 0 ? s->nnz : 1));
    int64_t at = 0;
    for (int32_t sl = 0; sl < ns; ++sl) {
        s->slice_pointers[sl] = (int32_t)at;
        const int64_t base = (int64_t)sl * C;
        const int32_t rows = (int32_t)((base + C <= n) ? C : n - base);
        for (int32_t j = 0; j < s->slice_widths[sl]; ++j) {
            /* active lanes of column j: rows of this slice longer than j
             * (rows are sorted, so they are lanes 0..act-1) */
            int32_t act = 0;
            for (int32_t r = 0; r < rows; ++r) act += kv[base + r].len > j;
            const int64_t v = s->vop_pointers[sl] + j;
            s->vop_lengths[v] = (uint8_t)(act - 1);
            int32_t mx = 0, mn = n;
            for (int32_t r = 0; r < act; ++r) {
                int32_t c = col[row_ptr[s->row_order[base + r]] + j];
                s->column_indices[at++] = c;
                if (c > mx) mx = c;
                if (c < mn) mn = c;
            }
            s->max_index[v] = mx;
            s->min_index[v] = mn;
        }
    }
    s->slice_pointers[ns] = (int32_t)at;
    free(kv);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* textbook BFS                                                            */

static void min_id_parents(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                           const int32_t *level, int32_t *parent)
{
    for (int32_t v = 0; v < n; ++v) {
        parent[v] = -1;
        if (level[v] <= 0) continue;
        int32_t best = -1;
        for (int64_t k = row_ptr[v]; k < row_ptr[v + 1]; ++k) {
            int32_t u = col[k];
            if (level[u] == level[v] - 1 && (best < 0 || u < best)) best = u;
        }
        parent[v] = best;
    }
    parent[root] = root;
}

int orc_bfs_queue(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                  int32_t *level, int32_t *parent)
{
    if (root < 0 || root >= n) return -1;
    /* out-edges: transpose of W */
    int64_t m = row_ptr[n];
    int64_t *tp = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *tc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!tp || !tc || !q) { free(tp); free(tc); free(q); return -2; }
    for (int64_t k = 0; k < m; ++k) tp[col[k] + 1]++;
    for (int32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    memcpy(cur, tp, sizeof(int64_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v)
        for (int64_t k = row_ptr[v]; k < row_ptr[v + 1]; ++k) tc[cur[col[k]]++] = v;
    free(cur);
    for (int32_t v = 0; v < n; ++v) level[v] = -1;
    level[root] = 0;
    int64_t head = 0, tail = 0;
    q[tail++] = root;
    int32_t maxl = 0;
    while (head < tail) {
        int32_t u = q[head++];
        for (int64_t k = tp[u]; k < tp[u + 1]; ++k) {
            int32_t v = tc[k];
            if (level[v] < 0) {
                level[v] = level[u] + 1;
                if (level[v] > maxl) maxl = level[v];
                q[tail++] = v;
            }
        }
    }
    free(tp); free(tc); free(q);
    min_id_parents(n, row_ptr, col, root, level, parent);
    return maxl + 1;
}

/* ---------------------------------------------------------------------- */
/* bfs/src/bfs_ref.h:26-67 restated with a boolean frontier: the reference's
 * doubles only matter through "!= 0", and its x holds non-negative path
 * counts, so x[i] != 0 <=> i is in the frontier.                          */

int orc_bfs_dense_ref(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                      int32_t *frontier, int32_t *level)
{
    if (root < 0 || root >= n) return -1;
    uint8_t *x = (uint8_t *)calloc((size_t)n, 1), *y = (uint8_t *)calloc((size_t)n, 1),
            *mask = (uint8_t *)calloc((size_t)n, 1);
    if (!x || !y || !mask) { free(x); free(y); free(mask); return -2; }
    for (int32_t v = 0; v < n; ++v) level[v] = -1;
    x[root] = 1;
    mask[root] = 1;
    level[root] = 0;
    int32_t nnzx = 1, iter = 0;
    do {
        for (int32_t i = 0; i < n; ++i) { /* kernel_csr_spmv */
            uint8_t acc = 0;
            for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) acc |= x[col[k]];
            y[i] = acc;
        }
        frontier[iter] = nnzx;
        int32_t cnt = 0; /* spmv_check */
        for (int32_t i = 0; i < n; ++i) {
            x[i] = 0;
            if (!mask[i] && y[i]) {
                mask[i] = 1;
                x[i] = 1;
                level[i] = iter + 1;
                ++cnt;
            }
        }
        nnzx = cnt;
        ++iter;
    } while (nnzx != 0);
    free(x); free(y); free(mask);
    return iter;
}

/* ---------------------------------------------------------------------- */
/* ALBBA level loop restated (bfs/src/bfs_SELLCSv2.c:126-765)              */

static int32_t *task_ranges(const orc_sell *s, int32_t tasks)
{
    /* bfs/src/bfs_SELLCSv2.c:61-85: greedy split of slices by stored entries */
    int32_t *tp = (int32_t *)malloc(sizeof(int32_t) * ((size_t)tasks + 1));
    int64_t total = s->slice_pointers[s->nslices];
    int64_t per = (total + tasks - 1) / tasks;
    int64_t acc = 0;
    int32_t t = 0;
    tp[t++] = 0;
    for (int32_t sl = 0; sl < s->nslices; ++sl) {
        if (acc >= per && t < tasks) { tp[t++] = sl; acc = 0; }
        acc += s->slice_pointers[sl + 1] - s->slice_pointers[sl];
    }
    for (int32_t i = t; i <= tasks; ++i) tp[i] = s->nslices;
    return tp;
}

int orc_albba_bfs(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                  int32_t *frontier_out, int32_t *mode_out, int64_t *bys_out, int64_t *byc_out,
                  int32_t capacity)
{
    return orc_albba_bfs_levels(s, root, tasks, level, parent, frontier_out, mode_out, bys_out, byc_out,
                                capacity, INT32_MAX);
}

static int albba_loop(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                      int32_t *frontier_out, int32_t *mode_out, int64_t *bys_out, int64_t *byc_out,
                      int32_t capacity, int32_t max_levels, int f9);

int orc_albba_bfs_levels(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                         int32_t *frontier_out, int32_t *mode_out, int64_t *bys_out, int64_t *byc_out,
                         int32_t capacity, int32_t max_levels)
{
    return albba_loop(s, root, tasks, level, parent, frontier_out, mode_out, bys_out, byc_out, capacity,
                      max_levels, 0);
}

int orc_albba_bfs_f9(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                     int32_t *frontier_out, int32_t *mode_out, int32_t capacity)
{
    return albba_loop(s, root, tasks, level, parent, frontier_out, mode_out, NULL, NULL, capacity, INT32_MAX,
                      1);
}

static int albba_loop(const orc_sell *s, int32_t root, int32_t tasks, int32_t *level, int32_t *parent,
                      int32_t *frontier_out, int32_t *mode_out, int64_t *bys_out, int64_t *byc_out,
                      int32_t capacity, int32_t max_levels, int f9)
{
    const int32_t n = s->n, C = s->C, ns = s->nslices;
    if (root < 0 || root >= n || tasks < 1) return -1;
    const int32_t T = (int32_t)(n * 0.04); /* bfs_SELLCSv2.c:282-284 */
    int32_t *pos_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *mask_x = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    int32_t *mask_y = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    int32_t *colidx_x = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *colidx_y = (int32_t *)malloc(sizeof(int32_t) * (size_t)n); /* per-task segments */
    int32_t *bucket = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    uint8_t *visited = (uint8_t *)calloc((size_t)ns * C, 1);   /* by position */
    int32_t *smax = (int32_t *)calloc((size_t)ns, sizeof(int32_t));
    int32_t *smin = (int32_t *)malloc(sizeof(int32_t) * (size_t)ns);
    int32_t *tp = task_ranges(s, tasks);
    int32_t *tcnt = (int32_t *)calloc((size_t)tasks + 1, sizeof(int32_t));
    int64_t *tbs = (int64_t *)calloc((size_t)tasks, sizeof(int64_t));
    int64_t *tbc = (int64_t *)calloc((size_t)tasks, sizeof(int64_t));
    if (!pos_of || !mask_x || !mask_y || !colidx_x || !colidx_y || !bucket || !visited || !smax ||
        !smin || !tp || !tcnt || !tbs || !tbc)
        return -2;
    for (int32_t p = 0; p < n; ++p) pos_of[s->row_order[p]] = p;
    /* entries the F9 offsets leave unwritten hold a visited vertex (the
     * root): a stale, harmless frontier entry rather than undefined memory */
    for (int32_t i = 0; i < n; ++i) colidx_x[i] = root;
    for (int32_t sl = 0; sl < ns; ++sl) { /* bfs_SELLCSv2.c:170-182 */
        smin[sl] = n;
        for (int32_t v = s->vop_pointers[sl]; v < s->vop_pointers[sl + 1]; ++v) {
            if (s->max_index[v] > smax[sl]) smax[sl] = s->max_index[v];
            if (s->min_index[v] < smin[sl]) smin[sl] = s->min_index[v];
        }
    }
    for (int32_t v = 0; v < n; ++v) { level[v] = -1; parent[v] = -1; }
    level[root] = 0;
    parent[root] = root;
    visited[pos_of[root]] = 1;
    mask_x[root] = 1;
    colidx_x[0] = root;
    int32_t nnzx = 1, iter = 0;
    do {
        const int pull = nnzx > T;
        int32_t x_min = n, x_max = 0;
        if (!pull) { /* bfs_SELLCSv2.c:302-317 */
            for (int32_t i = 0; i < nnzx; ++i) {
                if (colidx_x[i] > x_max) x_max = colidx_x[i];
                if (colidx_x[i] < x_min) x_min = colidx_x[i];
                bucket[colidx_x[i]] = i;
            }
        }
        if (iter < capacity) {
            if (frontier_out) frontier_out[iter] = nnzx;
            if (mode_out) mode_out[iter] = pull;
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t t = 0; t < tasks; ++t) {
            int32_t cnt = 0;
            int64_t bys = 0, byc = 0;
            const int64_t seg = (int64_t)tp[t] * C; /* task output segment */
            for (int32_t sl = tp[t]; sl < tp[t + 1]; ++sl) {
                const int64_t base = (int64_t)sl * C;
                const int32_t lanes = (int32_t)(base + C <= n ? C : n - base);
                if (!pull && (smax[sl] < x_min || smin[sl] > x_max)) { ++bys; continue; }
                int32_t nvis = 0;
                for (int32_t r = 0; r < lanes; ++r) nvis += visited[base + r];
                if (nvis == lanes) { if (pull) ++bys; continue; }
                int32_t hit[256];
                for (int32_t r = 0; r < lanes; ++r) hit[r] = -1; /* min hit id */
                int64_t at = s->slice_pointers[sl];
                for (int32_t j = 0; j < s->slice_widths[sl]; ++j) {
                    const int32_t v = s->vop_pointers[sl] + j;
                    const int32_t act = s->vop_lengths[v] + 1;
                    if (!pull && (s->max_index[v] < x_min || s->min_index[v] > x_max)) {
                        ++byc;
                        at += act;
                        continue;
                    }
                    for (int32_t r = 0; r < act; ++r) {
                        const int32_t c = s->column_indices[at + r];
                        const int in = pull ? (mask_x[c] > 0) : (colidx_x[bucket[c]] == c);
                        if (in && (hit[r] < 0 || c < hit[r])) hit[r] = c;
                    }
                    at += act;
                }
                for (int32_t r = 0; r < lanes; ++r) {
                    if (hit[r] < 0 || visited[base + r]) continue;
                    visited[base + r] = 1;
                    const int32_t vid = s->row_order[base + r];
                    level[vid] = iter + 1;
                    parent[vid] = hit[r];
                    if (pull) mask_y[vid] = 1;
                    else colidx_y[seg + cnt] = vid;
                    ++cnt;
                }
            }
            tcnt[t] = cnt;
            tbs[t] = bys;
            tbc[t] = byc;
        }
        int64_t bys = 0, byc = 0;
        int32_t nnzy = 0;
        for (int32_t t = 0; t < tasks; ++t) { nnzy += tcnt[t]; bys += tbs[t]; byc += tbc[t]; }
        if (iter < capacity) {
            if (bys_out) bys_out[iter] = bys;
            if (byc_out) byc_out[iter] = byc;
        }
        /* frontier conversions, bfs_SELLCSv2.c:472-525 */
        if (pull) {
            if (nnzy > T) {
                memcpy(mask_x, mask_y, sizeof(int32_t) * (size_t)n); /* spmv_merge_res */
            } else if (f9) {
                /* the reference's dense2sparse (bfs_SELLCSv2.c:488-499,
                 * bfs_par_kernels.c:754-787): task t scans the ids
                 * [tp[t] << 8, tp[t + 1] << 8) -- its slices' positions when
                 * C = 256 -- and writes them from the prefix of the tasks'
                 * counts, which count its slices' rows by position.  Where a
                 * sigma window crosses a task boundary the two differ: a task
                 * overwrites the next one's first entries (tasks run in order
                 * here) and leaves stale entries in its own range. */
                int32_t off = 0;
                for (int32_t t = 0; t < tasks; ++t) {
                    const int64_t lo = (int64_t)tp[t] << 8;
                    const int64_t hi = ((int64_t)tp[t + 1] << 8) < n ? ((int64_t)tp[t + 1] << 8) : n;
                    int32_t k = off;
                    for (int64_t v = lo; v < hi; ++v)
                        if (mask_y[v] > 0 && k < n) colidx_x[k++] = (int32_t)v;
                    off += tcnt[t];
                }
            } else {
                /* dense2sparse over ids, offsets from counts of the same ids */
                int32_t k = 0;
                for (int32_t v = 0; v < n; ++v)
                    if (mask_y[v] > 0) colidx_x[k++] = v;
            }
        } else {
            if (nnzy > T) { /* sparse2dense: mask_x keeps its stale entries */
                for (int32_t t = 0; t < tasks; ++t)
                    for (int32_t i = 0; i < tcnt[t]; ++i) mask_x[colidx_y[(int64_t)tp[t] * C + i]] = 1;
            } else { /* merge_res */
                int32_t k = 0;
                for (int32_t t = 0; t < tasks; ++t)
                    for (int32_t i = 0; i < tcnt[t]; ++i) colidx_x[k++] = colidx_y[(int64_t)tp[t] * C + i];
            }
        }
        nnzx = nnzy;
        if (nnzx > T) memset(mask_y, 0, sizeof(int32_t) * (size_t)n); /* reset */
        ++iter;
    } while (nnzx != 0 && iter < max_levels);
    free(pos_of); free(mask_x); free(mask_y); free(colidx_x); free(colidx_y); free(bucket);
    free(visited); free(smax); free(smin); free(tp); free(tcnt); free(tbs); free(tbc);
    return iter;
}

/* ---------------------------------------------------------------------- */

int orc_validate(int32_t n, const int64_t *row_ptr, const int32_t *col, int32_t root,
                 const int32_t *level, const int32_t *parent)
{
    if (root < 0 || root >= n || level[root] != 0 || parent[root] != root) return -1;
    for (int32_t v = 0; v < n; ++v) {
        if (level[v] < 0) {
            if (parent[v] != -1) return -2;
            continue;
        }
        if (v == root) continue;
        int32_t p = parent[v];
        if (p < 0 || p >= n || level[p] < 0 || level[p] != level[v] - 1) return -2;
        int found = 0, best = -1;
        for (int64_t k = row_ptr[v]; k < row_ptr[v + 1]; ++k) {
            int32_t u = col[k];
            if (u == p) found = 1;
            if (level[u] >= 0 && level[u] < level[v] - 1) return -4; /* v reachable sooner */
            if (level[u] == level[v] - 1 && (best < 0 || u < best)) best = u;
        }
        if (!found) return -3;
        if (best != p) return -6;
    }
    /* an edge u -> v (u in row v) from a reached u must reach v */
    for (int32_t v = 0; v < n; ++v)
        for (int64_t k = row_ptr[v]; k < row_ptr[v + 1]; ++k)
            if (level[col[k]] >= 0 && level[v] < 0) return -5;
    return 0;
}
