/*
 * albba_bfs -- command-line driver over libalbba.so.
 *
 * Default mode replays the reference application's sequence
 * (bfs/src/main.c:86-213): read_Dmatrix, values := 1.0, transpose to the
 * working matrix with convert_csr_to_csc, run_custom_bfs_test from vertex 0,
 * then print the frontier sizes and the reference's GTEPS
 * (bfs/src/report_results.c:36) -- using only the reference's entry points.
 *
 *   albba_bfs -m graph.mtx|graph.cbd [-w sigma] [-s 1] [-v 1]
 *
 * -s 1 orders the rows by their first column before the sigma-window sort,
 * as the reference's -s 1 (bfs/src/main.c:44-83, sort_flag).
 *
 * With -r the albba_* API runs instead: BFS from the given roots with a
 * chosen switch policy, results checked as BFS trees on the host.
 *
 *   albba_bfs -m graph.cbd -r 0,17,42 [-l selector.txt] [-C 64] [-w 4096] [-s 1]
 */
#include "albba.h"

#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int reference_sequence(char *path, int sort_flag)
{
    SparseMatrixCSR a;
    memset(&a, 0, sizeof a);
    a.name = path;
    a.nrows = -1;
    read_Dmatrix(&a.nrows, &a.ncolumns, &a.nnz, &a.row_pointers, &a.column_indices, &a.values,
                 &a.isSymmetric, a.name);
    if (a.nrows < 0 || !a.row_pointers) {
        fprintf(stderr, "could not load %s\n", path);
        return 1;
    }
    printf("input matrix A: ( %i, %i ) nnz = %i\n", a.nrows, a.ncolumns, a.nnz);
    for (int i = 0; i < a.nnz; ++i) a.values[i] = 1.0;
    SparseMatrixCSC t;
    convert_csr_to_csc(&a, &t);
    SparseMatrixCSR w = a;
    w.row_pointers = t.column_pointers;
    w.column_indices = t.row_indices;
    w.values = t.values;
    w.nrows = t.ncolumns;
    w.ncolumns = t.nrows;
    const int n = w.nrows;
    elem_t *bfs_x = calloc((size_t)n + 1, sizeof(elem_t));
    int *colidx_x = calloc((size_t)n + 1, sizeof(int));
    int *frontier = calloc((size_t)n + 1, sizeof(int));
    double *iter_times = calloc((size_t)n + 1, sizeof(double));
    int nnzx = 1, num_vex = 1, num_iter = 0;
    double conv_time = 0, bfs_time = 0;
    bfs_x[0] = 1;
    run_custom_bfs_test(&w, bfs_x, colidx_x, &nnzx, frontier, &num_vex, &conv_time, iter_times,
                        &bfs_time, &num_iter, true, sort_flag != 0);
    printf("iteration = %d, bfs runtime = %.3f ms, conversion = %.3f ms\n", num_iter, bfs_time,
           conv_time);
    printf("frontier:");
    for (int i = 0; i < num_iter; ++i) printf(" %d", frontier[i]);
    printf("\nvertices reached = %d\nGTEPS = %.4f\n", num_vex,
           w.nnz * 1.0 / (bfs_time * 1000000.0));
    free(bfs_x); free(colidx_x); free(frontier); free(iter_times);
    return 0;
}

static int api_sequence(const char *path, const char *roots, const char *selector, int C,
                        int sigma, int sort_flag)
{
    albba_csr g;
    int rc = albba_read_graph(path, &g);
    if (rc) { fprintf(stderr, "%s\n", albba_last_error()); return 1; }
    albba_sellcs *s = NULL;
    albba_ctx *ctx = NULL;
    albba_graph *G = NULL;
    albba_selector *sel = NULL;
    if ((rc = albba_sellcs_build_ex(&g, C, sigma, 1, sort_flag ? ALBBA_SELL_SORT_FIRST_COL : 0, &s)) ||
        (rc = albba_ctx_create(0, &ctx)) ||
        (rc = albba_graph_upload(ctx, s, &G))) {
        fprintf(stderr, "%s\n", albba_last_error());
        return 1;
    }
    albba_bfs_opts o;
    albba_bfs_opts_default(&o);
    if (selector) {
        if ((rc = albba_selector_create(ctx, 1, &sel)) || (rc = albba_selector_load(sel, selector))) {
            fprintf(stderr, "%s\n", albba_last_error());
            return 1;
        }
        o.switch_policy = ALBBA_SWITCH_LEARNED;
        o.selector = sel;
    }
    int32_t *level = malloc(sizeof(int32_t) * (size_t)g.n);
    int32_t *parent = malloc(sizeof(int32_t) * (size_t)g.n);
    char *list = strdup(roots), *save = NULL;
    int bad = 0;
    for (char *tok = strtok_r(list, ",", &save); tok; tok = strtok_r(NULL, ",", &save)) {
        int32_t root = atoi(tok);
        albba_bfs_stats st;
        memset(&st, 0, sizeof st);
        if ((rc = albba_bfs(ctx, G, root, &o, &st)) || (rc = albba_bfs_results(ctx, G, level, parent))) {
            fprintf(stderr, "root %d: %s\n", root, albba_last_error());
            bad = 1;
            continue;
        }
        /* host check: every reached vertex's parent is an in-neighbour one level up */
        int ok = level[root] == 0 && parent[root] == root;
        for (int32_t v = 0; ok && v < g.n; ++v) {
            if (level[v] <= 0) continue;
            int seen = 0;
            for (int64_t k = g.row_ptr[v]; k < g.row_ptr[v + 1]; ++k)
                if (g.col[k] == parent[v]) seen = 1;
            ok = seen && level[parent[v]] == level[v] - 1;
        }
        printf("root %d: levels %d reached %lld time %.3f ms GTEPS %.2f tree %s\n", root,
               st.num_levels, (long long)st.reached, st.bfs_ms,
               (double)g.nnz / (st.bfs_ms * 1e6), ok ? "ok" : "INVALID");
        bad |= !ok;
    }
    free(list); free(level); free(parent);
    albba_selector_free(sel);
    albba_graph_free(G);
    albba_sellcs_free(s);
    albba_ctx_destroy(ctx);
    albba_csr_free(&g);
    return bad;
}

int main(int argc, char **argv)
{
    char *path = NULL, *roots = NULL, *selector = NULL;
    int C = 64, sigma = 4096, sort_flag = 0, opt;
    while ((opt = getopt(argc, argv, "m:w:v:r:l:C:s:p:")) != -1) {
        switch (opt) {
        case 'm': path = optarg; break;
        case 'w': sigma_window = sigma = atoi(optarg); break;
        case 'v': verify = atoi(optarg); break;
        case 'p': perf_analysis = atoi(optarg); break;
        case 'r': roots = optarg; break;
        case 'l': selector = optarg; break;
        case 'C': C = atoi(optarg); break;
        case 's': sort_flag = atoi(optarg); break;
        default:
            fprintf(stderr, "usage: %s -m <graph.mtx|.cbd> [-w sigma] [-s 0|1] [-r r0,r1,...] [-l selector]\n",
                    argv[0]);
            return 2;
        }
    }
    if (!path) {
        fprintf(stderr, "You must specify a valid input method (-m)\n");
        return 2;
    }
    return roots ? api_sequence(path, roots, selector, C, sigma, sort_flag) : reference_sequence(path, sort_flag);
}
