/*
 * Thin C harness around the reference's own portable C sources, compiled in
 * place from /root/reference/bfs/src into oracle/_ref/libalbba_ref.so by
 * oracle/Makefile.  TEST INFRASTRUCTURE ONLY: it exists to pin the oracle and
 * to write golden fixtures; the product never loads it.
 *
 * The reference's app defines these globals in bfs/src/main.c:14-22; the
 * harness defines them because main.c itself pulls in the VE-only kernels.
 */
#include "common.h"
#include "sparse_matrix.h"
#include "sorting_utils.h"
#include "utils.h"
#include "biio.h"
#include "bfs_ref.h"

size_t align_size = 64;
int perf_analysis = 0;
int verify = 0;
int bypass_record = 0;

/* bfs/src/biio.h:127 */
int refh_read(char *path, int *nr, int *nc, int *nnz, int **rp, int **ci, double **val, int *sym)
{
    *nr = *nc = *nnz = -1;
    *rp = *ci = NULL;
    *val = NULL;
    *sym = -1;
    read_Dmatrix(nr, nc, nnz, rp, ci, val, sym, path);
    return (*rp && *ci) ? 0 : -1;
}

void refh_free(void *p) { free(p); }

/* bfs/src/sparse_matrix.c:230 */
void refh_csc(int nr, int nc, int nnz, int *rp, int *ci, int **cp_out, int **ri_out)
{
    SparseMatrixCSR a;
    memset(&a, 0, sizeof a);
    a.nrows = nr;
    a.ncolumns = nc;
    a.nnz = nnz;
    a.row_pointers = rp;
    a.column_indices = ci;
    a.values = (elem_t *)malloc(sizeof(elem_t) * (size_t)(nnz > 0 ? nnz : 1));
    for (int k = 0; k < nnz; ++k) a.values[k] = 1.0;
    a.name = "harness";
    SparseMatrixCSC t;
    convert_csr_to_csc(&a, &t);
    free(a.values);
    free(t.values);
    *cp_out = t.column_pointers;
    *ri_out = t.row_indices;
}

/* The builder sequence of sellcs_preprocess (bfs/src/bfs_SELLCSv2.c:34-59),
 * calling the reference's own functions. */
int refh_sell(int n, int nnz, int *rp, int *ci, int C, int sigma, int sort_flag, int *nslices,
              int **row_order, int **widths, int **vptr, unsigned char **vlen, int **sptr,
              int **colidx, int **maxi, int **mini)
{
    SparseMatrixCSR a;
    memset(&a, 0, sizeof a);
    a.nrows = n;
    a.ncolumns = n;
    a.nnz = nnz;
    a.row_pointers = rp;
    a.column_indices = ci;
    a.values = (elem_t *)malloc(sizeof(elem_t) * (size_t)(nnz > 0 ? nnz : 1));
    for (int k = 0; k < nnz; ++k) a.values[k] = 1.0;
    a.name = "harness";
    SparseMatrixSELLCS s;
    memset(&s, 0, sizeof s);
    s.C = C;
    s.sigma = sigma;
    int padded;
    int *rs;
    if (sort_flag) {
        s.row_order = sort_by_1st_colidx(a.row_pointers, a.column_indices, n);
        rs = get_rows_size_sorted(s.row_order, a.row_pointers, n, &padded);
        get_order_by_row_size_radix_sorted(s.row_order, rs, n, sigma);
    } else {
        rs = get_rows_size(a.row_pointers, n, &padded);
        s.row_order = get_order_by_row_size_radix(rs, n, sigma);
    }
    convert_csr_to_sellcs_dfc(&a, &s, rs, s.row_order, 0);
    free(rs);
    free(a.values);
    free(s.values);
    *nslices = s.nslices;
    *row_order = s.row_order;
    *widths = s.slice_widths;
    *vptr = s.vop_pointers;
    *vlen = s.vop_lengths;
    *sptr = s.slice_pointers;
    *colidx = s.column_indices;
    *maxi = s.max_index;
    *mini = s.min_index;
    return 0;
}

/* bfs/src/bfs_ref.h:26 from vertex 0 (the reference's only root), with the
 * state main.c:196-205 prepares. frontier[] gets num_vex_ref. */
int refh_bfs_ref(int n, int nnz, int *rp, int *ci, int *frontier, int *iters, int *sum_vex)
{
    SparseMatrixCSR a;
    memset(&a, 0, sizeof a);
    a.nrows = n;
    a.ncolumns = n;
    a.nnz = nnz;
    a.row_pointers = rp;
    a.column_indices = ci;
    a.values = (elem_t *)malloc(sizeof(elem_t) * (size_t)(nnz > 0 ? nnz : 1));
    for (int k = 0; k < nnz; ++k) a.values[k] = 1.0;
    a.name = "harness";
    elem_t *x = (elem_t *)calloc((size_t)n, sizeof(elem_t));
    int *colidx_x = (int *)calloc((size_t)n, sizeof(int));
    int nnzx = 1;
    x[0] = 1;
    colidx_x[0] = 0;
    int iter_ref = 0, sum_vex_ref = 1;
    run_bfs_ref(&a, x, colidx_x, &nnzx, &iter_ref, &sum_vex_ref, frontier);
    *iters = iter_ref;
    *sum_vex = sum_vex_ref;
    free(a.values);
    free(x);
    free(colidx_x);
    return 0;
}
