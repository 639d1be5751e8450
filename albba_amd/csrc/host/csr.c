/*
 * CSR utilities: the reference transpose, the BFS working matrix, .cbd
 * writing, and the library's error/allocation helpers.
 */
#include "internal.h"

#include <time.h>

static __thread char g_err[512];

int albba_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *albba_last_error(void) { return g_err; }

const char *albba_version(void) { return "albba-mi355x 0.1"; }

void *albba_calloc(size_t count, size_t size)
{
    return calloc(count ? count : 1, size ? size : 1);
}

void *albba_malloc(size_t bytes) { return malloc(bytes ? bytes : 1); }

double albba_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* bfs/src/sparse_matrix.c:230-290: count entries per column, exclusive scan,
 * then scatter rows in increasing order so every column lists its rows
 * ascending.  Arrays are malloc'ed (the reference uses aligned_alloc). */
void convert_csr_to_csc(const SparseMatrixCSR *csr, SparseMatrixCSC *csc)
{
    const int nc = csr->ncolumns, nnz = csr->nnz;
    int *cp = (int *)calloc((size_t)nc + 1, sizeof(int));
    int *ri = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    elem_t *va = (elem_t *)malloc(sizeof(elem_t) * ((size_t)nnz + 1));
    int *cur = (int *)malloc(sizeof(int) * ((size_t)nc + 1));
    if (!cp || !ri || !va || !cur) {
        fprintf(stderr, "Memory Allocation Error: could not allocate SparseMatrixCSC. Application will exit.");
        exit(1);
    }
    for (int k = 0; k < nnz; ++k) cp[csr->column_indices[k] + 1]++;
    for (int c = 0; c < nc; ++c) cp[c + 1] += cp[c];
    memcpy(cur, cp, sizeof(int) * (size_t)nc);
    for (int r = 0; r < csr->nrows; ++r)
        for (int k = csr->row_pointers[r]; k < csr->row_pointers[r + 1]; ++k) {
            int at = cur[csr->column_indices[k]]++;
            ri[at] = r;
            va[at] = csr->values ? csr->values[k] : 1.0;
        }
    free(cur);
    csc->name = csr->name;
    csc->nrows = csr->nrows;
    csc->ncolumns = nc;
    csc->nnz = nnz;
    csc->column_pointers = cp;
    csc->row_indices = ri;
    csc->values = va;
}

int albba_read_graph(const char *path, albba_csr *out)
{
    if (!path || !out) return albba_fail(ALBBA_EINVAL, "read_graph: null argument");
    size_t L = strlen(path);
    if (L < 3 || (strcmp(path + L - 3, "cbd") != 0 && strcmp(path + L - 3, "mtx") != 0))
        return albba_fail(ALBBA_EINVAL, "read_graph: %s is neither .cbd nor .mtx", path);
    int nr = -1, nc = -1, nnz = -1, sym = 0;
    int *rp = NULL, *ci = NULL;
    double *va = NULL;
    char *name = strdup(path);
    read_Dmatrix(&nr, &nc, &nnz, &rp, &ci, &va, &sym, name);
    free(name);
    if (!rp || !ci || nr < 0 || nnz < 0) {
        free(rp); free(ci); free(va);
        return albba_fail(ALBBA_EIO, "read_graph: could not load %s", path);
    }
    free(va);
    if (nr != nc) {
        free(rp); free(ci);
        return albba_fail(ALBBA_EINVAL, "read_graph: matrix is %d x %d, BFS needs a square one", nr, nc);
    }
    for (int k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= nc) {
            free(rp); free(ci);
            return albba_fail(ALBBA_EIO, "read_graph: column index out of range");
        }
    /* working matrix W = A^T (bfs/src/main.c:117-126): row i lists the
     * in-neighbours of i, ascending */
    int64_t *wp = (int64_t *)albba_calloc((size_t)nc + 1, sizeof(int64_t));
    int32_t *wc = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    int64_t *cur = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)nc + 1));
    if (!wp || !wc || !cur) {
        free(rp); free(ci); free(wp); free(wc); free(cur);
        return albba_fail(ALBBA_ENOMEM, "read_graph: working matrix");
    }
    for (int k = 0; k < nnz; ++k) wp[ci[k] + 1]++;
    for (int c = 0; c < nc; ++c) wp[c + 1] += wp[c];
    memcpy(cur, wp, sizeof(int64_t) * (size_t)nc);
    for (int r = 0; r < nr; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) wc[cur[ci[k]]++] = r;
    free(cur);
    free(rp);
    free(ci);
    out->n = nc;
    out->nnz = nnz;
    out->row_ptr = wp;
    out->col = wc;
    out->symmetric = sym;
    return ALBBA_OK;
}

int albba_write_cbd(const char *path, const albba_csr *g)
{
    if (!path || !g) return albba_fail(ALBBA_EINVAL, "write_cbd: null argument");
    if (g->nnz > INT32_MAX) return albba_fail(ALBBA_ERANGE, "write_cbd: nnz beyond int32");
    int *rp = (int *)albba_malloc(sizeof(int) * ((size_t)g->n + 1));
    double *va = (double *)albba_malloc(sizeof(double) * (size_t)(g->nnz > 0 ? g->nnz : 1));
    if (!rp || !va) { free(rp); free(va); return albba_fail(ALBBA_ENOMEM, "write_cbd"); }
    for (int32_t i = 0; i <= g->n; ++i) rp[i] = (int)g->row_ptr[i];
    for (int64_t k = 0; k < g->nnz; ++k) va[k] = 1.0;
    char *name = strdup(path);
    binary_write_csr(g->n, g->n, (int)g->nnz, rp, (int *)g->col, va, name);
    free(name);
    free(rp);
    free(va);
    FILE *f = fopen(path, "rb");
    if (!f) return albba_fail(ALBBA_EIO, "write_cbd: could not write %s", path);
    fclose(f);
    return ALBBA_OK;
}
