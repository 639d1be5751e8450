/*
 * SELL-C-sigma builders.
 *
 * Row ordering (both builders) follows bfs/src/sorting_utils.c:173-253: row
 * lengths are sorted descending inside consecutive windows of sigma rows by a
 * stable sort (the reference uses an LSD radix sort with reversed bucket
 * writes, which is stable and descending), so equal lengths keep their input
 * order.  Active-lane counts follow bfs/src/utils.c:198-241 (column j of a
 * slice has act_j = #rows with length > j lanes, stored as act_j - 1) and the
 * packing follows bfs/src/sparse_matrix.c:523-623 (column-major per slice,
 * lanes 0..act_j-1 of column j contiguous).
 *
 * sellcs_preprocess() is the reference-compatible builder: same struct, no
 * padding, column indices left in vertex-id space.
 *
 * albba_sellcs_build() is the layout the GPU kernels read: C rows per slice
 * (16..128), each column padded to `pad` int32 entries (pad = 32 gives
 * 128-byte aligned column starts), column indices renumbered to positions so
 * frontier and visited bitmaps share one index space, rows listing neighbours
 * in ascending original id, plus the out-edge (push) CSR over positions.
 */
#include "internal.h"

#include <omp.h>

/* ---- stable descending window sort ------------------------------------- */

void albba_sort_window_desc(int32_t *key, int32_t *id, int32_t lo, int32_t hi, int32_t *tk,
                            int32_t *ti)
{
    /* bottom-up merge sort on [lo,hi): stable, descending by key */
    const int32_t len = hi - lo;
    if (len < 2) return;
    int32_t *sk = key + lo, *si = id + lo;
    for (int32_t r = 0; r < len; r += 16) {
        int32_t e = r + 16 < len ? r + 16 : len;
        for (int32_t i = r + 1; i < e; ++i) {
            int32_t k = sk[i], v = si[i], j = i - 1;
            while (j >= r && sk[j] < k) { sk[j + 1] = sk[j]; si[j + 1] = si[j]; --j; }
            sk[j + 1] = k;
            si[j + 1] = v;
        }
    }
    int32_t *ak = sk, *ai = si, *bk = tk, *bi = ti;
    for (int32_t w = 16; w < len; w *= 2) {
        for (int32_t a = 0; a < len; a += 2 * w) {
            int32_t m = a + w < len ? a + w : len, e = a + 2 * w < len ? a + 2 * w : len;
            int32_t i = a, j = m, o = a;
            while (i < m && j < e) {
                if (ak[j] > ak[i]) { bk[o] = ak[j]; bi[o++] = ai[j++]; }
                else { bk[o] = ak[i]; bi[o++] = ai[i++]; }
            }
            while (i < m) { bk[o] = ak[i]; bi[o++] = ai[i++]; }
            while (j < e) { bk[o] = ak[j]; bi[o++] = ai[j++]; }
        }
        int32_t *t = ak; ak = bk; bk = t;
        t = ai; ai = bi; bi = t;
    }
    if (ak != sk) {
        memcpy(sk, ak, sizeof(int32_t) * (size_t)len);
        memcpy(si, ai, sizeof(int32_t) * (size_t)len);
    }
}

int albba_sigma_order(const int32_t *len, int32_t n, int32_t sigma, int32_t *order,
                      int32_t *sorted_len)
{
    if (sigma < 1) return albba_fail(ALBBA_EINVAL, "sigma must be >= 1");
    for (int32_t i = 0; i < n; ++i) { order[i] = i; sorted_len[i] = len[i]; }
    const int64_t nwin = ((int64_t)n + sigma - 1) / sigma;
    int fail = 0;
#pragma omp parallel
    {
        int32_t cap = sigma < n ? sigma : n;
        int32_t *tk = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        int32_t *ti = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        if (!tk || !ti) {
#pragma omp atomic write
            fail = 1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int64_t w = 0; w < nwin; ++w) {
                int32_t lo = (int32_t)(w * sigma);
                int32_t hi = (int32_t)((int64_t)lo + sigma < n ? lo + sigma : n);
                /* sort (sorted_len, order) pairs; work on window-relative temps */
                albba_sort_window_desc(sorted_len + lo, order + lo, 0, hi - lo, tk, ti);
            }
        }
        free(tk);
        free(ti);
    }
    return fail ? albba_fail(ALBBA_ENOMEM, "sigma sort buffers") : ALBBA_OK;
}

/* ---- reference-compatible builder -------------------------------------- */

void sellcs_preprocess(const SparseMatrixCSR *csr, SparseMatrixSELLCS *out, bool sort_flag)
{
    const int n = csr->nrows, C = out->C, sigma = out->sigma;
    if (C < 1 || C > 256 || sigma < 1 || n < 0 || (sigma % C != 0 && sigma < n)) {
        fprintf(stderr, "sellcs_preprocess: need 1 <= C <= 256 and sigma a multiple of C\n");
        exit(1);
    }
    int32_t *len = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *order = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *slen = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!len || !order || !slen) {
        fprintf(stderr, "Memory Allocation Error: could not allocate rows_size. Application will exit.");
        exit(1);
    }
    if (sort_flag) {
        /* bfs/src/sorting_utils.c:255-297: global stable descending sort by
         * each row's first column index, then the sigma-window length sort.
         * An empty row reads the next stored entry, as the reference does;
         * past the last entry the key is 0. */
        int32_t *key = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int32_t *t1 = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int32_t *t2 = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        if (!key || !t1 || !t2) exit(1);
        for (int i = 0; i < n; ++i) {
            int at = csr->row_pointers[i];
            key[i] = at < csr->nnz ? csr->column_indices[at] : 0;
            order[i] = i;
        }
        albba_sort_window_desc(key, order, 0, n, t1, t2);
        for (int i = 0; i < n; ++i)
            slen[i] = csr->row_pointers[order[i] + 1] - csr->row_pointers[order[i]];
        /* window sort of (slen, order) */
        const int32_t cap = sigma < n ? sigma : n;
        int32_t *tk = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        int32_t *ti = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        for (int64_t lo = 0; lo < n; lo += sigma) {
            int32_t hi = (int32_t)(lo + sigma < n ? lo + sigma : n);
            albba_sort_window_desc(slen + lo, order + lo, 0, hi - (int32_t)lo, tk, ti);
        }
        free(tk); free(ti); free(key); free(t1); free(t2);
    } else {
        for (int i = 0; i < n; ++i) len[i] = csr->row_pointers[i + 1] - csr->row_pointers[i];
        if (albba_sigma_order(len, n, sigma, order, slen) != ALBBA_OK) exit(1);
    }
    free(len);

    const int ns = (n + C - 1) / C;
    int *widths = (int *)albba_malloc(sizeof(int) * (size_t)(ns > 0 ? ns : 1));
    int *vptr = (int *)albba_malloc(sizeof(int) * ((size_t)ns + 2));
    int *sptr = (int *)albba_malloc(sizeof(int) * ((size_t)ns + 1));
    if (!widths || !vptr || !sptr) exit(1);
    int nv = 0;
    for (int s = 0; s < ns; ++s) {
        int w = 0;
        for (int r = s * C; r < n && r < (s + 1) * C; ++r) w = slen[r] > w ? slen[r] : w;
        widths[s] = w;
        vptr[s] = nv;
        nv += w;
    }
    vptr[ns] = nv;
    uint8_t *vlen = (uint8_t *)albba_malloc((size_t)(nv > 0 ? nv : 1));
    int *maxi = (int *)albba_calloc((size_t)(nv > 0 ? nv : 1), sizeof(int));
    int *mini = (int *)albba_calloc((size_t)(nv > 0 ? nv : 1), sizeof(int));
    if (!vlen || !maxi || !mini) exit(1);
    int total = 0;
    for (int s = 0; s < ns; ++s) {
        /* rows of a slice are sorted descending: walking from the shortest
         * row up, row r is the last one longer than every column it opens */
        int rows = (s + 1) * C <= n ? C : n - s * C;
        int prev = 0;
        sptr[s] = total;
        for (int r = rows - 1; r >= 0; --r) {
            int L = slen[s * C + r];
            for (int j = prev; j < L; ++j) {
                vlen[vptr[s] + j] = (uint8_t)r;
                total += r + 1;
            }
            if (L > prev) prev = L;
        }
    }
    sptr[ns] = total;
    int *colidx = (int *)albba_malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
    elem_t *vals = (elem_t *)albba_malloc(sizeof(elem_t) * (size_t)(total > 0 ? total : 1));
    if (!colidx || !vals) exit(1);
#pragma omp parallel for schedule(dynamic, 32)
    for (int s = 0; s < ns; ++s) {
        int at = sptr[s], base = s * C;
        for (int j = 0; j < widths[s]; ++j) {
            int act = vlen[vptr[s] + j] + 1, mx = 0, mn = n;
            for (int r = 0; r < act; ++r) {
                int k = csr->row_pointers[order[base + r]] + j;
                int c = csr->column_indices[k];
                colidx[at] = c;
                vals[at] = csr->values ? csr->values[k] : 1.0;
                mx = c > mx ? c : mx;
                mn = c < mn ? c : mn;
                ++at;
            }
            maxi[vptr[s] + j] = mx;
            mini[vptr[s] + j] = mn;
        }
    }
    free(slen);
    out->name = csr->name;
    out->values = vals;
    out->column_indices = colidx;
    out->nrows = n;
    out->ncolumns = csr->ncolumns;
    out->nnz = csr->nnz;
    out->nslices = ns;
    out->slice_widths = widths;
    out->slice_pointers = sptr;
    out->row_order = order;
    out->vop_lengths = vlen;
    out->vop_pointers = vptr;
    out->slice_widths_pointers = NULL;
    out->max_index = maxi;
    out->min_index = mini;
}

void free_sellcs_struct(SparseMatrixSELLCS *s)
{
    if (!s) return;
    free(s->values);
    free(s->column_indices);
    free(s->slice_pointers);
    free(s->slice_widths);
    free(s->row_order);
    free(s->vop_lengths);
    free(s->vop_pointers);
    free(s->max_index);
    free(s->min_index);
    s->values = NULL;
    s->column_indices = NULL;
    s->slice_pointers = s->slice_widths = s->row_order = s->vop_pointers = NULL;
    s->vop_lengths = NULL;
    s->max_index = s->min_index = NULL;
}

/* ---- device layout ---------------------------------------------------- */

void albba_sellcs_free(albba_sellcs *s)
{
    if (!s) return;
    free(s->row_order);
    free(s->pos_of);
    free(s->slice_widths);
    free(s->slice_pointers);
    free(s->vop_pointers);
    free(s->vop_lengths);
    free(s->column_indices);
    free(s->row_lengths);
    free(s->push_ptr);
    free(s->push_col);
    free(s->pull_ptr);
    free(s->pull_col);
    free(s);
}

static int rows_sorted(const albba_csr *g)
{
    int ok = 1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(&& : ok)
    for (int32_t i = 0; i < g->n; ++i)
        for (int64_t k = g->row_ptr[i] + 1; k < g->row_ptr[i + 1]; ++k)
            if (g->col[k - 1] > g->col[k]) ok = 0;
    return ok;
}

static int cmp_i32(const void *x, const void *y)
{
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

static inline int64_t pad_up(int64_t x, int64_t p) { return (x + p - 1) / p * p; }

/* The row order before the sigma-window sort under ALBBA_SELL_SORT_FIRST_COL
 * (bfs/src/sorting_utils.c:255-281): rows by first column index, descending,
 * ties in ascending id.  col holds rows in ascending order, so a row's first
 * entry is its smallest in-neighbour; an empty row reads the next stored
 * entry as the reference does, and 0 past the last one. */
static int first_col_order(const albba_csr *g, const int32_t *col, int32_t *pre)
{
    const int32_t n = g->n;
    int32_t *key = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    int32_t *tk = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    int32_t *ti = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    if (!key || !tk || !ti) { free(key); free(tk); free(ti); return ALBBA_ENOMEM; }
    for (int32_t v = 0; v < n; ++v) {
        const int64_t at = g->row_ptr[v];
        key[v] = at < g->nnz ? col[at] : 0;
        pre[v] = v;
    }
    albba_sort_window_desc(key, pre, 0, n, tk, ti);
    free(key); free(tk); free(ti);
    return ALBBA_OK;
}

int albba_sellcs_build(const albba_csr *g, int C, int sigma, int pad, albba_sellcs **out)
{
    return albba_sellcs_build_ex(g, C, sigma, pad, 0, out);
}

int albba_sellcs_build_ex(const albba_csr *g, int C, int sigma, int pad, int flags, albba_sellcs **out)
{
    if (!g || !out) return albba_fail(ALBBA_EINVAL, "sellcs_build: null argument");
    if (flags & ~ALBBA_SELL_SORT_FIRST_COL)
        return albba_fail(ALBBA_EINVAL, "sellcs_build: unknown flags %#x", (unsigned)flags);
    if (C != 16 && C != 32 && C != 64 && C != 128)
        return albba_fail(ALBBA_EINVAL, "sellcs_build: C must be 16, 32, 64 or 128 (got %d)", C);
    if (pad != 1 && pad != 16 && pad != 32)
        return albba_fail(ALBBA_EINVAL, "sellcs_build: pad must be 1, 16 or 32 (got %d)", pad);
    if (sigma < 1 || (sigma % C != 0 && sigma < g->n))
        return albba_fail(ALBBA_EINVAL, "sellcs_build: sigma (%d) must be a multiple of C (%d) or >= n", sigma, C);
    if (g->n < 1) return albba_fail(ALBBA_EINVAL, "sellcs_build: empty graph");
    if (g->nnz > INT32_MAX) return albba_fail(ALBBA_ERANGE, "sellcs_build: nnz beyond int32");
    const int32_t n = g->n;
    const int32_t ns = (int32_t)(((int64_t)n + C - 1) / C);
    const int64_t npos = (int64_t)ns * C;

    /* rows must list neighbours in ascending id for the min-id parent rule */
    const int32_t *gcol = g->col;
    int32_t *sorted_copy = NULL;
    if (!rows_sorted(g)) {
        sorted_copy = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(g->nnz > 0 ? g->nnz : 1));
        if (!sorted_copy) return albba_fail(ALBBA_ENOMEM, "sellcs_build: row copy");
        memcpy(sorted_copy, g->col, sizeof(int32_t) * (size_t)g->nnz);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t i = 0; i < n; ++i)
            qsort(sorted_copy + g->row_ptr[i], (size_t)(g->row_ptr[i + 1] - g->row_ptr[i]),
                  sizeof(int32_t), cmp_i32);
        gcol = sorted_copy;
    }

    albba_sellcs *s = (albba_sellcs *)albba_calloc(1, sizeof(albba_sellcs));
    if (!s) { free(sorted_copy); return albba_fail(ALBBA_ENOMEM, "sellcs_build"); }
    s->n = n; s->C = C; s->sigma = sigma; s->pad = pad; s->nslices = ns; s->nnz = g->nnz;
    s->order_flags = flags;
    int32_t *len = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    s->row_order = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    s->pos_of = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    s->row_lengths = (int32_t *)albba_calloc((size_t)npos, sizeof(int32_t));
    s->slice_widths = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)ns);
    s->slice_pointers = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)ns + 1));
    s->vop_pointers = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)ns + 1));
    if (!len || !s->row_order || !s->pos_of || !s->row_lengths || !s->slice_widths ||
        !s->slice_pointers || !s->vop_pointers) {
        free(len); free(sorted_copy); albba_sellcs_free(s);
        return albba_fail(ALBBA_ENOMEM, "sellcs_build: metadata");
    }
    int rc = ALBBA_OK;
    if (flags & ALBBA_SELL_SORT_FIRST_COL) {
        /* window sort over the first-column order: position p of the
         * pre-order holds row pre[p] (pos_of is free until the end) */
        int32_t *pre = s->pos_of;
        if ((rc = first_col_order(g, gcol, pre))) {
            free(len); free(sorted_copy); albba_sellcs_free(s);
            return albba_fail(rc, "sellcs_build: first-column order");
        }
        for (int32_t i = 0; i < n; ++i) len[i] = (int32_t)(g->row_ptr[pre[i] + 1] - g->row_ptr[pre[i]]);
        rc = albba_sigma_order(len, n, sigma, s->row_order, s->row_lengths);
        for (int32_t p = 0; !rc && p < n; ++p) len[p] = pre[s->row_order[p]];
        if (!rc) memcpy(s->row_order, len, sizeof(int32_t) * (size_t)n);
    } else {
        for (int32_t i = 0; i < n; ++i) len[i] = (int32_t)(g->row_ptr[i + 1] - g->row_ptr[i]);
        rc = albba_sigma_order(len, n, sigma, s->row_order, s->row_lengths);
    }
    free(len);
    if (rc) { free(sorted_copy); albba_sellcs_free(s); return rc; }
    for (int32_t p = 0; p < n; ++p) s->pos_of[s->row_order[p]] = p;

    /* widths, vop pointers, padded slice sizes */
    int64_t *ssize = (int64_t *)albba_calloc((size_t)ns + 1, sizeof(int64_t));
    if (!ssize) { free(sorted_copy); albba_sellcs_free(s); return albba_fail(ALBBA_ENOMEM, "sellcs_build"); }
    int64_t nv = 0;
    for (int32_t sl = 0; sl < ns; ++sl) {
        int32_t w = s->row_lengths[(int64_t)sl * C]; /* rows sorted desc within the slice */
        s->slice_widths[sl] = w;
        s->vop_pointers[sl] = nv;
        nv += w;
    }
    s->vop_pointers[ns] = nv;
    s->nvops = nv;
    s->vop_lengths = (uint8_t *)albba_malloc((size_t)(nv > 0 ? nv : 1));
    if (!s->vop_lengths) { free(ssize); free(sorted_copy); albba_sellcs_free(s); return albba_fail(ALBBA_ENOMEM, "sellcs_build"); }
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t sl = 0; sl < ns; ++sl) {
        const int32_t *sl_len = s->row_lengths + (int64_t)sl * C;
        const int32_t w = s->slice_widths[sl];
        uint8_t *vl = s->vop_lengths + s->vop_pointers[sl];
        /* rows are sorted descending: act_j = #rows with len > j; walk rows
         * from the last (shortest) to the first */
        int32_t prev = 0;
        int64_t tot = 0;
        for (int32_t r = C - 1; r >= 0; --r) {
            int32_t L = sl_len[r];
            for (int32_t j = prev; j < L; ++j) {
                vl[j] = (uint8_t)r; /* act - 1 = r: rows 0..r have length > j */
                tot += pad_up(r + 1, pad);
            }
            if (L > prev) prev = L;
        }
        (void)w;
        ssize[sl] = tot;
    }
    s->slice_pointers[0] = 0;
    for (int32_t sl = 0; sl < ns; ++sl) s->slice_pointers[sl + 1] = s->slice_pointers[sl] + ssize[sl];
    free(ssize);
    s->padded_nnz = s->slice_pointers[ns];
    if (s->padded_nnz > INT32_MAX) {
        free(sorted_copy); albba_sellcs_free(s);
        return albba_fail(ALBBA_ERANGE, "sellcs_build: padded size beyond int32");
    }
    s->column_indices = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(s->padded_nnz > 0 ? s->padded_nnz : 1));
    if (!s->column_indices) { free(sorted_copy); albba_sellcs_free(s); return albba_fail(ALBBA_ENOMEM, "sellcs_build: columns"); }
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t sl = 0; sl < ns; ++sl) {
        int64_t at = s->slice_pointers[sl];
        const int64_t base = (int64_t)sl * C;
        for (int32_t j = 0; j < s->slice_widths[sl]; ++j) {
            int32_t act = s->vop_lengths[s->vop_pointers[sl] + j] + 1;
            for (int32_t r = 0; r < act; ++r) {
                int32_t v = s->row_order[base + r];
                s->column_indices[at + r] = s->pos_of[gcol[g->row_ptr[v] + j]];
            }
            int64_t padded = pad_up(act, pad);
            for (int32_t r = act; r < padded; ++r) s->column_indices[at + r] = -1;
            at += padded;
        }
    }

    /* push CSR over positions = transpose of the layout matrix */
    s->push_ptr = (int64_t *)albba_calloc((size_t)npos + 1, sizeof(int64_t));
    s->push_col = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(g->nnz > 0 ? g->nnz : 1));
    if (!s->push_ptr || !s->push_col) { free(sorted_copy); albba_sellcs_free(s); return albba_fail(ALBBA_ENOMEM, "sellcs_build: push csr"); }
    if (g->symmetric == 1) {
        /* A == A^T: the out-edges of position p are its row, relabelled */
        for (int64_t p = 0; p < npos; ++p)
            s->push_ptr[p + 1] = s->push_ptr[p] + (p < n ? s->row_lengths[p] : 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t p = 0; p < n; ++p) {
            int32_t v = s->row_order[p];
            int64_t o = s->push_ptr[p];
            for (int64_t k = g->row_ptr[v]; k < g->row_ptr[v + 1]; ++k) s->push_col[o++] = s->pos_of[gcol[k]];
        }
    } else {
        for (int32_t v = 0; v < n; ++v)
            for (int64_t k = g->row_ptr[v]; k < g->row_ptr[v + 1]; ++k) s->push_ptr[s->pos_of[gcol[k]] + 1]++;
        for (int64_t p = 0; p < npos; ++p) s->push_ptr[p + 1] += s->push_ptr[p];
        int64_t *cur = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)npos);
        if (!cur) { free(sorted_copy); albba_sellcs_free(s); return albba_fail(ALBBA_ENOMEM, "sellcs_build: push csr"); }
        memcpy(cur, s->push_ptr, sizeof(int64_t) * (size_t)npos);
        for (int32_t p = 0; p < n; ++p) {
            int32_t v = s->row_order[p];
            for (int64_t k = g->row_ptr[v]; k < g->row_ptr[v + 1]; ++k) s->push_col[cur[s->pos_of[gcol[k]]]++] = p;
        }
        free(cur);
        /* row-major copy of the layout's own rows for the pull's long rows */
        s->pull_ptr = (int64_t *)albba_calloc((size_t)npos + 1, sizeof(int64_t));
        s->pull_col = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(g->nnz > 0 ? g->nnz : 1));
        if (!s->pull_ptr || !s->pull_col) { free(sorted_copy); albba_sellcs_free(s); return albba_fail(ALBBA_ENOMEM, "sellcs_build: pull csr"); }
        for (int64_t p = 0; p < npos; ++p)
            s->pull_ptr[p + 1] = s->pull_ptr[p] + (p < n ? s->row_lengths[p] : 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t p = 0; p < n; ++p) {
            int32_t v = s->row_order[p];
            int64_t o = s->pull_ptr[p];
            for (int64_t k = g->row_ptr[v]; k < g->row_ptr[v + 1]; ++k) s->pull_col[o++] = s->pos_of[gcol[k]];
        }
    }
    free(sorted_copy);
    *out = s;
    return ALBBA_OK;
}
