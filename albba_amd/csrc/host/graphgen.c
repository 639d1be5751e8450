/*
 * Seeded synthetic graphs and root selection.
 *
 * The reference has no generator (it only loads SuiteSparse files,
 * bfs/src/main.c:88-99), so the specification here is this library's own and
 * is fixed by tests through edge-list checksums and degree statistics:
 *
 *   RNG      r(seed, i) = splitmix64_finaliser(seed + (i+1) * 0x9E3779B97F4A7C15)
 *   RMAT     edge e (0 <= e < edgefactor * 2^scale) draws one 64-bit value per
 *            two recursion levels, k = e * ceil(scale/2) + level/2; the high
 *            32 bits serve the even level, the low 32 bits the odd level;
 *            u = bits / 2^32 picks quadrant (0,0) if u < a, (0,1) if u < a+b,
 *            (1,0) if u < a+b+c, else (1,1); level 0 sets the top bit.
 *   ER       M = round(n * mean_degree / 2) edges, edge e has endpoints
 *            r(seed, 2e) mod n and r(seed, 2e+1) mod n.
 *   grid     vertex (i, j) has id i * cols + j; horizontal edge e = i*(cols-1)+j
 *            and vertical edge e = H + i*cols + j (H = rows*(cols-1)) are removed
 *            when u01(r(seed, e)) < p_remove.
 *   permute  Fisher-Yates over [0,n): for i = n-1 .. 1 swap perm[i] with
 *            perm[r(perm_seed, i) mod (i+1)]; vertex v becomes perm[v].
 *   then     every edge (u,v), u != v, is stored as u->v and v->u; each row is
 *            sorted ascending and duplicates are removed.
 */
#include "internal.h"

#include <omp.h>

uint64_t albba_rng_u64(uint64_t seed, uint64_t idx) { return albba_rng_at(seed, idx); }

void albba_csr_free(albba_csr *g)
{
    if (!g) return;
    free(g->row_ptr);
    free(g->col);
    g->row_ptr = NULL;
    g->col = NULL;
    g->n = 0;
    g->nnz = 0;
}

int32_t *albba_make_perm(int32_t n, uint64_t perm_seed)
{
    int32_t *perm = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!perm) return NULL;
    for (int32_t i = 0; i < n; ++i) perm[i] = i;
    for (int64_t i = (int64_t)n - 1; i > 0; --i) {
        int64_t j = (int64_t)(albba_rng_at(perm_seed, (uint64_t)i) % (uint64_t)(i + 1));
        int32_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    return perm;
}

static void isort(int32_t *a, int64_t len)
{
    for (int64_t i = 1; i < len; ++i) {
        int32_t x = a[i];
        int64_t j = i - 1;
        while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; --j; }
        a[j + 1] = x;
    }
}

static int cmp_i32(const void *x, const void *y)
{
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

static void sort_row(int32_t *a, int64_t len)
{
    if (len < 48) isort(a, len);
    else qsort(a, (size_t)len, sizeof(int32_t), cmp_i32);
}

/* Symmetric CSR from an undirected edge list whose endpoints are already in
 * [0,n).  Self loops dropped, both directions stored, rows sorted, deduped. */
static int build_sym_csr(int32_t n, int64_t ne, const int32_t *src, const int32_t *dst,
                         const int32_t *perm, albba_csr *out)
{
    int64_t *deg = (int64_t *)albba_calloc((size_t)n + 1, sizeof(int64_t));
    if (!deg) return albba_fail(ALBBA_ENOMEM, "degree array");
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        int32_t u = perm ? perm[src[e]] : src[e];
        int32_t v = perm ? perm[dst[e]] : dst[e];
        if (u == v) continue;
#pragma omp atomic
        deg[u]++;
#pragma omp atomic
        deg[v]++;
    }
    int64_t *ptr = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)n + 1));
    if (!ptr) { free(deg); return albba_fail(ALBBA_ENOMEM, "row pointers"); }
    ptr[0] = 0;
    for (int32_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + deg[i];
    int64_t total = ptr[n];
    int32_t *col = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(total > 0 ? total : 1));
    if (!col) { free(deg); free(ptr); return albba_fail(ALBBA_ENOMEM, "columns"); }
    memcpy(deg, ptr, sizeof(int64_t) * (size_t)n); /* deg[] becomes the fill cursor */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        int32_t u = perm ? perm[src[e]] : src[e];
        int32_t v = perm ? perm[dst[e]] : dst[e];
        if (u == v) continue;
        int64_t a, b;
#pragma omp atomic capture
        a = deg[u]++;
#pragma omp atomic capture
        b = deg[v]++;
        col[a] = v;
        col[b] = u;
    }
    /* sort + dedup each row in place; deg[i] = new length */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) {
        int32_t *r = col + ptr[i];
        int64_t len = ptr[i + 1] - ptr[i];
        sort_row(r, len);
        int64_t k = 0;
        for (int64_t j = 0; j < len; ++j)
            if (k == 0 || r[j] != r[k - 1]) r[k++] = r[j];
        deg[i] = k;
    }
    int64_t *nptr = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)n + 1));
    if (!nptr) { free(deg); free(ptr); free(col); return albba_fail(ALBBA_ENOMEM, "row pointers"); }
    nptr[0] = 0;
    for (int32_t i = 0; i < n; ++i) nptr[i + 1] = nptr[i] + deg[i];
    int64_t m = nptr[n];
    int32_t *ncol = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    if (!ncol) { free(deg); free(ptr); free(col); free(nptr); return albba_fail(ALBBA_ENOMEM, "columns"); }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i)
        memcpy(ncol + nptr[i], col + ptr[i], sizeof(int32_t) * (size_t)deg[i]);
    free(deg);
    free(ptr);
    free(col);
    out->n = n;
    out->nnz = m;
    out->row_ptr = nptr;
    out->col = ncol;
    out->symmetric = 1;
    return ALBBA_OK;
}

int albba_gen_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                   uint64_t perm_seed, albba_csr *out)
{
    if (!out || scale < 1 || scale > 30 || edgefactor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0 + 1e-12)
        return albba_fail(ALBBA_EINVAL, "rmat: bad parameters");
    const int32_t n = (int32_t)1 << scale;
    const int64_t ne = (int64_t)edgefactor << scale;
    if (2 * ne > INT32_MAX) return albba_fail(ALBBA_ERANGE, "rmat: nnz beyond int32");
    int32_t *src = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)ne);
    int32_t *dst = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)ne);
    if (!src || !dst) { free(src); free(dst); return albba_fail(ALBBA_ENOMEM, "rmat edges"); }
    const double ab = a + b, abc = a + b + c;
    const int draws = (scale + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        uint32_t u = 0, v = 0;
        uint64_t r = 0;
        for (int lv = 0; lv < scale; ++lv) {
            if ((lv & 1) == 0) r = albba_rng_at(seed, (uint64_t)e * (uint64_t)draws + (uint64_t)(lv >> 1));
            uint32_t bits = (lv & 1) ? (uint32_t)r : (uint32_t)(r >> 32);
            double x = (double)bits * (1.0 / 4294967296.0);
            uint32_t bu, bv;
            if (x < a) { bu = 0; bv = 0; }
            else if (x < ab) { bu = 0; bv = 1; }
            else if (x < abc) { bu = 1; bv = 0; }
            else { bu = 1; bv = 1; }
            u |= bu << (scale - 1 - lv);
            v |= bv << (scale - 1 - lv);
        }
        src[e] = (int32_t)u;
        dst[e] = (int32_t)v;
    }
    int32_t *perm = albba_make_perm(n, perm_seed);
    if (!perm) { free(src); free(dst); return albba_fail(ALBBA_ENOMEM, "permutation"); }
    int rc = build_sym_csr(n, ne, src, dst, perm, out);
    free(src); free(dst); free(perm);
    return rc;
}

int albba_gen_er(int32_t n, double mean_degree, uint64_t seed, uint64_t perm_seed, albba_csr *out)
{
    if (!out || n < 1 || mean_degree < 0) return albba_fail(ALBBA_EINVAL, "er: bad parameters");
    const int64_t ne = (int64_t)((double)n * mean_degree / 2.0 + 0.5);
    if (2 * ne > INT32_MAX) return albba_fail(ALBBA_ERANGE, "er: nnz beyond int32");
    int32_t *src = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    int32_t *dst = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    if (!src || !dst) { free(src); free(dst); return albba_fail(ALBBA_ENOMEM, "er edges"); }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        src[e] = (int32_t)(albba_rng_at(seed, 2 * (uint64_t)e) % (uint64_t)n);
        dst[e] = (int32_t)(albba_rng_at(seed, 2 * (uint64_t)e + 1) % (uint64_t)n);
    }
    int32_t *perm = albba_make_perm(n, perm_seed);
    if (!perm) { free(src); free(dst); return albba_fail(ALBBA_ENOMEM, "permutation"); }
    int rc = build_sym_csr(n, ne, src, dst, perm, out);
    free(src); free(dst); free(perm);
    return rc;
}

int albba_gen_grid(int32_t rows, int32_t cols, double p_remove, uint64_t seed,
                   uint64_t perm_seed, albba_csr *out)
{
    if (!out || rows < 1 || cols < 1 || (int64_t)rows * cols > INT32_MAX || p_remove < 0 ||
        p_remove > 1)
        return albba_fail(ALBBA_EINVAL, "grid: bad parameters");
    const int32_t n = rows * cols;
    const int64_t H = (int64_t)rows * (cols - 1), V = (int64_t)(rows - 1) * cols;
    const int64_t ne = H + V;
    int32_t *src = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    int32_t *dst = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    if (!src || !dst) { free(src); free(dst); return albba_fail(ALBBA_ENOMEM, "grid edges"); }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        int32_t u, v;
        if (e < H) {
            int64_t i = e / (cols - 1), j = e % (cols - 1);
            u = (int32_t)(i * cols + j);
            v = u + 1;
        } else {
            int64_t k = e - H;
            int64_t i = k / cols, j = k % cols;
            u = (int32_t)(i * cols + j);
            v = u + cols;
        }
        if (albba_u01(albba_rng_at(seed, (uint64_t)e)) < p_remove) v = u; /* dropped as a self loop */
        src[e] = u;
        dst[e] = v;
    }
    int32_t *perm = albba_make_perm(n, perm_seed);
    if (!perm) { free(src); free(dst); return albba_fail(ALBBA_ENOMEM, "permutation"); }
    int rc = build_sym_csr(n, ne, src, dst, perm, out);
    free(src); free(dst); free(perm);
    return rc;
}

/* ---- roots ------------------------------------------------------------ */

static int32_t uf_find(int32_t *p, int32_t x)
{
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

int albba_pick_roots(const albba_csr *g, int count, uint64_t root_seed, int largest_cc,
                     int32_t *roots)
{
    if (!g || !roots || count < 0) return albba_fail(ALBBA_EINVAL, "pick_roots: bad arguments");
    const int32_t n = g->n;
    uint8_t *ok = (uint8_t *)albba_calloc((size_t)n + 1, 1);
    if (!ok) return albba_fail(ALBBA_ENOMEM, "pick_roots");
    for (int32_t v = 0; v < n; ++v) ok[v] = (g->row_ptr[v + 1] > g->row_ptr[v]);
    if (largest_cc) {
        int32_t *p = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int64_t *sz = (int64_t *)albba_calloc((size_t)n + 1, sizeof(int64_t));
        if (!p || !sz) { free(p); free(sz); free(ok); return albba_fail(ALBBA_ENOMEM, "pick_roots"); }
        for (int32_t v = 0; v < n; ++v) p[v] = v;
        for (int32_t v = 0; v < n; ++v)
            for (int64_t k = g->row_ptr[v]; k < g->row_ptr[v + 1]; ++k) {
                int32_t a = uf_find(p, v), b = uf_find(p, g->col[k]);
                if (a != b) { if (a < b) p[b] = a; else p[a] = b; }
            }
        for (int32_t v = 0; v < n; ++v) sz[uf_find(p, v)]++;
        /* largest component; ties go to the smallest representative */
        int32_t best = -1;
        for (int32_t v = 0; v < n; ++v)
            if (p[v] == v && (best < 0 || sz[v] > sz[best])) best = v;
        for (int32_t v = 0; v < n; ++v) ok[v] = ok[v] && uf_find(p, v) == best;
        free(p);
        free(sz);
    }
    int64_t eligible = 0;
    for (int32_t v = 0; v < n; ++v) eligible += ok[v];
    int written = 0;
    if (eligible <= count) {
        for (int32_t v = 0; v < n && written < count; ++v)
            if (ok[v]) roots[written++] = v;
    } else {
        for (uint64_t k = 0; written < count; ++k) {
            int32_t v = (int32_t)(albba_rng_at(root_seed, k) % (uint64_t)n);
            if (ok[v] == 1) {
                ok[v] = 2; /* taken */
                roots[written++] = v;
            }
        }
    }
    free(ok);
    return written;
}
