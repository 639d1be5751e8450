/*
 * BFS driver: device graph upload, the level loop with its direction switch,
 * and the reference-compatible run_custom_bfs_test.
 *
 * Level loop (bfs/src/bfs_SELLCSv2.c:294-765 restated for the GPU): every
 * level either pulls (SpMV over unvisited slices) or pushes (SpMSpV from the
 * frontier queue); the frontier lives in a position-space bitmap (always)
 * and a queue (after a push, or rebuilt from the bitmap when a push follows
 * a pull).  The host reads one small counter block per level; all per-vertex
 * work stays on the device.  The reference's switch (bfs_SELLCSv2.c:282-290,
 * 302, 344, 356) pulls when the frontier holds more than (int)(0.04 n)
 * vertices; it is the default policy, the fallback of the learned selector
 * and the yardstick for its lost time.
 */
#include "internal.h"
#include "../albd.h"
#include "device_int.h"
#include "selector_int.h"

#include <math.h>

int sigma_window = 16384;
int perf_analysis = 0;
int verify = 0;

/* ---- context ---- */

int albba_device_count(int *count)
{
    if (!count) return albba_fail(ALBBA_EINVAL, "device_count: null");
    int rc = albd_device_count(count);
    if (rc) return albba_fail(rc, "%s", albd_error());
    return ALBBA_OK;
}

int albba_ctx_create(int device, albba_ctx **out)
{
    if (!out) return albba_fail(ALBBA_EINVAL, "ctx_create: null");
    albba_ctx *c = (albba_ctx *)albba_calloc(1, sizeof(albba_ctx));
    if (!c) return albba_fail(ALBBA_ENOMEM, "ctx_create");
    int rc = albd_ctx_create(device, &c->d);
    if (rc) {
        free(c);
        return albba_fail(rc, "%s", albd_error());
    }
    c->device = device;
    *out = c;
    return ALBBA_OK;
}

void albba_ctx_destroy(albba_ctx *c)
{
    if (!c) return;
    if (getenv("ALBBA_VERBOSE")) {
        unsigned long hits = 0, misses = 0;
        albd_mail_stats(c->d, &hits, &misses);
        fprintf(stderr, "[albba] counter mailbox: %lu posts seen, %lu fallbacks\n", hits, misses);
    }
    albd_ctx_destroy(c->d);
    free(c);
}

int albba_ctx_synchronize(albba_ctx *c)
{
    if (!c) return albba_fail(ALBBA_EINVAL, "synchronize: null");
    int rc = albd_sync(c->d);
    return rc ? albba_fail(rc, "%s", albd_error()) : ALBBA_OK;
}

albd_ctx *albba_ctx_dev(albba_ctx *c) { return c ? c->d : NULL; }

int albba_bench_copy(albba_ctx *c, int64_t bytes, int reps, float *ms)
{
    if (!c || !ms || bytes < 16 || reps < 1) return albba_fail(ALBBA_EINVAL, "bench_copy: bad argument");
    const size_t words = (size_t)(bytes / 16) * 2;
    void *a = NULL, *b = NULL;
    int rc = albd_malloc(c->d, words * 8, &a);
    if (!rc) rc = albd_malloc(c->d, words * 8, &b);
    /* one untimed copy touches both buffers first */
    if (!rc) rc = albd_fill_u64(c->d, a, 0x5a5a5a5a5a5a5a5aull, words);
    if (!rc) rc = albd_copy_u64(c->d, b, a, words);
    if (!rc) rc = albd_event_record(c->d, 4);
    for (int r = 0; !rc && r < reps; ++r) rc = albd_copy_u64(c->d, b, a, words);
    if (!rc) rc = albd_event_record(c->d, 5);
    float t = 0;
    if (!rc) rc = albd_event_elapsed(c->d, 4, 5, &t);
    if (!rc) rc = albd_sync(c->d);
    char buf[256] = "";
    if (rc) snprintf(buf, sizeof buf, "%s", albd_error());
    albd_free(c->d, a);
    albd_free(c->d, b);
    if (rc) return albba_fail(rc, "bench_copy: %s", buf);
    *ms = t / (float)reps;
    return ALBBA_OK;
}

/* ---- upload ---- */

static void *dalloc(albba_graph *G, size_t bytes, int *rc)
{
    void *p = NULL;
    if (*rc) return NULL;
    *rc = albd_malloc(G->ctx->d, bytes, &p);
    if (!*rc) G->device_bytes += (int64_t)bytes;
    return p;
}

static int upload_padded_i64(albba_ctx *c, int64_t *dst, const int64_t *src, int64_t n_src,
                             int64_t n_dst)
{
    int64_t *tmp = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)n_dst);
    if (!tmp) return ALBBA_ENOMEM;
    memcpy(tmp, src, sizeof(int64_t) * (size_t)n_src);
    for (int64_t i = n_src; i < n_dst; ++i) tmp[i] = src[n_src - 1];
    int rc = albd_h2d(c->d, dst, tmp, sizeof(int64_t) * (size_t)n_dst);
    free(tmp);
    return rc;
}

static int upload_padded_i32(albba_ctx *c, int32_t *dst, const int32_t *src, int64_t n_src,
                             int64_t n_dst, int32_t fill)
{
    int32_t *tmp = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n_dst);
    if (!tmp) return ALBBA_ENOMEM;
    memcpy(tmp, src, sizeof(int32_t) * (size_t)n_src);
    for (int64_t i = n_src; i < n_dst; ++i) tmp[i] = fill;
    int rc = albd_h2d(c->d, dst, tmp, sizeof(int32_t) * (size_t)n_dst);
    free(tmp);
    return rc;
}

void albba_graph_free(albba_graph *G)
{
    if (!G) return;
    albd_ctx *d = G->ctx->d;
    albd_free(d, (void *)G->g.slice_ptr);
    albd_free(d, (void *)G->g.slice_width);
    albd_free(d, (void *)G->g.vop_ptr);
    albd_free(d, (void *)G->g.vop_len);
    albd_free(d, (void *)G->g.colidx);
    albd_free(d, (void *)G->g.row_order);
    albd_free(d, (void *)G->g.row_len);
    if (G->g.pull_ptr != G->g.push_ptr) {
        albd_free(d, (void *)G->g.pull_ptr);
        albd_free(d, (void *)G->g.pull_col);
    }
    albd_free(d, (void *)G->g.push_ptr);
    albd_free(d, (void *)G->g.push_col);
    albd_free(d, (void *)G->g.haslen);
    albd_free(d, (void *)G->g.pos_of);
    albd_free(d, (void *)G->g.row_off);
    albd_free(d, G->st.lp);
    albd_free(d, G->st.visited);
    albd_free(d, G->st.front);
    albd_free(d, G->st.next);
    albd_free(d, G->st.summary);
    albd_free(d, G->st.queue);
    albd_free(d, G->st.queue_next);
    albd_free(d, G->st.prefix);
    albd_free(d, G->st.block_sums);
    albd_free(d, G->st.pmin);
    albd_free(d, G->st.level);
    albd_free(d, G->st.parent);
    albd_free(d, G->st.active);
    albd_free(d, G->st.active_next);
    albd_free(d, G->st.ctr);
    albd_free(d, G->st.slots);
    albd_free(d, G->st.level_log);
    albd_free(d, G->st.pkey);
    albd_free(d, G->st.persist);
    albd_free(d, G->scratch_visited);
    albd_free(d, G->scratch_ctr);
    free(G->pos_of);
    free(G->host_log);
    free(G->word_entries);
    free(G->list_copy);
    free(G);
}

/* The BFS state of a graph whose layout arrays are on the device (shapes
 * from G->g): allocate, build haslen, fill the keys, and keep the host copy
 * of pos_of.  Sets the error message on failure. */
static int graph_state_alloc(albba_graph *G)
{
    albd_graph *g = &G->g;
    const int64_t nwords = g->nwords, npos = g->npos;
    const int32_t n = g->n;
    int rc = 0;
    g->haslen = (const uint64_t *)dalloc(G, sizeof(uint64_t) * (size_t)nwords, &rc);
    albd_state *st = &G->st;
    const int64_t nsum = (nwords + 63) / 64 + 1;
    st->visited = (uint64_t *)dalloc(G, sizeof(uint64_t) * (size_t)nwords, &rc);
    st->front = (uint64_t *)dalloc(G, sizeof(uint64_t) * (size_t)nwords, &rc);
    st->next = (uint64_t *)dalloc(G, sizeof(uint64_t) * (size_t)nwords, &rc);
    st->summary = (uint64_t *)dalloc(G, sizeof(uint64_t) * (size_t)nsum, &rc);
    st->queue = (int32_t *)dalloc(G, sizeof(int32_t) * (size_t)npos, &rc);
    st->queue_next = (int32_t *)dalloc(G, sizeof(int32_t) * (size_t)npos, &rc);
    st->prefix = (int64_t *)dalloc(G, sizeof(int64_t) * (size_t)(npos + 1), &rc);
    st->block_sums = (int64_t *)dalloc(G, sizeof(int64_t) * (ALBD_SCAN_BLOCKS + 1), &rc);
    st->pmin = (uint32_t *)dalloc(G, sizeof(uint32_t) * (size_t)npos, &rc);
    st->lp = (unsigned long long *)dalloc(G, sizeof(unsigned long long) * (size_t)npos, &rc);
    st->level = (int32_t *)dalloc(G, sizeof(int32_t) * (size_t)n, &rc);
    st->parent = (int32_t *)dalloc(G, sizeof(int32_t) * (size_t)n, &rc);
    st->active = (int32_t *)dalloc(G, sizeof(int32_t) * (size_t)nwords, &rc);
    st->active_next = (int32_t *)dalloc(G, sizeof(int32_t) * (size_t)nwords, &rc);
    st->ctr = (albd_counters *)dalloc(G, sizeof(albd_counters), &rc);
    st->slots = (albd_slot *)dalloc(G, sizeof(albd_slot) * ALBD_NSLOT, &rc);
    st->level_log = (int64_t *)dalloc(G, sizeof(int64_t) * 4 * ALBD_PERSIST_LEVELS, &rc);
    st->pkey = (unsigned long long *)dalloc(G, sizeof(unsigned long long) * (size_t)npos, &rc);
    st->persist = (albd_persist *)dalloc(G, sizeof(albd_persist), &rc);
    G->scratch_visited = dalloc(G, sizeof(uint64_t) * (size_t)nwords, &rc);
    G->scratch_ctr = dalloc(G, sizeof(albd_counters), &rc);
    g->row_off = g->win >= 1 && g->win <= 65536
                     ? (const uint16_t *)dalloc(G, sizeof(uint16_t) * (size_t)(n > 0 ? n : 1), &rc)
                     : NULL;
    if (rc) return albba_fail(rc, "%s", albd_error());
    if ((rc = albd_build_haslen(G->ctx->d, g, (uint64_t *)g->haslen)) ||
        (g->row_off && (rc = albd_build_row_off(G->ctx->d, g, (uint16_t *)g->row_off))) ||
        (rc = albd_fill_u64(G->ctx->d, st->pkey, ~0ull, (size_t)npos)) ||
        (rc = albd_fill_u32(G->ctx->d, st->pmin, 0xFFFFFFFFu, (size_t)npos)) ||
        (rc = albd_sync(G->ctx->d)))
        return albba_fail(rc, "%s", albd_error());
    G->pos_of = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)n);
    G->host_log = (int64_t *)albba_malloc(sizeof(int64_t) * 4 * ALBD_PERSIST_LEVELS);
    if (!G->pos_of || !G->host_log) return albba_fail(ALBBA_ENOMEM, "host arrays");
    if ((rc = albd_d2h(G->ctx->d, G->pos_of, g->pos_of, sizeof(int32_t) * (size_t)n)))
        return albba_fail(rc, "%s", albd_error());
    return ALBBA_OK;
}

int albba_graph_upload(albba_ctx *ctx, const albba_sellcs *s, albba_graph **out)
{
    if (!ctx || !s || !out) return albba_fail(ALBBA_EINVAL, "graph_upload: null argument");
    albba_graph *G = (albba_graph *)albba_calloc(1, sizeof(albba_graph));
    if (!G) return albba_fail(ALBBA_ENOMEM, "graph_upload");
    G->ctx = ctx;
    G->n = s->n;
    G->sigma = s->sigma;
    G->order_flags = s->order_flags;
    G->nnz = s->nnz;
    const int64_t npos_raw = (int64_t)s->nslices * s->C;
    const int64_t nwords = (npos_raw + 63) / 64;
    const int64_t npos = nwords * 64;
    const int64_t ns_alloc = (npos + s->C - 1) / s->C;
    albd_graph *g = &G->g;
    g->n = s->n;
    g->C = s->C;
    g->pad = s->pad;
    g->sparse_lanes = 8;
    g->win = (s->order_flags & ALBBA_SELL_SORT_FIRST_COL) ? 0 : (s->sigma < s->n ? s->sigma : (s->n > 0 ? s->n : 1));
    g->nslices = s->nslices;
    g->nwords = (int32_t)nwords;
    g->npos = npos;
    g->nnz = s->nnz;
    g->padded_nnz = s->padded_nnz;
    g->nvops = s->nvops;
    int rc = 0;
    g->slice_ptr = (const int64_t *)dalloc(G, sizeof(int64_t) * (size_t)(ns_alloc + 1), &rc);
    g->slice_width = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)ns_alloc, &rc);
    g->vop_ptr = (const int64_t *)dalloc(G, sizeof(int64_t) * (size_t)(ns_alloc + 1), &rc);
    g->vop_len = (const uint8_t *)dalloc(G, (size_t)(s->nvops > 0 ? s->nvops : 1), &rc);
    g->colidx = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)(s->padded_nnz > 0 ? s->padded_nnz : 1), &rc);
    g->row_order = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)npos, &rc);
    g->row_len = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)npos, &rc);
    g->push_ptr = (const int64_t *)dalloc(G, sizeof(int64_t) * (size_t)(npos + 1), &rc);
    g->push_col = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)(s->nnz > 0 ? s->nnz : 1), &rc);
    if (s->pull_ptr) {
        g->pull_ptr = (const int64_t *)dalloc(G, sizeof(int64_t) * (size_t)(npos + 1), &rc);
        g->pull_col = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)(s->nnz > 0 ? s->nnz : 1), &rc);
    } else {
        g->pull_ptr = g->push_ptr;
        g->pull_col = g->push_col;
    }
    g->pos_of = (const int32_t *)dalloc(G, sizeof(int32_t) * (size_t)(s->n > 0 ? s->n : 1), &rc);
    if (rc) {
        int code = rc;
        char buf[512];
        snprintf(buf, sizeof buf, "%s", albd_error());
        albba_graph_free(G);
        return albba_fail(code, "graph_upload: %s", buf);
    }
    /* host copies, padded to the device shapes */
    if ((rc = upload_padded_i64(ctx, (int64_t *)g->slice_ptr, s->slice_pointers, s->nslices + 1, ns_alloc + 1)) ||
        (rc = upload_padded_i32(ctx, (int32_t *)g->slice_width, s->slice_widths, s->nslices, ns_alloc, 0)) ||
        (rc = upload_padded_i64(ctx, (int64_t *)g->vop_ptr, s->vop_pointers, s->nslices + 1, ns_alloc + 1)) ||
        (rc = albd_h2d(ctx->d, (void *)g->vop_len, s->vop_lengths, (size_t)s->nvops)) ||
        (rc = albd_h2d(ctx->d, (void *)g->colidx, s->column_indices, sizeof(int32_t) * (size_t)s->padded_nnz)) ||
        (rc = upload_padded_i32(ctx, (int32_t *)g->row_order, s->row_order, s->n, npos, -1)) ||
        (rc = upload_padded_i32(ctx, (int32_t *)g->row_len, s->row_lengths, npos_raw, npos, 0)) ||
        (rc = upload_padded_i64(ctx, (int64_t *)g->push_ptr, s->push_ptr, npos_raw + 1, npos + 1)) ||
        (rc = albd_h2d(ctx->d, (void *)g->push_col, s->push_col, sizeof(int32_t) * (size_t)s->nnz)) ||
        (s->pull_ptr && (rc = upload_padded_i64(ctx, (int64_t *)g->pull_ptr, s->pull_ptr, npos_raw + 1, npos + 1))) ||
        (s->pull_ptr && (rc = albd_h2d(ctx->d, (void *)g->pull_col, s->pull_col, sizeof(int32_t) * (size_t)s->nnz))) ||
        (rc = albd_h2d(ctx->d, (void *)g->pos_of, s->pos_of, sizeof(int32_t) * (size_t)s->n))) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s", rc == ALBBA_ENOMEM ? "host staging buffer" : albd_error());
        albba_graph_free(G);
        return albba_fail(rc, "graph_upload: %s", buf);
    }
    if ((rc = graph_state_alloc(G))) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s", albba_last_error());
        albba_graph_free(G);
        return albba_fail(rc, "graph_upload: %s", buf);
    }
    *out = G;
    return ALBBA_OK;
}

static void *build_alloc(void *arg, size_t bytes)
{
    int rc = 0;
    void *p = dalloc((albba_graph *)arg, bytes, &rc);
    return rc ? NULL : p;
}

int albba_graph_build(albba_ctx *ctx, const albba_csr *cg, int C, int sigma, int pad, albba_graph **out,
                      albba_build_report *report)
{
    return albba_graph_build_ex(ctx, cg, C, sigma, pad, 0, out, report);
}

int albba_graph_build_ex(albba_ctx *ctx, const albba_csr *cg, int C, int sigma, int pad, int flags,
                         albba_graph **out, albba_build_report *report)
{
    if (!ctx || !cg || !out) return albba_fail(ALBBA_EINVAL, "graph_build: null argument");
    if (flags & ~ALBBA_SELL_SORT_FIRST_COL)
        return albba_fail(ALBBA_EINVAL, "graph_build: unknown flags %#x", (unsigned)flags);
    if (C != 16 && C != 32 && C != 64 && C != 128)
        return albba_fail(ALBBA_EINVAL, "graph_build: C must be 16, 32, 64 or 128 (got %d)", C);
    if (pad != 1 && pad != 16 && pad != 32)
        return albba_fail(ALBBA_EINVAL, "graph_build: pad must be 1, 16 or 32 (got %d)", pad);
    if (cg->n < 1 || !cg->row_ptr || (cg->nnz > 0 && !cg->col))
        return albba_fail(ALBBA_EINVAL, "graph_build: empty graph");
    if (sigma < 1 || (sigma % C != 0 && sigma < cg->n))
        return albba_fail(ALBBA_EINVAL, "graph_build: sigma (%d) must be a multiple of C (%d) or >= n", sigma, C);
    if (cg->nnz < 0 || cg->nnz > INT32_MAX) return albba_fail(ALBBA_ERANGE, "graph_build: nnz beyond int32");
    /* the kernels index col through row_ptr: check it before the upload */
    if (cg->row_ptr[0] != 0 || cg->row_ptr[cg->n] != cg->nnz)
        return albba_fail(ALBBA_EINVAL, "graph_build: row_ptr must run from 0 to nnz");
    int monotone = 1;
#pragma omp parallel for schedule(static) reduction(&& : monotone)
    for (int32_t i = 0; i < cg->n; ++i)
        if (cg->row_ptr[i + 1] < cg->row_ptr[i]) monotone = 0;
    if (!monotone) return albba_fail(ALBBA_EINVAL, "graph_build: row_ptr decreases");
    albba_graph *G = (albba_graph *)albba_calloc(1, sizeof(albba_graph));
    if (!G) return albba_fail(ALBBA_ENOMEM, "graph_build");
    G->ctx = ctx;
    G->n = cg->n;
    G->nnz = cg->nnz;
    G->sigma = sigma;
    G->order_flags = flags;
    albd_build_input in;
    memset(&in, 0, sizeof in);
    in.row_ptr = cg->row_ptr;
    in.col = cg->col;
    in.sort_first = (flags & ALBBA_SELL_SORT_FIRST_COL) != 0;
    in.nnz = cg->nnz;
    in.n = cg->n;
    in.C = C;
    in.sigma = sigma;
    in.pad = pad;
    in.symmetric = cg->symmetric == 1;
    albd_build_report r;
    memset(&r, 0, sizeof r);
    int rc = albd_build_graph(ctx->d, &in, &G->g, build_alloc, G, &r);
    if (rc) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s", albd_error());
        albba_graph_free(G);
        return albba_fail(rc, "graph_build: %s", buf);
    }
    if ((rc = graph_state_alloc(G))) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s", albba_last_error());
        albba_graph_free(G);
        return albba_fail(rc, "graph_build: %s", buf);
    }
    if (report) {
        report->upload_ms = r.upload_ms;
        report->build_ms = r.build_ms;
        report->rows_sorted = r.rows_sorted;
        report->reserved_ = 0;
    }
    *out = G;
    return ALBBA_OK;
}

/* Stream `bytes` of f (from its current offset) into device memory at dst
 * through two pinned buffers: reading one while the other's copy runs. */
#define LOAD_CHUNK ((size_t)64 << 20)
static int stream_to_device(albd_ctx *d, FILE *f, void *dst, size_t bytes, void *buf[2])
{
    size_t done = 0;
    int k = 0, rc;
    while (done < bytes) {
        const size_t len = bytes - done < LOAD_CHUNK ? bytes - done : LOAD_CHUNK;
        /* the copy that last used this buffer must have finished */
        if ((rc = albd_event_wait(d, 4 + k))) return rc;
        if (fread(buf[k], 1, len, f) != len) return albba_fail(ALBBA_EIO, "graph_load: short read");
        if ((rc = albd_h2d_async(d, (char *)dst + done, buf[k], len)) || (rc = albd_event_record(d, 4 + k)))
            return rc;
        done += len;
        k ^= 1;
    }
    return 0;
}

int albba_graph_load(albba_ctx *ctx, const char *path, int C, int sigma, int pad, albba_graph **out,
                     albba_load_report *report)
{
    if (!ctx || !path || !out) return albba_fail(ALBBA_EINVAL, "graph_load: null argument");
    if (C != 16 && C != 32 && C != 64 && C != 128)
        return albba_fail(ALBBA_EINVAL, "graph_load: C must be 16, 32, 64 or 128 (got %d)", C);
    if (pad != 1 && pad != 16 && pad != 32)
        return albba_fail(ALBBA_EINVAL, "graph_load: pad must be 1, 16 or 32 (got %d)", pad);
    const size_t L = strlen(path);
    if (L < 4 || strcmp(path + L - 4, ".cbd") != 0)
        return albba_fail(ALBBA_EINVAL, "graph_load: %s is not a .cbd file (load .mtx with albba_read_graph)", path);
    const double t0 = albba_now_ms();
    FILE *f = fopen(path, "rb");
    if (!f) return albba_fail(ALBBA_EIO, "graph_load: cannot open %s", path);
    /* bfs/src/biio.h:21-71: int32 nrows, ncols, nnz; int32 row_ptr[nrows+1];
     * int32 col[nnz]; double val[nnz] (not read: the BFS ignores values) */
    int32_t hdr[3];
    if (fread(hdr, sizeof(int32_t), 3, f) != 3) { fclose(f); return albba_fail(ALBBA_EIO, "graph_load: short header"); }
    const int32_t n = hdr[0];
    const int64_t nnz = hdr[2];
    if (hdr[0] != hdr[1] || n < 1 || nnz < 0) {
        fclose(f);
        return albba_fail(ALBBA_EINVAL, "graph_load: %d x %d matrix with %lld entries", hdr[0], hdr[1], (long long)nnz);
    }
    if (sigma < 1 || (sigma % C != 0 && sigma < n)) {
        fclose(f);
        return albba_fail(ALBBA_EINVAL, "graph_load: sigma (%d) must be a multiple of C (%d) or >= n", sigma, C);
    }
    albd_ctx *d = ctx->d;
    int32_t *rp32 = NULL, *col = NULL;
    void *buf[2] = {NULL, NULL};
    int rc = albd_malloc(d, sizeof(int32_t) * ((size_t)n + 1), (void **)&rp32);
    if (!rc) rc = albd_malloc(d, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1), (void **)&col);
    if (!rc) rc = albd_pinned_alloc(d, LOAD_CHUNK, &buf[0]);
    if (!rc) rc = albd_pinned_alloc(d, LOAD_CHUNK, &buf[1]);
    /* a fresh pair of events: waiting on one never recorded returns at once */
    if (!rc) rc = albd_event_record(d, 4);
    if (!rc) rc = albd_event_record(d, 5);
    if (!rc) rc = stream_to_device(d, f, rp32, sizeof(int32_t) * ((size_t)n + 1), buf);
    if (!rc) rc = stream_to_device(d, f, col, sizeof(int32_t) * (size_t)nnz, buf);
    if (!rc) rc = albd_sync(d);
    fclose(f);
    albd_pinned_free(d, buf[0]);
    albd_pinned_free(d, buf[1]);
    const double t1 = albba_now_ms();
    int64_t *wp = NULL;
    int32_t *wc = NULL;
    int bad = 0;
    if (!rc) rc = albd_transpose_csr(d, rp32, col, n, nnz, &wp, &wc, &bad);
    albd_free(d, rp32);
    albd_free(d, col);
    if (rc || bad) {
        char buf2[512];
        snprintf(buf2, sizeof buf2, "%s", bad ? "row pointers or column indices out of range" :
                 (rc == ALBBA_EIO ? albba_last_error() : albd_error()));
        albd_free(d, wp);
        albd_free(d, wc);
        return albba_fail(bad ? ALBBA_EINVAL : rc, "graph_load: %s", buf2);
    }
    const double t2 = albba_now_ms();
    albba_graph *G = (albba_graph *)albba_calloc(1, sizeof(albba_graph));
    if (!G) { albd_free(d, wp); albd_free(d, wc); return albba_fail(ALBBA_ENOMEM, "graph_load"); }
    G->ctx = ctx;
    G->n = n;
    G->nnz = nnz;
    G->sigma = sigma;
    albd_build_input in;
    memset(&in, 0, sizeof in);
    in.row_ptr = wp;
    in.col = wc;
    in.nnz = nnz;
    in.n = n;
    in.C = C;
    in.sigma = sigma;
    in.pad = pad;
    in.symmetric = 0; /* the .cbd path leaves isSymmetric unset: W is built as directed */
    in.on_device = 1;
    albd_build_report r;
    memset(&r, 0, sizeof r);
    rc = albd_build_graph(d, &in, &G->g, build_alloc, G, &r);
    albd_free(d, wp);
    albd_free(d, wc);
    if (rc) {
        char buf2[512];
        snprintf(buf2, sizeof buf2, "%s", albd_error());
        albba_graph_free(G);
        return albba_fail(rc, "graph_load: %s", buf2);
    }
    if ((rc = graph_state_alloc(G))) {
        char buf2[512];
        snprintf(buf2, sizeof buf2, "%s", albba_last_error());
        albba_graph_free(G);
        return albba_fail(rc, "graph_load: %s", buf2);
    }
    if (report) {
        report->read_ms = (float)(t1 - t0);
        report->prep_ms = (float)(t2 - t1);
        report->build_ms = r.build_ms;
        report->total_ms = (float)(albba_now_ms() - t0);
    }
    *out = G;
    return ALBBA_OK;
}

static int bits_for_ids(int64_t n)
{
    int b = 1;
    while (b < 62 && ((int64_t)1 << b) < n) ++b;
    return b;
}

int albba_graph_generate(albba_ctx *ctx, const albba_gen_params *p, int C, int sigma, int pad,
                         albba_graph **out, albba_csr *host_copy, albba_gen_report *report)
{
    if (!ctx || !p || !out) return albba_fail(ALBBA_EINVAL, "graph_generate: null argument");
    if (C != 16 && C != 32 && C != 64 && C != 128)
        return albba_fail(ALBBA_EINVAL, "graph_generate: C must be 16, 32, 64 or 128 (got %d)", C);
    if (pad != 1 && pad != 16 && pad != 32)
        return albba_fail(ALBBA_EINVAL, "graph_generate: pad must be 1, 16 or 32 (got %d)", pad);
    /* the host generators' parameter rules (host/graphgen.c) */
    albd_gen_spec g;
    memset(&g, 0, sizeof g);
    g.kind = p->kind;
    g.seed = p->seed;
    if (p->kind == ALBBA_GEN_RMAT) {
        if (p->scale < 1 || p->scale > 30 || p->edgefactor < 1 || p->a < 0 || p->b < 0 || p->c < 0 ||
            p->a + p->b + p->c > 1.0 + 1e-12)
            return albba_fail(ALBBA_EINVAL, "graph_generate: bad RMAT parameters");
        g.scale = p->scale;
        g.n = (int32_t)1 << p->scale;
        g.ne = (int64_t)p->edgefactor << p->scale;
        g.a = p->a;
        g.ab = p->a + p->b;
        g.abc = p->a + p->b + p->c;
        g.bits = p->scale;
    } else if (p->kind == ALBBA_GEN_ER) {
        if (p->n < 1 || p->mean_degree < 0) return albba_fail(ALBBA_EINVAL, "graph_generate: bad ER parameters");
        g.n = p->n;
        g.ne = (int64_t)((double)p->n * p->mean_degree / 2.0 + 0.5);
        g.bits = bits_for_ids(g.n);
    } else if (p->kind == ALBBA_GEN_GRID) {
        if (p->rows < 1 || p->cols < 1 || (int64_t)p->rows * p->cols > INT32_MAX || p->p_remove < 0 ||
            p->p_remove > 1)
            return albba_fail(ALBBA_EINVAL, "graph_generate: bad grid parameters");
        g.rows = p->rows;
        g.cols = p->cols;
        g.n = p->rows * p->cols;
        g.H = (int64_t)p->rows * (p->cols - 1);
        g.ne = g.H + (int64_t)(p->rows - 1) * p->cols;
        g.p_remove = p->p_remove;
        g.bits = bits_for_ids(g.n);
    } else {
        return albba_fail(ALBBA_EINVAL, "graph_generate: unknown kind %d", p->kind);
    }
    if (2 * g.ne > INT32_MAX) return albba_fail(ALBBA_ERANGE, "graph_generate: nnz beyond int32");
    if (sigma < 1 || (sigma % C != 0 && sigma < g.n))
        return albba_fail(ALBBA_EINVAL, "graph_generate: sigma (%d) must be a multiple of C (%d) or >= n", sigma, C);
    const double t0 = albba_now_ms();
    int32_t *perm = albba_make_perm(g.n, p->perm_seed);
    if (!perm) return albba_fail(ALBBA_ENOMEM, "graph_generate: permutation");
    const double t1 = albba_now_ms();
    albd_ctx *d = ctx->d;
    int64_t *rp = NULL, m = 0;
    int32_t *col = NULL;
    int rc = albd_gen_csr(d, &g, perm, &rp, &col, &m);
    free(perm);
    if (rc) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s", albd_error());
        return albba_fail(rc, "graph_generate: %s", buf);
    }
    const double t2 = albba_now_ms();
    if (host_copy) {
        memset(host_copy, 0, sizeof *host_copy);
        int64_t *hrp = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)g.n + 1));
        int32_t *hcol = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
        if (!hrp || !hcol) {
            free(hrp); free(hcol); albd_free(d, rp); albd_free(d, col);
            return albba_fail(ALBBA_ENOMEM, "graph_generate: host copy");
        }
        if ((rc = albd_d2h(d, hrp, rp, sizeof(int64_t) * ((size_t)g.n + 1))) ||
            (rc = albd_d2h(d, hcol, col, sizeof(int32_t) * (size_t)m))) {
            char buf[512];
            snprintf(buf, sizeof buf, "%s", albd_error());
            free(hrp); free(hcol); albd_free(d, rp); albd_free(d, col);
            return albba_fail(rc, "graph_generate: %s", buf);
        }
        host_copy->n = g.n;
        host_copy->nnz = m;
        host_copy->row_ptr = hrp;
        host_copy->col = hcol;
        host_copy->symmetric = 1;
    }
    albba_graph *G = (albba_graph *)albba_calloc(1, sizeof(albba_graph));
    if (!G) {
        albd_free(d, rp); albd_free(d, col);
        if (host_copy) albba_csr_free(host_copy);
        return albba_fail(ALBBA_ENOMEM, "graph_generate");
    }
    G->ctx = ctx;
    G->n = g.n;
    G->nnz = m;
    G->sigma = sigma;
    albd_build_input in;
    memset(&in, 0, sizeof in);
    in.row_ptr = rp;
    in.col = col;
    in.nnz = m;
    in.n = g.n;
    in.C = C;
    in.sigma = sigma;
    in.pad = pad;
    in.symmetric = 1;
    in.on_device = 1;
    albd_build_report r;
    memset(&r, 0, sizeof r);
    rc = albd_build_graph(d, &in, &G->g, build_alloc, G, &r);
    albd_free(d, rp);
    albd_free(d, col);
    char buf[512] = "";
    if (rc) snprintf(buf, sizeof buf, "%s", albd_error());
    else if ((rc = graph_state_alloc(G))) snprintf(buf, sizeof buf, "%s", albba_last_error());
    if (!rc) {
        if (report) {
            report->perm_ms = (float)(t1 - t0);
            report->gen_ms = (float)(t2 - t1);
            report->build_ms = r.build_ms;
            report->total_ms = (float)(albba_now_ms() - t0);
        }
        *out = G;
        return ALBBA_OK;
    }
    albba_graph_free(G);
    if (host_copy) albba_csr_free(host_copy);
    return albba_fail(rc, "graph_generate: %s", buf);
}

int albba_graph_layout(albba_ctx *ctx, const albba_graph *G, albba_sellcs **out)
{
    if (!ctx || !G || !out || G->ctx != ctx) return albba_fail(ALBBA_EINVAL, "graph_layout: bad argument");
    const albd_graph *g = &G->g;
    const int64_t ns = g->nslices, npos_raw = ns * g->C;
    albba_sellcs *s = (albba_sellcs *)albba_calloc(1, sizeof(albba_sellcs));
    if (!s) return albba_fail(ALBBA_ENOMEM, "graph_layout");
    s->n = g->n;
    s->C = g->C;
    s->sigma = G->sigma;
    s->order_flags = G->order_flags;
    s->pad = g->pad;
    s->nslices = g->nslices;
    s->nnz = g->nnz;
    s->padded_nnz = g->padded_nnz;
    s->nvops = g->nvops;
    const int directed = g->pull_ptr != g->push_ptr;
    s->row_order = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)g->n);
    s->pos_of = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)g->n);
    s->slice_widths = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(ns > 0 ? ns : 1));
    s->slice_pointers = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)(ns + 1));
    s->vop_pointers = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)(ns + 1));
    s->vop_lengths = (uint8_t *)albba_malloc((size_t)(g->nvops > 0 ? g->nvops : 1));
    s->column_indices = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(g->padded_nnz > 0 ? g->padded_nnz : 1));
    s->row_lengths = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)npos_raw);
    s->push_ptr = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)(npos_raw + 1));
    s->push_col = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(g->nnz > 0 ? g->nnz : 1));
    if (directed) {
        s->pull_ptr = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)(npos_raw + 1));
        s->pull_col = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)(g->nnz > 0 ? g->nnz : 1));
    }
    if (!s->row_order || !s->pos_of || !s->slice_widths || !s->slice_pointers || !s->vop_pointers ||
        !s->vop_lengths || !s->column_indices || !s->row_lengths || !s->push_ptr || !s->push_col ||
        (directed && (!s->pull_ptr || !s->pull_col))) {
        albba_sellcs_free(s);
        return albba_fail(ALBBA_ENOMEM, "graph_layout");
    }
    albd_ctx *d = ctx->d;
    int rc = 0;
    if ((rc = albd_d2h(d, s->row_order, g->row_order, sizeof(int32_t) * (size_t)g->n)) ||
        (rc = albd_d2h(d, s->pos_of, g->pos_of, sizeof(int32_t) * (size_t)g->n)) ||
        (rc = albd_d2h(d, s->slice_widths, g->slice_width, sizeof(int32_t) * (size_t)ns)) ||
        (rc = albd_d2h(d, s->slice_pointers, g->slice_ptr, sizeof(int64_t) * (size_t)(ns + 1))) ||
        (rc = albd_d2h(d, s->vop_pointers, g->vop_ptr, sizeof(int64_t) * (size_t)(ns + 1))) ||
        (rc = albd_d2h(d, s->vop_lengths, g->vop_len, (size_t)g->nvops)) ||
        (rc = albd_d2h(d, s->column_indices, g->colidx, sizeof(int32_t) * (size_t)g->padded_nnz)) ||
        (rc = albd_d2h(d, s->row_lengths, g->row_len, sizeof(int32_t) * (size_t)npos_raw)) ||
        (rc = albd_d2h(d, s->push_ptr, g->push_ptr, sizeof(int64_t) * (size_t)(npos_raw + 1))) ||
        (rc = albd_d2h(d, s->push_col, g->push_col, sizeof(int32_t) * (size_t)g->nnz)) ||
        (directed && (rc = albd_d2h(d, s->pull_ptr, g->pull_ptr, sizeof(int64_t) * (size_t)(npos_raw + 1)))) ||
        (directed && (rc = albd_d2h(d, s->pull_col, g->pull_col, sizeof(int32_t) * (size_t)g->nnz)))) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s", albd_error());
        albba_sellcs_free(s);
        return albba_fail(rc, "graph_layout: %s", buf);
    }
    *out = s;
    return ALBBA_OK;
}

int64_t albba_graph_device_bytes(const albba_graph *G) { return G ? G->device_bytes : 0; }

int albba_graph_size(const albba_graph *G, int32_t *n, int64_t *nnz)
{
    if (!G || !n || !nnz) return albba_fail(ALBBA_EINVAL, "graph_size: null argument");
    *n = G->n;
    *nnz = G->nnz;
    return ALBBA_OK;
}

/* ---- level loop ---- */

void albba_bfs_opts_default(albba_bfs_opts *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->switch_policy = ALBBA_SWITCH_HEURISTIC;
    o->fixed_choice = ALBBA_CHOICE_PULL_SB;
    o->timing_reps = 3;
    o->small_frontier_edges = 65536;
}

/* k_level_begin flags; the select kernel moves the queue/list counters from
 * one level to the next and zeroes the slots */
enum { LB_SLOTS = 1, LB_SMALL = 8 };
int albd_level_begin_flags(albd_ctx *c, const albd_graph *g, const albd_state *st, int flags);

typedef struct loop_state {
    int have_queue;       /* st.queue holds the frontier                */
    int prev_push;        /* last level pushed (queue_next pending)     */
    int list_valid;       /* st.active is a superset of unvisited words */
    int list_pending;     /* last level built a list (length pending)   */
    int64_t list_len;     /* host copy of the st.active length          */
} loop_state;

typedef struct level_sums {
    int64_t next_count, next_edges, next_words, scanned, words, bypassed, col_bypassed;
    int64_t qnext_len, active_len, queue_len, push_edges;
} level_sums;

static void sum_counters(const albd_counters *h, level_sums *o)
{
    /* the select kernel has summed the slot counters of the level */
    o->next_count = (int64_t)h->frontier;
    o->next_edges = (int64_t)h->frontier_edges;
    o->next_words = (int64_t)h->frontier_words;
    o->scanned = (int64_t)h->scanned;
    o->words = (int64_t)h->words;
    o->bypassed = (int64_t)h->bypassed;
    o->col_bypassed = (int64_t)h->col_bypassed;
    o->qnext_len = (int64_t)h->qnext_len;
    o->active_len = (int64_t)h->active_len;
    o->queue_len = (int64_t)h->queue_len;
    o->push_edges = (int64_t)h->push_edges;
}

/* Launch the kernels of one level (no pointer swaps).  With `mark` set, HIP
 * events 6 and 7 bracket the main kernel (pull, or push expansion). */
static int run_level(albba_graph *G, const loop_state *ls, int choice, int32_t level,
                     int64_t frontier, int mark)
{
    albd_ctx *d = G->ctx->d;
    /* slots too: trial runs of one level start from zero each time */
    int rc = albd_level_begin_flags(d, &G->g, &G->st, LB_SLOTS);
    if (rc) return rc;
    if (choice == ALBBA_CHOICE_PUSH) {
        if (!ls->have_queue && (rc = albd_front_to_queue(d, &G->g, &G->st))) return rc;
        if ((rc = albd_push_scan(d, &G->g, &G->st, frontier))) return rc;
        if (mark && (rc = albd_event_record(d, 6))) return rc;
        if ((rc = albd_push_expand(d, &G->g, &G->st))) return rc;
        if (mark && (rc = albd_event_record(d, 7))) return rc;
        return albd_push_finalize(d, &G->g, &G->st, level, INT64_MAX);
    }
    int variant = ALBD_PULL_ALL;
    if (choice == ALBBA_CHOICE_PULL_SB) variant = ALBD_PULL_LIST;
    if (choice == ALBBA_CHOICE_PULL_CB) variant = ALBD_PULL_LIST_CB;
    if (variant != ALBD_PULL_ALL &&
        (rc = albd_build_active(d, &G->g, &G->st, ls->list_valid, ls->list_len, level)))
        return rc;
    if (variant == ALBD_PULL_LIST_CB && (rc = albd_build_summary(d, &G->g, &G->st))) return rc;
    if (mark && (rc = albd_event_record(d, 6))) return rc;
    /* a rebuilt list is never longer than the list it was compacted from */
    const int64_t bound = ls->list_valid ? ls->list_len : G->g.nwords;
    if ((rc = albd_pull(d, &G->g, &G->st, variant, level, bound))) return rc;
    return mark ? albd_event_record(d, 7) : 0;
}

/* Choice-independent preparation of the next level, launched behind the
 * select kernel so that it runs while the host waits for the choice: zero
 * `next` and the slots, move the queue/list counters, compact the
 * unvisited-word list (slice bypass) and build the frontier summary (column
 * bypass).  A list that the chosen kernel does not use is dropped. */
static int prep_level(albba_graph *G, const loop_state *ls, int32_t level)
{
    /* one launch sized for every word (the list length is on the device) */
    return albd_prep(G->ctx->d, &G->g, &G->st, ls->list_valid, albd_cb_supported(&G->g), level);
}

/* The select kernel of level sel->level and that level's preparation
 * (prep_level) in one launch; the counters come to the mailbox (seq). */
static int select_prep(albba_graph *G, const loop_state *ls, const albd_select_args *sel, unsigned int *seq)
{
    return albd_select_prep_post(G->ctx->d, &G->g, &G->st, sel, seq, ls->list_valid, albd_cb_supported(&G->g));
}

/* The kernels of one level after prep_level (no pointer swaps); a push
 * lists the next frontier's queue while its frontier edges are at most
 * build_max. */
/* The stored entries of the words the list pull of `level` scanned: the
 * column entries a full scan of those slices reads.  The pull's word list
 * and its length are copied back (after the pull, before the next level
 * rebuilds them) and summed over a per-word table made once per graph.  For
 * the per-level accounting only (time_levels). */
static int swept_entries(albba_graph *G, int32_t level, int64_t *out)
{
    albd_ctx *d = G->ctx->d;
    const int64_t nw = G->g.nwords;
    int rc;
    if (!G->word_entries) {
        int64_t *dev = NULL;
        G->word_entries = (int64_t *)albba_malloc(sizeof(int64_t) * (size_t)nw);
        G->list_copy = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)nw);
        rc = !G->word_entries || !G->list_copy ? ALBBA_ENOMEM : 0;
        if (!rc && !(rc = albd_malloc(d, sizeof(int64_t) * (size_t)nw, (void **)&dev)) &&
            !(rc = albd_word_entries(d, &G->g, dev)))
            rc = albd_d2h(d, G->word_entries, dev, sizeof(int64_t) * (size_t)nw);
        albd_free(d, dev);
        if (rc) {
            free(G->word_entries);
            free(G->list_copy);
            G->word_entries = NULL;
            G->list_copy = NULL;
            return rc == ALBBA_ENOMEM ? albba_fail(rc, "bfs: swept entries: no memory")
                                      : albba_fail(rc, "bfs: swept entries: %s", albd_error());
        }
    }
    unsigned long long cnt = 0;
    if ((rc = albd_d2h(d, &cnt, &G->st.ctr->list_cnt[level % 3], sizeof cnt)))
        return albba_fail(rc, "bfs: swept entries: %s", albd_error());
    if (cnt > (unsigned long long)nw) cnt = (unsigned long long)nw;
    if (cnt && (rc = albd_d2h(d, G->list_copy, G->st.active_next, sizeof(int32_t) * (size_t)cnt)))
        return albba_fail(rc, "bfs: swept entries: %s", albd_error());
    int64_t sum = 0;
    for (unsigned long long k = 0; k < cnt; ++k) sum += G->word_entries[G->list_copy[k]];
    *out = sum;
    return 0;
}

static int run_main(albba_graph *G, const loop_state *ls, int choice, int32_t level,
                    int64_t frontier, int mark, int64_t build_max)
{
    albd_ctx *d = G->ctx->d;
    int rc;
    if (choice == ALBBA_CHOICE_PUSH) {
        if (!ls->have_queue && (rc = albd_front_to_queue(d, &G->g, &G->st))) return rc;
        if ((rc = albd_push_scan(d, &G->g, &G->st, frontier))) return rc;
        if (mark && (rc = albd_event_record(d, 6))) return rc;
        if ((rc = albd_push_expand(d, &G->g, &G->st))) return rc;
        if (mark && (rc = albd_event_record(d, 7))) return rc;
        return albd_push_finalize(d, &G->g, &G->st, level, build_max);
    }
    int variant = ALBD_PULL_ALL;
    if (choice == ALBBA_CHOICE_PULL_SB) variant = ALBD_PULL_LIST;
    if (choice == ALBBA_CHOICE_PULL_CB) variant = ALBD_PULL_LIST_CB;
    if (mark && (rc = albd_event_record(d, 6))) return rc;
    const int64_t bound = ls->list_valid ? ls->list_len : G->g.nwords;
    if ((rc = albd_pull(d, &G->g, &G->st, variant, level, bound))) return rc;
    return mark ? albd_event_record(d, 7) : 0;
}

/* Push levels in the single-workgroup kernel, starting with the level whose
 * frontier is st.queue (or the front bitmap after a pull).  Returns the
 * number of levels run in *nlev; afterwards the state is that of one
 * finished push level (queue_next holds the last frontier's queue). */
static int run_small(albba_graph *G, const loop_state *ls, int32_t level, int64_t frontier,
                     int64_t edge_cap, int64_t vertex_cap, int32_t blocks, int mark, int prepped,
                     int *nlev)
{
    albd_ctx *d = G->ctx->d;
    (void)prepped; /* the counters moved in the select kernel either way */
    int rc = albd_level_begin_flags(d, &G->g, &G->st, LB_SMALL);
    if (rc) return rc;
    if (!ls->have_queue && (rc = albd_front_to_queue(d, &G->g, &G->st))) return rc;
    if (mark && (rc = albd_event_record(d, 6))) return rc;
    rc = albd_push_small(d, &G->g, &G->st, level, edge_cap, vertex_cap, ALBD_PERSIST_LEVELS,
                         blocks, 1u << 22);
    if (rc) return rc;
    if (mark && (rc = albd_event_record(d, 7))) return rc;
    if ((rc = albd_read_counters(d, &G->st))) return rc;
    const albd_counters *hc = albd_host_counters(d);
    if (hc->persist_error) {
        /* a grid barrier gave up: the keys may be left set */
        albd_fill_u64(d, G->st.pkey, ~0ull, (size_t)G->g.npos);
        albd_sync(d);
        return albba_fail(ALBBA_EHIP, "small-frontier kernel: grid barrier timed out");
    }
    *nlev = hc->persist_levels;
    if (hc->persist_in_queue) {
        /* the last frontier's queue is in st.queue: swap so that it is in
         * queue_next, where a finished push level leaves it */
        int32_t *q = G->st.queue;
        G->st.queue = G->st.queue_next;
        G->st.queue_next = q;
    }
    return 0;
}

/* Pointer swaps and bookkeeping after a committed level (list length from
 * the level's counters: commit_counts); queue_built: a pushed level listed
 * the next frontier's queue. */
static void commit_pointers(albba_graph *G, loop_state *ls, int choice, int queue_built)
{
    uint64_t *t = G->st.front;
    G->st.front = G->st.next;
    G->st.next = t;
    if (ls->list_pending) {
        /* the list built one level earlier is now st.active (see below) */
        ls->list_pending = 0;
    }
    if (choice == ALBBA_CHOICE_PUSH) {
        int32_t *q = G->st.queue;
        G->st.queue = G->st.queue_next;
        G->st.queue_next = q;
        ls->have_queue = queue_built;
        ls->prev_push = 1;
    } else {
        ls->have_queue = 0;
        ls->prev_push = 0;
        if (choice != ALBBA_CHOICE_PULL) {
            int32_t *a = G->st.active;
            G->st.active = G->st.active_next;
            G->st.active_next = a;
            ls->list_valid = 1;
            ls->list_pending = 1;
        }
    }
}

/* After the select kernel (which hands a used unvisited-word list over as
 * the current one): the current list's length bounds the next lists. */
static void commit_counts(loop_state *ls, int choice, const level_sums *sums)
{
    if (choice == ALBBA_CHOICE_PULL_SB || choice == ALBBA_CHOICE_PULL_CB) ls->list_len = sums->active_len;
}

static int choice_ok(const albba_graph *G, int c)
{
    if (c < 0 || c >= ALBBA_NUM_CHOICES) return 0;
    if (c == ALBBA_CHOICE_PULL_CB && !albd_cb_supported(&G->g)) return 0;
    return 1;
}

static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* events of a BFS: its start and end (a second pair lets the next BFS of a
 * sequence start before the previous one's time is read) */
typedef struct bfs_events {
    int start, end;
    int defer; /* leave bfs_ms to the caller, which reads the pair later */
} bfs_events;

static int bfs_run(albba_ctx *ctx, albba_graph *G, int32_t root, const albba_bfs_opts *opts_in,
                   albba_bfs_stats *stats, bfs_events ev);

/* pmin and the small-frontier keys are all ones at rest (the kernels restore
 * what they use); after a failed BFS they are refilled before the next one */
static int bfs_checked(albba_ctx *ctx, albba_graph *G, int32_t root, const albba_bfs_opts *opts_in,
                       albba_bfs_stats *stats, bfs_events ev)
{
    if (ctx && G && G->ctx == ctx && G->needs_reset) {
        int rc = albd_fill_u32(ctx->d, G->st.pmin, 0xFFFFFFFFu, (size_t)G->g.npos);
        if (!rc) rc = albd_fill_u64(ctx->d, G->st.pkey, ~0ull, (size_t)G->g.npos);
        if (!rc) rc = albd_sync(ctx->d);
        if (rc) return albba_fail(rc, "bfs: %s", albd_error());
        G->needs_reset = 0;
    }
    int rc = bfs_run(ctx, G, root, opts_in, stats, ev);
    if (rc && rc != ALBBA_EINVAL && G) G->needs_reset = 1;
    return rc;
}

int albba_bfs(albba_ctx *ctx, albba_graph *G, int32_t root, const albba_bfs_opts *opts_in,
              albba_bfs_stats *stats)
{
    const bfs_events ev = {0, 1, 0};
    return bfs_checked(ctx, G, root, opts_in, stats, ev);
}

int albba_bfs_multi(albba_ctx *ctx, albba_graph *G, const int32_t *roots, int32_t nroots,
                    const albba_bfs_opts *opts, int32_t nopts, albba_bfs_stats *stats)
{
    if (!ctx || !G || G->ctx != ctx) return albba_fail(ALBBA_EINVAL, "bfs_multi: bad context/graph");
    if (nroots < 0 || (nroots > 0 && !roots))
        return albba_fail(ALBBA_EINVAL, "bfs_multi: %d roots", nroots);
    if (nopts != 1 && nopts != nroots)
        return albba_fail(ALBBA_EINVAL, "bfs_multi: nopts %d must be 1 or nroots %d", nopts, nroots);
    if (!opts) return albba_fail(ALBBA_EINVAL, "bfs_multi: no options");
    for (int32_t i = 0; i < nroots; ++i)
        if (roots[i] < 0 || roots[i] >= G->n)
            return albba_fail(ALBBA_EINVAL, "bfs_multi: root %d out of range [0,%d)", roots[i], G->n);
    albd_ctx *d = ctx->d;
    int rc = 0;
    for (int32_t i = 0; i < nroots && !rc; ++i) {
        /* BFS i uses pair i % 2; its level loop waited on mailboxes written
         * after BFS i-1 ended (one stream), so that pair is complete */
        const bfs_events ev = {(i & 1) ? 8 : 0, (i & 1) ? 9 : 1, 1};
        rc = bfs_checked(ctx, G, roots[i], nopts == 1 ? opts : opts + i, stats ? stats + i : NULL, ev);
        if (!rc && i > 0 && stats) {
            const int s = (i & 1) ? 0 : 8;
            float ms = 0;
            if ((rc = albd_event_elapsed(d, s, s + 1, &ms))) rc = albba_fail(rc, "bfs_multi: %s", albd_error());
            stats[i - 1].bfs_ms = ms;
        }
    }
    if (!rc && nroots > 0) {
        const int s = ((nroots - 1) & 1) ? 8 : 0;
        float ms = 0;
        if (stats && !(rc = albd_event_elapsed(d, s, s + 1, &ms))) stats[nroots - 1].bfs_ms = ms;
        if (!rc) rc = albd_sync(d);
        if (rc) rc = albba_fail(rc, "bfs_multi: %s", albd_error());
    }
    return rc;
}

/* The script's choice for `level`, or -1 past its end or where it names an
 * unavailable choice (the heuristic's then). */
static int script_at(const albba_graph *G, const albba_bfs_opts *o, int32_t level)
{
    if (o->switch_policy != ALBBA_SWITCH_SCRIPT || !o->script || level < 0 || level >= o->script_len)
        return -1;
    return choice_ok(G, o->script[level]) ? o->script[level] : -1;
}

static int bfs_run(albba_ctx *ctx, albba_graph *G, int32_t root, const albba_bfs_opts *opts_in,
                   albba_bfs_stats *stats, bfs_events ev)
{
    if (!ctx || !G || G->ctx != ctx) return albba_fail(ALBBA_EINVAL, "bfs: bad context/graph");
    if (root < 0 || root >= G->n) return albba_fail(ALBBA_EINVAL, "bfs: root %d out of range [0,%d)", root, G->n);
    albba_bfs_opts opts;
    if (opts_in) opts = *opts_in;
    else albba_bfs_opts_default(&opts);
    const int policy = opts.switch_policy;
    if (policy < ALBBA_SWITCH_HEURISTIC || policy > ALBBA_SWITCH_BEST)
        return albba_fail(ALBBA_EINVAL, "bfs: unknown switch policy %d", policy);
    if (policy == ALBBA_SWITCH_FIXED && !choice_ok(G, opts.fixed_choice))
        return albba_fail(ALBBA_EINVAL, "bfs: fixed choice %d not available", opts.fixed_choice);
    if (policy == ALBBA_SWITCH_LEARNED && !opts.selector)
        return albba_fail(ALBBA_EINVAL, "bfs: learned policy needs a selector");
    albd_ctx *d = ctx->d;
    const int reps = opts.timing_reps > 0 ? opts.timing_reps : 1;
    const int32_t rp = G->pos_of[root];
    int valid_mask = 0;
    for (int c = 0; c < ALBBA_NUM_CHOICES; ++c) valid_mask |= choice_ok(G, c) << c;
    albd_select_args sel = {policy == ALBBA_SWITCH_LEARNED ? albba_selector_dparams(opts.selector) : NULL,
                            0, ALBBA_CHOICE_PUSH, valid_mask, 1, script_at(G, &opts, 0)};

    /* the heuristic pushes while F <= (int)(n * 0.04) */
    const int64_t push_vertex_cap = (int64_t)(int)((double)G->n * 0.04);
    /* A push level lists the next frontier's queue only while its frontier
     * edges (which bound the next frontier) stay at most queue_max: past it
     * the next level is likely a pull, which needs no queue (a push after it
     * lists the bitmap instead).  The heuristic pushes exactly while the
     * frontier is at most push_vertex_cap; the learned selector turns to
     * pulls earlier, at the ratio its training chose (see
     * albba_selector_set_queue_ratio). */
    const float qratio = policy == ALBBA_SWITCH_LEARNED ? albba_selector_queue_ratio(opts.selector) : 0.0f;
    const int64_t learned_qmax = qratio > 0.0f ? (int64_t)ceil((double)qratio * (double)G->n) : push_vertex_cap;
    const int64_t queue_max_policy = policy == ALBBA_SWITCH_LEARNED ? learned_qmax : push_vertex_cap;
    /* a scripted level lists the queue exactly when the script pushes next
     * (-1: never; past the script the heuristic's rule) */
#define QUEUE_MAX(L)                                                                                   \
    (policy != ALBBA_SWITCH_SCRIPT || script_at(G, &opts, (L) + 1) < 0                                 \
         ? queue_max_policy                                                                            \
         : (script_at(G, &opts, (L) + 1) == ALBBA_CHOICE_PUSH ? INT64_MAX : (int64_t)-1))
    loop_state ls = {1, 0, 0, 0, 0};
    int small_streak = 0; /* consecutive plain push levels that stayed small */
    int32_t level = 0;
    int rc;
    /* The heuristic, learned and scripted policies prepare every level
     * speculatively beside the select kernel (select_prep), which also
     * writes their choice (sel_choice: the MLP's or the script's, else the
     * heuristic's); the others prepare each level after their choice
     * (run_level). */
    const int spec = policy == ALBBA_SWITCH_HEURISTIC || policy == ALBBA_SWITCH_LEARNED ||
                     policy == ALBBA_SWITCH_SCRIPT;
    const int use_sel = policy != ALBBA_SWITCH_HEURISTIC; /* kernels queued ahead read sel_choice */
    unsigned int seq = 0;
    /* kernels already queued for the coming level: the choice whose pull is
     * queued, or -1; whether the push kernels are */
    int queued_pull = -1, queued_push = 0;
    /* the vertex-order pass queued behind the last posted level (it runs only
     * if that level's select kernel finds the frontier empty), and the
     * frontier of the level before the current one */
    int vo_queued = 0;
    int64_t frontier_prev = 0;
    /* init, then the select kernel turns the root into level 0's frontier */
    if ((rc = albd_event_record(d, ev.start)) || (rc = albd_bfs_init(d, &G->g, &G->st, rp, root)))
        return albba_fail(rc, "bfs: %s", albd_error());
    if (spec)
        rc = select_prep(G, &ls, &sel, &seq) ||
             /* the root level is a push as a rule; the small-frontier
              * kernel needs two push levels before it */
             (!opts.time_levels && (queued_push = 1, albd_push_spec(d, &G->g, &G->st, 0, use_sel,
                                                                    QUEUE_MAX(0)))) ||
             albd_mail_wait(d, &G->st, seq);
    else
        rc = albd_select_read(d, &G->g, &G->st, &sel);
    if (rc) return albba_fail(ALBBA_EHIP, "bfs: %s", albd_error());
    const albd_counters *hc = albd_host_counters(d);
    int64_t frontier = (int64_t)hc->frontier;

    while (frontier > 0) {
        float feats[ALBBA_NUM_FEATURES];
        memcpy(feats, hc->features, sizeof feats);
        /* heuristic, learned and scripted choices come from the select
         * kernel (the learned and scripted ones already fall back to the
         * heuristic) */
        int choice = use_sel ? hc->sel_choice : hc->heur_choice;
        float choice_ms[ALBBA_NUM_CHOICES] = {0, 0, 0, 0};
        const int64_t queue_max = QUEUE_MAX(level);
        if (policy == ALBBA_SWITCH_FIXED) {
            choice = opts.fixed_choice;
        } else if (policy == ALBBA_SWITCH_BEST) {
            /* time every choice on a snapshot of the state, commit the fastest */
            if ((rc = albd_d2d(d, G->scratch_visited, G->st.visited, sizeof(uint64_t) * (size_t)G->g.nwords)) ||
                (rc = albd_d2d(d, G->scratch_ctr, G->st.ctr, sizeof(albd_counters))))
                return albba_fail(rc, "bfs: %s", albd_error());
            float best = INFINITY;
            for (int c = 0; c < ALBBA_NUM_CHOICES; ++c) {
                if (!choice_ok(G, c)) { choice_ms[c] = INFINITY; continue; }
                float t[16];
                int nr = reps < 16 ? reps : 16;
                for (int r = 0; r < nr; ++r) {
                    /* a choice 4x slower than the best so far cannot win:
                     * one trial is enough to rank it */
                    if (r == 1 && t[0] > 4.0f * best) { nr = 1; break; }
                    /* the state after the select kernel, the preparation the
                     * deployed loop queues behind it for every choice (zero
                     * next, unvisited-word list, frontier summary), then the
                     * choice's own kernels -- the only part that is timed, as
                     * only that part differs between choices when deployed */
                    if ((rc = albd_d2d(d, G->st.visited, G->scratch_visited, sizeof(uint64_t) * (size_t)G->g.nwords)) ||
                        (rc = albd_d2d(d, G->st.ctr, G->scratch_ctr, sizeof(albd_counters))) ||
                        (rc = albd_level_begin_flags(d, &G->g, &G->st, LB_SLOTS)) || (rc = prep_level(G, &ls, level)) ||
                        (rc = albd_event_record(d, 4)) || (rc = run_main(G, &ls, c, level, frontier, 0, INT64_MAX)) ||
                        (rc = albd_event_record(d, 5)) || (rc = albd_event_elapsed(d, 4, 5, &t[r])))
                        return albba_fail(rc, "bfs: %s", albd_error());
                }
                qsort(t, (size_t)nr, sizeof(float), cmp_float);
                choice_ms[c] = t[nr / 2];
                if (choice_ms[c] < best) { best = choice_ms[c]; choice = c; }
            }
            if ((rc = albd_d2d(d, G->st.visited, G->scratch_visited, sizeof(uint64_t) * (size_t)G->g.nwords)) ||
                (rc = albd_d2d(d, G->st.ctr, G->scratch_ctr, sizeof(albd_counters))))
                return albba_fail(rc, "bfs: %s", albd_error());
        }
        /* A pushed level with a small frontier, after two more such levels,
         * starts a run of push levels in the persistent kernel; it continues
         * while the frontier stays within the edge cap and within the
         * heuristic's push threshold (bfs/src/bfs_SELLCSv2.c:282-284), so the
         * heuristic's choices are unchanged.  The persistent launch costs
         * more than one level: it pays off on slowly growing frontiers
         * (high-diameter graphs), which the streak detects, not on the one or
         * two small levels of a power-law BFS. */
        const int64_t first_edges = (int64_t)hc->frontier_edges;
        const int small_ok = choice == ALBBA_CHOICE_PUSH && opts.small_frontier_edges > 0 &&
                             (policy == ALBBA_SWITCH_HEURISTIC || policy == ALBBA_SWITCH_LEARNED) &&
                             first_edges <= opts.small_frontier_edges && frontier <= push_vertex_cap;
        /* (never with push kernels queued: they are queued only while
         * small_streak < 2) */
        const int small = small_ok && small_streak >= 2 && !queued_push;
        small_streak = small_ok && !small ? small_streak + 1 : 0;
        int nlev = 1;
        int queue_built = 1;
        if (opts.time_levels && (rc = albd_event_record(d, 2))) return albba_fail(rc, "bfs: %s", albd_error());
        if (small)
            rc = run_small(G, &ls, level, frontier, opts.small_frontier_edges, push_vertex_cap,
                           opts.small_frontier_blocks, opts.time_levels, spec, &nlev);
        else if (spec) {
            const int ran = choice == queued_pull ||
                            (queued_push && choice == ALBBA_CHOICE_PUSH && frontier <= ALBD_SPEC_PUSH_MAX);
            /* the finalize kernel applies queue_max to the select kernel's
             * edge count, queued ahead or not */
            queue_built = choice != ALBBA_CHOICE_PUSH || first_edges <= queue_max;
            if (!ran) rc = run_main(G, &ls, choice, level, frontier, opts.time_levels, queue_max);
        }
        else
            rc = run_level(G, &ls, choice, level, frontier, opts.time_levels);
        if (rc) return albba_fail(rc, "bfs: %s", albd_error());
        if (opts.time_levels && (rc = albd_event_record(d, 3))) return albba_fail(rc, "bfs: %s", albd_error());
        int64_t swept = -1;
        if (opts.time_levels && stats && stats->swept && level < stats->capacity && !small &&
            (choice == ALBBA_CHOICE_PULL_SB || choice == ALBBA_CHOICE_PULL_CB) &&
            (rc = swept_entries(G, level, &swept)))
            return rc;
        if (small && nlev > 1 && stats &&
            (rc = albd_d2h(d, G->host_log, G->st.level_log, sizeof(int64_t) * 4 * (size_t)(nlev - 1))))
            return albba_fail(rc, "bfs: %s", albd_error());
        sel.level = level + nlev;
        sel.prev_choice = choice;
        sel.first = 0;
        sel.forced = script_at(G, &opts, level + nlev);
        if (spec) {
            /* swap to the next level's state, post the select kernel with
             * the next level's preparation beside it, then wait for the
             * counters */
            commit_pointers(G, &ls, choice, queue_built);
            if ((rc = select_prep(G, &ls, &sel, &seq))) return albba_fail(rc, "bfs: %s", albd_error());
            /* Queue the likely next kernels now, so that the device does
             * not wait for the host to read the choice: after a pull the
             * same pull, as a rule; after a push a push while it listed the
             * next queue (queue_max), else the pull the policy takes for
             * large frontiers.  The kernels read the choice themselves and
             * return at once if it differs; the host then queues the chosen
             * kernels behind them.  Not with per-level events, which must
             * bracket the chosen kernels, and no push while the
             * small-frontier kernel may take the next level. */
            queued_pull = -1;
            queued_push = 0;
            int guess = choice;
            if (choice == ALBBA_CHOICE_PUSH && !small) {
                if (first_edges > queue_max)
                    guess = policy == ALBBA_SWITCH_LEARNED && albd_cb_supported(&G->g) ? ALBBA_CHOICE_PULL_CB
                                                                                        : ALBBA_CHOICE_PULL_SB;
                else if (small_streak >= 2)
                    guess = -1;
            }
            if (sel.forced >= 0) {
                /* a scripted level is known: queue it, except a push that
                 * would need the queue listed from the bitmap first */
                guess = sel.forced;
                if (guess == ALBBA_CHOICE_PUSH && !(choice == ALBBA_CHOICE_PUSH && !small && queue_built))
                    guess = -1;
            }
            if (guess == ALBBA_CHOICE_PUSH && !opts.time_levels) {
                if ((rc = albd_push_spec(d, &G->g, &G->st, level + nlev, use_sel,
                                         QUEUE_MAX(level + nlev))))
                    return albba_fail(rc, "bfs: %s", albd_error());
                queued_push = 1;
            }
            if (guess >= 0 && guess != ALBBA_CHOICE_PUSH && !opts.time_levels) {
                const int variant = guess == ALBBA_CHOICE_PULL_SB
                                        ? ALBD_PULL_LIST
                                        : (guess == ALBBA_CHOICE_PULL_CB ? ALBD_PULL_LIST_CB : ALBD_PULL_ALL);
                /* lists only shrink: the last length the host knows bounds
                 * the list the preparation builds */
                if ((rc = albd_pull_spec(d, &G->g, &G->st, variant, level + nlev,
                                         ls.list_valid && ls.list_len > 0 ? ls.list_len : G->g.nwords,
                                         use_sel)))
                    return albba_fail(rc, "bfs: %s", albd_error());
                queued_pull = guess;
            }
            /* A power-law BFS ends with a level of a few vertices after one of
             * thousands (RMAT s24: 1-141 after 1 222-42 840): after such a
             * level the vertex-order pass goes behind the next level's
             * kernels too, so that the device does not wait for the host to
             * read that the frontier is empty.  A wrong guess costs one
             * kernel that returns at once. */
            vo_queued = 0;
            if (!opts.time_levels && !small && nlev == 1 && frontier <= 1024 && 64 * frontier < frontier_prev) {
                if ((rc = albd_to_vertex_order(d, &G->g, &G->st, 1)))
                    return albba_fail(rc, "bfs: %s", albd_error());
                vo_queued = 1;
            }
            if ((rc = albd_mail_wait(d, &G->st, seq))) return albba_fail(rc, "bfs: %s", albd_error());
        } else if ((rc = albd_select_read(d, &G->g, &G->st, &sel))) {
            return albba_fail(rc, "bfs: %s", albd_error());
        }
        level_sums sums;
        sum_counters(hc, &sums);
        if (stats && level < stats->capacity) {
            float lms = 0, kms = 0;
            if (opts.time_levels) {
                albd_event_elapsed(d, 2, 3, &lms);
                albd_event_elapsed(d, 6, 7, &kms);
            }
            for (int j = 0; j < nlev && level + j < stats->capacity; ++j) {
                const int32_t lv = level + j;
                const int push = choice == ALBBA_CHOICE_PUSH;
                /* levels after the first of a small-frontier run take their
                 * sizes from the kernel's level log */
                const int64_t fj = j == 0 ? frontier : G->host_log[4 * (j - 1)];
                if (stats->frontier) stats->frontier[lv] = (int32_t)fj;
                if (stats->choice) stats->choice[lv] = choice;
                if (stats->features) {
                    float *f = stats->features + (size_t)lv * ALBBA_NUM_FEATURES;
                    if (j == 0) memcpy(f, feats, sizeof feats);
                    else memset(f, 0, sizeof feats);
                }
                if (stats->choice_ms) memcpy(stats->choice_ms + (size_t)lv * ALBBA_NUM_CHOICES, choice_ms, sizeof choice_ms);
                if (stats->level_ms) stats->level_ms[lv] = lms / (float)nlev;
                if (stats->kernel_ms) stats->kernel_ms[lv] = kms / (float)nlev;
                if (small) {
                    if (stats->entries) stats->entries[lv] = j == 0 ? first_edges : G->host_log[4 * (j - 1) + 1];
                    if (stats->words) stats->words[lv] = fj;
                } else {
                    if (stats->entries) stats->entries[lv] = push ? sums.push_edges : sums.scanned;
                    if (stats->words) stats->words[lv] = push ? sums.queue_len : sums.words;
                }
                /* words the pull did not scan: off the unvisited-word list,
                 * or on it but visited or empty by now */
                if (stats->bypassed) stats->bypassed[lv] = push || small ? 0 : (int64_t)G->g.nwords - sums.words;
                if (stats->col_bypassed) stats->col_bypassed[lv] = push || small ? 0 : sums.col_bypassed;
                if (stats->swept) stats->swept[lv] = j == 0 ? swept : -1;
            }
        }
        if (!spec) commit_pointers(G, &ls, choice, 1);
        commit_counts(&ls, choice, &sums);
        frontier_prev = frontier;
        frontier = (int64_t)hc->frontier;
        level += nlev;
    }
    float ms = 0;
    if ((!vo_queued && (rc = albd_to_vertex_order(d, &G->g, &G->st, 0))) || (rc = albd_event_record(d, ev.end)) ||
        (!ev.defer && (rc = albd_event_elapsed(d, ev.start, ev.end, &ms))))
        return albba_fail(rc, "bfs: %s", albd_error());
    if (stats) {
        stats->num_levels = level;
        stats->reached = (int64_t)hc->reached;
        stats->edges_in_component = (int64_t)hc->reached_edges;
        if (!ev.defer) stats->bfs_ms = ms;
    }
    return ALBBA_OK;
}
#undef QUEUE_MAX

int albba_bfs_results(albba_ctx *ctx, albba_graph *G, int32_t *level, int32_t *parent)
{
    if (!ctx || !G) return albba_fail(ALBBA_EINVAL, "bfs_results: null");
    int rc = 0;
    if (level && (rc = albd_d2h(ctx->d, level, G->st.level, sizeof(int32_t) * (size_t)G->n)))
        return albba_fail(rc, "bfs_results: %s", albd_error());
    if (parent && (rc = albd_d2h(ctx->d, parent, G->st.parent, sizeof(int32_t) * (size_t)G->n)))
        return albba_fail(rc, "bfs_results: %s", albd_error());
    return ALBBA_OK;
}

/* ---- bfs/src/bfs.h:13 compatibility ---- */

void run_custom_bfs_test(const SparseMatrixCSR *matrix, elem_t *bfs_x, int *colidx_x, int *nnzx,
                         int *frontier, int *num_vex, double *conv_time, double *iter_times,
                         double *bfs_time, int *num_iter, bool bit_flag, bool sort_flag)
{
    (void)bfs_x;
    (void)colidx_x;
    (void)bit_flag; /* the frontier is always a bitmap here */
    const int n = matrix->nrows;
    albba_csr g;
    g.n = n;
    g.nnz = matrix->nnz;
    g.symmetric = 0;
    g.row_ptr = (int64_t *)albba_malloc(sizeof(int64_t) * ((size_t)n + 1));
    g.col = matrix->column_indices;
    if (!g.row_ptr) { fprintf(stderr, "run_custom_bfs_test: out of memory\n"); exit(1); }
    for (int i = 0; i <= n; ++i) g.row_ptr[i] = matrix->row_pointers[i];
    int sigma = sigma_window < 64 ? 64 : (sigma_window / 64) * 64;
    double t0 = albba_now_ms();
    albba_sellcs *s = NULL;
    /* sort_flag: the reference's -s 1 row order (bfs/src/bfs_SELLCSv2.c:40-45) */
    int rc = albba_sellcs_build_ex(&g, 64, sigma, 1, sort_flag ? ALBBA_SELL_SORT_FIRST_COL : 0, &s);
    *conv_time = albba_now_ms() - t0;
    free(g.row_ptr);
    albba_ctx *ctx = NULL;
    albba_graph *G = NULL;
    if (rc || (rc = albba_ctx_create(0, &ctx)) || (rc = albba_graph_upload(ctx, s, &G))) {
        fprintf(stderr, "run_custom_bfs_test: %s\n", albba_last_error());
        exit(1);
    }
    albba_sellcs_free(s);
    int cap = n > 0 ? n : 1;
    float *lms = (float *)albba_calloc((size_t)cap, sizeof(float));
    albba_bfs_stats st;
    memset(&st, 0, sizeof st);
    st.capacity = cap;
    st.frontier = frontier;
    st.level_ms = lms;
    albba_bfs_opts o;
    albba_bfs_opts_default(&o);
    o.time_levels = 1;
    if ((rc = albba_bfs(ctx, G, 0, &o, &st))) {
        fprintf(stderr, "run_custom_bfs_test: %s\n", albba_last_error());
        exit(1);
    }
    *num_iter = st.num_levels;
    *num_vex = (int)st.reached;
    *bfs_time = st.bfs_ms;
    if (iter_times)
        for (int i = 0; i < st.num_levels && i < cap; ++i) iter_times[i] = lms[i];
    *nnzx = 0;
    free(lms);
    albba_graph_free(G);
    albba_ctx_destroy(ctx);
}
