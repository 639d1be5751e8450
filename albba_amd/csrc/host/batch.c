/*
 * Batched BFS host side: per-instance device scratch, the split of the
 * resident workgroups into teams, one persistent launch, per-instance
 * results from the device clock.
 *
 * Team split (split_teams): an instance's work is its stored entries plus
 * its vertices; instances worth at least two workgroups of the fair share get
 * a team of their own, the others share one-workgroup teams, largest first
 * onto the least loaded team.
 */
#include "internal.h"
#include "device_int.h"
#include "selector_int.h"

#define CTL_WORDS 32

struct albba_batch {
    albba_ctx *ctx;
    int32_t count;
    int32_t wcap;                 /* resident workgroups of the batch kernel */
    albba_graph **graphs;
    albd_batch_item *items;       /* host copies (pinned, head of `stage`) */
    unsigned long long *ctl;      /* device: count * CTL_WORDS */
    unsigned long long *hctl;     /* host mirror (pinned) */
    void **allocs;                /* device allocations to free */
    int32_t nallocs;
    int32_t *team_of;             /* instance -> team, team sizes of the last run */
    int32_t *team_size_of;
    /* host work arrays of a run, sized at creation for `wcap` workgroups, so
     * that a run allocates no host memory */
    int32_t *order, *share, *cnt, *team_first, *team_size, *team_ptr, *team_items, *wg_team;
    double *load;
    /* pinned staging of a run's device tables: the instance descriptors, then
     * (at tab_off bytes) the team tables */
    void *stage;
    size_t tab_off, stage_bytes;
};

/* device bytes of a run's tables: instance descriptors, then the team tables
 * (W workgroups, T teams) at a 256-byte boundary */
static size_t tab_words(int32_t N, int32_t W, int32_t T) { return (size_t)W + 3 * (size_t)T + 1 + (size_t)N; }
static size_t tab_offset(int32_t N) { return (sizeof(albd_batch_item) * (size_t)N + 255) / 256 * 256; }

static void *bdalloc(albba_batch *b, size_t bytes, int *rc)
{
    void *p = NULL;
    if (*rc) return NULL;
    *rc = albd_malloc(b->ctx->d, bytes ? bytes : 8, &p);
    if (!*rc) b->allocs[b->nallocs++] = p;
    return p;
}

void albba_batch_free(albba_batch *b)
{
    if (!b) return;
    for (int32_t i = 0; i < b->nallocs; ++i) albd_free(b->ctx->d, b->allocs[i]);
    albd_pinned_free(b->ctx->d, b->stage);
    albd_pinned_free(b->ctx->d, b->hctl);
    free(b->allocs);
    free(b->graphs);
    free(b->team_of);
    free(b->team_size_of);
    free(b->order); free(b->share); free(b->cnt); free(b->team_first); free(b->team_size);
    free(b->team_ptr); free(b->team_items); free(b->wg_team); free(b->load);
    free(b);
}

int albba_batch_create(albba_ctx *ctx, albba_graph *const *graphs, int32_t count, albba_batch **out)
{
    if (!ctx || !graphs || count <= 0 || !out) return albba_fail(ALBBA_EINVAL, "batch_create: bad argument");
    for (int32_t i = 0; i < count; ++i)
        if (!graphs[i] || graphs[i]->ctx != ctx)
            return albba_fail(ALBBA_EINVAL, "batch_create: graph %d missing or on another context", i);
    int wcap = 0;
    int rc = albd_batch_capacity(ctx->d, &wcap);
    if (rc) return albba_fail(rc, "batch_create: %s", albd_error());
    if (wcap < 1) return albba_fail(ALBBA_EHIP, "batch_create: no resident workgroup");
    albba_batch *b = (albba_batch *)albba_calloc(1, sizeof(albba_batch));
    if (!b) return albba_fail(ALBBA_ENOMEM, "batch_create");
    b->ctx = ctx;
    b->count = count;
    b->wcap = wcap;
    const size_t N = (size_t)count, W = (size_t)wcap, T = N < W ? N : W;
    b->graphs = (albba_graph **)albba_calloc(N, sizeof(albba_graph *));
    b->allocs = (void **)albba_calloc(N * 12 + 1, sizeof(void *));
    b->team_of = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->team_size_of = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->order = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->share = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->cnt = (int32_t *)albba_calloc(W + 1, sizeof(int32_t));
    b->load = (double *)albba_calloc(W, sizeof(double));
    b->team_first = (int32_t *)albba_calloc(T, sizeof(int32_t));
    b->team_size = (int32_t *)albba_calloc(T, sizeof(int32_t));
    b->team_ptr = (int32_t *)albba_calloc(T + 1, sizeof(int32_t));
    b->team_items = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->wg_team = (int32_t *)albba_calloc(W, sizeof(int32_t));
    b->tab_off = tab_offset(count);
    b->stage_bytes = b->tab_off + sizeof(int32_t) * tab_words(count, wcap, (int32_t)T);
    if (!b->graphs || !b->allocs || !b->team_of || !b->team_size_of || !b->order || !b->share ||
        !b->cnt || !b->load || !b->team_first || !b->team_size || !b->team_ptr || !b->team_items ||
        !b->wg_team) {
        albba_batch_free(b);
        return albba_fail(ALBBA_ENOMEM, "batch_create");
    }
    if ((rc = albd_pinned_alloc(ctx->d, b->stage_bytes, &b->stage)) ||
        (rc = albd_pinned_alloc(ctx->d, sizeof(unsigned long long) * CTL_WORDS * N, (void **)&b->hctl))) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s", albd_error());
        albba_batch_free(b);
        return albba_fail(rc, "batch_create: %s", buf);
    }
    b->items = (albd_batch_item *)b->stage;
    memset(b->stage, 0, b->stage_bytes);
    b->ctl = (unsigned long long *)bdalloc(b, sizeof(unsigned long long) * CTL_WORDS * (size_t)count, &rc);
    for (int32_t i = 0; i < count && !rc; ++i) {
        albba_graph *G = graphs[i];
        b->graphs[i] = G;
        albd_batch_item *it = &b->items[i];
        const size_t nw = (size_t)G->g.nwords, np = (size_t)G->g.npos, n = (size_t)G->g.n;
        it->g = G->g;
        it->visited = (uint64_t *)bdalloc(b, 8 * nw, &rc);
        for (int k = 0; k < 3; ++k) it->bm[k] = (uint64_t *)bdalloc(b, 8 * nw, &rc);
        it->found = (int32_t *)bdalloc(b, 4 * np, &rc);
        it->key = (unsigned long long *)bdalloc(b, 8 * np, &rc);
        it->level = (int32_t *)bdalloc(b, 4 * n, &rc);
        it->parent = (int32_t *)bdalloc(b, 4 * n, &rc);
        if (!rc) rc = albd_fill_u64(ctx->d, it->key, ~0ull, np);
        it->ctl = b->ctl + (size_t)i * CTL_WORDS;
        it->mapping = ALBD_MAP_LANE_ROW;
    }
    if (!rc) rc = albd_sync(ctx->d);
    if (rc) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s", albd_error());
        albba_batch_free(b);
        return albba_fail(rc, "batch_create: %s", buf);
    }
    *out = b;
    return ALBBA_OK;
}

static double inst_cost(const albba_graph *G) { return (double)G->g.nnz + (double)G->g.n; }

/* Split W workgroups into teams and give every instance a team.  Fills
 * team_first/team_size (teams), team_ptr/team_items (instances per team),
 * wg_team (workgroups); returns the number of teams.
 *
 * An instance's fair share is W * cost / total.  Instances whose share is at
 * least 2 workgroups get a team of their own of that size (at most 64, and
 * at most half of W together); the other instances go, largest first, to the
 * least loaded of the remaining one-workgroup teams. */
static int32_t split_teams(albba_batch *b, int32_t W, int32_t *team_first, int32_t *team_size,
                           int32_t *team_ptr, int32_t *team_items, int32_t *wg_team, int32_t *inst_team)
{
    const int32_t N = b->count;
    double total = 0;
    for (int32_t i = 0; i < N; ++i) total += inst_cost(b->graphs[i]);
    if (total <= 0) total = 1;
    int32_t *order = b->order, *share = b->share, *cnt = b->cnt;
    double *load = b->load;
    memset(load, 0, sizeof(double) * (size_t)W);
    memset(cnt, 0, sizeof(int32_t) * ((size_t)W + 1));
    for (int32_t i = 0; i < N; ++i) {
        order[i] = i;
        double s = (double)W * inst_cost(b->graphs[i]) / total;
        share[i] = s >= 64 ? 64 : (int32_t)s;
    }
    for (int32_t i = 1; i < N; ++i) { /* insertion sort by cost, descending */
        int32_t x = order[i], j = i - 1;
        while (j >= 0 && inst_cost(b->graphs[order[j]]) < inst_cost(b->graphs[x])) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = x;
    }
    /* own teams for the big instances, within half of the workgroups and
     * leaving one workgroup for every remaining instance when N is small */
    int32_t T = 0, used = 0, nbig = 0;
    for (int32_t k = 0; k < N; ++k) {
        const int32_t i = order[k];
        const int32_t rest = N - k - 1;          /* instances still to place */
        int32_t s = share[i];
        if (s < 2) break;
        const int32_t room = W - used - (rest < W / 2 ? rest : W / 2);
        if (s > room) s = room;
        if (s < 2 || used + s > (N <= W ? W : W / 2)) break;
        team_size[T] = s;
        inst_team[i] = T;
        cnt[T + 1] = 1;
        ++T;
        used += s;
        ++nbig;
    }
    /* one-workgroup teams for the rest, largest instance first */
    const int32_t small_teams = W - used < N - nbig ? W - used : N - nbig;
    const int32_t T0 = T;
    for (int32_t t = 0; t < small_teams; ++t) team_size[T0 + t] = 1;
    T += small_teams;
    for (int32_t k = nbig; k < N; ++k) {
        const int32_t i = order[k];
        int32_t t = T0;
        for (int32_t j = T0 + 1; j < T; ++j)
            if (load[j] < load[t]) t = j;
        load[t] += inst_cost(b->graphs[i]);
        inst_team[i] = t;
        ++cnt[t + 1];
    }
    for (int32_t t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
    for (int32_t t = 0; t <= T; ++t) team_ptr[t] = cnt[t];
    for (int32_t k = 0; k < N; ++k) {
        const int32_t i = order[k];
        team_items[cnt[inst_team[i]]++] = i;
    }
    int32_t wg = 0;
    for (int32_t t = 0; t < T; ++t) {
        team_first[t] = wg;
        for (int32_t s = 0; s < team_size[t]; ++s) wg_team[wg++] = t;
    }
    return T;
}

/* One run.  The device tables (instance descriptors, team tables) go to
 * fresh device memory, or with `noalloc` to the library's fixed scratch
 * array (albd_batch_scratch); the host side uses the batch's own arrays, so
 * the noalloc form allocates nothing. */
static int batch_run(albba_batch *b, const int32_t *roots, const albba_bfs_opts *opts, const int32_t *mapping,
                     int32_t workgroups, albba_batch_result *results, float *total_ms, int noalloc)
{
    const char *what = noalloc ? "batch_run_noalloc" : "batch_run";
    if (!b || !roots) return albba_fail(ALBBA_EINVAL, "%s: bad argument", what);
    albba_bfs_opts o;
    if (opts) o = *opts;
    else albba_bfs_opts_default(&o);
    if (o.switch_policy != ALBBA_SWITCH_HEURISTIC && o.switch_policy != ALBBA_SWITCH_LEARNED)
        return albba_fail(ALBBA_EINVAL, "%s: policy must be heuristic or learned", what);
    if (o.switch_policy == ALBBA_SWITCH_LEARNED && !o.selector)
        return albba_fail(ALBBA_EINVAL, "%s: learned policy needs a selector", what);
    const int32_t N = b->count;
    for (int32_t i = 0; i < N; ++i) {
        if (roots[i] < 0 || roots[i] >= b->graphs[i]->n)
            return albba_fail(ALBBA_EINVAL, "%s: root %d of instance %d out of range", what, roots[i], i);
        if (mapping && mapping[i] != ALBBA_MAP_LANE_ROW && mapping[i] != ALBBA_MAP_WAVE_ROW)
            return albba_fail(ALBBA_EINVAL, "%s: mapping %d of instance %d", what, mapping[i], i);
    }
    albd_ctx *d = b->ctx->d;
    const int32_t W = workgroups > 0 && workgroups < b->wcap ? workgroups : b->wcap;
    int32_t *inst_team = b->team_of;
    const int32_t T = split_teams(b, W, b->team_first, b->team_size, b->team_ptr, b->team_items, b->wg_team,
                                  inst_team);
    if (T <= 0) return albba_fail(ALBBA_EHIP, "%s: no team", what);
    int32_t used = 0;
    for (int32_t t = 0; t < T; ++t) used += b->team_size[t];
    for (int32_t i = 0; i < N; ++i) {
        b->items[i].root = roots[i];
        b->items[i].mapping = mapping ? mapping[i] : ALBD_MAP_LANE_ROW;
        b->team_size_of[i] = b->team_size[inst_team[i]];
    }
    const size_t words = tab_words(N, W, T), bytes = b->tab_off + sizeof(int32_t) * words;
    /* device copies of the tables */
    void *dev = NULL;
    int rc = 0;
    if (noalloc) {
        size_t cap = 0;
        if ((rc = albd_batch_scratch(d, &dev, &cap))) return albba_fail(rc, "%s: %s", what, albd_error());
        if (N > ALBBA_BATCH_NOALLOC_MAX || bytes > cap)
            return albba_fail(ALBBA_ERANGE, "%s: %d instances need %zu bytes of tables; the scratch holds "
                              "%zu bytes, at most %d instances", what, N, bytes, cap, ALBBA_BATCH_NOALLOC_MAX);
    } else if ((rc = albd_malloc(d, bytes, &dev))) {
        return albba_fail(rc, "%s: %s", what, albd_error());
    }
    int32_t *h_tab = (int32_t *)((char *)b->stage + b->tab_off), *p = h_tab;
    memcpy(p, b->wg_team, sizeof(int32_t) * (size_t)used);
    p += W;
    memcpy(p, b->team_first, sizeof(int32_t) * (size_t)T);
    p += T;
    memcpy(p, b->team_size, sizeof(int32_t) * (size_t)T);
    p += T;
    memcpy(p, b->team_ptr, sizeof(int32_t) * ((size_t)T + 1));
    p += T + 1;
    memcpy(p, b->team_items, sizeof(int32_t) * (size_t)N);
    albd_batch_args a;
    memset(&a, 0, sizeof a);
    const int32_t *dt = (const int32_t *)((char *)dev + b->tab_off);
    a.items = (const albd_batch_item *)dev;
    a.wg_team = dt;
    a.team_first = dt + W;
    a.team_size = dt + W + T;
    a.team_ptr = dt + W + 2 * T;
    a.team_items = dt + W + 3 * T + 1;
    a.params = o.switch_policy == ALBBA_SWITCH_LEARNED ? albba_selector_dparams(o.selector) : NULL;
    a.spin_limit = 1u << 24;
    float ms = 0;
    /* descriptors and tables in one copy from the pinned staging buffer */
    (void)((rc = albd_h2d(d, dev, b->stage, bytes)) ||
           (rc = albd_fill_u64(d, b->ctl, 0, (size_t)N * CTL_WORDS)) || (rc = albd_event_record(d, 4)) ||
           (rc = albd_batch_launch(d, &a, used)) || (rc = albd_event_record(d, 5)) ||
           (rc = albd_event_elapsed(d, 4, 5, &ms)) ||
           (rc = albd_d2h(d, b->hctl, b->ctl, sizeof(unsigned long long) * (size_t)N * CTL_WORDS)));
    char buf[256] = "";
    if (rc) snprintf(buf, sizeof buf, "%s", albd_error());
    if (!noalloc) albd_free(d, dev);
    if (rc) return albba_fail(rc, "%s: %s", what, buf);
    const double tpm = albd_clock_ticks_per_ms(d);
    long long t_first = 0;
    int any_error = 0;
    for (int32_t i = 0; i < N; ++i) {
        const long long *c = (const long long *)(b->hctl + (size_t)i * CTL_WORDS);
        if (i == 0 || c[ALBD_RES_T0] < t_first) t_first = c[ALBD_RES_T0];
    }
    for (int32_t i = 0; i < N; ++i) {
        const long long *c = (const long long *)(b->hctl + (size_t)i * CTL_WORDS);
        const int err = c[ALBD_RES_ERROR] != 0 || (b->hctl[(size_t)i * CTL_WORDS + 2] & 0xffffffffull) != 0 ||
                        c[ALBD_RES_T1] == 0;
        any_error |= err;
        if (results) {
            albba_batch_result *r = &results[i];
            memset(r, 0, sizeof *r);
            r->error = err;
            r->num_levels = (int32_t)c[ALBD_RES_LEVELS];
            r->reached = c[ALBD_RES_REACHED];
            r->edges_in_component = c[ALBD_RES_EDGES];
            r->ms = err ? 0.0 : (double)(c[ALBD_RES_T1] - c[ALBD_RES_T0]) / tpm;
            r->done_ms = err ? 0.0 : (double)(c[ALBD_RES_T1] - t_first) / tpm;
            r->push_levels = (int32_t)c[ALBD_RES_PUSHES];
            r->pull_levels = (int32_t)c[ALBD_RES_PULLS];
            r->team_size = b->team_size_of[i];
            r->team = inst_team[i];
        }
    }
    if (total_ms) *total_ms = ms;
    if (any_error) {
        /* a team gave up waiting: its keys may be left set */
        for (int32_t i = 0; i < N; ++i)
            albd_fill_u64(d, b->items[i].key, ~0ull, (size_t)b->items[i].g.npos);
        albd_sync(d);
        return albba_fail(ALBBA_EHIP, "%s: a team barrier timed out", what);
    }
    return ALBBA_OK;
}

int albba_batch_run(albba_batch *b, const int32_t *roots, const albba_bfs_opts *opts,
                    const int32_t *mapping, int32_t workgroups, albba_batch_result *results,
                    float *total_ms)
{
    return batch_run(b, roots, opts, mapping, workgroups, results, total_ms, 0);
}

int albba_batch_run_noalloc(albba_batch *b, const int32_t *roots, const albba_bfs_opts *opts,
                            const int32_t *mapping, int32_t workgroups, albba_batch_result *results,
                            float *total_ms)
{
    return batch_run(b, roots, opts, mapping, workgroups, results, total_ms, 1);
}

int albba_batch_results(albba_batch *b, int32_t i, int32_t *level, int32_t *parent)
{
    if (!b || i < 0 || i >= b->count) return albba_fail(ALBBA_EINVAL, "batch_results: bad instance");
    const size_t n = (size_t)b->items[i].g.n;
    int rc = 0;
    if (level) rc = albd_d2h(b->ctx->d, level, b->items[i].level, 4 * n);
    if (!rc && parent) rc = albd_d2h(b->ctx->d, parent, b->items[i].parent, 4 * n);
    return rc ? albba_fail(rc, "batch_results: %s", albd_error()) : ALBBA_OK;
}
