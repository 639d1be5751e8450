/*
 * Batched BFS host side: per-instance device scratch, the split of the
 * resident workgroups into teams, one persistent launch, per-instance
 * results from the device clock.
 *
 * Team split, first run (split_teams): an instance's work is its stored
 * entries plus its vertices; instances worth at least two workgroups of the
 * fair share get a team of their own, the others share one-workgroup teams,
 * largest first onto the least loaded team.
 *
 * Later runs (split_teams_hist) use the level count L and run time t of
 * every instance in the first run (the share split).  A level costs a team
 * of s workgroups lat(s) = LAT_A + LAT_B / s microseconds of latency however
 * little it holds (a 4-neighbour grid of 4.1 M vertices run alone: 47, 28,
 * 18, 15, 14 and 14 us per level with teams of 1, 2, 4, 8, 16 and 32
 * workgroups); the rest of t, measured at the old team size s0, is work that
 * divides over the team: t(s) = L lat(s) + (t - L lat(s0)) k s0 / s, where
 * k = (the graph's stored entries) / (the entries of the root's component),
 * at most 64, covers a later root in a larger component.  A bisection finds
 * the shortest makespan M for which every instance fits: an instance with
 * t(1) <= M / 4 shares a one-workgroup team (first-fit decreasing into
 * teams of total <= M / 2, a margin for the estimates), any other gets its
 * own team of the smallest s <= 8 with t(s) <= M (larger teams gain little
 * per level and poll their barrier harder), never smaller than the own team
 * it had in the share split, and the teams must fit the workgroups.  Workgroups left over go one at a time to the own team whose
 * instance would still take longest.  Other roots on the same graphs take
 * other level counts (±25 % on the largest grids), so the split fits them
 * less well than it fits the roots it was measured on.
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
    /* the previous run: per-instance levels, run time (us) and team size */
    int has_hist;
    int32_t *hist_levels, *hist_team, *own;
    double *hist_us, *hist_scale, *est;
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
    free(b->hist_levels); free(b->hist_team); free(b->own); free(b->hist_us); free(b->hist_scale); free(b->est);
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
    b->hist_levels = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->hist_team = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->own = (int32_t *)albba_calloc(N, sizeof(int32_t));
    b->hist_us = (double *)albba_calloc(N, sizeof(double));
    b->hist_scale = (double *)albba_calloc(N, sizeof(double));
    b->est = (double *)albba_calloc(N, sizeof(double));
    b->tab_off = tab_offset(count);
    b->stage_bytes = b->tab_off + sizeof(int32_t) * tab_words(count, wcap, (int32_t)T);
    if (!b->graphs || !b->allocs || !b->team_of || !b->team_size_of || !b->order || !b->share ||
        !b->cnt || !b->load || !b->team_first || !b->team_size || !b->team_ptr || !b->team_items ||
        !b->wg_team || !b->hist_levels || !b->hist_team || !b->own || !b->hist_us || !b->hist_scale || !b->est) {
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

/* per-level latency of a team of s workgroups, microseconds */
#define LAT_A 12.0
#define LAT_B 35.0
#define TEAM_MAX 8
/* Other roots take other level counts and the estimates come from one run
 * under its own contention: only instances estimated at most SHARE_FRAC of
 * the makespan share workgroups, in teams filled to BIN_FRAC of it. */
#define SHARE_FRAC 0.25
#define BIN_FRAC 0.5

static double lat_us(int32_t s) { return LAT_A + LAT_B / (double)s; }

/* instance i's estimated run time (us) on a team of s workgroups; the work
 * part is scaled up by hist_scale when the measured root's component held
 * few of the graph's entries (another root may reach the rest) */
static double est_us(const albba_batch *b, int32_t i, int32_t s)
{
    const double L = (double)b->hist_levels[i];
    const int32_t s0 = b->hist_team[i] > 0 ? b->hist_team[i] : 1;
    double work = b->hist_us[i] - L * lat_us(s0);
    if (work < 0) work = 0;
    return L * lat_us(s) + work * b->hist_scale[i] * (double)s0 / (double)s;
}

/* the smallest team size <= TEAM_MAX with instance i done within M, or 0;
 * never below the instance's own team in the share split (a work-heavy
 * instance slows by far more than its team shrinks: an RMAT s20 instance
 * that took 4.6 ms on 12 workgroups took 218 ms on one beside busy grids) */
static int32_t team_for(const albba_batch *b, int32_t i, double M)
{
    const int32_t s0 = b->hist_team[i] >= 2 ? b->hist_team[i] : 1;
    for (int32_t s = s0; s <= (s0 > TEAM_MAX ? s0 : TEAM_MAX); ++s)
        if (est_us(b, i, s) <= M) return s;
    return 0;
}

/* First fit of the instances idx[0..n) (longest first) into one-workgroup
 * teams of total time <= M; load[] gets the teams' totals and team_of (if
 * given) each instance's team.  Returns the number of teams, or -1 when more
 * than max_teams would be needed. */
static int32_t pack_small(const albba_batch *b, const int32_t *idx, int32_t n, double M, double *load,
                          int32_t max_teams, int32_t *team_of)
{
    int32_t nt = 0;
    for (int32_t k = 0; k < n; ++k) {
        const double t = b->est[idx[k]];
        int32_t j = 0;
        while (j < nt && load[j] + t > M) ++j;
        if (j == nt) {
            if (nt == max_teams) return -1;
            load[nt++] = 0;
        }
        load[j] += t;
        if (team_of) team_of[idx[k]] = j;
    }
    return nt;
}

/* Can every instance finish within M?  own[i] gets the own-team size (0: a
 * one-workgroup team); small[] the others, longest first; returns the
 * number of workgroups needed, or W + 1. */
static int32_t fit(albba_batch *b, int32_t W, double M, int32_t *small, int32_t *nsmall)
{
    const int32_t N = b->count;
    int32_t own_total = 0, ns = 0;
    for (int32_t k = 0; k < N; ++k) {
        const int32_t i = b->order[k];
        b->own[i] = 0;
        if (b->est[i] <= SHARE_FRAC * M && b->hist_team[i] < 2) {
            small[ns++] = i;
            continue;
        }
        const int32_t s = team_for(b, i, M);
        if (s == 0 || (own_total += s) > W) return W + 1;
        b->own[i] = s;
    }
    *nsmall = ns;
    const int32_t nt = pack_small(b, small, ns, BIN_FRAC * M, b->load, W - own_total, NULL);
    return nt < 0 ? W + 1 : own_total + nt;
}

/* The team split from the previous run's levels and times (see the top of
 * this file); same outputs as split_teams. */
static int32_t split_teams_hist(albba_batch *b, int32_t W, int32_t *team_first, int32_t *team_size,
                                int32_t *team_ptr, int32_t *team_items, int32_t *wg_team, int32_t *inst_team)
{
    const int32_t N = b->count;
    int32_t *order = b->order, *cnt = b->cnt, *small = b->share;
    double lo = 0, hi = 0;
    for (int32_t i = 0; i < N; ++i) {
        b->est[i] = est_us(b, i, 1);
        order[i] = i;
        const double best = est_us(b, i, TEAM_MAX) < b->est[i] ? est_us(b, i, TEAM_MAX) : b->est[i];
        if (best > lo) lo = best;
        hi += b->est[i];
    }
    for (int32_t i = 1; i < N; ++i) { /* insertion sort by one-workgroup time, descending */
        int32_t x = order[i], j = i - 1;
        while (j >= 0 && b->est[order[j]] < b->est[x]) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = x;
    }
    if (hi < lo) hi = lo;
    int32_t nsmall = 0;
    /* hi: every instance on one workgroup one after the other always fits */
    for (int it = 0; it < 40 && hi - lo > 1e-3 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (fit(b, W, mid, small, &nsmall) <= W) hi = mid;
        else lo = mid;
    }
    if (fit(b, W, hi, small, &nsmall) > W) return 0;
    /* own teams, longest instance first, then the one-workgroup teams */
    int32_t T = 0, used = 0;
    for (int32_t k = 0; k < N; ++k) {
        const int32_t i = order[k];
        if (!b->own[i]) continue;
        team_size[T] = b->own[i];
        inst_team[i] = T++;
        used += b->own[i];
    }
    /* left-over workgroups to the own team that would still take longest */
    int32_t spare = W - used - pack_small(b, small, nsmall, BIN_FRAC * hi, b->load, W - used, NULL);
    while (spare > 0 && T > 0) {
        int32_t tb = -1;
        double worst = 0;
        for (int32_t k = 0; k < N; ++k) {
            const int32_t i = order[k];
            if (!b->own[i]) continue;
            const int32_t t = inst_team[i], s = team_size[t];
            const double e = est_us(b, i, s);
            if (s < TEAM_MAX && est_us(b, i, s + 1) < e && e > worst) {
                worst = e;
                tb = t;
            }
        }
        if (tb < 0) break;
        ++team_size[tb];
        --spare;
    }
    const int32_t ns_teams = pack_small(b, small, nsmall, BIN_FRAC * hi, b->load, W - used, b->own);
    for (int32_t k = 0; k < nsmall; ++k) inst_team[small[k]] = T + b->own[small[k]];
    for (int32_t t = 0; t < ns_teams; ++t) team_size[T + t] = 1;
    T += ns_teams;
    memset(cnt, 0, sizeof(int32_t) * ((size_t)T + 1));
    for (int32_t i = 0; i < N; ++i) ++cnt[inst_team[i] + 1];
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

void albba_batch_forget(albba_batch *b)
{
    if (b) b->has_hist = 0;
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
    int32_t T = 0;
    if (b->has_hist)
        T = split_teams_hist(b, W, b->team_first, b->team_size, b->team_ptr, b->team_items, b->wg_team,
                             inst_team);
    const int share_split = T <= 0;
    if (share_split)
        T = split_teams(b, W, b->team_first, b->team_size, b->team_ptr, b->team_items, b->wg_team,
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
    if (!any_error) {
        /* The next runs size their teams from the share-split run only.
         * Times measured under a split made from them shift with the
         * contention that split created: feeding them back made later
         * splits swing between 100 and 360 ms on the 256-instance mix, and
         * raising each instance's level count to the most seen over new
         * roots (to cover farther-reaching roots) gave the grids more
         * workgroups than helped (100, 116, 141 ms over three new root
         * sets against 98, 107, 125 without). */
        if (share_split)
            for (int32_t i = 0; i < N; ++i) {
                const long long *c = (const long long *)(b->hctl + (size_t)i * CTL_WORDS);
                b->hist_levels[i] = (int32_t)c[ALBD_RES_LEVELS];
                b->hist_us[i] = 1e3 * (double)(c[ALBD_RES_T1] - c[ALBD_RES_T0]) / tpm;
                b->hist_team[i] = b->team_size_of[i];
                /* all the graph's entries over the root component's, at most
                 * 64: a root in a small component says little about one in
                 * the large component */
                const double comp = c[ALBD_RES_EDGES] > 0 ? (double)c[ALBD_RES_EDGES] : 1.0;
                const double sc = (double)b->graphs[i]->g.nnz / comp;
                b->hist_scale[i] = sc < 1.0 ? 1.0 : (sc > 64.0 ? 64.0 : sc);
            }
        b->has_hist = 1;
    }
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
