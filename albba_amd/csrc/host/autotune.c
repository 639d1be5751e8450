/*
 * Autotuner: the layout of one graph by measurement on the device.
 *
 * Search space: slice height C in {32, 64, 128} (one 64-bit frontier word is
 * half a slice, one slice or two slices: the pull kernel's wave covers 64
 * positions in every case), sorting window sigma in {C, 256, 4096, n}
 * (sigma = C keeps the vertex order of every slice, sigma >= n sorts all rows
 * by length), and, for the batched BFS, a lane or a whole wave per pulled row.
 * Each layout is timed end to end on the caller's roots and policy, so the
 * choice accounts for the slice padding, the bypass rates and the locality
 * of the level/parent writes together.
 */
#include "internal.h"
#include "device_int.h"

static int time_layout(albba_ctx *ctx, const albba_csr *g, int C, int sigma, const int32_t *roots,
                       int32_t nroots, const albba_bfs_opts *opts, double *ms_out,
                       albba_graph **keep)
{
    albba_sellcs *s = NULL;
    albba_graph *G = NULL;
    int rc = albba_sellcs_build(g, C, sigma, 1, &s);
    if (!rc) rc = albba_graph_upload(ctx, s, &G);
    albba_sellcs_free(s);
    if (rc) return rc;
    double total = 0;
    albba_bfs_stats st;
    memset(&st, 0, sizeof st);
    rc = albba_bfs(ctx, G, roots[0], opts, &st); /* warm-up */
    for (int32_t i = 0; !rc && i < nroots; ++i) {
        memset(&st, 0, sizeof st);
        rc = albba_bfs(ctx, G, roots[i], opts, &st);
        total += st.bfs_ms;
    }
    if (rc || !keep) albba_graph_free(G);
    else *keep = G;
    if (!rc) *ms_out = total / (double)nroots;
    return rc;
}

int albba_autotune(albba_ctx *ctx, const albba_csr *g, const int32_t *roots, int32_t nroots,
                   const albba_bfs_opts *opts, const albba_tune_config *candidates, int32_t ncand,
                   albba_tune_config *results, albba_tune_config *best)
{
    if (!ctx || !g || !roots || nroots <= 0 || !best)
        return albba_fail(ALBBA_EINVAL, "autotune: bad argument");
    for (int32_t i = 0; i < nroots; ++i)
        if (roots[i] < 0 || roots[i] >= g->n)
            return albba_fail(ALBBA_EINVAL, "autotune: root %d out of range", roots[i]);
    albba_tune_config def[12];
    if (!candidates) {
        static const int Cs[3] = {32, 64, 128};
        ncand = 0;
        for (int a = 0; a < 3; ++a) {
            const int sig[4] = {Cs[a], 256, 4096, g->n > Cs[a] ? g->n : Cs[a]};
            for (int k = 0; k < 4; ++k) {
                memset(&def[ncand], 0, sizeof def[ncand]);
                def[ncand].C = Cs[a];
                def[ncand].sigma = sig[k];
                ++ncand;
            }
        }
        candidates = def;
    }
    if (ncand <= 0) return albba_fail(ALBBA_EINVAL, "autotune: no candidate");
    int ibest = -1;
    albba_tune_config local;
    memset(&local, 0, sizeof local);
    for (int32_t c = 0; c < ncand; ++c) {
        albba_tune_config r = candidates[c];
        r.ms = 0;
        r.batch_ms = 0;
        r.mapping = ALBBA_MAP_LANE_ROW;
        r.status = time_layout(ctx, g, r.C, r.sigma, roots, nroots, opts, &r.ms, NULL);
        if (results) results[c] = r;
        if (!r.status && (ibest < 0 || r.ms < local.ms)) {
            ibest = c;
            local = r;
        }
    }
    if (ibest < 0) return albba_fail(ALBBA_EINVAL, "autotune: no candidate layout could run");
    /* pull mapping of the batched BFS on the chosen layout */
    albba_graph *G = NULL;
    double ms = 0;
    int rc = time_layout(ctx, g, local.C, local.sigma, roots, 1, opts, &ms, &G);
    if (rc) return rc;
    albba_graph **gs = (albba_graph **)albba_malloc(sizeof(albba_graph *) * (size_t)nroots);
    int32_t *map = (int32_t *)albba_malloc(sizeof(int32_t) * (size_t)nroots);
    albba_batch *b = NULL;
    if (!gs || !map) rc = albba_fail(ALBBA_ENOMEM, "autotune");
    for (int32_t i = 0; !rc && i < nroots; ++i) gs[i] = G;
    if (!rc) rc = albba_batch_create(ctx, gs, nroots, &b);
    albba_bfs_opts bo;
    albba_bfs_opts_default(&bo);
    if (opts && opts->switch_policy == ALBBA_SWITCH_LEARNED && opts->selector) {
        bo.switch_policy = ALBBA_SWITCH_LEARNED;
        bo.selector = opts->selector;
    }
    double best_bms = -1;
    for (int m = ALBBA_MAP_LANE_ROW; !rc && m <= ALBBA_MAP_WAVE_ROW; ++m) {
        for (int32_t i = 0; i < nroots; ++i) map[i] = m;
        float t = 0, tw = 0;
        rc = albba_batch_run(b, roots, &bo, map, 0, NULL, &tw); /* warm-up */
        if (!rc) rc = albba_batch_run(b, roots, &bo, map, 0, NULL, &t);
        if (!rc && (best_bms < 0 || t < best_bms)) {
            best_bms = t;
            local.mapping = m;
            local.batch_ms = t;
        }
    }
    albba_batch_free(b);
    albba_graph_free(G);
    free(gs);
    free(map);
    if (rc) return rc;
    if (results) {
        results[ibest].mapping = local.mapping;
        results[ibest].batch_ms = local.batch_ms;
    }
    *best = local;
    return ALBBA_OK;
}
