// Batched BFS: many instances (graphs and roots of uneven size) in one
// persistent launch of co-resident workgroups.  The workgroups are split into teams; a team runs the
// complete level loop of its instances one after the other, with one team
// barrier per level and the direction of every level chosen on the device
// (the reference heuristic, bfs/src/bfs_SELLCSv2.c:282-284, or the learned
// selector's MLP), so no instance ever waits for the host.
//
// Exactness, as in the small-frontier kernel: every position has a 64-bit
// key (level << 32 | parent id), all ones at rest.  A push edge u -> v at
// level L does atomicMin(key[v], L << 32 | id(u)); a pull lane that finds
// its row's first frontier neighbour u does the same.  The key keeps the
// smallest (level, parent) pair, so it holds the BFS level and the min-id
// parent no matter in which order the team's waves run, and the old value
// names the one discoverer.  The visited bitmap is only a filter; the
// frontier bitmaps are exact (a pull relies on them for the min-id parent).
//
// Frontier bitmaps B_L (vertices of level L) rotate over three buffers: at
// level L the team reads B_L, sets B_{L+1} and clears B_{L-1}, which becomes
// B_{L+2}.  A push level lists its discoveries (found list), which is the
// next level's queue and lets the bitmap be cleared two levels later without
// a full pass; a pulled bitmap is cleared by a full pass.  Discoverers set
// the visited bits at once: a push needs no visited filter (the key decides),
// a pull reads visited as the exact set of vertices found so far.
#include "dev_common.h"
#include "select_dev.h"
#include "../../../include/albba.h"

namespace {

constexpr int kBB = 512;              // threads per workgroup
constexpr int kBWaves = kBB / 64;
constexpr int kBChunk = kBB;          // queue entries staged per pass
constexpr int kBPer = 4;              // edges per thread per window
constexpr int kBWin = kBPer * kBB;    // edges per window

struct BLds {
    int64_t pre[kBChunk];
    int64_t start[kBChunk];
    uint32_t ou[kBChunk];
    int64_t scan[kBWaves];
    int32_t found[kBWin];
    unsigned int nf;
    unsigned long long bf;
    unsigned long long red[4][kBWaves];
    unsigned int s_cnt[kBWaves];
    unsigned long long s_base;
    int ok, choice;
    float params[ALBBA_MLP_NPARAM];
    float xs[ALBBA_MLP_IN], h1[ALBBA_MLP_H1], h2[ALBBA_MLP_H2], ys[ALBBA_MLP_OUT];
};

__device__ __forceinline__ unsigned int *err_flag(const albd_batch_item &it)
{
    return reinterpret_cast<unsigned int *>(&it.ctl[2]);
}

__device__ __forceinline__ long long *ring(const albd_batch_item &it, int L)
{
    return reinterpret_cast<long long *>(it.ctl) + 4 + 4 * (L & 3);
}

// per-wave counts -> one atomic per workgroup into the level's ring slot
__device__ void add_counts(BLds &S, long long *slot, unsigned long long f, unsigned long long e,
                           unsigned long long w)
{
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    f = wave_sum(f);
    e = wave_sum(e);
    w = wave_sum(w);
    if (lane == 0) {
        S.red[0][wid] = f;
        S.red[1][wid] = e;
        S.red[3][wid] = w;
    }
    __syncthreads();
    if (threadIdx.x < 4 && threadIdx.x != 2) {
        unsigned long long t = 0;
        for (int i = 0; i < kBWaves; ++i) t += S.red[threadIdx.x][i];
        if (t) atomicAdd(reinterpret_cast<unsigned long long *>(&slot[threadIdx.x]), t);
    }
    __syncthreads();
}

// the LDS list of a push window -> the instance's found list at
// found[base + ...] (one reservation per window on the level's counter)
__device__ void flush_found(BLds &S, const albd_batch_item &it, unsigned long long *counter, int64_t base)
{
    __syncthreads();
    const unsigned int nf = S.nf;
    if (threadIdx.x == 0) S.bf = nf ? atomicAdd(counter, (unsigned long long)nf) : 0;
    __syncthreads();
    if (threadIdx.x == 0) S.nf = 0;
    const unsigned long long bf = S.bf;
    for (unsigned int i = threadIdx.x; i < nf; i += kBB) st_coherent(&it.found[base + bf + i], S.found[i]);
    __syncthreads();
}

// One BFS of one instance by a team of T workgroups (this one is `rank`).
// The found list holds, level after level, the vertices each push found and
// the rebuilt frontier of a push that follows a pull: the frontier queue of
// level L is the range [fs_cur, fe_cur).  Returns false when a team barrier
// gave up.
__device__ bool run_instance(BLds &S, const albd_batch_item &it, int rank, int T, bool learned,
                             uint32_t spin_limit)
{
    const albd_graph &g = it.g;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid >> 6;
    const int64_t gtid = (int64_t)rank * kBB + tid, gthreads = (int64_t)T * kBB;
    const int64_t nwords = g.nwords;
    unsigned long long *bar = &it.ctl[0];
    unsigned int *err = err_flag(it);
    const bool leader = rank == 0 && tid == 0;
    if (leader) st_coherent(reinterpret_cast<long long *>(&it.ctl[ALBD_RES_T0]), (long long)wall_clock64());
    for (int64_t i = gtid; i < nwords; i += gthreads) {
        st_coherent(&it.visited[i], (uint64_t)0);
        st_coherent(&it.bm[0][i], (uint64_t)0);
        st_coherent(&it.bm[1][i], (uint64_t)0);
        st_coherent(&it.bm[2][i], (uint64_t)0);
    }
    if (tid == 0) S.nf = 0;
    if (!team_sync(bar, err, T, spin_limit, &S.ok)) return false;
    const int32_t rp = g.pos_of[it.root];
    const int64_t rdeg = g.push_ptr[rp + 1] - g.push_ptr[rp];
    if (leader) {
        // the root: level 0, its own parent, frontier B_0, found entry 0
        const uint64_t bit = 1ull << (rp & 63);
        st_coherent(&it.key[rp], (unsigned long long)(uint32_t)it.root);
        atomicOr(reinterpret_cast<unsigned long long *>(&it.visited[rp >> 6]), bit);
        atomicOr(reinterpret_cast<unsigned long long *>(&it.bm[0][rp >> 6]), bit);
        st_coherent(&it.found[0], rp);
    }
    if (!team_sync(bar, err, T, spin_limit, &S.ok)) return false;

    // team-uniform loop state
    unsigned long long F = 1, mf = (unsigned long long)rdeg, nzw = 1, F_prev = 0;
    unsigned long long reached = 1, redges = (unsigned long long)rdeg;
    int prev_choice = ALBBA_CHOICE_PUSH;
    int64_t fs_cur = 0, fe_cur = 1, fs_prev = 0, fe_prev = 0; // found ranges of B_L, B_{L-1}
    bool have_queue = true;                                   // B_L's range holds its vertices
    bool cur_by_pull = false, prev_by_pull = false;           // how B_L, B_{L-1} were made
    long long pushes = 0, pulls = 0;
    const long long thr = (long long)(int)((double)g.n * 0.04);
    int L = 0;
    for (; F > 0; ++L) {
        uint64_t *bcur = it.bm[L % 3], *bnext = it.bm[(L + 1) % 3], *bprev = it.bm[(L + 2) % 3];
        long long *slot = ring(it, L);
        // direction of this level (the same in every workgroup of the team).
        // A batched pull visits every word of the instance with the team's few
        // workgroups, so the learned selector is consulted only when the
        // frontier is large by the heuristic's count or its out-edges reach
        // 1/16 of the unexplored edges; otherwise the level pushes.
        int choice = (long long)F > thr ? ALBBA_CHOICE_PULL_SB : ALBBA_CHOICE_PUSH;
        const unsigned long long unexplored =
            (unsigned long long)g.nnz > redges ? (unsigned long long)g.nnz - redges : 0ull;
        if (learned && ((long long)F > thr || 16 * mf >= unexplored)) {
            if (tid < 64)
                level_features(S.xs, g.n, g.nnz, nwords, F, mf, nzw, reached, redges, L, prev_choice,
                               F_prev);
            __syncthreads();
            const int c = mlp_argmin(S.params, S.xs, S.h1, S.h2, S.ys, 0xF, &S.choice);
            if (c >= 0) choice = c;
        }
        const bool push = choice == ALBBA_CHOICE_PUSH;
        // the slot two levels ahead is free: its last readers passed a barrier
        if (leader)
            for (int f = 0; f < 4; ++f) st_coherent(ring(it, L + 2) + f, 0ll);
        // B_{L-1} goes back to zero (it becomes B_{L+2})
        if (L >= 1) {
            if (prev_by_pull) {
                for (int64_t i = gtid; i < nwords; i += gthreads) st_coherent(&bprev[i], (uint64_t)0);
            } else {
                for (int64_t i = fs_prev + gtid; i < fe_prev; i += gthreads)
                    st_coherent(&bprev[ld_coherent(&it.found[i]) >> 6], (uint64_t)0);
            }
        }
        const unsigned long long lkey = (unsigned long long)(uint32_t)(L + 1) << 32;
        unsigned long long my_f = 0, my_e = 0, my_w = 0;
        if (push) {
            ++pushes;
            if (!have_queue) {
                // B_L came from a pull: list its vertices from the bitmap,
                // one reservation per block step
                for (int64_t wb = (int64_t)rank * kBB; wb < nwords; wb += gthreads) {
                    const int64_t w = wb + tid;
                    const uint64_t word = w < nwords ? ld_coherent(&bcur[w]) : 0ull;
                    unsigned long long wtot;
                    const unsigned long long ex =
                        wave_exclusive_scan<unsigned long long>((unsigned long long)__popcll(word), &wtot);
                    if (lane == 0) S.s_cnt[wid] = (unsigned int)wtot;
                    __syncthreads();
                    if (tid == 0) {
                        unsigned long long tot = 0;
                        for (int i = 0; i < kBWaves; ++i) tot += S.s_cnt[i];
                        S.s_base = tot ? atomicAdd(reinterpret_cast<unsigned long long *>(&slot[2]), tot) : 0ull;
                    }
                    __syncthreads();
                    unsigned long long o = (unsigned long long)fe_cur + S.s_base + ex;
                    for (int i = 0; i < wid; ++i) o += S.s_cnt[i];
                    __syncthreads();
                    uint64_t kb = word;
                    while (kb) {
                        const int b = __ffsll((unsigned long long)kb) - 1;
                        kb &= kb - 1;
                        st_coherent(&it.found[o++], (int32_t)(w * 64 + b));
                    }
                }
                if (!team_sync(bar, err, T, spin_limit, &S.ok)) return false;
                fs_cur = fe_cur;
                fe_cur += (int64_t)ld_coherent(&slot[2]);
            }
            // expansion: workgroup `rank` takes queue entries [q0, q1)
            const int64_t qlen = fe_cur - fs_cur;
            const int64_t per = (qlen + T - 1) / T;
            const int64_t q0 = fs_cur + (int64_t)rank * per;
            const int64_t q1 = q0 + per < fe_cur ? q0 + per : fe_cur;
            for (int64_t base = q0; base < q1; base += kBChunk) { // uniform in the workgroup
                const int cnt_e = (int)(q1 - base < kBChunk ? q1 - base : kBChunk);
                int64_t a0 = 0, e0 = 0;
                uint32_t o0 = 0;
                if (tid < cnt_e) {
                    const int32_t u = ld_coherent(&it.found[base + tid]);
                    a0 = g.push_ptr[u];
                    e0 = g.push_ptr[u + 1];
                    o0 = (uint32_t)g.row_order[u];
                }
                int64_t Te;
                const int64_t ex = block_scan_excl<kBB>(e0 - a0, &Te, S.scan);
                S.pre[tid] = ex;
                S.start[tid] = a0;
                S.ou[tid] = o0;
                __syncthreads();
                for (int64_t wb = 0; wb < Te; wb += kBWin) { // uniform
                    int64_t e = wb + (int64_t)tid * kBPer;
                    const int64_t e_end = e + kBPer < Te ? e + kBPer : Te;
                    int lo = 0;
                    if (e < e_end) {
                        int hi = cnt_e;
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (S.pre[mid] <= e) lo = mid;
                            else hi = mid;
                        }
                    }
                    int32_t v[kBPer];
                    uint32_t ou[kBPer];
                    bool okv[kBPer];
#pragma unroll
                    for (int i = 0; i < kBPer; ++i) {
                        okv[i] = e < e_end;
                        v[i] = 0;
                        ou[i] = 0;
                        if (okv[i]) {
                            while (lo + 1 < cnt_e && S.pre[lo + 1] <= e) ++lo;
                            v[i] = g.push_col[S.start[lo] + (e - S.pre[lo])];
                            ou[i] = S.ou[lo];
                            ++e;
                        }
                    }
                    // the key decides: no visited filter on the way
                    bool won[kBPer];
#pragma unroll
                    for (int i = 0; i < kBPer; ++i) {
                        won[i] = false;
                        if (okv[i]) won[i] = atomicMin(&it.key[v[i]], lkey | ou[i]) == ~0ull;
                    }
#pragma unroll
                    for (int i = 0; i < kBPer; ++i) {
                        if (won[i]) {
                            const uint64_t bit = 1ull << (v[i] & 63);
                            const unsigned long long old =
                                atomicOr(reinterpret_cast<unsigned long long *>(&bnext[v[i] >> 6]), bit);
                            atomicOr(reinterpret_cast<unsigned long long *>(&it.visited[v[i] >> 6]), bit);
                            my_w += old == 0 ? 1 : 0;
                            my_e += (unsigned long long)(g.push_ptr[v[i] + 1] - g.push_ptr[v[i]]);
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kBPer; ++i) {
                        const unsigned int pf = wave_append(&S.nf, won[i]);
                        if (won[i]) S.found[pf] = v[i];
                    }
                    flush_found(S, it, reinterpret_cast<unsigned long long *>(&slot[0]), fe_cur);
                }
                __syncthreads(); // pre/start/ou are restaged
            }
        } else {
            ++pulls;
            // pull: workgroup `rank` takes words [w0, w1) in runs of 64; wave
            // wid takes every kBWaves-th run.  The wave's lanes test the 64
            // words of a run at once (rows with a length, not visited: every
            // vertex found so far has its visited bit, set by its
            // discoverer), then the wave works through the words that have
            // such rows; their unvisited lanes look for their row's first
            // neighbour in B_L (rows list neighbours by ascending id)
            const int64_t per = (nwords + T - 1) / T;
            const int64_t w0 = (int64_t)rank * per, w1 = w0 + per < nwords ? w0 + per : nwords;
            for (int64_t run = w0 + (int64_t)wid * 64; run < w1; run += (int64_t)kBWaves * 64) {
                const int64_t wl = run + lane;
                uint64_t unres_l = 0;
                if (wl < w1) {
                    const uint64_t hl = g.haslen[wl];
                    if (hl) unres_l = hl & ~ld_coherent(&it.visited[wl]);
                }
                uint64_t todo_w = __ballot(unres_l != 0);
                while (todo_w) {
                    const int kw = __ffsll((unsigned long long)todo_w) - 1;
                    todo_w &= todo_w - 1;
                    const int64_t w = run + kw;
                    const uint64_t unres = __shfl(unres_l, kw, 64);
                    const int64_t p = w * 64 + lane;
                    int32_t par = -1;
                    if (it.mapping == ALBD_MAP_LANE_ROW) {
                        if ((unres >> lane) & 1ull) {
                            const int64_t r1 = g.pull_ptr[p + 1];
                            for (int64_t k = g.pull_ptr[p]; k < r1; ++k) {
                                const int32_t u = g.pull_col[k];
                                if ((ld_coherent(&bcur[u >> 6]) >> (u & 63)) & 1ull) {
                                    par = g.row_order[u];
                                    break;
                                }
                            }
                        }
                    } else {
                        // the wave scans one row at a time, 64 entries per step
                        uint64_t todo = unres;
                        while (todo) {
                            const int r = __ffsll((unsigned long long)todo) - 1;
                            todo &= todo - 1;
                            const int64_t rs = g.pull_ptr[w * 64 + r], re = g.pull_ptr[w * 64 + r + 1];
                            for (int64_t k = rs; k < re; k += 64) {
                                const int64_t kk = k + lane;
                                bool h = false;
                                int32_t u = 0;
                                if (kk < re) {
                                    u = g.pull_col[kk];
                                    h = (ld_coherent(&bcur[u >> 6]) >> (u & 63)) & 1ull;
                                }
                                const uint64_t hm = __ballot(h);
                                if (hm) {
                                    const int32_t first = __shfl(u, __ffsll((unsigned long long)hm) - 1, 64);
                                    if (lane == r) par = g.row_order[first];
                                    break;
                                }
                            }
                        }
                    }
                    bool won = false;
                    if (par >= 0) won = atomicMin(&it.key[p], lkey | (uint32_t)par) == ~0ull;
                    const uint64_t fm = __ballot(won);
                    if (fm) {
                        if (lane == 0) {
                            const unsigned long long old =
                                atomicOr(reinterpret_cast<unsigned long long *>(&bnext[w]), fm);
                            atomicOr(reinterpret_cast<unsigned long long *>(&it.visited[w]), fm);
                            my_w += old == 0 ? 1 : 0;
                        }
                        if (won) {
                            my_f += 1;
                            my_e += (unsigned long long)(g.push_ptr[p + 1] - g.push_ptr[p]);
                        }
                    }
                }
            }
        }
        add_counts(S, slot, my_f, my_e, my_w);
        if (!team_sync(bar, err, T, spin_limit, &S.ok)) return false;
        // B_{L+1} is complete: its size, out-edges, words and (push) list
        mf = (unsigned long long)ld_coherent(&slot[1]);
        nzw = (unsigned long long)ld_coherent(&slot[3]);
        fs_prev = fs_cur;
        fe_prev = fe_cur;
        prev_by_pull = cur_by_pull;
        F_prev = F;
        if (push) {
            // the level's appends, counted in its own ring slot: a
            // workgroup past the barrier may already append for the next
            // level while another still reads this one's count
            F = (unsigned long long)ld_coherent(&slot[0]);
            fs_cur = fe_cur;
            fe_cur += (int64_t)F;
        } else {
            F = (unsigned long long)ld_coherent(&slot[0]);
            fs_cur = fe_cur; // a pulled level lists nothing
        }
        have_queue = push;
        cur_by_pull = !push;
        reached += F;
        redges += mf;
        prev_choice = push ? ALBBA_CHOICE_PUSH : ALBBA_CHOICE_PULL;
    }
    // results in vertex order from the keys; the keys go back to all ones
    for (int64_t v = gtid; v < g.n; v += gthreads) {
        const int32_t p = g.pos_of[v];
        const unsigned long long k = ld_coherent(&it.key[p]);
        const bool hit = k != ~0ull;
        it.level[v] = hit ? (int32_t)(k >> 32) : -1;
        it.parent[v] = hit ? (int32_t)(uint32_t)k : -1;
        if (hit) st_coherent(&it.key[p], ~0ull);
    }
    if (!team_sync(bar, err, T, spin_limit, &S.ok)) return false;
    if (leader) {
        long long *res = reinterpret_cast<long long *>(it.ctl);
        st_coherent(&res[ALBD_RES_LEVELS], (long long)L);
        st_coherent(&res[ALBD_RES_REACHED], (long long)reached);
        st_coherent(&res[ALBD_RES_EDGES], (long long)redges);
        st_coherent(&res[ALBD_RES_PUSHES], pushes);
        st_coherent(&res[ALBD_RES_PULLS], pulls);
        st_coherent(&res[ALBD_RES_T1], (long long)wall_clock64());
    }
    return true;
}

__global__ __launch_bounds__(kBB, 4) void k_bfs_batch(albd_batch_args a)
{
    __shared__ BLds S;
    if (a.params) mlp_to_lds<kBB>(S.params, a.params);
    __syncthreads();
    const int t = a.wg_team[blockIdx.x];
    const int rank = (int)blockIdx.x - a.team_first[t], T = a.team_size[t];
    for (int j = a.team_ptr[t]; j < a.team_ptr[t + 1]; ++j) {
        const albd_batch_item &it = a.items[a.team_items[j]];
        if (!run_instance(S, it, rank, T, a.params != nullptr, a.spin_limit)) {
            if (threadIdx.x == 0)
                st_coherent(reinterpret_cast<long long *>(it.ctl) + ALBD_RES_ERROR, 1ll);
            break;
        }
    }
}

// The device tables of albba_batch_run_noalloc (instance descriptors, team
// tables) live here instead of in memory allocated per run.  Each device
// holds its own copy at its own address; host code must use that address
// (from hipGetSymbolAddress), never the array's name, which in host code is
// only a stand-in.  Every run on a device shares the array: runs that
// overlap on one device race.
static __device__ __align__(256) unsigned char scratch[ALBD_BATCH_SCRATCH_BYTES];
static_assert(ALBBA_BATCH_NOALLOC_MAX * sizeof(albd_batch_item) + 256 +
                      sizeof(int32_t) * (4096 + 3 * ALBBA_BATCH_NOALLOC_MAX + 1 + ALBBA_BATCH_NOALLOC_MAX) <=
                  ALBD_BATCH_SCRATCH_BYTES,
              "the scratch holds the tables of the largest no-alloc batch (up to 4096 workgroups)");

constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices]; // device address of `scratch`, per device

} // namespace

extern "C" {

int albd_batch_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return albd_set_error_msg(ALBBA_ENODEV, "batch scratch: device index beyond the table");
    if (__atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE)) return 0;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return albd_set_error_msg(ALBBA_EHIP, "batch scratch: no device address");
    __atomic_store_n(&g_scratch_ptr[device], p, __ATOMIC_RELEASE);
    return 0;
}

int albd_batch_scratch(albd_ctx *c, void **ptr, size_t *bytes)
{
    const int device = albd_ctx_device(c);
    void *p = device >= 0 && device < kMaxDevices ? __atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE)
                                                  : nullptr;
    if (!p) return albd_set_error_msg(ALBBA_EHIP, "batch scratch: not set up on this device");
    *ptr = p;
    *bytes = ALBD_BATCH_SCRATCH_BYTES;
    return 0;
}

int albd_batch_capacity(albd_ctx *c, int *workgroups)
{
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bfs_batch, kBB, 0));
    *workgroups = per_cu * albd_num_cus(c);
    return 0;
}

int albd_batch_launch(albd_ctx *c, const albd_batch_args *a, int32_t workgroups)
{
    // at most albd_batch_capacity workgroups, all resident at once; an
    // ordinary launch on the context's stream, which starts once the
    // stream's earlier kernels have finished (a team barrier that waited in
    // vain would give up after spin_limit polls, never hang)
    int cap = 0;
    if (int rc = albd_batch_capacity(c, &cap)) return rc;
    if (workgroups < 1 || workgroups > cap)
        return albd_set_error_msg(ALBBA_EINVAL, "batch launch: more workgroups than can be resident");
    hipLaunchKernelGGL(k_bfs_batch, dim3(workgroups), dim3(kBB), 0, albd_stream(c), *a);
    HIPCHK(hipGetLastError());
    return 0;
}

double albd_clock_ticks_per_ms(albd_ctx *c)
{
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, albd_ctx_device(c)) != hipSuccess ||
        khz <= 0)
        return 100000.0; // the 100 MHz constant clock of CDNA
    return (double)khz;
}

} // extern "C"
