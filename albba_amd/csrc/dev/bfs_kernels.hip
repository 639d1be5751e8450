// BFS kernels for the SELL-C-sigma layout on CDNA4 (gfx950, wave64).
//
// Index space: every bitmap, queue and per-position array is indexed by the
// layout POSITION p (row p of the SELL matrix is vertex row_order[p]); only
// level[] and parent[] are indexed by vertex id.  One wave owns one 64-bit
// word of positions: w covers positions 64w .. 64w+63, i.e. 64/C slices when
// C <= 64 and half a slice when C == 128, so visited/next words of a pull are
// written without atomics.
//
// Pull (bfs/src/bfs_par_kernels.c:128-198, OR-AND SpMV with visited-slice
// bypass): every unvisited lane walks its row's columns in order and stops at
// the first neighbour in the frontier.  Rows list neighbours by ascending
// vertex id, so that first hit is the min-id parent.  The wave leaves a word
// as soon as every lane is resolved or out of columns.
//
// Push (the SpMSpV slot of bfs/src/bfs_par_kernels.c:566-656, done here as a
// true frontier-driven expansion): edges of the frontier queue are split into
// 64-edge chunks via a degree prefix sum; each edge into a vertex unvisited
// at the start of the level does atomicMin(pmin[v], id(u)) and sets v in the
// next bitmap; a word pass then writes level/parent and builds the next queue.
#include "dev_common.h"
#include "../../../include/albba.h"

#include <climits>

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

struct PullArgs {
    albd_graph g;
    albd_state st;
    int32_t level;
    const int32_t *items;               // LIST variants: word list
    const unsigned long long *nitems;   // LIST variants: its length (device)
    int32_t sum_words32;                // CB: LDS summary size in 32-bit words
};

__device__ __forceinline__ void slot_add(albd_counters *ctr, unsigned long long cnt,
                                         unsigned long long edges, unsigned long long words,
                                         unsigned long long scanned,
                                         unsigned long long words_scanned = 0,
                                         unsigned long long bypassed = 0)
{
    albd_slot *s = &ctr->slot[blockIdx.x % ALBD_NSLOT];
    if (cnt) atomicAdd(&s->next_count, cnt);
    if (edges) atomicAdd(&s->next_edges, edges);
    if (words) atomicAdd(&s->next_words, words);
    if (scanned) atomicAdd(&s->scanned, scanned);
    if (words_scanned) atomicAdd(&s->words, words_scanned);
    if (bypassed) atomicAdd(&s->bypassed, bypassed);
}

// ---------------------------------------------------------------------------
// pull

template <int C, int VAR, int U>
__global__ __launch_bounds__(kBlock) void k_pull(PullArgs a)
{
    extern __shared__ uint32_t lds_sum[];
    const albd_graph &g = a.g;
    const int lane = lane_id();
    const int wib = threadIdx.x >> 6;

    if constexpr (VAR == ALBD_PULL_LIST_CB) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.st.summary);
        for (int i = threadIdx.x; i < a.sum_words32; i += kBlock) lds_sum[i] = src[i];
        __syncthreads();
    }
    const int64_t nitems = (VAR == ALBD_PULL_ALL) ? (int64_t)g.nwords : (int64_t)(*a.nitems);
    // XCD-aware split: blocks are dispatched round-robin over the 8 XCDs, so
    // blocks with the same blockIdx % 8 share an L2; give each XCD one
    // contiguous eighth of the items so writes into a sigma window stay in
    // one L2.
    const int xcd = blockIdx.x & 7;
    const int nbx = ((int)gridDim.x - xcd + 7) >> 3;
    const int bx = blockIdx.x >> 3;
    const int64_t lo = nitems * xcd / 8, hi = nitems * (xcd + 1) / 8;
    const int64_t stride = (int64_t)nbx * kWavesPerBlock;
    const int padm = g.pad - 1;
    constexpr int SUB = C < 64 ? C : 64;
    const int grp = (lane / SUB) * SUB;
    const int sl = lane % SUB;

    unsigned long long cnt = 0, nzw = 0, scanned = 0, edges = 0, wscan = 0, wbyp = 0;

    for (int64_t it = lo + (int64_t)bx * kWavesPerBlock + wib; it < hi; it += stride) {
        const int32_t w = (VAR == ALBD_PULL_ALL) ? (int32_t)it : a.items[it];
        const uint64_t vis = a.st.visited[w];
        uint64_t unres = g.haslen[w] & ~vis;
        if (unres == 0) {
            ++wbyp; // slice bypass: every lane with a row is visited
            continue;
        }
        ++wscan;
        const uint64_t start = unres;
        const int64_t p = (int64_t)w * 64 + lane;
        int s, l;
        if constexpr (C <= 64) {
            s = (int)(p / C);
            l = (int)(p % C);
        } else {
            s = w >> 1;
            l = ((w & 1) << 6) + lane;
        }
        const int width = g.slice_width[s];
        int64_t colptr = g.slice_ptr[s];
        const int64_t vp = g.vop_ptr[s];
        int parent_pos = -1;
        bool done = false;
        for (int j0 = 0; !done; j0 += SUB) {
            const int jj = j0 + sl;
            const int actv = (jj < width) ? (int)g.vop_len[vp + jj] + 1 : 0;
#pragma unroll 1
            for (int k = 0; k < SUB; k += U) {
                int act[U], cidx[U];
                bool av[U], hit[U];
                int64_t cp = colptr;
                const bool mine = (unres >> lane) & 1ull;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    act[u] = __shfl(actv, grp + k + u, 64);
                    av[u] = mine && (l < act[u]);
                    cidx[u] = av[u] ? __builtin_nontemporal_load(&g.colidx[cp + l]) : 0;
                    cp += (act[u] + padm) & ~padm;
                }
                const uint64_t m0 = __ballot(av[0]);
                if (m0 == 0) {
                    done = true;
                    break;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    bool h = false;
                    if (av[u]) {
                        const int c = cidx[u];
                        const int wd = c >> 6;
                        bool maybe = true;
                        if constexpr (VAR == ALBD_PULL_LIST_CB)
                            maybe = (lds_sum[wd >> 5] >> (wd & 31)) & 1u;
                        if (maybe) h = (a.st.front[wd] >> (c & 63)) & 1ull;
                    }
                    hit[u] = h;
                }
                scanned += __popcll(m0);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool h = hit[u] && ((unres >> lane) & 1ull);
                    if (h) parent_pos = cidx[u];
                    unres &= ~__ballot(h);
                    if (u > 0) scanned += __popcll(__ballot(av[u]));
                }
                colptr = cp;
                if (unres == 0) {
                    done = true;
                    break;
                }
            }
        }
        const uint64_t found = start & ~unres;
        if (found) {
            if (lane == 0) {
                a.st.next[w] = found;
                a.st.visited[w] = vis | found;
            }
            if ((found >> lane) & 1ull) {
                const int o = g.row_order[p];
                a.st.level[o] = a.level + 1;
                a.st.parent[o] = g.row_order[parent_pos];
                edges += (unsigned long long)(g.push_ptr[p + 1] - g.push_ptr[p]);
            }
            cnt += __popcll(found);
            nzw += 1;
        }
    }
    edges = wave_sum(edges);
    if (lane == 0) slot_add(a.st.ctr, cnt, edges, nzw, scanned, wscan, wbyp);
}

// ---------------------------------------------------------------------------
// word passes

__global__ __launch_bounds__(kBlock) void k_haslen(const int32_t *row_len, uint64_t *haslen,
                                                   int32_t nwords)
{
    const int64_t gt = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t w = gt >> 6;
    const bool pred = (w < nwords) && row_len[gt] > 0;
    const uint64_t m = __ballot(pred);
    if (lane_id() == 0 && w < nwords) haslen[w] = m;
}

__global__ __launch_bounds__(kBlock) void k_bfs_reset(albd_state st, int32_t n, int64_t npos,
                                                      int32_t nwords)
{
    const int64_t total = npos > n ? npos : n;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kBlock) {
        if (i < n) {
            st.level[i] = -1;
            st.parent[i] = -1;
        }
        if (i < npos) st.pmin[i] = 0xFFFFFFFFu;
        if (i < nwords) {
            st.visited[i] = 0;
            st.front[i] = 0;
            st.next[i] = 0;
        }
    }
    if (blockIdx.x == 0) {
        unsigned long long *c = reinterpret_cast<unsigned long long *>(st.ctr);
        for (int i = threadIdx.x; i < (int)(sizeof(albd_counters) / 8); i += kBlock) c[i] = 0;
    }
}

__global__ void k_bfs_root(albd_graph g, albd_state st, int32_t rp, int32_t ro)
{
    st.level[ro] = 0;
    st.parent[ro] = ro;
    const uint64_t bit = 1ull << (rp & 63);
    st.visited[rp >> 6] = bit;
    st.front[rp >> 6] = bit;
    if (g.push_ptr[rp + 1] > g.push_ptr[rp]) {
        st.queue[0] = rp;
        st.ctr->queue_len = 1;
    } else {
        st.ctr->queue_len = 0;
    }
}

enum { LB_QUEUE_FROM_NEXT = 1, LB_ZERO_QUEUE = 2, LB_ACTIVE_FROM_NEXT = 4 };

__global__ __launch_bounds__(kBlock) void k_level_begin(albd_state st, int32_t nwords, int flags)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nwords;
         i += (int64_t)gridDim.x * kBlock)
        st.next[i] = 0;
    if (blockIdx.x == 0) {
        unsigned long long *c = reinterpret_cast<unsigned long long *>(st.ctr->slot);
        for (int i = threadIdx.x; i < (int)(sizeof(st.ctr->slot) / 8); i += kBlock) c[i] = 0;
        if (threadIdx.x == 0) {
            if (flags & LB_QUEUE_FROM_NEXT) st.ctr->queue_len = st.ctr->qnext_len;
            if (flags & LB_ZERO_QUEUE) st.ctr->queue_len = 0;
            if (flags & LB_ACTIVE_FROM_NEXT) st.ctr->active_len = st.ctr->active_next;
            st.ctr->qnext_len = 0;
            st.ctr->active_next = 0;
        }
    }
}

// unvisited-word list: from every word (FROM_LIST = false) or from the
// previous list (visited only grows, so an old list is a superset)
template <bool FROM_LIST>
__global__ __launch_bounds__(kBlock) void k_active(albd_graph g, albd_state st)
{
    const int lane = lane_id();
    const int64_t n_in = FROM_LIST ? (int64_t)st.ctr->active_len : (int64_t)g.nwords;
    const int64_t ngroups = (n_in + 63) >> 6;
    const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t grp = gw; grp < ngroups; grp += nw) {
        const int64_t i = grp * 64 + lane;
        int32_t w = -1;
        bool keep = false;
        if (i < n_in) {
            w = FROM_LIST ? st.active[i] : (int32_t)i;
            keep = (g.haslen[w] & ~st.visited[w]) != 0;
        }
        const uint64_t m = __ballot(keep);
        if (m == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&st.ctr->active_next, (unsigned long long)__popcll(m));
        base = __shfl(base, 0, 64);
        if (keep) st.active_next[base + __popcll(m & lanemask_lt(lane))] = w;
    }
}

__global__ __launch_bounds__(kBlock) void k_summary(albd_graph g, albd_state st)
{
    const int lane = lane_id();
    const int64_t ngroups = ((int64_t)g.nwords + 63) >> 6;
    const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t grp = gw; grp < ngroups; grp += nw) {
        const int64_t w = grp * 64 + lane;
        const bool nz = (w < g.nwords) && st.front[w] != 0;
        const uint64_t m = __ballot(nz);
        if (lane == 0) st.summary[grp] = m;
    }
}

// Words -> queue of positions with out-degree > 0.  FINALIZE (after a push):
// the words are st.next; merge them into visited, write level/parent from
// pmin and count the new frontier.  Otherwise the words are st.front.
template <bool FINALIZE>
__global__ __launch_bounds__(kBlock) void k_words_to_queue(albd_graph g, albd_state st,
                                                           int32_t level)
{
    const int lane = lane_id();
    const uint64_t *words = FINALIZE ? st.next : st.front;
    int32_t *q = FINALIZE ? st.queue_next : st.queue;
    unsigned long long *qlen = FINALIZE ? &st.ctr->qnext_len : &st.ctr->queue_len;
    const int64_t ngroups = ((int64_t)g.nwords + 63) >> 6;
    const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    unsigned long long cnt = 0, nzw = 0, edges = 0;
    for (int64_t grp = gw; grp < ngroups; grp += nw) {
        const int64_t w = grp * 64 + lane;
        const uint64_t word = (w < g.nwords) ? words[w] : 0ull;
        if (FINALIZE && word) {
            st.visited[w] |= word; // only this lane touches word w here
            cnt += __popcll(word);
            nzw += 1;
        }
        const uint64_t nzmask = __ballot(word != 0);
        if (nzmask == 0) continue;
        uint64_t keepw = 0; // lane k holds the keep mask of word grp*64+k
        uint64_t mm = nzmask;
        while (mm) {
            const int k = __ffsll((unsigned long long)mm) - 1;
            mm &= mm - 1;
            const uint64_t bits = __shfl(word, k, 64);
            const int64_t pp = (grp * 64 + k) * 64 + lane;
            bool keep = false;
            if ((bits >> lane) & 1ull) {
                const int64_t d = g.push_ptr[pp + 1] - g.push_ptr[pp];
                keep = d > 0;
                if (FINALIZE) {
                    const int o = g.row_order[pp];
                    st.level[o] = level + 1;
                    st.parent[o] = (int32_t)st.pmin[pp];
                    edges += (unsigned long long)d;
                }
            }
            const uint64_t km = __ballot(keep);
            if (lane == k) keepw = km;
        }
        unsigned long long total;
        const unsigned long long excl =
            wave_exclusive_scan<unsigned long long>((unsigned long long)__popcll(keepw), &total);
        if (total == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(qlen, total);
        base = __shfl(base, 0, 64);
        uint64_t km = __ballot(keepw != 0);
        while (km) {
            const int k = __ffsll((unsigned long long)km) - 1;
            km &= km - 1;
            const uint64_t kw = __shfl(keepw, k, 64);
            const unsigned long long off = __shfl(excl, k, 64);
            if ((kw >> lane) & 1ull)
                q[base + off + __popcll(kw & lanemask_lt(lane))] = (int32_t)((grp * 64 + k) * 64 + lane);
        }
    }
    if (FINALIZE) {
        cnt = wave_sum(cnt);
        nzw = wave_sum(nzw);
        edges = wave_sum(edges);
        if (lane == 0) slot_add(st.ctr, cnt, edges, nzw, 0);
    }
}

// ---------------------------------------------------------------------------
// push: degree prefix over the queue

__device__ __forceinline__ int64_t out_degree(const albd_graph &g, int32_t u)
{
    return g.push_ptr[u + 1] - g.push_ptr[u];
}

// exclusive scan of one value per thread across a 256-thread block
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t *total,
                                                        int64_t *lds /* >= 5 */)
{
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    int64_t wt;
    int64_t ex = wave_exclusive_scan<int64_t>(v, &wt);
    if (lane == 0) lds[wid] = wt;
    __syncthreads();
    int64_t off = 0, tot = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
        if (i < wid) off += lds[i];
        tot += lds[i];
    }
    __syncthreads();
    *total = tot;
    return ex + off;
}

__global__ __launch_bounds__(kBlock) void k_scan_small(albd_graph g, albd_state st)
{
    __shared__ int64_t lds[8];
    const int64_t qlen = (int64_t)st.ctr->queue_len;
    int64_t carry = 0;
    for (int64_t base = 0; base < qlen; base += kBlock) {
        const int64_t i = base + threadIdx.x;
        const int64_t d = (i < qlen) ? out_degree(g, st.queue[i]) : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(d, &tot, lds);
        if (i < qlen) st.prefix[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        st.prefix[qlen] = carry;
        st.ctr->push_edges = carry;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_p1(albd_graph g, albd_state st)
{
    __shared__ int64_t lds[8];
    const int64_t qlen = (int64_t)st.ctr->queue_len;
    const int64_t cs = (qlen + gridDim.x - 1) / gridDim.x;
    const int64_t lo = cs * blockIdx.x, hi = lo + cs < qlen ? lo + cs : qlen;
    int64_t sum = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const int64_t d = out_degree(g, st.queue[i]);
        st.prefix[i] = d;
        sum += d;
    }
    int64_t tot;
    (void)block_exclusive_scan(sum, &tot, lds);
    if (threadIdx.x == 0) st.block_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_p2(albd_state st, int32_t nblocks)
{
    __shared__ int64_t lds[8];
    const int64_t qlen = (int64_t)st.ctr->queue_len;
    int64_t carry = 0;
    for (int base = 0; base < nblocks; base += kBlock) {
        const int i = base + threadIdx.x;
        const int64_t v = i < nblocks ? st.block_sums[i] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(v, &tot, lds);
        if (i < nblocks) st.block_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        st.block_sums[nblocks] = carry;
        st.prefix[qlen] = carry;
        st.ctr->push_edges = carry;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_p3(albd_state st)
{
    __shared__ int64_t lds[8];
    const int64_t qlen = (int64_t)st.ctr->queue_len;
    const int64_t cs = (qlen + gridDim.x - 1) / gridDim.x;
    const int64_t lo = cs * blockIdx.x, hi = lo + cs < qlen ? lo + cs : qlen;
    int64_t carry = st.block_sums[blockIdx.x];
    for (int64_t base = lo; base < hi; base += kBlock) {
        const int64_t i = base + threadIdx.x;
        const int64_t d = i < hi ? st.prefix[i] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(d, &tot, lds);
        if (i < hi) st.prefix[i] = carry + ex;
        carry += tot;
    }
}

// ---------------------------------------------------------------------------
// push: expansion over 64-edge chunks

__global__ __launch_bounds__(kBlock) void k_push_expand(albd_graph g, albd_state st)
{
    const int lane = lane_id();
    const int64_t qlen = (int64_t)st.ctr->queue_len;
    if (qlen == 0) return;
    const int64_t E = st.prefix[qlen];
    const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    const int64_t nchunks = (E + 63) >> 6;
    const int64_t per = (nchunks + nw - 1) / nw;
    const int64_t c_lo = gw * per, c_hi = c_lo + per < nchunks ? c_lo + per : nchunks;
    if (c_lo >= c_hi) return;
    // first vertex: largest v with prefix[v] <= e0 (uniform search, broadcast loads)
    int64_t e0 = c_lo << 6;
    int64_t lo = 0, hi = qlen;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (st.prefix[mid] <= e0) lo = mid;
        else hi = mid;
    }
    int64_t v0 = lo;
    for (int64_t ch = c_lo; ch < c_hi; ++ch) {
        e0 = ch << 6;
        const int64_t e = e0 + lane;
        // v0 holds edge e0 and queue entries have degree >= 1, so the chunk's
        // edges belong to vertices v0 .. v0+63
        const int64_t Pk = (v0 + 1 + lane <= qlen) ? st.prefix[v0 + 1 + lane] : LLONG_MAX;
        const int64_t P0 = st.prefix[v0];
        int t = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int64_t probe = __shfl(Pk, t + step - 1, 64);
            if (probe <= e) t += step;
        }
        const int64_t prevP = __shfl(Pk, t > 0 ? t - 1 : 0, 64);
        const int64_t ui = v0 + t;
        if (e < E) {
            const int64_t start = t == 0 ? P0 : prevP;
            const int32_t u = st.queue[ui];
            const uint32_t ou = (uint32_t)g.row_order[u];
            const int32_t v = g.push_col[g.push_ptr[u] + (e - start)];
            if (!((st.visited[v >> 6] >> (v & 63)) & 1ull)) {
                atomicMin(&st.pmin[v], ou);
                atomicOr(reinterpret_cast<unsigned long long *>(&st.next[v >> 6]),
                         1ull << (v & 63));
            }
        }
        // the next chunk starts at edge e0 + 64; its vertex is v0 plus the
        // number of prefix entries v0+1 .. v0+64 at or below that edge (the
        // lanes hold exactly those 64 entries, so this count may reach 64)
        v0 += __popcll(__ballot(Pk <= e0 + 64));
    }
}

int grid_persistent(albd_ctx *c, int blocks_per_cu)
{
    int g = albd_num_cus(c) * blocks_per_cu;
    return g < 8 ? 8 : g;
}

int grid_for_groups(albd_ctx *c, int64_t groups)
{
    int64_t blocks = (groups + kWavesPerBlock - 1) / kWavesPerBlock;
    int64_t cap = (int64_t)albd_num_cus(c) * 8;
    if (blocks > cap) blocks = cap;
    return blocks < 1 ? 1 : (int)blocks;
}

template <int C>
int launch_pull_c(albd_ctx *c, const PullArgs &a, int variant, size_t lds)
{
    hipStream_t s = albd_stream(c);
    constexpr int U = 2;
    if (variant == ALBD_PULL_ALL) {
        hipLaunchKernelGGL((k_pull<C, ALBD_PULL_ALL, U>), dim3(grid_persistent(c, 8)), dim3(kBlock), 0, s, a);
    } else if (variant == ALBD_PULL_LIST) {
        hipLaunchKernelGGL((k_pull<C, ALBD_PULL_LIST, U>), dim3(grid_persistent(c, 8)), dim3(kBlock), 0, s, a);
    } else {
        int per_cu = (int)(160 * 1024 / (lds + 1024));
        if (per_cu > 8) per_cu = 8;
        if (per_cu < 1) per_cu = 1;
        hipLaunchKernelGGL((k_pull<C, ALBD_PULL_LIST_CB, U>), dim3(grid_persistent(c, per_cu)), dim3(kBlock), lds, s, a);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" {

int albd_build_haslen(albd_ctx *c, const albd_graph *g, uint64_t *haslen)
{
    const int64_t threads = (int64_t)g->nwords * 64;
    hipLaunchKernelGGL(k_haslen, dim3((unsigned)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       albd_stream(c), g->row_len, haslen, g->nwords);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_bfs_init(albd_ctx *c, const albd_graph *g, const albd_state *st, int32_t rp, int32_t ro)
{
    hipStream_t s = albd_stream(c);
    int64_t total = g->npos > g->n ? g->npos : g->n;
    int64_t blocks = (total + kBlock - 1) / kBlock;
    int64_t cap = (int64_t)albd_num_cus(c) * 16;
    hipLaunchKernelGGL(k_bfs_reset, dim3((unsigned)(blocks > cap ? cap : blocks)), dim3(kBlock), 0, s,
                       *st, g->n, g->npos, g->nwords);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_bfs_root, dim3(1), dim3(1), 0, s, *g, *st, rp, ro);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_level_begin_flags(albd_ctx *c, const albd_graph *g, const albd_state *st, int flags)
{
    int64_t blocks = ((int64_t)g->nwords + kBlock - 1) / kBlock;
    int64_t cap = (int64_t)albd_num_cus(c) * 8;
    hipLaunchKernelGGL(k_level_begin, dim3((unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks))),
                       dim3(kBlock), 0, albd_stream(c), *st, g->nwords, flags);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_level_begin(albd_ctx *c, const albd_graph *g, const albd_state *st)
{
    return albd_level_begin_flags(c, g, st, 0);
}

int albd_build_active(albd_ctx *c, const albd_graph *g, albd_state *st, int from_list,
                      int64_t prev_len)
{
    hipStream_t s = albd_stream(c);
    const int64_t n_in = from_list ? prev_len : g->nwords;
    const int grid = grid_for_groups(c, (n_in + 63) / 64);
    if (from_list)
        hipLaunchKernelGGL(k_active<true>, dim3(grid), dim3(kBlock), 0, s, *g, *st);
    else
        hipLaunchKernelGGL(k_active<false>, dim3(grid), dim3(kBlock), 0, s, *g, *st);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_cb_supported(const albd_graph *g)
{
    // summary bitmap (one bit per frontier word) must fit the LDS budget
    const int64_t bytes = (((int64_t)g->nwords + 63) / 64) * 8;
    return bytes <= 64 * 1024;
}

int albd_build_summary(albd_ctx *c, const albd_graph *g, const albd_state *st)
{
    const int grid = grid_for_groups(c, ((int64_t)g->nwords + 63) / 64);
    hipLaunchKernelGGL(k_summary, dim3(grid), dim3(kBlock), 0, albd_stream(c), *g, *st);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_pull(albd_ctx *c, const albd_graph *g, const albd_state *st, int variant, int32_t level,
              int64_t items_bound)
{
    (void)items_bound;
    PullArgs a;
    a.g = *g;
    a.st = *st;
    a.level = level;
    a.items = st->active_next;
    a.nitems = &st->ctr->active_next;
    const int64_t sum_words64 = ((int64_t)g->nwords + 63) / 64;
    a.sum_words32 = (int32_t)(sum_words64 * 2);
    size_t lds = 0;
    if (variant == ALBD_PULL_LIST_CB) {
        if (!albd_cb_supported(g)) return albd_set_error_msg(ALBBA_EINVAL, "column bypass: graph too large for the LDS summary");
        lds = (size_t)a.sum_words32 * 4;
    }
    switch (g->C) {
    case 16: return launch_pull_c<16>(c, a, variant, lds);
    case 32: return launch_pull_c<32>(c, a, variant, lds);
    case 64: return launch_pull_c<64>(c, a, variant, lds);
    case 128: return launch_pull_c<128>(c, a, variant, lds);
    default: return albd_set_error_msg(ALBBA_EINVAL, "pull: unsupported slice height");
    }
}

int albd_front_to_queue(albd_ctx *c, const albd_graph *g, const albd_state *st)
{
    const int grid = grid_for_groups(c, ((int64_t)g->nwords + 63) / 64);
    hipLaunchKernelGGL(k_words_to_queue<false>, dim3(grid), dim3(kBlock), 0, albd_stream(c), *g, *st, 0);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_push_scan(albd_ctx *c, const albd_graph *g, const albd_state *st, int64_t qlen)
{
    hipStream_t s = albd_stream(c);
    if (qlen <= 8192) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kBlock), 0, s, *g, *st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(k_scan_p1, dim3(ALBD_SCAN_BLOCKS), dim3(kBlock), 0, s, *g, *st);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_scan_p2, dim3(1), dim3(kBlock), 0, s, *st, (int32_t)ALBD_SCAN_BLOCKS);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_scan_p3, dim3(ALBD_SCAN_BLOCKS), dim3(kBlock), 0, s, *st);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_push_expand(albd_ctx *c, const albd_graph *g, const albd_state *st)
{
    hipLaunchKernelGGL(k_push_expand, dim3(grid_persistent(c, 8)), dim3(kBlock), 0, albd_stream(c), *g, *st);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_push_finalize(albd_ctx *c, const albd_graph *g, const albd_state *st, int32_t level)
{
    const int grid = grid_for_groups(c, ((int64_t)g->nwords + 63) / 64);
    hipLaunchKernelGGL(k_words_to_queue<true>, dim3(grid), dim3(kBlock), 0, albd_stream(c), *g, *st, level);
    HIPCHK(hipGetLastError());
    return 0;
}

} // extern "C"
