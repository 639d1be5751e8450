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
// at the start of the level does atomicMin(pmin[v], id(u)); the edge that
// finds pmin[v] still all ones sets v in the next bitmap; a word pass then
// writes level/parent and builds the next queue.
#include "dev_common.h"
#include "select_dev.h"
#include "../../../include/albba.h"

#include <climits>

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

// A kernel queued before its level's direction is known runs only when the
// select kernel chose it: spec < 0 runs always; otherwise the learned (1) or
// the heuristic (0) choice must be `choice`, and the frontier nonempty and at
// most fmax.  The select kernel wrote those counters before this kernel
// started, and nothing in between changes them, so every thread of every
// block decides alike.
struct Spec {
    int32_t spec, choice;
    unsigned long long fmax;
};

__device__ __forceinline__ bool spec_go(const albd_counters *ctr, const Spec &sp)
{
    if (sp.spec < 0) return true;
    const unsigned long long f = ctr->frontier;
    const int c = sp.spec ? ctr->sel_choice : ctr->heur_choice;
    return __builtin_amdgcn_readfirstlane(f != 0 && f <= sp.fmax && c == sp.choice ? 1 : 0) != 0;
}

constexpr Spec kNoSpec{-1, 0, 0ull};

struct PullArgs {
    albd_graph g;
    albd_state st;
    int32_t level;
    const int32_t *items;               // LIST variants: word list
    const unsigned long long *nitems;   // LIST variants: its length (device)
    int32_t sum_words32;                // CB: LDS summary size in 32-bit words
    int32_t per_block;                  // consecutive items per block
    // queued before the level's direction is known: -1 no; else run only if
    // the select kernel's choice (1: the learned one, 0: the heuristic's) is
    // spec_choice and the frontier is not empty
    int32_t spec;
    int32_t spec_choice;
};

__device__ __forceinline__ void slot_add(albd_slot *slots, unsigned long long cnt,
                                         unsigned long long edges, unsigned long long words,
                                         unsigned long long scanned,
                                         unsigned long long words_scanned = 0,
                                         unsigned long long bypassed = 0,
                                         unsigned long long col_bypassed = 0)
{
    albd_slot *s = &slots[(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) % ALBD_NSLOT];
    if (cnt) atomicAdd(&s->next_count, cnt);
    if (edges) atomicAdd(&s->next_edges, edges);
    if (words) atomicAdd(&s->next_words, words);
    if (scanned) atomicAdd(&s->scanned, scanned);
    if (words_scanned) atomicAdd(&s->words, words_scanned);
    if (col_bypassed) atomicAdd(&s->col_bypassed, col_bypassed);
    if (bypassed) atomicAdd(&s->bypassed, bypassed);
}

// ---------------------------------------------------------------------------
// pull

// The column-bypass pull stages a frontier summary of up to 32 KB in LDS per
// block; its blocks are 1024 threads so that two of them (32 waves) share a
// CU, where 256-thread blocks would be held to 16 waves by the LDS.
constexpr int kCBBlock = 1024;

template <int VAR> constexpr int pull_block() { return VAR == ALBD_PULL_LIST_CB ? kCBBlock : kBlock; }

// The pull of one level (NT threads a block).
template <int C, int VAR, int NT>
__device__ __forceinline__ void pull_body(const PullArgs &a, uint32_t *lds_sum)
{
    const albd_graph &g = a.g;
    const int lane = lane_id();

    if constexpr (VAR == ALBD_PULL_LIST_CB) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.st.summary);
        for (int i = threadIdx.x; i < a.sum_words32; i += NT) lds_sum[i] = src[i];
        __syncthreads();
    }
    const int64_t nitems = (VAR == ALBD_PULL_ALL) ? (int64_t)g.nwords : (int64_t)(*a.nitems);
    // Word cost is heavy-tailed (the first slice of every sigma window holds
    // its hubs), so work is balanced twice over without global atomics: the
    // hardware dispatcher hands out blocks of per_block consecutive words,
    // and the block's waves take words one at a time through an LDS counter.
    // Blocks are dispatched round-robin over the 8 XCDs (blockIdx % 8); the
    // block -> range map is XCD-major, so each XCD sweeps its own contiguous
    // stretch of words and the level/parent writes of a sigma window merge in
    // one L2.  gridDim is a multiple of 8 and was sized from an upper bound of
    // the item count, so the share of a block is taken from the real count.
    __shared__ int next_item;
    if (threadIdx.x == 0) next_item = 0;
    __syncthreads();
    const int64_t vb = (int64_t)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    int64_t per = (nitems + gridDim.x - 1) / gridDim.x;
    if (per < 1) per = 1;
    const int64_t b0 = vb * per;
    const int64_t b1 = b0 + per < nitems ? b0 + per : nitems;
    const int padm = g.pad - 1;
    constexpr int SUB = C < 64 ? C : 64;
    const uint64_t gmask = SUB == 64 ? ~0ull : (((1ull << SUB) - 1) << ((lane / SUB) * SUB));

    const bool sym = g.pull_col == g.push_col;
    unsigned long long cnt = 0, nzw = 0, scanned = 0, edges = 0, wscan = 0, wbyp = 0;
    // column bypass: this lane's column entries whose frontier word the
    // summary shows empty (their gathers are skipped)
    unsigned int cbs = 0;
    // One word ahead: while a word's columns are scanned, the next word is
    // claimed, its list entry read, and then its row lengths, visited word
    // and slice pointer loaded, so that a word starts with its first column
    // loads rather than two dependent round trips.  A wave holds one word
    // ahead only, which bounds what it can keep from the block's other waves.
    auto claim = [&]() -> int64_t {
        int k = 0;
        if (lane == 0) k = atomicAdd(&next_item, 1);
        return b0 + __shfl(k, 0, 64);
    };
    auto word_of = [&](int64_t it) -> int32_t {
        if (it >= b1) return -1;
        return (VAR == ALBD_PULL_ALL) ? (int32_t)it : a.items[it];
    };
    struct Head {
        int len, len_o;
        uint64_t vis;
        int64_t cp;
    };
    // row lengths stream like the columns: non-temporal, so that the
    // frontier and visited bitmaps keep their L2 lines
    auto load_head = [&](int32_t w) -> Head {
        Head h{0, 0, ~0ull, 0};
        if (w < 0) return h;
        const int64_t p = (int64_t)w * 64 + lane;
        if constexpr (C <= 64) {
            h.len = __builtin_nontemporal_load(&g.row_len[p]);
            h.cp = g.slice_ptr[p / C];
        } else {
            const int64_t s = w >> 1;
            h.len = __builtin_nontemporal_load(&g.row_len[p]);
            h.len_o = __builtin_nontemporal_load(&g.row_len[s * 128 + (((w & 1) ^ 1) << 6) + lane]);
            h.cp = g.slice_ptr[s];
        }
        h.vis = a.st.visited[w];
        return h;
    };
    int32_t w = word_of(claim());
    Head h = load_head(w);
    while (w >= 0) {
        const int32_t w1 = word_of(claim());
        Head h1{0, 0, ~0ull, 0};
        bool ahead = false; // h1 loaded (wave-uniform)
        {
            const int64_t p = (int64_t)w * 64 + lane;
            // lane row lengths replace per-column active-lane counts: rows of
            // a slice are sorted by length, so column j holds the lanes whose
            // row is longer than j and its size is a ballot away
            int l;
            if constexpr (C <= 64)
                l = (int)(p % C);
            else
                l = ((w & 1) << 6) + lane;
            const int len = h.len, len_o = h.len_o;
            const uint64_t vis = h.vis;
            int64_t cp = h.cp;
            uint64_t unres = __ballot(len > 0) & ~vis;
            if (unres == 0) {
                ++wbyp; // slice bypass: every lane with a row is visited
            } else {
                ++wscan;
                const uint64_t start = unres;
                int parent_pos = -1;
                int j = 0, step = 1;
                // columns are loaded 1, 2, then 4 at a time: most lanes of a
                // dense level hit in the first column, the survivors tend to
                // run long
                for (;;) {
                    const bool mine = (unres >> lane) & 1ull;
                    const uint64_t live = __ballot(mine && j < len);
                    // few live lanes left: their rows go to the row phase below
                    if (__popcll(live) <= g.sparse_lanes) break;
                    int cidx[4];
                    bool av[4], hit[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        av[u] = false;
                        cidx[u] = 0;
                        if (u < step) {
                            const int col = j + u;
                            int act;
                            if constexpr (C <= 64)
                                act = __popcll(__ballot(col < len) & gmask);
                            else
                                act = __popcll(__ballot(col < len)) + __popcll(__ballot(col < len_o));
                            av[u] = mine && col < len;
                            // streaming: non-temporal, so that the frontier
                            // and visited bitmaps stay in L2
                            if (av[u]) cidx[u] = __builtin_nontemporal_load(&g.colidx[cp + l]);
                            cp += (act + padm) & ~padm;
                        }
                    }
                    // the next word's head travels with this word's first
                    // columns
                    if (!ahead) {
                        h1 = load_head(w1);
                        ahead = true;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        bool hh = false;
                        if (av[u]) {
                            const int c = cidx[u];
                            const int wd = c >> 6;
                            bool maybe = true;
                            if constexpr (VAR == ALBD_PULL_LIST_CB) {
                                maybe = (lds_sum[wd >> 5] >> (wd & 31)) & 1u;
                                cbs += maybe ? 0u : 1u;
                            }
                            if (maybe) hh = (a.st.front[wd] >> (c & 63)) & 1ull;
                        }
                        hit[u] = hh;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (u < step) {
                            const bool hh = hit[u] && ((unres >> lane) & 1ull);
                            if (hh) parent_pos = cidx[u];
                            unres &= ~__ballot(hh);
                            scanned += __popcll(__ballot(av[u]));
                        }
                    }
                    j += step;
                    if (unres == 0) break;
                    step = step < 4 ? 2 * step : 4;
                }
                if (!ahead) {
                    h1 = load_head(w1);
                    ahead = true;
                }
                // Row phase: the r <= sparse_lanes remaining rows (columns j..
                // onward) are scanned from the row-major copy side by side,
                // each by a group of G = 64/2^ceil(log2 r) lanes reading G
                // consecutive entries per step; the first hit in row order is
                // the min id.
                const uint64_t rem = __ballot(((unres >> lane) & 1ull) && j < len);
                const int r = __popcll(rem);
                if (r) {
                    int G = 64;
                    while (G * r > 64) G >>= 1;
                    const int gid = lane / G, gl = lane % G;
                    const uint64_t gmask_r = (G == 64) ? ~0ull : (((1ull << G) - 1) << (gid * G));
                    int L = -1; // the gid-th remaining lane
                    {
                        uint64_t t = rem;
                        for (int k = 0; k < gid && t; ++k) t &= t - 1;
                        if (gid < r) L = __ffsll((unsigned long long)t) - 1;
                    }
                    const int rl = __shfl(len, L < 0 ? 0 : L, 64);
                    const int64_t rs = L >= 0 ? g.pull_ptr[(int64_t)w * 64 + L] : 0;
                    int got = -1;
                    for (int k = j;; k += G) {
                        const bool live_g = L >= 0 && got < 0 && k < rl;
                        if (__ballot(live_g) == 0) break;
                        const int e = k + gl;
                        bool hh = false;
                        int c = 0;
                        if (live_g && e < rl) {
                            c = g.pull_col[rs + e];
                            const int wd = c >> 6;
                            bool maybe = true;
                            if constexpr (VAR == ALBD_PULL_LIST_CB) {
                                maybe = (lds_sum[wd >> 5] >> (wd & 31)) & 1u;
                                cbs += maybe ? 0u : 1u;
                            }
                            if (maybe) hh = (a.st.front[wd] >> (c & 63)) & 1ull;
                        }
                        scanned += __popcll(__ballot(live_g && e < rl));
                        const uint64_t hm = __ballot(hh) & gmask_r;
                        const int src = hm ? __ffsll((unsigned long long)hm) - 1 : lane;
                        const int first = __shfl(c, src, 64);
                        if (hm && got < 0) got = first;
                    }
                    for (int q = 0; q < r; ++q) {
                        const int Lq = __shfl(L, q * G, 64);
                        const int gq = __shfl(got, q * G, 64);
                        if (gq >= 0) {
                            unres &= ~(1ull << Lq);
                            if (lane == Lq) parent_pos = gq;
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
                        // position order: the lanes of a word write one run
                        a.st.lp[p] = pack_lp(a.level + 1, g.row_order[parent_pos]);
                        // out-degree: the row's length when the matrix is
                        // symmetric (the pull rows are the push rows)
                        edges += sym ? (unsigned long long)len
                                     : (unsigned long long)(g.push_ptr[p + 1] - g.push_ptr[p]);
                    }
                    cnt += __popcll(found);
                    nzw += 1;
                }
            }
        }
        if (!ahead) h1 = load_head(w1);
        w = w1;
        h = h1;
    }
    edges = wave_sum(edges);
    unsigned long long cb_total = 0;
    if constexpr (VAR == ALBD_PULL_LIST_CB) cb_total = wave_sum((unsigned long long)cbs);
    if (lane == 0) slot_add(a.st.slots, cnt, edges, nzw, scanned, wscan, wbyp, cb_total);
}

template <int C, int VAR>
__global__ __launch_bounds__(pull_block<VAR>()) void k_pull(PullArgs a)
{
    extern __shared__ uint32_t lds_sum[];
    if (!spec_go(a.st.ctr, Spec{a.spec, a.spec_choice, ~0ull})) return;
    pull_body<C, VAR, pull_block<VAR>()>(a, lds_sum);
}


// ---------------------------------------------------------------------------
// word passes

// One reservation per block in a global list: every thread calls it with its
// wave's count (lane 0's value counts); returns the wave's first slot.  Same
// address device atomics serialise, so the word passes take one per block
// step rather than one per wave.
template <int NW = kWavesPerBlock>
__device__ __forceinline__ unsigned long long block_reserve(unsigned long long *counter,
                                                           unsigned int wave_count,
                                                           unsigned int *s_cnt,
                                                           unsigned long long *s_base)
{
    const int wid = threadIdx.x >> 6;
    if (lane_id() == 0) s_cnt[wid] = wave_count;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int tot = 0;
        for (int i = 0; i < NW; ++i) tot += s_cnt[i];
        *s_base = tot ? atomicAdd(counter, (unsigned long long)tot) : 0ull;
    }
    __syncthreads();
    unsigned long long off = *s_base;
    for (int i = 0; i < wid; ++i) off += s_cnt[i];
    __syncthreads(); // s_cnt is reused by the next call
    return off;
}

__global__ __launch_bounds__(kBlock) void k_haslen(const int32_t *row_len, uint64_t *haslen,
                                                   int32_t nwords)
{
    const int64_t gt = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t w = gt >> 6;
    const bool pred = (w < nwords) && row_len[gt] > 0;
    const uint64_t m = __ballot(pred);
    if (lane_id() == 0 && w < nwords) haslen[w] = m;
}

// The state of a BFS from position rp (vertex ro) in one launch: the bitmaps
// and the slot and counter blocks zeroed, except what the root sets -- its
// visited and frontier bits, its level/parent entry, the queue holding it
// (when it has out-edges), and slot 0 seeded as the frontier of "level -1",
// which the first select kernel sums.  lp holds stale values outside the
// visited set (the vertex-order pass masks them), and pmin is all ones at
// rest (the push finalize restores every entry it consumed).
// each row's vertex id as an offset inside its window (two bytes where
// row_order takes four: the vertex-order pass streams these)
__global__ __launch_bounds__(kBlock) void k_row_off(albd_graph g, uint16_t *row_off)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < g.n) row_off[p] = (uint16_t)(g.row_order[p] - (int32_t)((p / g.win) * g.win));
}

__global__ __launch_bounds__(kBlock) void k_bfs_reset(albd_graph g, albd_state st, int32_t rp, int32_t ro)
{
    const int64_t deg = g.push_ptr[rp + 1] - g.push_ptr[rp];
    const int32_t rw = rp >> 6;
    const uint64_t bit = 1ull << (rp & 63);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < g.nwords;
         i += (int64_t)gridDim.x * kBlock) {
        const uint64_t v = i == rw ? bit : 0ull;
        st.visited[i] = v;
        st.front[i] = v;
        st.next[i] = 0;
    }
    static_assert(offsetof(albd_slot, next_count) == 0 && offsetof(albd_slot, next_edges) == 8 &&
                      offsetof(albd_slot, next_words) == 16,
                  "slot 0 seeded by word");
    unsigned long long *slots = reinterpret_cast<unsigned long long *>(st.slots);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
         i < (int64_t)(ALBD_NSLOT * sizeof(albd_slot) / 8); i += (int64_t)gridDim.x * kBlock)
        slots[i] = i == 0 || i == 2 ? 1ull : (i == 1 ? (unsigned long long)deg : 0ull);
    if (blockIdx.x == 0) {
        constexpr int kQ = (int)(offsetof(albd_counters, queue_len) / 8);
        unsigned long long *c = reinterpret_cast<unsigned long long *>(st.ctr);
        for (int i = threadIdx.x; i < (int)(sizeof(albd_counters) / 8); i += kBlock)
            c[i] = i == kQ && deg > 0 ? 1ull : 0ull;
        if (threadIdx.x == 0) {
            st.lp[rp] = pack_lp(0, ro);
            if (deg > 0) st.queue[0] = rp;
        }
    }
}

// End of a level: the frontier it produced, the running totals, the
// selector features (level_features) and the policy's choice for the next
// level.  One block.
constexpr int kSelectBlock = 256;

__device__ __forceinline__ void select_body(const albd_graph &g, const albd_state &st, const albd_select_args &sel,
                                            albd_counters *mail, unsigned int seq)
{
    // the counter block up to the mailbox sequence word, as 64-bit words
    constexpr int kCtrWords = (int)(offsetof(albd_counters, seq) / 8);
    static_assert(kCtrWords <= 64, "one wave moves the counter block");
    __shared__ float xs[ALBBA_MLP_IN], h1[ALBBA_MLP_H1], h2[ALBBA_MLP_H2], ys[ALBBA_MLP_OUT];
    __shared__ float s_params[ALBBA_MLP_NPARAM];
    __shared__ unsigned long long s_red[7][kSelectBlock / 64];
    __shared__ union {
        albd_counters c;
        unsigned long long w[sizeof(albd_counters) / 8];
    } s_ctr;
    const int lane = lane_id(), wid = threadIdx.x >> 6, tid = threadIdx.x;
    unsigned long long *ctr_words = reinterpret_cast<unsigned long long *>(st.ctr);
    // Every load first (the counter block, the weights, the slots), so that
    // they travel together; then the block sums the slots, and wave 0 alone
    // computes the features, runs the MLP and writes the counter block back
    // (one lane per word: thread 0 alone would chain one load per word).
    constexpr int kPer = ALBD_NSLOT / kSelectBlock;
    static_assert(ALBD_NSLOT % kSelectBlock == 0, "slots split evenly over the block");
    static_assert(sizeof(albd_slot) == 64, "slot = 4 x 16 counter bytes");
    // a used unvisited-word list (slot (level - 1) % 3, built by the
    // previous level's preparation) becomes the current list
    const bool list_used = !sel.first && (sel.prev_choice == ALBBA_CHOICE_PULL_SB ||
                                          sel.prev_choice == ALBBA_CHOICE_PULL_CB);
    unsigned long long cw = 0, used_len = 0;
    if (tid < kCtrWords) cw = ctr_words[tid];
    if (tid == 0 && list_used) used_len = st.ctr->list_cnt[(sel.level + 2) % 3];
    // the slot that the level after this one fills starts from zero
    if (tid == 0) st.ctr->list_cnt[(sel.level + 1) % 3] = 0;
    float pr[mlp_per_thread<kSelectBlock>()];
    if (sel.params) mlp_load<kSelectBlock>(pr, sel.params);
    ulonglong2 v[kPer][4];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(&st.slots[tid + k * kSelectBlock]);
        v[k][0] = p[0];
        v[k][1] = p[1];
        v[k][2] = p[2];
        v[k][3] = p[3];
    }
    // read once, by this thread only: the next level starts from zero
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        ulonglong2 *p = reinterpret_cast<ulonglong2 *>(&st.slots[tid + k * kSelectBlock]);
        p[0] = p[1] = p[2] = p[3] = make_ulonglong2(0ull, 0ull);
    }
    if (tid < kCtrWords) s_ctr.w[tid] = cw;
    if (sel.params) mlp_store<kSelectBlock>(s_params, pr);
    unsigned long long F = 0, mf = 0, nzw = 0, scn = 0, wds = 0, byp = 0, cbp = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        F += v[k][0].x;   // next_count
        mf += v[k][0].y;  // next_edges
        nzw += v[k][1].x; // next_words
        scn += v[k][1].y; // scanned
        wds += v[k][2].x; // words
        byp += v[k][2].y; // bypassed
        cbp += v[k][3].x; // col_bypassed
    }
    F = wave_sum(F);
    mf = wave_sum(mf);
    nzw = wave_sum(nzw);
    scn = wave_sum(scn);
    wds = wave_sum(wds);
    byp = wave_sum(byp);
    cbp = wave_sum(cbp);
    if (lane == 0) {
        s_red[0][wid] = F;
        s_red[1][wid] = mf;
        s_red[2][wid] = nzw;
        s_red[3][wid] = scn;
        s_red[4][wid] = wds;
        s_red[5][wid] = byp;
        s_red[6][wid] = cbp;
    }
    __syncthreads();
    if (wid != 0) return;
    albd_counters &c = s_ctr.c;
    // the totals in every lane of wave 0
    F = mf = nzw = scn = wds = byp = cbp = 0;
    for (int i = 0; i < kSelectBlock / 64; ++i) {
        F += s_red[0][i];
        mf += s_red[1][i];
        nzw += s_red[2][i];
        scn += s_red[3][i];
        wds += s_red[4][i];
        byp += s_red[5][i];
        cbp += s_red[6][i];
    }
    const unsigned long long reached = c.reached + F;
    const unsigned long long redges = c.reached_edges + mf;
    level_features(xs, g.n, g.nnz, g.nwords, F, mf, nzw, reached, redges, sel.level, sel.prev_choice,
                   c.frontier);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int choice = sel.params ? mlp_argmin_wave(s_params, xs, h1, h2, ys, sel.valid_mask) : -1;
    if (lane == 0) {
        c.frontier = F;
        c.frontier_edges = mf;
        c.frontier_words = nzw;
        c.scanned = scn;
        c.words = wds;
        c.bypassed = byp;
        c.col_bypassed = cbp;
        c.reached = reached;
        c.reached_edges = redges;
        const int heur = heuristic_choice(F, g.n);
        c.heur_choice = heur;
        c.sel_choice = sel.forced >= 0 ? sel.forced : (choice >= 0 ? choice : heur);
        for (int i = 0; i < ALBBA_MLP_IN; ++i) c.features[i] = xs[i];
        // hand the level's lists over to the next level: a pushed level's
        // queue becomes the queue (after a pull the queue is rebuilt from
        // the bitmap, from zero); a used unvisited-word list becomes the list
        if (!sel.first) c.queue_len = sel.prev_choice == ALBBA_CHOICE_PUSH ? c.qnext_len : 0ull;
        c.qnext_len = 0;
        if (list_used) c.active_len = used_len;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < kCtrWords) {
        const unsigned long long w = s_ctr.w[lane];
        ctr_words[lane] = w;
        if (mail) reinterpret_cast<unsigned long long *>(mail)[lane] = w;
    }
    if (mail) {
        // the whole block to host memory, then the sequence number with a
        // system-scope release: the host sees the data once it sees seq
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) __hip_atomic_store(&mail->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ __launch_bounds__(kSelectBlock) void k_select(albd_graph g, albd_state st, albd_select_args sel,
                                                         albd_counters *mail, unsigned int seq)
{
    select_body(g, st, sel, mail, seq);
}

// Zero st.next (the select kernel has moved the counters and zeroed the
// slots).  LB_SLOTS: zero the slots too (trial runs of one level);
// LB_SMALL: the small-frontier kernel follows; zero its grid state, level
// 0's counters and the error flag here instead of in launches of their own.
enum { LB_SLOTS = 1, LB_SMALL = 8 };

__global__ __launch_bounds__(kBlock) void k_level_begin(albd_state st, int32_t nwords, int flags)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nwords;
         i += (int64_t)gridDim.x * kBlock)
        st.next[i] = 0;
    if (flags & LB_SLOTS) {
        unsigned long long *slots = reinterpret_cast<unsigned long long *>(st.slots);
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
             i < (int64_t)(ALBD_NSLOT * sizeof(albd_slot) / 8); i += (int64_t)gridDim.x * kBlock)
            slots[i] = 0;
    }
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            if (flags & LB_SMALL) {
                st.persist->barrier = 0;
                st.persist->error = 0;
                for (int f = 0; f < 4; ++f) st.level_log[f] = 0;
                st.ctr->persist_error = 0;
                // the levels run inside that kernel skip the select kernels
                // that clear the list-count slots
                for (int k = 0; k < 3; ++k) st.ctr->list_cnt[k] = 0;
            }
        }
    }
}

// Where the list preparation of level L finds its counts: the input list's
// length in active_len (in_slot < 0: the select kernel of level L has
// already handed it over) or in list_cnt[in_slot] (the select kernel runs in
// the same launch); the output count in list_cnt[out_slot = L % 3].
struct ListSlots {
    int32_t in_slot, out_slot;
};

// unvisited-word list: from every word (FROM_LIST = false) or from the
// previous list (visited only grows, so an old list is a superset).  Blocks
// bid of nb.
template <bool FROM_LIST>
__device__ __forceinline__ void active_body(const albd_graph &g, const albd_state &st, ListSlots ls, int bid,
                                            int nb, unsigned int *s_cnt, unsigned long long *s_base)
{
    constexpr int kGroups = 4; // 64-item groups per wave per step
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const int64_t n_in =
        FROM_LIST ? (int64_t)(ls.in_slot < 0 ? st.ctr->active_len : st.ctr->list_cnt[ls.in_slot]) : (int64_t)g.nwords;
    unsigned long long *out_cnt = &st.ctr->list_cnt[ls.out_slot];
    const int64_t step = (int64_t)kBlock * kGroups;
    for (int64_t base = (int64_t)bid * step; base < n_in; base += (int64_t)nb * step) {
        int32_t w[kGroups];
        uint64_t m[kGroups];
        unsigned int cnt = 0;
#pragma unroll
        for (int k = 0; k < kGroups; ++k) {
            const int64_t i = base + (int64_t)(wid * kGroups + k) * 64 + lane;
            bool keep = false;
            w[k] = -1;
            if (i < n_in) {
                w[k] = FROM_LIST ? st.active[i] : (int32_t)i;
                keep = (g.haslen[w[k]] & ~st.visited[w[k]]) != 0;
            }
            m[k] = __ballot(keep);
            cnt += (unsigned int)__popcll(m[k]);
        }
        unsigned long long off = block_reserve(out_cnt, cnt, s_cnt, s_base);
#pragma unroll
        for (int k = 0; k < kGroups; ++k) {
            if ((m[k] >> lane) & 1ull) st.active_next[off + __popcll(m[k] & lanemask_lt(lane))] = w[k];
            off += (unsigned long long)__popcll(m[k]);
        }
    }
}

__device__ __forceinline__ void summary_body(const albd_graph &g, const albd_state &st, int bid, int nb)
{
    const int lane = lane_id();
    const int64_t ngroups = ((int64_t)g.nwords + 63) >> 6;
    const int64_t gw = ((int64_t)bid * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)nb * kBlock) >> 6;
    for (int64_t grp = gw; grp < ngroups; grp += nw) {
        const int64_t w = grp * 64 + lane;
        const bool nz = (w < g.nwords) && st.front[w] != 0;
        const uint64_t m = __ballot(nz);
        if (lane == 0) st.summary[grp] = m;
    }
}

template <bool FROM_LIST>
__global__ __launch_bounds__(kBlock) void k_active(albd_graph g, albd_state st, ListSlots ls)
{
    __shared__ unsigned int s_cnt[kWavesPerBlock];
    __shared__ unsigned long long s_base;
    active_body<FROM_LIST>(g, st, ls, blockIdx.x, gridDim.x, s_cnt, &s_base);
}

__global__ __launch_bounds__(kBlock) void k_summary(albd_graph g, albd_state st)
{
    summary_body(g, st, blockIdx.x, gridDim.x);
}

// The next level's preparation: zero st.next, compact the unvisited-word
// list, build the frontier summary.  The three parts touch disjoint data.
// Blocks bid of nb.
template <bool FROM_LIST, bool SUMMARY>
__device__ __forceinline__ void prep_body(const albd_graph &g, const albd_state &st, ListSlots ls, int bid, int nb)
{
    __shared__ unsigned int s_cnt[kWavesPerBlock];
    __shared__ unsigned long long s_base;
    for (int64_t i = (int64_t)bid * kBlock + threadIdx.x; i < g.nwords; i += (int64_t)nb * kBlock)
        st.next[i] = 0;
    active_body<FROM_LIST>(g, st, ls, bid, nb, s_cnt, &s_base);
    if constexpr (SUMMARY) summary_body(g, st, bid, nb);
}

template <bool FROM_LIST, bool SUMMARY>
__global__ __launch_bounds__(kBlock) void k_prep(albd_graph g, albd_state st, ListSlots ls)
{
    prep_body<FROM_LIST, SUMMARY>(g, st, ls, blockIdx.x, gridDim.x);
}

// The select kernel of level L (block 0) and level L's preparation (the
// other blocks) in one launch, so that neither waits for the other.  They
// share no data: the preparation reads the input list's length from its
// slot rather than from active_len, writes its count into a slot the select
// kernel does not touch, and zeroes st.next, which the select kernel does
// not read; the counters the select kernel sums are in st.slots.
template <bool FROM_LIST, bool SUMMARY>
__global__ __launch_bounds__(kBlock) void k_select_prep(albd_graph g, albd_state st, albd_select_args sel,
                                                        albd_counters *mail, unsigned int seq, ListSlots ls)
{
    static_assert(kSelectBlock == kBlock, "one block shape for both parts");
    if (blockIdx.x == 0) select_body(g, st, sel, mail, seq);
    else prep_body<FROM_LIST, SUMMARY>(g, st, ls, blockIdx.x - 1, gridDim.x - 1);
}

// Words -> queue of positions with out-degree > 0.  FINALIZE (after a push):
// the words are st.next; merge them into visited, write level/parent from
// pmin and count the new frontier; the queue only while the level's
// frontier edges (the select kernel's count) are at most build_max: a larger
// push likely leads to a pull, which needs none, and a push after it lists
// the bitmap.  Otherwise the words are st.front.
//
// 1024-thread blocks: the list reservation is one same-address atomic per
// block, and those serialise (256-thread blocks made the reservations the
// larger part of a big push level's finalize).
constexpr int kQBlock = 1024;
constexpr int kQWaves = kQBlock / 64;

template <bool FINALIZE>
__global__ __launch_bounds__(kQBlock) void k_words_to_queue(albd_graph g, albd_state st,
                                                            int32_t level, Spec sp, long long build_max)
{
    __shared__ unsigned int s_cnt[kQWaves];
    __shared__ unsigned long long s_base;
    if (!spec_go(st.ctr, sp)) return;
    const bool build = !FINALIZE || (long long)st.ctr->frontier_edges <= build_max;
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const uint64_t *words = FINALIZE ? st.next : st.front;
    int32_t *q = FINALIZE ? st.queue_next : st.queue;
    unsigned long long *qlen = FINALIZE ? &st.ctr->qnext_len : &st.ctr->queue_len;
    const int64_t ngroups = ((int64_t)g.nwords + 63) >> 6;
    unsigned long long cnt = 0, nzw = 0, edges = 0;
    // block-uniform steps of one 64-word group per wave
    for (int64_t gb = (int64_t)blockIdx.x * kQWaves; gb < ngroups; gb += (int64_t)gridDim.x * kQWaves) {
        const int64_t grp = gb + wid;
        const int64_t w = grp * 64 + lane;
        const uint64_t word = (grp < ngroups && w < g.nwords) ? words[w] : 0ull;
        if (FINALIZE && word) {
            st.visited[w] |= word; // only this lane touches word w here
            cnt += __popcll(word);
            nzw += 1;
        }
        const uint64_t nzmask = __ballot(word != 0);
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
                    st.lp[pp] = pack_lp(level + 1, (int32_t)st.pmin[pp]);
                    st.pmin[pp] = 0xFFFFFFFFu;
                    edges += (unsigned long long)d;
                }
            }
            const uint64_t km = __ballot(keep);
            if (lane == k) keepw = km;
        }
        if (!build) continue; // kernel-uniform
        unsigned long long total;
        const unsigned long long excl =
            wave_exclusive_scan<unsigned long long>((unsigned long long)__popcll(keepw), &total);
        const unsigned long long base = block_reserve<kQWaves>(qlen, (unsigned int)total, s_cnt, &s_base);
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
        if (lane == 0) slot_add(st.slots, cnt, edges, nzw, 0);
    }
}

// level and parent from position order to vertex order, once per BFS.
// Positions [k*win, (k+1)*win) hold the vertices of the same id range (the
// sigma windows), so a block takes one window: it streams the window's row
// offsets (row_off), lp and visited words in position order, scatters level and
// parent into LDS by vertex, and writes both out in vertex order.  Every
// global access is a coalesced stream.  LDS: 8 bytes per window row.
constexpr int kMaxWin = 8192;

// 512-thread blocks: a window's LDS arrays (32 KB at 4096 rows) allow four
// blocks per CU, which then hold the CU's 32 wave slots
constexpr int kVoBlock = 512;

__global__ __launch_bounds__(kVoBlock) void k_to_vertex_order_win(albd_graph g, albd_state st, int if_done)
{
    // queued before the host knew the BFS had ended: run only if the last
    // select kernel found the frontier empty
    if (if_done && st.ctr->frontier != 0) return;
    extern __shared__ int32_t s_win[];
    int32_t *s_level = s_win, *s_parent = s_win + g.win;
    constexpr int kV = 4; // positions in flight per thread
    const int64_t nwin = ((int64_t)g.n + g.win - 1) / g.win;
    for (int64_t w = blockIdx.x; w < nwin; w += gridDim.x) {
        const int64_t lo = w * g.win;
        const int32_t len = (int32_t)((int64_t)g.n - lo < g.win ? (int64_t)g.n - lo : g.win);
        // unreached vertices keep -1; only reached positions load their
        // offset and lp (a window's empty rows, its last positions, are
        // never reached unless one is the root: 43 % of the positions of an
        // RMAT s22 graph, whose lines are then not read)
        for (int32_t i = threadIdx.x; i < len; i += kVoBlock) {
            s_level[i] = -1;
            s_parent[i] = -1;
        }
        __syncthreads();
        for (int32_t i0 = threadIdx.x; i0 < len; i0 += kVoBlock * kV) {
            bool reached[kV];
#pragma unroll
            for (int k = 0; k < kV; ++k) {
                const int32_t i = i0 + k * kVoBlock;
                const int64_t p = lo + (i < len ? i : 0);
                reached[k] = i < len && ((st.visited[p >> 6] >> (p & 63)) & 1ull);
            }
            int32_t o[kV];
            unsigned long long x[kV];
#pragma unroll
            for (int k = 0; k < kV; ++k) {
                const int64_t p = lo + i0 + k * kVoBlock;
                o[k] = reached[k] ? (int32_t)g.row_off[p] : 0;
                x[k] = reached[k] ? st.lp[p] : 0ull;
            }
#pragma unroll
            for (int k = 0; k < kV; ++k) {
                if (reached[k] && o[k] < len) {
                    s_level[o[k]] = (int32_t)(x[k] >> 32);
                    s_parent[o[k]] = (int32_t)(uint32_t)x[k];
                }
            }
        }
        __syncthreads();
        for (int32_t i = threadIdx.x; i < len; i += kVoBlock) {
            st.level[lo + i] = s_level[i];
            st.parent[lo + i] = s_parent[i];
        }
        __syncthreads(); // the LDS arrays are reused by the next window
    }
}

// The same for windows longer than the LDS arrays: a block takes 4096
// consecutive vertex ids and gathers lp through pos_of.
constexpr int kVertexChunk = 4096;

__global__ __launch_bounds__(kBlock) void k_to_vertex_order(albd_graph g, albd_state st, int if_done)
{
    if (if_done && st.ctr->frontier != 0) return;
    constexpr int kV = 4; // vertices in flight per thread
    for (int64_t c0 = (int64_t)blockIdx.x * kVertexChunk; c0 < g.n; c0 += (int64_t)gridDim.x * kVertexChunk) {
        for (int64_t v0 = c0 + threadIdx.x; v0 < c0 + kVertexChunk && v0 < g.n; v0 += kBlock * kV) {
            int32_t p[kV];
#pragma unroll
            for (int k = 0; k < kV; ++k) {
                const int64_t v = v0 + k * kBlock;
                p[k] = (v < g.n && v < c0 + kVertexChunk) ? g.pos_of[v] : 0;
            }
            uint64_t vis[kV];
            unsigned long long x[kV];
#pragma unroll
            for (int k = 0; k < kV; ++k) {
                vis[k] = st.visited[p[k] >> 6];
                x[k] = st.lp[p[k]];
            }
#pragma unroll
            for (int k = 0; k < kV; ++k) {
                const int64_t v = v0 + k * kBlock;
                if (v >= g.n || v >= c0 + kVertexChunk) break;
                const bool reached = (vis[k] >> (p[k] & 63)) & 1ull;
                st.level[v] = reached ? (int32_t)(x[k] >> 32) : -1;
                st.parent[v] = reached ? (int32_t)(uint32_t)x[k] : -1;
            }
        }
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

__global__ __launch_bounds__(kBlock) void k_scan_small(albd_graph g, albd_state st, Spec sp)
{
    __shared__ int64_t lds[8];
    if (!spec_go(st.ctr, sp)) return;
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

__global__ __launch_bounds__(kBlock) void k_push_expand(albd_graph g, albd_state st, Spec sp)
{
    if (!spec_go(st.ctr, sp)) return;
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
                // pmin is all ones at rest: the first edge into v this level
                // sees that value back and is the one that sets v's bit
                if (atomicMin(&st.pmin[v], ou) == 0xFFFFFFFFu)
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

// ---------------------------------------------------------------------------
// small-frontier push levels in one persistent launch
//
// High-diameter graphs spend thousands of levels with a few thousand
// frontier vertices; a launch sequence and a host round trip per level would
// dominate, and one workgroup alone is bound by its CU's gather rate.  A
// grid of G co-resident workgroups runs the push levels back to
// back with one grid barrier per level, until the next frontier has more
// than edge_cap out-edges or more than vertex_cap vertices.
//
// Exactness without a second barrier per level: every position p has a
// 64-bit key, all ones outside this kernel.  An edge u -> v at level L does
// atomicMin(key[v], L << 32 | id(u)).  The old value tells the discoverer
// (all ones: first at this level), and the key keeps the smallest (level,
// parent id) pair, i.e. the min-id parent at v's BFS level: keys of earlier
// levels are smaller and never change.  Two filters save atomics: the
// visited bitmap as it was when the kernel started, and the key itself when
// it already holds an earlier level (keys only decrease, so a stale copy
// costs an atomic, never a wrong answer).  Neither is written during the
// launch: the vertices it finds get their level/parent words, visited bits
// and keys back to all ones when it ends.
//
// Cross-workgroup data (found list, queues, keys, counters) moves through
// device-coherent loads, stores and atomics; every wave waits for its own
// memory operations before the barrier.

constexpr int kSmallBlock = 1024;
constexpr int kSmallWaves = kSmallBlock / 64;

constexpr int kSmallChunk = kSmallBlock;      // queue entries staged per pass
constexpr int kSmallPer = 4;                  // edges per thread per window
constexpr int kSmallWin = kSmallPer * kSmallBlock;

struct SmallArgs {
    albd_graph g;
    albd_state st;
    int32_t level0;
    int32_t max_levels;
    int64_t edge_cap;
    int64_t vertex_cap;
    uint32_t spin_limit;
};

__global__ __launch_bounds__(kSmallBlock) void k_push_small(SmallArgs a)
{
    const albd_graph &g = a.g;
    const albd_state &st = a.st;
    __shared__ int64_t s_pre[kSmallChunk];
    __shared__ int64_t s_start[kSmallChunk];
    __shared__ uint32_t s_ou[kSmallChunk];
    __shared__ int64_t s_scan[kSmallWaves];
    __shared__ int32_t s_found[kSmallWin];
    __shared__ int32_t s_queue[kSmallWin];
    __shared__ unsigned int s_nf, s_nq;
    __shared__ unsigned long long s_bf, s_bq, s_sum;
    __shared__ int s_ok;
    albd_persist *P = st.persist;
    // cumulative list of the vertices this launch finds, level after level
    int32_t *found = reinterpret_cast<int32_t *>(st.prefix);
    int32_t *qa = st.queue, *qb = st.queue_next;
    int64_t qlen = (int64_t)st.ctr->queue_len;
    const int tid = threadIdx.x, lane = lane_id();
    const unsigned int G = gridDim.x, w = blockIdx.x;
    const int64_t gtid = (int64_t)w * kSmallBlock + tid, gthreads = (int64_t)G * kSmallBlock;
    if (tid == 0) {
        s_nf = 0;
        s_nq = 0;
        s_sum = 0;
    }
    __syncthreads();
    int64_t fo = 0; // found[0, fo): the earlier levels of this launch
    bool ok = true;
    for (int k = 0; ok; ++k) {
        const uint32_t lvl = (uint32_t)(a.level0 + k + 1);
        // found, edges, queue entries of level k, and the level's count in
        // the found list (appends go to found[fo + ...]).  Per level: a
        // workgroup past the barrier may already append for the next level
        // while another still reads this level's counts.
        int64_t *cnt = st.level_log + 4 * k;
        // expansion: workgroup w takes queue entries [q0, q1)
        const int64_t per = (qlen + G - 1) / G;
        const int64_t q0 = (int64_t)w * per, q1 = q0 + per < qlen ? q0 + per : qlen;
        unsigned long long my_edges = 0;
        for (int64_t base = q0; base < q1; base += kSmallChunk) { // uniform in the workgroup
            const int cnt_e = (int)(q1 - base < kSmallChunk ? q1 - base : kSmallChunk);
            int64_t a0 = 0, e0 = 0;
            uint32_t o0 = 0;
            if (tid < cnt_e) {
                const int32_t u = ld_coherent(&qa[base + tid]);
                a0 = g.push_ptr[u];
                e0 = g.push_ptr[u + 1];
                o0 = (uint32_t)g.row_order[u];
            }
            int64_t T;
            const int64_t ex = block_scan_excl<kSmallBlock>(e0 - a0, &T, s_scan);
            s_pre[tid] = ex;
            s_start[tid] = a0;
            s_ou[tid] = o0;
            __syncthreads();
            for (int64_t wb = 0; wb < T; wb += kSmallWin) { // uniform
                int64_t e = wb + (int64_t)tid * kSmallPer;
                const int64_t e_end = e + kSmallPer < T ? e + kSmallPer : T;
                int lo = 0;
                if (e < e_end) {
                    int hi = cnt_e;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s_pre[mid] <= e) lo = mid;
                        else hi = mid;
                    }
                }
                int32_t v[kSmallPer];
                uint32_t ou[kSmallPer];
                bool okv[kSmallPer];
#pragma unroll
                for (int i = 0; i < kSmallPer; ++i) {
                    okv[i] = e < e_end;
                    v[i] = 0;
                    ou[i] = 0;
                    if (okv[i]) {
                        while (lo + 1 < cnt_e && s_pre[lo + 1] <= e) ++lo;
                        v[i] = g.push_col[s_start[lo] + (e - s_pre[lo])];
                        ou[i] = s_ou[lo];
                        ++e;
                    }
                }
                uint64_t vis[kSmallPer];
                unsigned long long pk[kSmallPer];
                int64_t d0[kSmallPer], d1[kSmallPer];
                // the filters and the out-degree together: one round trip
                // ahead of the atomic instead of one after it
#pragma unroll
                for (int i = 0; i < kSmallPer; ++i) {
                    vis[i] = okv[i] ? st.visited[v[i] >> 6] : ~0ull;
                    pk[i] = okv[i] ? st.pkey[v[i]] : 0ull;
                    d0[i] = okv[i] ? g.push_ptr[v[i]] : 0;
                    d1[i] = okv[i] ? g.push_ptr[v[i] + 1] : 0;
                }
                bool won[kSmallPer];
#pragma unroll
                for (int i = 0; i < kSmallPer; ++i) {
                    won[i] = false;
                    // unvisited at the launch and not found at an earlier
                    // level of it
                    if (!((vis[i] >> (v[i] & 63)) & 1ull) && (uint32_t)(pk[i] >> 32) >= lvl) {
                        const unsigned long long key =
                            ((unsigned long long)lvl << 32) | (unsigned long long)ou[i];
                        const unsigned long long old = atomicMin(&st.pkey[v[i]], key);
                        won[i] = old == ~0ull;
                    }
                }
                int64_t dv[kSmallPer];
#pragma unroll
                for (int i = 0; i < kSmallPer; ++i)
                    dv[i] = won[i] ? d1[i] - d0[i] : 0;
#pragma unroll
                for (int i = 0; i < kSmallPer; ++i) {
                    const unsigned int pf = wave_append(&s_nf, won[i]);
                    if (won[i]) s_found[pf] = v[i];
                    const unsigned int pq = wave_append(&s_nq, dv[i] > 0);
                    if (dv[i] > 0) s_queue[pq] = v[i];
                    my_edges += (unsigned long long)dv[i];
                }
                __syncthreads();
                // flush the window's lists: one reservation per list
                const unsigned int nf = s_nf, nq = s_nq;
                if (tid == 0) {
                    s_bf = nf ? atomicAdd(reinterpret_cast<unsigned long long *>(&cnt[3]), (unsigned long long)nf)
                              : 0;
                    s_bq = nq ? atomicAdd(reinterpret_cast<unsigned long long *>(&cnt[2]),
                                          (unsigned long long)nq)
                              : 0;
                }
                __syncthreads();
                const unsigned long long bf = s_bf, bq = s_bq;
                for (unsigned int i = tid; i < nf; i += kSmallBlock) st_coherent(&found[fo + bf + i], s_found[i]);
                for (unsigned int i = tid; i < nq; i += kSmallBlock) st_coherent(&qb[bq + i], s_queue[i]);
                __syncthreads();
                if (tid == 0) {
                    s_nf = 0;
                    s_nq = 0;
                }
            }
            __syncthreads(); // s_pre and friends are restaged
        }
        my_edges = wave_sum(my_edges);
        if (lane == 0 && my_edges) atomicAdd(&s_sum, my_edges);
        __syncthreads();
        if (tid == 0) {
            if (s_sum) atomicAdd(reinterpret_cast<unsigned long long *>(&cnt[1]), s_sum);
            s_sum = 0;
            // the next level's counters, unused until after the barrier
            if (w == 0 && k + 1 < a.max_levels)
                for (int f = 0; f < 4; ++f) st_coherent(&cnt[4 + f], (int64_t)0);
        }
        ok = team_sync(&P->barrier, &P->error, G, a.spin_limit, &s_ok);
        if (!ok) break;
        const int64_t na = ld_coherent(&cnt[3]);
        const int64_t total = fo + na;
        const int64_t edges = ld_coherent(&cnt[1]);
        const int64_t nq = ld_coherent(&cnt[2]);
        if (w == 0 && tid == 0) st_coherent(&cnt[0], na);
        const bool last = na == 0 || edges > a.edge_cap || na > a.vertex_cap ||
                          k + 1 >= a.max_levels;
        if (last) {
            // this level's vertices: level, parent, visited and next bits;
            // then every key set by this launch goes back to all ones
            unsigned long long my_words = 0;
            for (int64_t i = fo + gtid; i < total; i += gthreads) {
                const int32_t v = ld_coherent(&found[i]);
                const unsigned long long key = ld_coherent(&st.pkey[v]);
                st.lp[v] = key;
                const uint64_t bit = 1ull << (v & 63);
                atomicOr(reinterpret_cast<unsigned long long *>(&st.visited[v >> 6]), bit);
                const unsigned long long old =
                    atomicOr(reinterpret_cast<unsigned long long *>(&st.next[v >> 6]), bit);
                my_words += old == 0 ? 1 : 0;
                st.pkey[v] = ~0ull;
            }
            // the earlier levels' vertices: level, parent and visited too
            for (int64_t i = gtid; i < fo; i += gthreads) {
                const int32_t v = ld_coherent(&found[i]);
                st.lp[v] = ld_coherent(&st.pkey[v]); // the same (level, parent) packing
                atomicOr(reinterpret_cast<unsigned long long *>(&st.visited[v >> 6]), 1ull << (v & 63));
                st.pkey[v] = ~0ull;
            }
            my_words = wave_sum(my_words);
            if (lane == 0 && my_words) atomicAdd(&st.slots[1 + (w % (ALBD_NSLOT - 1))].next_words, my_words);
            if (w == 0 && tid == 0) {
                st.slots[0].next_count = (unsigned long long)na;
                st.slots[0].next_edges = (unsigned long long)edges;
                st.ctr->qnext_len = (unsigned long long)nq;
                st.ctr->persist_levels = k + 1;
                st.ctr->persist_in_queue = qb == st.queue ? 1 : 0;
            }
            break;
        }
        // an intermediate level: the select kernel sees only the last one,
        // so these counts go straight into the running totals
        if (w == 0 && tid == 0) {
            st.ctr->reached += (unsigned long long)na;
            st.ctr->reached_edges += (unsigned long long)edges;
        }
        fo = total;
        int32_t *t = qa;
        qa = qb;
        qb = t;
        qlen = nq;
    }
    if (!ok && tid == 0) st.ctr->persist_error = 1;
}

int grid_persistent(albd_ctx *c, int blocks_per_cu)
{
    int g = albd_num_cus(c) * blocks_per_cu;
    return g < 8 ? 8 : g;
}

// blocks of k_words_to_queue (1024 threads) for `groups` 64-word groups
int grid_for_qgroups(albd_ctx *c, int64_t groups)
{
    int64_t blocks = (groups + kQWaves - 1) / kQWaves;
    int64_t cap = (int64_t)albd_num_cus(c) * 2;
    if (blocks > cap) blocks = cap;
    return blocks < 1 ? 1 : (int)blocks;
}

int grid_for_groups(albd_ctx *c, int64_t groups)
{
    int64_t blocks = (groups + kWavesPerBlock - 1) / kWavesPerBlock;
    int64_t cap = (int64_t)albd_num_cus(c) * 8;
    if (blocks > cap) blocks = cap;
    return blocks < 1 ? 1 : (int)blocks;
}

template <int C>
int launch_pull_c(albd_ctx *c, PullArgs a, int variant, size_t lds, int64_t items)
{
    hipStream_t s = albd_stream(c);
    // blocks of per_block words; the grid is a multiple of 8 (one share per XCD)
    int64_t blocks;
    if (variant == ALBD_PULL_LIST_CB) {
        // the LDS summary is staged once per block: fewer, longer blocks
        int per_cu = (int)(160 * 1024 / (lds + 1024));
        per_cu = per_cu > 2 ? 2 : (per_cu < 1 ? 1 : per_cu);
        const int64_t resident = (int64_t)albd_num_cus(c) * per_cu;
        a.per_block = (int32_t)((items + resident - 1) / resident);
        if (a.per_block < 16) a.per_block = 16;
    } else {
        a.per_block = 16;
    }
    blocks = (items + a.per_block - 1) / a.per_block;
    blocks = (blocks + 7) / 8 * 8;
    if (blocks < 8) blocks = 8;
    if (variant == ALBD_PULL_ALL)
        hipLaunchKernelGGL((k_pull<C, ALBD_PULL_ALL>), dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    else if (variant == ALBD_PULL_LIST)
        hipLaunchKernelGGL((k_pull<C, ALBD_PULL_LIST>), dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_pull<C, ALBD_PULL_LIST_CB>), dim3((unsigned)blocks), dim3(kCBBlock), lds, s, a);
    HIPCHK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" {

int albd_build_row_off(albd_ctx *c, const albd_graph *g, uint16_t *row_off)
{
    if (g->n <= 0) return 0;
    if (g->win < 1 || g->win > 65536) return albd_set_error_msg(ALBBA_EINVAL, "row offsets: window too long");
    hipLaunchKernelGGL(k_row_off, dim3((unsigned)(((int64_t)g->n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       albd_stream(c), *g, row_off);
    HIPCHK(hipGetLastError());
    return 0;
}

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
    int64_t blocks = ((int64_t)g->nwords + kBlock - 1) / kBlock;
    int64_t cap = (int64_t)albd_num_cus(c) * 16;
    hipLaunchKernelGGL(k_bfs_reset, dim3((unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks))), dim3(kBlock),
                       0, s, *g, *st, rp, ro);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_to_vertex_order(albd_ctx *c, const albd_graph *g, const albd_state *st, int if_done)
{
    if (g->win >= 1 && g->win <= kMaxWin && g->row_off) {
        int64_t blocks = ((int64_t)g->n + g->win - 1) / g->win;
        int64_t cap = (int64_t)albd_num_cus(c) * 16;
        hipLaunchKernelGGL(k_to_vertex_order_win,
                           dim3((unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks))), dim3(kVoBlock),
                           sizeof(int32_t) * 2 * (size_t)g->win, albd_stream(c), *g, *st, if_done);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int64_t blocks = ((int64_t)g->n + kVertexChunk - 1) / kVertexChunk;
    int64_t cap = (int64_t)albd_num_cus(c) * 32;
    hipLaunchKernelGGL(k_to_vertex_order, dim3((unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks))),
                       dim3(kBlock), 0, albd_stream(c), *g, *st, if_done);
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
                      int64_t prev_len, int32_t level)
{
    hipStream_t s = albd_stream(c);
    const int64_t n_in = from_list ? prev_len : g->nwords;
    const int grid = grid_for_groups(c, (n_in + 63) / 64);
    const ListSlots ls{-1, level % 3};
    if (from_list)
        hipLaunchKernelGGL(k_active<true>, dim3(grid), dim3(kBlock), 0, s, *g, *st, ls);
    else
        hipLaunchKernelGGL(k_active<false>, dim3(grid), dim3(kBlock), 0, s, *g, *st, ls);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_cb_supported(const albd_graph *g)
{
    // summary bitmap (one bit per frontier word) must fit the LDS budget
    const int64_t bytes = (((int64_t)g->nwords + 63) / 64) * 8;
    return bytes <= 64 * 1024;
}

int albd_prep(albd_ctx *c, const albd_graph *g, const albd_state *st, int from_list, int summary, int32_t level)
{
    const int grid = grid_for_groups(c, ((int64_t)g->nwords + 63) / 64);
    hipStream_t s = albd_stream(c);
    const ListSlots ls{-1, level % 3};
    if (from_list) {
        if (summary) hipLaunchKernelGGL((k_prep<true, true>), dim3(grid), dim3(kBlock), 0, s, *g, *st, ls);
        else hipLaunchKernelGGL((k_prep<true, false>), dim3(grid), dim3(kBlock), 0, s, *g, *st, ls);
    } else {
        if (summary) hipLaunchKernelGGL((k_prep<false, true>), dim3(grid), dim3(kBlock), 0, s, *g, *st, ls);
        else hipLaunchKernelGGL((k_prep<false, false>), dim3(grid), dim3(kBlock), 0, s, *g, *st, ls);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_select_prep_post(albd_ctx *c, const albd_graph *g, const albd_state *st, const albd_select_args *sel,
                          unsigned int *seq, int from_list, int summary)
{
    albd_counters *mail = albd_mail_next(c, seq);
    // one select block in front of the preparation's grid
    const int grid = 1 + grid_for_groups(c, ((int64_t)g->nwords + 63) / 64);
    hipStream_t s = albd_stream(c);
    const bool list_used = !sel->first && (sel->prev_choice == ALBBA_CHOICE_PULL_SB ||
                                           sel->prev_choice == ALBBA_CHOICE_PULL_CB);
    const ListSlots ls{list_used ? (sel->level + 2) % 3 : -1, sel->level % 3};
    const unsigned int q = *seq;
    if (from_list) {
        if (summary)
            hipLaunchKernelGGL((k_select_prep<true, true>), dim3(grid), dim3(kBlock), 0, s, *g, *st, *sel, mail, q, ls);
        else
            hipLaunchKernelGGL((k_select_prep<true, false>), dim3(grid), dim3(kBlock), 0, s, *g, *st, *sel, mail, q, ls);
    } else {
        if (summary)
            hipLaunchKernelGGL((k_select_prep<false, true>), dim3(grid), dim3(kBlock), 0, s, *g, *st, *sel, mail, q, ls);
        else
            hipLaunchKernelGGL((k_select_prep<false, false>), dim3(grid), dim3(kBlock), 0, s, *g, *st, *sel, mail, q, ls);
    }
    HIPCHK(hipGetLastError());
    return 0;
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
    return albd_pull_spec(c, g, st, variant, level, items_bound, -1);
}

int albd_pull_spec(albd_ctx *c, const albd_graph *g, const albd_state *st, int variant, int32_t level,
                   int64_t items_bound, int spec)
{
    PullArgs a;
    a.g = *g;
    a.st = *st;
    a.level = level;
    a.spec = spec;
    a.spec_choice = variant == ALBD_PULL_ALL ? ALBBA_CHOICE_PULL
                                             : (variant == ALBD_PULL_LIST ? ALBBA_CHOICE_PULL_SB : ALBBA_CHOICE_PULL_CB);
    // the list level `level`'s preparation built, and its count
    a.items = st->active_next;
    a.nitems = &st->ctr->list_cnt[level % 3];
    a.per_block = 16;
    const int64_t sum_words64 = ((int64_t)g->nwords + 63) / 64;
    a.sum_words32 = (int32_t)(sum_words64 * 2);
    size_t lds = 0;
    if (variant == ALBD_PULL_LIST_CB) {
        if (!albd_cb_supported(g)) return albd_set_error_msg(ALBBA_EINVAL, "column bypass: graph too large for the LDS summary");
        lds = (size_t)a.sum_words32 * 4;
    }
    // items: every word, or an upper bound of the list length the device holds
    int64_t items = g->nwords;
    if (variant != ALBD_PULL_ALL && items_bound > 0 && items_bound < items) items = items_bound;
    switch (g->C) {
    case 16: return launch_pull_c<16>(c, a, variant, lds, items);
    case 32: return launch_pull_c<32>(c, a, variant, lds, items);
    case 64: return launch_pull_c<64>(c, a, variant, lds, items);
    case 128: return launch_pull_c<128>(c, a, variant, lds, items);
    default: return albd_set_error_msg(ALBBA_EINVAL, "pull: unsupported slice height");
    }
}

int albd_front_to_queue(albd_ctx *c, const albd_graph *g, const albd_state *st)
{
    const int grid = grid_for_qgroups(c, ((int64_t)g->nwords + 63) / 64);
    hipLaunchKernelGGL(k_words_to_queue<false>, dim3(grid), dim3(kQBlock), 0, albd_stream(c), *g, *st, 0, kNoSpec,
                       LLONG_MAX);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_push_scan(albd_ctx *c, const albd_graph *g, const albd_state *st, int64_t qlen)
{
    hipStream_t s = albd_stream(c);
    if (qlen <= ALBD_SPEC_PUSH_MAX) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kBlock), 0, s, *g, *st, kNoSpec);
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
    hipLaunchKernelGGL(k_push_expand, dim3(grid_persistent(c, 8)), dim3(kBlock), 0, albd_stream(c), *g, *st,
                       kNoSpec);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_push_small(albd_ctx *c, const albd_graph *g, const albd_state *st, int32_t level0,
                    int64_t edge_cap, int64_t vertex_cap, int32_t max_levels, int32_t blocks,
                    uint32_t spin_limit)
{
    if (max_levels > ALBD_PERSIST_LEVELS) max_levels = ALBD_PERSIST_LEVELS;
    if (max_levels < 1) max_levels = 1;
    // Every workgroup of the grid barrier must be resident at once: the grid
    // is at most what the occupancy allows on every CU together, and the
    // launch is an ordinary one on the context's stream, whose earlier
    // kernels have finished before it starts (stream order).  A workgroup
    // that could not become resident would only make the barrier give up
    // after spin_limit polls (ctr->persist_error), never hang.
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_push_small, kSmallBlock, 0));
    const int cap = per_cu * albd_num_cus(c);
    if (blocks <= 0) blocks = 64;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) return albd_set_error_msg(ALBBA_EHIP, "small-frontier kernel cannot be resident");
    // level 0's counters, the grid state and the error flag were zeroed by
    // the level-begin kernel (LB_SMALL)
    SmallArgs a{*g, *st, level0, max_levels, edge_cap, vertex_cap, spin_limit};
    hipLaunchKernelGGL(k_push_small, dim3(blocks), dim3(kSmallBlock), 0, albd_stream(c), a);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_select(albd_ctx *c, const albd_graph *g, const albd_state *st, const albd_select_args *sel)
{
    hipLaunchKernelGGL(k_select, dim3(1), dim3(kSelectBlock), 0, albd_stream(c), *g, *st, *sel,
                       (albd_counters *)nullptr, 0u);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_select_post(albd_ctx *c, const albd_graph *g, const albd_state *st,
                     const albd_select_args *sel, unsigned int *seq)
{
    albd_counters *mail = albd_mail_next(c, seq);
    hipLaunchKernelGGL(k_select, dim3(1), dim3(kSelectBlock), 0, albd_stream(c), *g, *st, *sel, mail,
                       *seq);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_select_read(albd_ctx *c, const albd_graph *g, const albd_state *st,
                     const albd_select_args *sel)
{
    unsigned int seq;
    if (int rc = albd_select_post(c, g, st, sel, &seq)) return rc;
    return albd_mail_wait(c, st, seq);
}

int albd_push_finalize(albd_ctx *c, const albd_graph *g, const albd_state *st, int32_t level, int64_t build_max)
{
    const int grid = grid_for_qgroups(c, ((int64_t)g->nwords + 63) / 64);
    hipLaunchKernelGGL(k_words_to_queue<true>, dim3(grid), dim3(kQBlock), 0, albd_stream(c), *g, *st, level,
                       kNoSpec, (long long)build_max);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_push_spec(albd_ctx *c, const albd_graph *g, const albd_state *st, int32_t level, int spec,
                   int64_t build_max)
{
    // the frontier bound is the one below which albd_push_scan takes the
    // one-block scan
    const Spec sp{spec, ALBBA_CHOICE_PUSH, (unsigned long long)ALBD_SPEC_PUSH_MAX};
    hipStream_t s = albd_stream(c);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kBlock), 0, s, *g, *st, sp);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_push_expand, dim3(grid_persistent(c, 8)), dim3(kBlock), 0, s, *g, *st, sp);
    HIPCHK(hipGetLastError());
    const int grid = grid_for_qgroups(c, ((int64_t)g->nwords + 63) / 64);
    hipLaunchKernelGGL(k_words_to_queue<true>, dim3(grid), dim3(kQBlock), 0, s, *g, *st, level, sp,
                       (long long)build_max);
    HIPCHK(hipGetLastError());
    return 0;
}

} // extern "C"
