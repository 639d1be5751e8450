// SELL-C-sigma layout built on the device from a CSR: the same arrays as the
// host builder (albba_sellcs_build), padded to the device shapes exactly as
// albba_graph_upload pads them.
//
// Steps (reference: bfs/src/bfs_SELLCSv2.c:34-59 sellcs_preprocess;
// bfs/src/sorting_utils.c:173-253 the sigma-window sort; bfs/src/utils.c:
// 198-241 the active lanes per column; bfs/src/sparse_matrix.c:523-623 the
// column packing):
//   1. row lengths, the longest row, column-range and row-order checks;
//   2. rows not in ascending id order are sorted (segmented radix sort);
//   3. the sigma-window order: one radix sort of 64-bit keys
//      (window, longest - length, id within the window) -- ascending keys give
//      descending lengths inside each window with ties in ascending id, the
//      order of the reference's stable descending radix sort;
//   4. slice widths (first row of each slice) and their scan (vop pointers);
//   5. per (slice, column): active lanes (a binary search in the slice's
//      descending lengths) and the padded column size, whose scan gives each
//      column's start and the slice pointers;
//   6. the row-major copy of the rows with columns as positions (the push
//      CSR when the matrix is symmetric, the pull CSR otherwise), written
//      coalesced, then the column-major layout from it: lanes are a slice's
//      rows, so each column is one coalesced store;
//   7. directed graphs: the push CSR (transpose over positions) by one radix
//      sort of (column, row) keys.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "../../../include/albba.h"
#include "dev_common.h"

namespace {

constexpr int kBB = 256; // threads per block of the builder kernels

__device__ __forceinline__ int64_t pad_up(int64_t x, int64_t p) { return (x + p - 1) / p * p; }

int grid_for(int64_t items, int per_block)
{
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}

// Walk the entries of the 64 rows held by the lanes of one wave (row i of
// the group: `start`, `len` in lane i), 64 entries per step, lanes taking
// consecutive entries: f(i, j, active) for entry j of row i.  Every lane
// calls f on every step (the shuffles inside need the whole wave); `active`
// says whether its (i, j) is an entry.
template <class F>
__device__ __forceinline__ void wave_rows_foreach(int32_t len, F &&f)
{
    const int lane = lane_id();
    unsigned long long tot;
    const unsigned long long ex = wave_exclusive_scan<unsigned long long>((unsigned long long)len, &tot);
    for (unsigned long long e0 = 0; e0 < tot; e0 += 64) {
        const unsigned long long e = e0 + (unsigned long long)lane;
        // the row of entry e: the last lane whose exclusive offset is <= e
        int lo = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const unsigned long long x = __shfl(ex, lo + step, 64);
            if (x <= e) lo += step;
        }
        const unsigned long long base = __shfl(ex, lo, 64);
        f(lo, (int64_t)(e - base), e < tot);
    }
}

// 1. lengths, longest row, checks.  Rows are in ascending order exactly
// when every descent col[k-1] > col[k] sits at the start of a row: one flat
// pass counts all descents (and checks the column range), one pass over the
// rows counts the descents at row starts.  flags: [0] longest row, [1] a
// column index outside [0, n), [2] descents, [3] descents at row starts.
__global__ __launch_bounds__(kBB) void k_rows(const int64_t *rp, const int32_t *col, int32_t n, int64_t nnz,
                                              int32_t *len, unsigned long long *flags)
{
    const int lane = lane_id();
    unsigned long long mx = 0, at_start = 0;
    for (int64_t r = (int64_t)blockIdx.x * kBB + threadIdx.x; r < n; r += (int64_t)gridDim.x * kBB) {
        const int64_t a = rp[r], b = rp[r + 1];
        len[r] = (int32_t)(b - a);
        mx = max(mx, (unsigned long long)(b - a));
        // a non-empty row's start is distinct from every other one's
        if (a > 0 && b > a && col[a - 1] > col[a]) ++at_start;
    }
    mx = max(mx, (unsigned long long)__shfl_xor(mx, 32, 64));
    mx = max(mx, (unsigned long long)__shfl_xor(mx, 16, 64));
    mx = max(mx, (unsigned long long)__shfl_xor(mx, 8, 64));
    mx = max(mx, (unsigned long long)__shfl_xor(mx, 4, 64));
    mx = max(mx, (unsigned long long)__shfl_xor(mx, 2, 64));
    mx = max(mx, (unsigned long long)__shfl_xor(mx, 1, 64));
    at_start = wave_sum(at_start);
    if (lane == 0) {
        if (mx) atomicMax(&flags[0], mx);
        if (at_start) atomicAdd(&flags[3], at_start);
    }
    (void)nnz;
}

__global__ __launch_bounds__(kBB) void k_entries(const int32_t *col, int32_t n, int64_t nnz,
                                                 unsigned long long *flags)
{
    const int lane = lane_id();
    unsigned long long desc = 0, bad = 0;
    for (int64_t k = (int64_t)blockIdx.x * kBB + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * kBB) {
        const int32_t c = col[k];
        if (c < 0 || c >= n) bad = 1;
        if (k > 0 && col[k - 1] > c) ++desc;
    }
    desc = wave_sum(desc);
    const uint64_t any_bad = __ballot(bad != 0);
    if (lane == 0) {
        if (desc) atomicAdd(&flags[2], desc);
        if (any_bad) atomicOr(&flags[1], 1ull);
    }
}

struct KeySpec {
    int32_t n, sigma, maxlen;
    int ib, lb; // bits of the id-in-window and length fields
};

// 3a. first-column order (ALBBA_SELL_SORT_FIRST_COL): key (n - 1 - first
// column) << 32 | id, so that ascending keys are the reference's stable
// descending sort by first column (bfs/src/sorting_utils.c:255-281); an
// empty row takes the next stored entry, 0 past the last one
__global__ __launch_bounds__(kBB) void k_first_keys(const int64_t *rp, const int32_t *col, int32_t n, int64_t nnz,
                                                    unsigned long long *keys)
{
    for (int64_t v = (int64_t)blockIdx.x * kBB + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBB) {
        const int64_t at = rp[v];
        const int32_t first = at < nnz ? col[at] : 0;
        keys[v] = ((unsigned long long)(uint32_t)(n - 1 - first) << 32) | (unsigned long long)v;
    }
}

__global__ __launch_bounds__(kBB) void k_low_ids(const unsigned long long *keys, int32_t n, int32_t *pre)
{
    for (int64_t r = (int64_t)blockIdx.x * kBB + threadIdx.x; r < n; r += (int64_t)gridDim.x * kBB)
        pre[r] = (int32_t)(uint32_t)keys[r];
}

// 3. sort keys and their decoding.  The window sort runs over the rows in
// the order pre[] (the first-column order), or in id order when pre is null:
// rank r is row pre[r], and a key carries the rank's offset in its window.
__global__ __launch_bounds__(kBB) void k_keys(const int32_t *len, const int32_t *pre, KeySpec k,
                                              unsigned long long *keys)
{
    for (int64_t r = (int64_t)blockIdx.x * kBB + threadIdx.x; r < k.n; r += (int64_t)gridDim.x * kBB) {
        const unsigned long long w = (unsigned long long)(r / k.sigma);
        const unsigned long long off = (unsigned long long)(r - (int64_t)w * k.sigma);
        const int32_t v = pre ? pre[r] : (int32_t)r;
        const unsigned long long inv = (unsigned long long)(k.maxlen - len[v]);
        keys[r] = (w << (k.lb + k.ib)) | (inv << k.ib) | off;
    }
}

__global__ __launch_bounds__(kBB) void k_decode(const unsigned long long *keys, const int32_t *pre, KeySpec k,
                                                int64_t npos, int32_t *row_order, int32_t *row_len,
                                                int32_t *pos_of)
{
    const unsigned long long offmask = k.ib ? (~0ull >> (64 - k.ib)) : 0ull;
    const unsigned long long lenmask = k.lb ? (~0ull >> (64 - k.lb)) : 0ull;
    for (int64_t p = (int64_t)blockIdx.x * kBB + threadIdx.x; p < npos; p += (int64_t)gridDim.x * kBB) {
        if (p < k.n) {
            const unsigned long long key = keys[p];
            const int64_t w = (int64_t)(key >> (k.lb + k.ib));
            const int64_t r = w * k.sigma + (int64_t)(key & offmask);
            const int32_t v = pre ? pre[r] : (int32_t)r;
            row_order[p] = v;
            row_len[p] = k.maxlen - (int32_t)((key >> k.ib) & lenmask);
            pos_of[v] = (int32_t)p;
        } else {
            row_order[p] = -1;
            row_len[p] = 0;
        }
    }
}

// identity keys (one window whose rows all have length 0, or n == 1)
__global__ __launch_bounds__(kBB) void k_iota_keys(KeySpec k, unsigned long long *keys)
{
    for (int64_t v = (int64_t)blockIdx.x * kBB + threadIdx.x; v < k.n; v += (int64_t)gridDim.x * kBB)
        keys[v] = (unsigned long long)v;
}

// 4. widths (int32 for the layout, int64 for the scan; zero past the slices)
__global__ __launch_bounds__(kBB) void k_widths(const int32_t *row_len, int32_t C, int64_t ns, int64_t ns_alloc,
                                                int32_t *width, int64_t *w64)
{
    for (int64_t s = (int64_t)blockIdx.x * kBB + threadIdx.x; s <= ns_alloc; s += (int64_t)gridDim.x * kBB) {
        const int32_t w = s < ns ? row_len[s * C] : 0;
        if (s < ns_alloc) width[s] = w;
        w64[s] = w;
    }
}

// 5. active lanes per column and padded column sizes
__global__ __launch_bounds__(kBB) void k_columns(const int32_t *row_len, const int64_t *vop_ptr, int64_t ns,
                                                 int32_t C, int32_t pad, int64_t nvops, uint8_t *vop_len,
                                                 int64_t *csize, int32_t *slice_of)
{
    for (int64_t q = (int64_t)blockIdx.x * kBB + threadIdx.x; q <= nvops; q += (int64_t)gridDim.x * kBB) {
        if (q == nvops) {
            csize[q] = 0;
            continue;
        }
        // slice s: the last with vop_ptr[s] <= q (vop_ptr ascending, ns + 1 entries)
        int64_t lo = 0, hi = ns; // vop_ptr[lo] <= q < vop_ptr[hi]
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (vop_ptr[mid] <= q) lo = mid;
            else hi = mid;
        }
        const int64_t j = q - vop_ptr[lo];
        // active lanes: rows of the slice longer than j (lengths descending)
        const int32_t *L = row_len + lo * C;
        int a = 0, b = C; // L[a-1] > j (or a == 0), L[b] <= j (or b == C)
        while (a < b) {
            const int m = (a + b) >> 1;
            if (L[m] > j) a = m + 1;
            else b = m;
        }
        vop_len[q] = (uint8_t)(a - 1);
        csize[q] = pad_up(a, pad);
        slice_of[q] = (int32_t)lo;
    }
}

__global__ __launch_bounds__(kBB) void k_slice_ptr(const int64_t *vop_ptr, const int64_t *cstart, int64_t ns_alloc,
                                                   int64_t *slice_ptr)
{
    for (int64_t s = (int64_t)blockIdx.x * kBB + threadIdx.x; s <= ns_alloc; s += (int64_t)gridDim.x * kBB)
        slice_ptr[s] = cstart[vop_ptr[s]];
}

// rows longer than this are copied by the long-row pass, a block per row
constexpr int32_t kLongRow = 4096;

struct RowArgs {
    const int64_t *rp;
    const int32_t *col;
    const int32_t *row_order, *row_len, *pos_of;
    // with power-of-two windows of at most 65 536 rows: a vertex's position
    // is its window's first id plus a 16-bit offset, and the random gathers
    // read this half-size table instead of pos_of (null otherwise)
    const uint16_t *pos_off;
    int32_t win_mask;
    const int64_t *rm_ptr;  // row-major copy: pointers over positions
    int32_t *rm_col;
    // 64-row groups holding more than kHeavyGroup entries are listed here
    // (heavy_stat[0]: their count, [1]: the most entries of one) and copied
    // by k_rowmajor_heavy
    int32_t *heavy_list;
    unsigned long long *heavy_stat;
    int32_t n;
};

__device__ __forceinline__ int32_t position_of(const RowArgs &a, int32_t v)
{
    return a.pos_off ? (v & ~a.win_mask) | (int32_t)a.pos_off[v] : a.pos_of[v];
}

__device__ __forceinline__ void copy_entry(const RowArgs &a, int64_t src, int64_t dst)
{
    const int32_t c = position_of(a, a.col[src]);
    a.rm_col[dst] = c;
}

// 6a. the row-major copy (columns as positions).  A wave takes 64
// consecutive positions (a group) and spreads their entries evenly over its
// lanes, so that neighbouring lanes read neighbouring entries of one row (a
// lane walking a row of its own would have every lane of every resident wave
// on a different line).  Four steps of 64 entries are handled together so
// that their position gathers are in flight at once; the streamed reads of
// the CSR columns and the stores of the copy are non-temporal, which leaves
// the caches to the gathered position table.  Rows longer than kLongRow are
// left to k_rowmajor_long.  The first groups of every sigma window hold its
// hubs, so groups dealt to waves by position would leave a few waves most of
// the entries (RMAT s24, waves striding by their count: the busiest wave 41
// times the mean); groups of more than kHeavyGroup entries are listed instead
// and copied in chunks of kHeavyChunk entries by k_rowmajor_heavy, and each
// wave takes a run of consecutive groups, which spreads a window's heavier
// groups over several waves.
constexpr unsigned long long kHeavyGroup = 2048;
constexpr int kHeavyChunk = 4096; // per block: kHeavyChunk / 4 per wave

// entries [e_lo, e_hi) of a group: ex is the lane rows' exclusive prefix of
// their lengths, start and rm the lane row's first CSR and row-major entry
__device__ __forceinline__ void copy_group(const RowArgs &a, int lane, unsigned long long ex,
                                           unsigned long long e_lo, unsigned long long e_hi, int64_t start,
                                           int64_t rm)
{
    constexpr int U = 4;
    for (unsigned long long e0 = e_lo; e0 < e_hi; e0 += 64 * U) {
        int64_t src[U], dst[U];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned long long e = e0 + (unsigned long long)(u * 64 + lane);
            int lo = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const unsigned long long x = __shfl(ex, lo + step, 64);
                if (x <= e) lo += step;
            }
            const int64_t j = (int64_t)(e - __shfl(ex, lo, 64));
            src[u] = __shfl(start, lo, 64) + j;
            dst[u] = __shfl(rm, lo, 64) + j;
            act[u] = e < e_hi;
        }
        int32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = act[u] ? __builtin_nontemporal_load(&a.col[src[u]]) : 0;
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = act[u] ? position_of(a, c[u]) : 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!act[u]) continue;
            __builtin_nontemporal_store(c[u], &a.rm_col[dst[u]]);
        }
    }
}

// the group at positions p0 .. p0 + 63 as seen by one wave: its lane's row
// (rows longer than kLongRow count as empty), the prefix and the total
struct Group {
    int64_t start, rm;
    unsigned long long ex, tot;
};

__device__ __forceinline__ Group load_group(const RowArgs &a, int64_t p0, int lane)
{
    Group gr{0, 0, 0, 0};
    const int64_t p = p0 + lane;
    int32_t l = 0;
    if (p < a.n) {
        l = a.row_len[p];
        if (l > kLongRow) l = 0;
        gr.start = a.rp[a.row_order[p]];
        gr.rm = a.rm_ptr[p];
    }
    gr.ex = wave_exclusive_scan<unsigned long long>((unsigned long long)l, &gr.tot);
    return gr;
}

__global__ __launch_bounds__(kBB) void k_rowmajor(RowArgs a)
{
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    const int64_t ngroups = ((int64_t)a.n + 63) / 64;
    const int64_t per = (ngroups + nwaves - 1) / nwaves;
    const int64_t g1 = (wave + 1) * per < ngroups ? (wave + 1) * per : ngroups;
    for (int64_t g = wave * per; g < g1; ++g) {
        const int64_t p0 = g * 64;
        const Group gr = load_group(a, p0, lane);
        if (gr.tot > kHeavyGroup) {
            if (lane == 0) {
                a.heavy_list[atomicAdd(&a.heavy_stat[0], 1ull)] = (int32_t)(p0 >> 6);
                atomicMax(&a.heavy_stat[1], gr.tot);
            }
            continue;
        }
        copy_group(a, lane, gr.ex, 0, gr.tot, gr.start, gr.rm);
    }
}

// 6a''. the listed groups: block (x, y) copies entries [x * kHeavyChunk,
// (x + 1) * kHeavyChunk) of group list[y], a quarter per wave
__global__ __launch_bounds__(kBB) void k_rowmajor_heavy(RowArgs a, const int32_t *list)
{
    constexpr int kPerWave = kHeavyChunk / (kBB / 64);
    const int lane = lane_id();
    const Group gr = load_group(a, (int64_t)list[blockIdx.y] * 64, lane);
    const unsigned long long lo = (unsigned long long)blockIdx.x * kHeavyChunk + (threadIdx.x >> 6) * kPerWave;
    if (lo >= gr.tot) return;
    const unsigned long long hi = lo + kPerWave < gr.tot ? lo + kPerWave : gr.tot;
    copy_group(a, lane, gr.ex, lo, hi, gr.start, gr.rm);
}

// the in-window offsets of the vertices' positions (power-of-two windows)
__global__ __launch_bounds__(kBB) void k_pos_off(const int32_t *pos_of, int32_t n, int32_t win_mask, uint16_t *out)
{
    for (int64_t v = (int64_t)blockIdx.x * kBB + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBB)
        out[v] = (uint16_t)(pos_of[v] & win_mask);
}

// positions of the rows longer than kLongRow
__global__ __launch_bounds__(kBB) void k_collect_long(const int32_t *row_len, int32_t n, int32_t *list,
                                                      unsigned int *count)
{
    for (int64_t p = (int64_t)blockIdx.x * kBB + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBB)
        if (row_len[p] > kLongRow) list[atomicAdd(count, 1u)] = (int32_t)p;
}

// 6a'. the long rows: block (x, y) copies entries [x * kLongRow, +kLongRow)
// of long row y, kLongRow / kBB independent entries per thread
__global__ __launch_bounds__(kBB) void k_rowmajor_long(RowArgs a, const int32_t *list)
{
    const int32_t p = list[blockIdx.y];
    const int32_t l = a.row_len[p];
    const int64_t j0 = (int64_t)blockIdx.x * kLongRow;
    if (j0 >= l) return;
    const int64_t start = a.rp[a.row_order[p]], rm = a.rm_ptr[p];
#pragma unroll 4
    for (int k = 0; k < kLongRow / kBB; ++k) {
        const int64_t j = j0 + k * kBB + threadIdx.x;
        if (j < l) copy_entry(a, start + j, rm + j);
    }
}

// 6b. the columns from the row-major copy.  A wave takes 64 consecutive
// (slice, column) entries; their metadata are loaded once, one entry per
// lane, and broadcast; the lanes are the slice's rows, so the stores of a
// column are coalesced and each lane's loads walk its own row.
__device__ __forceinline__ int64_t readlane64(int64_t v, int i)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, i);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), i);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__global__ __launch_bounds__(kBB) void k_colmajor(const int64_t *rm_ptr, const int32_t *rm_col,
                                                  const int32_t *slice_of, const int64_t *vop_ptr,
                                                  const uint8_t *vop_len, const int64_t *cstart, int64_t nvops,
                                                  int32_t C, int32_t *colidx)
{
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    const int nt = C > 64 ? 2 : 1; // lanes cover C rows
    for (int64_t q0 = wave * 64; q0 < nvops; q0 += nwaves * 64) {
        const int64_t q = q0 + lane;
        int32_t sl = 0, act = 0;
        int64_t j = 0, base = 0;
        if (q < nvops) {
            sl = slice_of[q];
            j = q - vop_ptr[sl];
            act = (int32_t)vop_len[q] + 1;
            base = cstart[q];
        }
        const int cnt = (int)min((int64_t)64, nvops - q0);
        int32_t cur = -1;
        int64_t rmp0 = 0, rmp1 = 0;
        for (int i = 0; i < cnt; ++i) {
            const int32_t s_i = __builtin_amdgcn_readlane(sl, i);
            if (s_i != cur) {
                cur = s_i;
                const int64_t r0 = (int64_t)s_i * C + lane;
                rmp0 = lane < C ? rm_ptr[r0] : 0;
                if (nt == 2) rmp1 = rm_ptr[r0 + 64];
            }
            const int64_t j_i = readlane64(j, i), b_i = readlane64(base, i);
            const int32_t a_i = __builtin_amdgcn_readlane(act, i);
            if (lane < a_i) colidx[b_i + lane] = rm_col[rmp0 + j_i];
            if (nt == 2 && lane + 64 < a_i) colidx[b_i + 64 + lane] = rm_col[rmp1 + j_i];
        }
    }
}

// 6b'. C = 64: the same columns, 16 at a time through LDS.  Load k of a
// 16-column step reads rows 4k .. 4k + 3, 16 consecutive entries of each
// (lane l: row 4k + l / 16, column l % 16 of the step), so that one load
// touches a few lines where one per row (lanes as rows) touches 64; the
// values go to a 16 x 65 tile (the odd stride spreads a load's 16-lane row
// groups over the banks) and come back one column per store, lanes as rows.
// Lanes past a column's active rows read entries they do not store (clamped
// to the copy).
constexpr int kColStep = 16;

__global__ __launch_bounds__(kBB) void k_colmajor64(const int64_t *rm_ptr, const int32_t *rm_col,
                                                    const int32_t *slice_of, const int64_t *vop_ptr,
                                                    const uint8_t *vop_len, const int64_t *cstart, int64_t nvops,
                                                    int64_t rm_len, int32_t *colidx)
{
    __shared__ int32_t tile[kBB / 64][kColStep * 65];
    int32_t *t = tile[threadIdx.x >> 6];
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    const int64_t last = rm_len > 0 ? rm_len - 1 : 0;
    const int vl = lane & (kColStep - 1), rsub = lane >> 4;
    for (int64_t q0 = wave * 64; q0 < nvops; q0 += nwaves * 64) {
        const int64_t q = q0 + lane;
        int32_t sl = 0, act = 0;
        int64_t j = 0, base = 0;
        if (q < nvops) {
            sl = slice_of[q];
            j = q - vop_ptr[sl];
            act = (int32_t)vop_len[q] + 1;
            base = cstart[q];
        }
        const int cnt = (int)min((int64_t)64, nvops - q0);
        for (int i0 = 0; i0 < cnt; i0 += kColStep) {
            // this lane's column of the step (past the wave's last: that one
            // again, not stored)
            const int v = i0 + vl < cnt ? i0 + vl : cnt - 1;
            const int64_t s_v = __shfl(sl, v, 64);
            const int64_t j_v = __shfl(j, v, 64);
            int64_t rp[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) rp[k] = rm_ptr[s_v * 64 + 4 * k + rsub];
            int32_t val[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int64_t e = rp[k] + j_v;
                val[k] = rm_col[e < last ? e : last];
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) t[vl * 65 + 4 * k + rsub] = val[k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int k = 0; k < kColStep; ++k) {
                const int i = i0 + k;
                if (i >= cnt) break;
                const int64_t b_i = readlane64(base, i);
                const int32_t a_i = __builtin_amdgcn_readlane(act, i);
                if (lane < a_i) colidx[b_i + lane] = t[k * 65 + lane];
            }
            // the next step rewrites the tile
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

__global__ __launch_bounds__(kBB) void k_widen_u32(const unsigned int *in, int64_t count, int64_t *out)
{
    for (int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBB)
        out[i] = in[i];
}

// 7. a transpose by one sort (the directed push CSR over positions, and
// W = A^T of a loaded file): entry (r, c) becomes the key c << shift | r, a
// radix sort orders the keys by column and, within a column, by row, and one
// pass splits them into the transpose's columns and row pointers.  Atomic
// placement into per-column cursors serialised on the hubs' columns (RMAT
// s22: 20 ms for the fill of one directed build) and still needed a sort of
// every row.
__global__ __launch_bounds__(kBB) void k_transpose_keys(const int64_t *rp, const int32_t *col, int32_t n,
                                                        int shift, unsigned long long *keys)
{
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    for (int64_t p0 = wave * 64; p0 < n; p0 += nwaves * 64) {
        const int64_t p = p0 + lane;
        int64_t start = 0;
        int32_t l = 0;
        if (p < n) {
            start = rp[p];
            l = (int32_t)(rp[p + 1] - start);
        }
        wave_rows_foreach(l, [&](int i, int64_t j, bool act) {
            const int64_t s = __shfl(start, i, 64);
            if (act)
                keys[s + j] = ((unsigned long long)(uint32_t)col[s + j] << shift) | (unsigned long long)(p0 + i);
        });
    }
}

// the sorted keys' rows are the transpose's entries; the first key of each
// column sets the row pointers of that column and of the empty columns
// before it (the last key those after it)
__global__ __launch_bounds__(kBB) void k_transpose_split(const unsigned long long *sorted, int64_t nnz, int shift,
                                                         int32_t ncols, int64_t *tp, int32_t *tc)
{
    const unsigned long long mask = (1ull << shift) - 1;
    for (int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kBB) {
        const unsigned long long k = sorted[i];
        tc[i] = (int32_t)(k & mask);
        const int64_t cc = (int64_t)(k >> shift);
        const int64_t cp = i == 0 ? -1 : (int64_t)(sorted[i - 1] >> shift);
        for (int64_t x = cp + 1; x <= cc; ++x) tp[x] = i;
        if (i == nnz - 1)
            for (int64_t x = cc + 1; x <= ncols; ++x) tp[x] = nnz;
    }
}

// int32 row pointers widened to int64; flags[0] set unless they run from 0
// to nnz without decreasing
__global__ __launch_bounds__(kBB) void k_widen_rp(const int32_t *rp32, int32_t n, int64_t nnz, int64_t *rp64,
                                                  unsigned long long *flags)
{
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x; i <= n; i += (int64_t)gridDim.x * kBB) {
        const int32_t v = rp32[i];
        rp64[i] = v;
        if (i == 0 && v != 0) bad = true;
        if (i == n && (int64_t)v != nnz) bad = true;
        if (i < n && rp32[i + 1] < v) bad = true;
    }
    if (__ballot(bad) != 0 && lane_id() == 0) atomicOr(&flags[0], 1ull);
}

int bits_for(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

// device temporaries of one build, freed together
struct Temps {
    static constexpr int kMax = 32;
    albd_ctx *c;
    void *p[kMax];
    int np = 0;
    explicit Temps(albd_ctx *ctx) : c(ctx) {}
    ~Temps()
    {
        for (int i = 0; i < np; ++i) albd_free(c, p[i]);
    }
    template <class T> int get(size_t count, T **out)
    {
        if (np == kMax) return albd_set_error_msg(ALBBA_EINVAL, "graph_build: too many temporaries");
        void *q = nullptr;
        const int rc = albd_malloc(c, count * sizeof(T), &q);
        if (rc) return rc;
        p[np++] = q;
        *out = static_cast<T *>(q);
        return 0;
    }
};

#define ROCCHK(x)                                                          \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) return albd_set_error(e_, #x, __FILE__, __LINE__); \
    } while (0)

int scan_i64(albd_ctx *c, Temps &T, const int64_t *in, int64_t *out, size_t count)
{
    size_t bytes = 0;
    ROCCHK(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, count, rocprim::plus<int64_t>(),
                                   albd_stream(c)));
    void *tmp = nullptr;
    int rc = T.get<unsigned char>(bytes ? bytes : 1, reinterpret_cast<unsigned char **>(&tmp));
    if (rc) return rc;
    ROCCHK(rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, count, rocprim::plus<int64_t>(),
                                   albd_stream(c)));
    return 0;
}

// the transpose of an nrows x ncols CSR (rp, col) into (tp[ncols + 1], tc),
// each transposed row listing its entries' rows in ascending order
int transpose_by_sort(albd_ctx *c, Temps &T, const int64_t *rp, const int32_t *col, int32_t nrows, int32_t ncols,
                      int64_t nnz, int64_t *tp, int32_t *tc)
{
    hipStream_t s = albd_stream(c);
    if (nnz <= 0) {
        HIPCHK(hipMemsetAsync(tp, 0, sizeof(int64_t) * ((size_t)ncols + 1), s));
        return 0;
    }
    const int shift = bits_for((uint64_t)(nrows > 1 ? nrows - 1 : 0));
    int end_bit = shift + bits_for((uint64_t)(ncols > 1 ? ncols - 1 : 0));
    if (end_bit < 1) end_bit = 1;
    unsigned long long *keys = nullptr, *sorted = nullptr;
    int rc;
    if ((rc = T.get(nnz, &keys)) || (rc = T.get(nnz, &sorted))) return rc;
    hipLaunchKernelGGL(k_transpose_keys, dim3(grid_for(nrows, kBB)), dim3(kBB), 0, s, rp, col, nrows, shift, keys);
    HIPCHK(hipGetLastError());
    size_t bytes = 0;
    ROCCHK(rocprim::radix_sort_keys(nullptr, bytes, keys, sorted, (size_t)nnz, 0, end_bit, s));
    unsigned char *tmp = nullptr;
    if ((rc = T.get(bytes ? bytes : 1, &tmp))) return rc;
    ROCCHK(rocprim::radix_sort_keys(tmp, bytes, keys, sorted, (size_t)nnz, 0, end_bit, s));
    hipLaunchKernelGGL(k_transpose_split, dim3(grid_for(nnz, kBB)), dim3(kBB), 0, s, sorted, nnz, shift, ncols, tp,
                       tc);
    HIPCHK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" int albd_build_graph(albd_ctx *c, const albd_build_input *in, albd_graph *g, albd_build_alloc alloc,
                                void *alloc_arg, albd_build_report *rep)
{
    const int32_t n = in->n, C = in->C, pad = in->pad;
    const int64_t nnz = in->nnz;
    const int32_t sigma = in->sigma >= n ? n : in->sigma;
    const int64_t ns = ((int64_t)n + C - 1) / C;
    const int64_t npos_raw = ns * C;
    const int64_t nwords = (npos_raw + 63) / 64;
    const int64_t npos = nwords * 64;
    const int64_t ns_alloc = (npos + C - 1) / C;
    hipStream_t s = albd_stream(c);
    HIPCHK(hipSetDevice(albd_ctx_device(c)));
    Temps T(c);
    int rc;

    // the CSR on the device
    const int64_t *rp = nullptr;
    const int32_t *col = nullptr;
    int64_t *rp_up = nullptr;
    int32_t *col_up = nullptr;
    if (in->on_device) {
        rp = in->row_ptr;
        col = in->col;
    } else {
        if ((rc = T.get(n + 1, &rp_up)) || (rc = T.get(nnz > 0 ? nnz : 1, &col_up))) return rc;
        rp = rp_up;
        col = col_up;
    }
    hipEvent_t ev[4];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    struct EvFree {
        hipEvent_t *e;
        ~EvFree()
        {
            for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
        }
    } evfree{ev};
    HIPCHK(hipEventRecord(ev[0], s));
    if (!in->on_device) {
        HIPCHK(hipMemcpyAsync(rp_up, in->row_ptr, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, s));
        if (nnz > 0)
            HIPCHK(hipMemcpyAsync(col_up, in->col, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipEventRecord(ev[1], s));

    // 1.
    int32_t *len = nullptr;
    unsigned long long *flags = nullptr;
    if ((rc = T.get(n, &len)) || (rc = T.get(4, &flags))) return rc;
    HIPCHK(hipMemsetAsync(flags, 0, 4 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_entries, dim3(grid_for(nnz, kBB * 8)), dim3(kBB), 0, s, col, n, nnz, flags);
    HIPCHK(hipGetLastError());
    unsigned long long hflags[4];
    HIPCHK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // every column index is checked before any of them is used as an index
    if (hflags[1]) return albd_set_error_msg(ALBBA_EINVAL, "graph_build: column index outside [0, n)");
    hipLaunchKernelGGL(k_rows, dim3(grid_for(n, kBB * 4)), dim3(kBB), 0, s, rp, col, n, nnz, len, flags);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int32_t maxlen = (int32_t)hflags[0];
    const bool unsorted_rows = hflags[2] > hflags[3];
    rep->rows_sorted = unsorted_rows ? 0 : 1;

    // 2.
    if (unsorted_rows && nnz > 0) {
        int32_t *sorted = nullptr;
        if ((rc = T.get(nnz, &sorted))) return rc;
        size_t bytes = 0;
        ROCCHK(rocprim::segmented_radix_sort_keys(nullptr, bytes, col, sorted, (unsigned int)nnz, (unsigned int)n,
                                                  rp, rp + 1, 0, 32, s));
        unsigned char *tmp = nullptr;
        if ((rc = T.get(bytes ? bytes : 1, &tmp))) return rc;
        ROCCHK(rocprim::segmented_radix_sort_keys(tmp, bytes, col, sorted, (unsigned int)nnz, (unsigned int)n, rp,
                                                  rp + 1, 0, 32, s));
        col = sorted;
    }

    // 3.
    g->n = n;
    g->C = C;
    g->pad = pad;
    g->sparse_lanes = 8;
    // no window holds its own id range after the first-column order
    g->win = in->sort_first ? 0 : (sigma > 0 ? sigma : 1);
    g->nslices = (int32_t)ns;
    g->nwords = (int32_t)nwords;
    g->npos = npos;
    g->nnz = nnz;
    int32_t *row_order = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)npos);
    int32_t *row_len = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)npos);
    int32_t *pos_of = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)n);
    int32_t *width = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)ns_alloc);
    int64_t *vop_ptr = (int64_t *)alloc(alloc_arg, sizeof(int64_t) * (size_t)(ns_alloc + 1));
    int64_t *slice_ptr = (int64_t *)alloc(alloc_arg, sizeof(int64_t) * (size_t)(ns_alloc + 1));
    int64_t *rm_ptr = (int64_t *)alloc(alloc_arg, sizeof(int64_t) * (size_t)(npos + 1));
    int32_t *rm_col = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    g->row_order = row_order;
    g->row_len = row_len;
    g->pos_of = pos_of;
    g->slice_width = width;
    g->vop_ptr = vop_ptr;
    g->slice_ptr = slice_ptr;
    if (in->symmetric == 1) {
        g->push_ptr = g->pull_ptr = rm_ptr;
        g->push_col = g->pull_col = rm_col;
    } else {
        g->pull_ptr = rm_ptr;
        g->pull_col = rm_col;
    }
    if (!row_order || !row_len || !pos_of || !width || !vop_ptr || !slice_ptr || !rm_ptr || !rm_col)
        return ALBBA_EDEVMEM;

    KeySpec ks;
    ks.n = n;
    ks.sigma = sigma;
    ks.maxlen = maxlen;
    ks.ib = bits_for((uint64_t)(sigma - 1));
    ks.lb = bits_for((uint64_t)maxlen);
    const int wb = bits_for((uint64_t)((n - 1) / sigma));
    const int end_bit = wb + ks.lb + ks.ib;
    if (end_bit > 64) return albd_set_error_msg(ALBBA_ERANGE, "graph_build: sort key beyond 64 bits");
    unsigned long long *keys = nullptr, *sorted_keys = nullptr;
    if ((rc = T.get(n, &keys)) || (rc = T.get(n, &sorted_keys))) return rc;
    int32_t *pre = nullptr;
    if (in->sort_first && n > 1) {
        if ((rc = T.get(n, &pre))) return rc;
        hipLaunchKernelGGL(k_first_keys, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, rp, col, n, nnz, keys);
        HIPCHK(hipGetLastError());
        size_t bytes = 0;
        const int fb = 32 + bits_for((uint64_t)(n - 1));
        ROCCHK(rocprim::radix_sort_keys(nullptr, bytes, keys, sorted_keys, (size_t)n, 0, fb, s));
        unsigned char *tmp = nullptr;
        if ((rc = T.get(bytes ? bytes : 1, &tmp))) return rc;
        ROCCHK(rocprim::radix_sort_keys(tmp, bytes, keys, sorted_keys, (size_t)n, 0, fb, s));
        hipLaunchKernelGGL(k_low_ids, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, sorted_keys, n, pre);
        HIPCHK(hipGetLastError());
    }
    if (maxlen > 0 && n > 1) {
        hipLaunchKernelGGL(k_keys, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, len, pre, ks, keys);
        HIPCHK(hipGetLastError());
        size_t bytes = 0;
        ROCCHK(rocprim::radix_sort_keys(nullptr, bytes, keys, sorted_keys, (size_t)n, 0, end_bit, s));
        unsigned char *tmp = nullptr;
        if ((rc = T.get(bytes ? bytes : 1, &tmp))) return rc;
        ROCCHK(rocprim::radix_sort_keys(tmp, bytes, keys, sorted_keys, (size_t)n, 0, end_bit, s));
    } else {
        // every row empty (keys would all tie): identity order
        ks.lb = 0;
        ks.ib = 32;
        hipLaunchKernelGGL(k_iota_keys, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, ks, sorted_keys);
        HIPCHK(hipGetLastError());
        ks.sigma = 0x7fffffff;
    }
    hipLaunchKernelGGL(k_decode, dim3(grid_for(npos, kBB)), dim3(kBB), 0, s, sorted_keys, pre, ks, npos,
                       row_order, row_len, pos_of);
    HIPCHK(hipGetLastError());

    // 4.
    int64_t *w64 = nullptr;
    if ((rc = T.get(ns_alloc + 1, &w64))) return rc;
    hipLaunchKernelGGL(k_widths, dim3(grid_for(ns_alloc + 1, kBB)), dim3(kBB), 0, s, row_len, C, ns, ns_alloc,
                       width, w64);
    HIPCHK(hipGetLastError());
    if ((rc = scan_i64(c, T, w64, vop_ptr, (size_t)(ns_alloc + 1)))) return rc;
    int64_t nvops = 0;
    HIPCHK(hipMemcpyAsync(&nvops, vop_ptr + ns, sizeof nvops, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    g->nvops = nvops;

    // 5.
    uint8_t *vop_len = (uint8_t *)alloc(alloc_arg, (size_t)(nvops > 0 ? nvops : 1));
    g->vop_len = vop_len;
    if (!vop_len) return ALBBA_EDEVMEM;
    int64_t *csize = nullptr, *cstart = nullptr;
    int32_t *slice_of = nullptr;
    if ((rc = T.get(nvops + 1, &csize)) || (rc = T.get(nvops + 1, &cstart)) || (rc = T.get(nvops + 1, &slice_of)))
        return rc;
    hipLaunchKernelGGL(k_columns, dim3(grid_for(nvops + 1, kBB)), dim3(kBB), 0, s, row_len, vop_ptr, ns, C, pad,
                       nvops, vop_len, csize, slice_of);
    HIPCHK(hipGetLastError());
    if ((rc = scan_i64(c, T, csize, cstart, (size_t)(nvops + 1)))) return rc;
    hipLaunchKernelGGL(k_slice_ptr, dim3(grid_for(ns_alloc + 1, kBB)), dim3(kBB), 0, s, vop_ptr, cstart, ns_alloc,
                       slice_ptr);
    HIPCHK(hipGetLastError());
    int64_t padded = 0;
    HIPCHK(hipMemcpyAsync(&padded, cstart + nvops, sizeof padded, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (padded > INT32_MAX) return albd_set_error_msg(ALBBA_ERANGE, "graph_build: padded size beyond int32");
    g->padded_nnz = padded;

    // 6.
    int32_t *colidx = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)(padded > 0 ? padded : 1));
    g->colidx = colidx;
    if (!colidx) return ALBBA_EDEVMEM;
    if (pad > 1 && padded > 0 && (rc = albd_fill_u32(c, colidx, 0xFFFFFFFFu, (size_t)padded))) return rc;
    int64_t *len64 = nullptr;
    if ((rc = T.get(npos + 1, &len64))) return rc;
    // row-major pointers: scan of the position-order lengths
    {
        unsigned int *lu = reinterpret_cast<unsigned int *>(row_len);
        hipLaunchKernelGGL(k_widen_u32, dim3(grid_for(npos, kBB)), dim3(kBB), 0, s, lu, npos, len64);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemsetAsync(len64 + npos, 0, sizeof(int64_t), s));
        if ((rc = scan_i64(c, T, len64, rm_ptr, (size_t)(npos + 1)))) return rc;
    }
    RowArgs ra;
    ra.rp = rp;
    ra.col = col;
    ra.row_order = row_order;
    ra.row_len = row_len;
    ra.pos_of = pos_of;
    ra.pos_off = nullptr;
    ra.win_mask = 0;
    if (!in->sort_first && sigma >= 1 && sigma <= 65536 && (sigma & (sigma - 1)) == 0 && n > 0) {
        uint16_t *po = nullptr;
        if ((rc = T.get(n, &po))) return rc;
        hipLaunchKernelGGL(k_pos_off, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, pos_of, n, sigma - 1, po);
        HIPCHK(hipGetLastError());
        ra.pos_off = po;
        ra.win_mask = sigma - 1;
    }
    ra.rm_ptr = rm_ptr;
    ra.rm_col = rm_col;
    ra.n = n;
    const int64_t ngroups = ((int64_t)n + 63) / 64;
    if ((rc = T.get(ngroups > 0 ? ngroups : 1, &ra.heavy_list)) || (rc = T.get(2, &ra.heavy_stat))) return rc;
    HIPCHK(hipMemsetAsync(ra.heavy_stat, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_rowmajor, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, ra);
    HIPCHK(hipGetLastError());
    {
        unsigned long long hs[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(hs, ra.heavy_stat, sizeof hs, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const unsigned int chunks = (unsigned int)((hs[1] + kHeavyChunk - 1) / kHeavyChunk);
        // grid y is limited to 65535: take the groups in batches
        for (unsigned long long y0 = 0; y0 < hs[0]; y0 += 65535ull) {
            const unsigned int ny = (unsigned int)(hs[0] - y0 < 65535ull ? hs[0] - y0 : 65535ull);
            hipLaunchKernelGGL(k_rowmajor_heavy, dim3(chunks, ny), dim3(kBB), 0, s, ra, ra.heavy_list + y0);
            HIPCHK(hipGetLastError());
        }
    }
    if (maxlen > kLongRow) {
        int32_t *list = nullptr;
        unsigned int *count = nullptr;
        if ((rc = T.get(n, &list)) || (rc = T.get(1, &count))) return rc;
        HIPCHK(hipMemsetAsync(count, 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_collect_long, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, row_len, n, list, count);
        HIPCHK(hipGetLastError());
        unsigned int nlong = 0;
        HIPCHK(hipMemcpyAsync(&nlong, count, sizeof nlong, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const unsigned int chunks = (unsigned int)((maxlen + kLongRow - 1) / kLongRow);
        // grid y is limited to 65535: take the long rows in groups
        for (unsigned int y0 = 0; y0 < nlong; y0 += 65535u) {
            const unsigned int ny = nlong - y0 < 65535u ? nlong - y0 : 65535u;
            hipLaunchKernelGGL(k_rowmajor_long, dim3(chunks, ny), dim3(kBB), 0, s, ra, list + y0);
            HIPCHK(hipGetLastError());
        }
    }
    if (nvops > 0) {
        if (C == 64)
            hipLaunchKernelGGL(k_colmajor64, dim3(grid_for(nvops, kBB)), dim3(kBB), 0, s, rm_ptr, rm_col, slice_of,
                               vop_ptr, vop_len, cstart, nvops, nnz, colidx);
        else
            hipLaunchKernelGGL(k_colmajor, dim3(grid_for(nvops, kBB)), dim3(kBB), 0, s, rm_ptr, rm_col, slice_of,
                               vop_ptr, vop_len, cstart, nvops, C, colidx);
        HIPCHK(hipGetLastError());
    }

    // 7.
    if (in->symmetric != 1) {
        int64_t *push_ptr = (int64_t *)alloc(alloc_arg, sizeof(int64_t) * (size_t)(npos + 1));
        int32_t *push_col = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
        g->push_ptr = push_ptr;
        g->push_col = push_col;
        if (!push_ptr || !push_col) return ALBBA_EDEVMEM;
        if ((rc = transpose_by_sort(c, T, rm_ptr, rm_col, n, (int32_t)npos, nnz, push_ptr, push_col))) return rc;
    }
    HIPCHK(hipEventRecord(ev[2], s));
    HIPCHK(hipEventSynchronize(ev[2]));
    float up = 0, bu = 0;
    HIPCHK(hipEventElapsedTime(&up, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&bu, ev[1], ev[2]));
    rep->upload_ms = up;
    rep->build_ms = bu;
    return 0;
}

extern "C" int albd_transpose_csr(albd_ctx *c, const int32_t *rp32, const int32_t *col, int32_t n, int64_t nnz,
                                  int64_t **wp_out, int32_t **wc_out, int *bad)
{
    *wp_out = nullptr;
    *wc_out = nullptr;
    *bad = 0;
    hipStream_t s = albd_stream(c);
    HIPCHK(hipSetDevice(albd_ctx_device(c)));
    Temps T(c);
    int rc;
    int64_t *rp64 = nullptr;
    unsigned long long *flags = nullptr;
    if ((rc = T.get(n + 1, &rp64)) || (rc = T.get(4, &flags))) return rc;
    HIPCHK(hipMemsetAsync(flags, 0, 4 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_widen_rp, dim3(grid_for((int64_t)n + 1, kBB)), dim3(kBB), 0, s, rp32, n, nnz, rp64, flags);
    HIPCHK(hipGetLastError());
    if (nnz > 0) {
        hipLaunchKernelGGL(k_entries, dim3(grid_for(nnz, kBB * 8)), dim3(kBB), 0, s, col, n, nnz, flags);
        HIPCHK(hipGetLastError());
    }
    unsigned long long hflags[4];
    HIPCHK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hflags[0] || hflags[1]) {
        *bad = 1;
        return 0;
    }
    int64_t *wp = nullptr;
    int32_t *wc = nullptr;
    if ((rc = albd_malloc(c, sizeof(int64_t) * (size_t)(n + 1), (void **)&wp))) return rc;
    if ((rc = albd_malloc(c, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1), (void **)&wc))) {
        albd_free(c, wp);
        return rc;
    }
    struct Out {
        albd_ctx *c;
        int64_t *wp;
        int32_t *wc;
        bool keep = false;
        ~Out()
        {
            if (!keep) {
                albd_free(c, wp);
                albd_free(c, wc);
            }
        }
    } out{c, wp, wc};
    // W's row c lists the rows r of A's entries (r, c), ascending
    if ((rc = transpose_by_sort(c, T, rp64, col, n, n, nnz, wp, wc))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    out.keep = true;
    *wp_out = wp;
    *wc_out = wc;
    return 0;
}
