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
//   6. the columns: a wave takes 64 consecutive positions and spreads their
//      entries evenly over its lanes; the row-major copy (push CSR when the
//      matrix is symmetric, pull CSR otherwise) is written in the same pass;
//   7. directed graphs: the push CSR (transpose over positions) by a degree
//      histogram, a scan, an atomic fill and a per-row sort.
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

// 1. lengths, longest row, checks
__global__ __launch_bounds__(kBB) void k_lengths(const int64_t *rp, const int32_t *col, int32_t n,
                                                 int32_t *len, unsigned int *flags)
{
    // flags[0]: longest row; flags[1]: a row out of ascending order;
    // flags[2]: a column index outside [0, n)
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    unsigned int mx = 0, unsorted = 0, bad = 0;
    for (int64_t r0 = wave * 64; r0 < n; r0 += nwaves * 64) {
        const int64_t r = r0 + lane;
        int64_t start = 0;
        int32_t l = 0;
        if (r < n) {
            start = rp[r];
            l = (int32_t)(rp[r + 1] - start);
            len[r] = l;
            mx = max(mx, (unsigned int)l);
        }
        wave_rows_foreach(l, [&](int i, int64_t j, bool act) {
            const int64_t s = __shfl(start, i, 64);
            if (act) {
                const int32_t c = col[s + j];
                if (c < 0 || c >= n) bad = 1;
                if (j > 0 && col[s + j - 1] > c) unsorted = 1;
            }
        });
    }
    const uint64_t any_unsorted = __ballot(unsorted != 0), any_bad = __ballot(bad != 0);
    // wave maximum by butterfly
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) mx = max(mx, (unsigned int)__shfl_xor(mx, d, 64));
    if (lane == 0) {
        if (mx) atomicMax(&flags[0], mx);
        if (any_unsorted) atomicOr(&flags[1], 1u);
        if (any_bad) atomicOr(&flags[2], 1u);
    }
}

struct KeySpec {
    int32_t n, sigma, maxlen;
    int ib, lb; // bits of the id-in-window and length fields
};

// 3. sort keys and their decoding
__global__ __launch_bounds__(kBB) void k_keys(const int32_t *len, KeySpec k, unsigned long long *keys)
{
    for (int64_t v = (int64_t)blockIdx.x * kBB + threadIdx.x; v < k.n; v += (int64_t)gridDim.x * kBB) {
        const unsigned long long w = (unsigned long long)(v / k.sigma);
        const unsigned long long off = (unsigned long long)(v - (int64_t)w * k.sigma);
        const unsigned long long inv = (unsigned long long)(k.maxlen - len[v]);
        keys[v] = (w << (k.lb + k.ib)) | (inv << k.ib) | off;
    }
}

__global__ __launch_bounds__(kBB) void k_decode(const unsigned long long *keys, KeySpec k, int64_t npos,
                                                int32_t *row_order, int32_t *row_len, int32_t *pos_of)
{
    const unsigned long long offmask = k.ib ? (~0ull >> (64 - k.ib)) : 0ull;
    const unsigned long long lenmask = k.lb ? (~0ull >> (64 - k.lb)) : 0ull;
    for (int64_t p = (int64_t)blockIdx.x * kBB + threadIdx.x; p < npos; p += (int64_t)gridDim.x * kBB) {
        if (p < k.n) {
            const unsigned long long key = keys[p];
            const int64_t w = (int64_t)(key >> (k.lb + k.ib));
            const int32_t v = (int32_t)(w * k.sigma + (int64_t)(key & offmask));
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
                                                 int64_t *csize)
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
    }
}

__global__ __launch_bounds__(kBB) void k_slice_ptr(const int64_t *vop_ptr, const int64_t *cstart, int64_t ns_alloc,
                                                   int64_t *slice_ptr)
{
    for (int64_t s = (int64_t)blockIdx.x * kBB + threadIdx.x; s <= ns_alloc; s += (int64_t)gridDim.x * kBB)
        slice_ptr[s] = cstart[vop_ptr[s]];
}

struct FillArgs {
    const int64_t *rp;
    const int32_t *col;
    const int32_t *row_order, *row_len, *pos_of;
    const int64_t *vop_ptr, *cstart;
    const int64_t *rm_ptr; // row-major copy: pointers over positions
    int32_t *colidx;
    int32_t *rm_col;
    unsigned int *push_cnt; // directed: out-degree histogram over positions, or null
    int32_t n;
    int32_t C;
    int32_t log2C;
};

// 6. columns and the row-major copy
__global__ __launch_bounds__(kBB) void k_fill(FillArgs a)
{
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    for (int64_t p0 = wave * 64; p0 < a.n; p0 += nwaves * 64) {
        const int64_t p = p0 + lane;
        int64_t start = 0, vp = 0, rm = 0;
        int32_t l = 0;
        if (p < a.n) {
            start = a.rp[a.row_order[p]];
            l = a.row_len[p];
            vp = a.vop_ptr[p >> a.log2C];
            rm = a.rm_ptr[p];
        }
        wave_rows_foreach(l, [&](int i, int64_t j, bool act) {
            const int64_t s = __shfl(start, i, 64), v0 = __shfl(vp, i, 64), r0 = __shfl(rm, i, 64);
            if (act) {
                const int64_t pp = p0 + i;
                const int32_t c = a.pos_of[a.col[s + j]];
                a.colidx[a.cstart[v0 + j] + (pp & (a.C - 1))] = c;
                a.rm_col[r0 + j] = c;
                if (a.push_cnt) atomicAdd(&a.push_cnt[c], 1u);
            }
        });
    }
}

__global__ __launch_bounds__(kBB) void k_widen_u32(const unsigned int *in, int64_t count, int64_t *out)
{
    for (int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBB)
        out[i] = in[i];
}

// 7. directed: the transpose over positions; entries land in any order and
// are sorted per row afterwards
__global__ __launch_bounds__(kBB) void k_push_fill(const int64_t *rm_ptr, const int32_t *rm_col, int32_t n,
                                                   const int64_t *push_ptr, unsigned int *cursor,
                                                   int32_t *push_col)
{
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kBB + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBB) >> 6;
    for (int64_t p0 = wave * 64; p0 < n; p0 += nwaves * 64) {
        const int64_t p = p0 + lane;
        int64_t start = 0;
        int32_t l = 0;
        if (p < n) {
            start = rm_ptr[p];
            l = (int32_t)(rm_ptr[p + 1] - start);
        }
        wave_rows_foreach(l, [&](int i, int64_t j, bool act) {
            const int64_t s = __shfl(start, i, 64);
            if (act) {
                const int32_t q = rm_col[s + j];
                const unsigned int at = atomicAdd(&cursor[q], 1u);
                push_col[push_ptr[q] + at] = (int32_t)(p0 + i);
            }
        });
    }
}

int ilog2_exact(int x)
{
    int b = 0;
    while ((1 << b) < x) ++b;
    return b;
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
    int64_t *rp = nullptr;
    int32_t *col = nullptr;
    if ((rc = T.get(n + 1, &rp)) || (rc = T.get(nnz > 0 ? nnz : 1, &col))) return rc;
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
    HIPCHK(hipMemcpyAsync(rp, in->row_ptr, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, s));
    if (nnz > 0) HIPCHK(hipMemcpyAsync(col, in->col, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(ev[1], s));

    // 1.
    int32_t *len = nullptr;
    unsigned int *flags = nullptr;
    if ((rc = T.get(n, &len)) || (rc = T.get(4, &flags))) return rc;
    HIPCHK(hipMemsetAsync(flags, 0, 4 * sizeof(unsigned int), s));
    hipLaunchKernelGGL(k_lengths, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, rp, col, n, len, flags);
    HIPCHK(hipGetLastError());
    unsigned int hflags[4];
    HIPCHK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hflags[2]) return albd_set_error_msg(ALBBA_EINVAL, "graph_build: column index outside [0, n)");
    const int32_t maxlen = (int32_t)hflags[0];
    rep->rows_sorted = hflags[1] ? 0 : 1;

    // 2.
    if (hflags[1] && nnz > 0) {
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
    if (maxlen > 0 && n > 1) {
        hipLaunchKernelGGL(k_keys, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, len, ks, keys);
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
    hipLaunchKernelGGL(k_decode, dim3(grid_for(npos, kBB)), dim3(kBB), 0, s, sorted_keys, ks, npos, row_order,
                       row_len, pos_of);
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
    if ((rc = T.get(nvops + 1, &csize)) || (rc = T.get(nvops + 1, &cstart))) return rc;
    hipLaunchKernelGGL(k_columns, dim3(grid_for(nvops + 1, kBB)), dim3(kBB), 0, s, row_len, vop_ptr, ns, C, pad,
                       nvops, vop_len, csize);
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
    unsigned int *cnt = nullptr;
    if (in->symmetric != 1) {
        if ((rc = T.get(npos + 1, &cnt))) return rc;
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned int) * (size_t)(npos + 1), s));
    }
    FillArgs fa;
    fa.rp = rp;
    fa.col = col;
    fa.row_order = row_order;
    fa.row_len = row_len;
    fa.pos_of = pos_of;
    fa.vop_ptr = vop_ptr;
    fa.cstart = cstart;
    fa.rm_ptr = rm_ptr;
    fa.colidx = colidx;
    fa.rm_col = rm_col;
    fa.push_cnt = cnt;
    fa.n = n;
    fa.C = C;
    fa.log2C = ilog2_exact(C);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, fa);
    HIPCHK(hipGetLastError());

    // 7.
    if (in->symmetric != 1) {
        int64_t *push_ptr = (int64_t *)alloc(alloc_arg, sizeof(int64_t) * (size_t)(npos + 1));
        int32_t *push_col = (int32_t *)alloc(alloc_arg, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
        g->push_ptr = push_ptr;
        g->push_col = push_col;
        if (!push_ptr || !push_col) return ALBBA_EDEVMEM;
        hipLaunchKernelGGL(k_widen_u32, dim3(grid_for(npos + 1, kBB)), dim3(kBB), 0, s, cnt, npos + 1, len64);
        HIPCHK(hipGetLastError());
        if ((rc = scan_i64(c, T, len64, push_ptr, (size_t)(npos + 1)))) return rc;
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned int) * (size_t)(npos + 1), s));
        int32_t *unsorted = nullptr;
        if ((rc = T.get(nnz > 0 ? nnz : 1, &unsorted))) return rc;
        hipLaunchKernelGGL(k_push_fill, dim3(grid_for(n, kBB)), dim3(kBB), 0, s, rm_ptr, rm_col, n, push_ptr, cnt,
                           unsorted);
        HIPCHK(hipGetLastError());
        if (nnz > 0) {
            size_t bytes = 0;
            ROCCHK(rocprim::segmented_radix_sort_keys(nullptr, bytes, unsorted, push_col, (unsigned int)nnz,
                                                      (unsigned int)npos, push_ptr, push_ptr + 1, 0, 32, s));
            unsigned char *tmp = nullptr;
            if ((rc = T.get(bytes ? bytes : 1, &tmp))) return rc;
            ROCCHK(rocprim::segmented_radix_sort_keys(tmp, bytes, unsorted, push_col, (unsigned int)nnz,
                                                      (unsigned int)npos, push_ptr, push_ptr + 1, 0, 32, s));
        }
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
