// Seeded synthetic graphs generated on the device: the same symmetric CSR,
// bit for bit, as the host generators (host/graphgen.c) -- rows ascending,
// both directions of every edge, self loops and duplicates removed, vertex
// ids permuted.  The reference has no generator (bfs/src/main.c:88-99 only
// loads files); the edge rules are this library's own, stated in
// host/graphgen.c, and are restated here from that statement:
//   RMAT  edge e draws one 64-bit value per two recursion levels from the
//         counter RNG, the high 32 bits for the even level, the low 32 bits for
//         the odd one; u = bits / 2^32 picks the quadrant against a, a+b, a+b+c;
//   ER    endpoints r(seed, 2e) mod n and r(seed, 2e+1) mod n;
//   grid  horizontal then vertical edges, removed when u01(r(seed, e)) < p.
// The permutation (a sequential Fisher-Yates) comes from the host.  On the
// device: one 64-bit key (u << bits) | v per direction (self loops become an
// all-ones key that sorts last), one radix sort, unique, then the row
// pointers by binary search.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "../../../include/albba.h"
#include "dev_common.h"

namespace {

constexpr int kGB = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rng_at(uint64_t seed, uint64_t idx)
{
    return mix64(seed + (idx + 1) * 0x9E3779B97F4A7C15ull);
}

__device__ __forceinline__ double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

int grid_for(int64_t items, int per_block)
{
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}

__global__ __launch_bounds__(kGB) void k_gen_keys(albd_gen_spec g, const int32_t *perm, unsigned long long *keys)
{
    const int draws = (g.scale + 1) / 2;
    for (int64_t e = (int64_t)blockIdx.x * kGB + threadIdx.x; e < g.ne; e += (int64_t)gridDim.x * kGB) {
        uint32_t u = 0, v = 0;
        if (g.kind == ALBBA_GEN_RMAT) {
            uint64_t r = 0;
            for (int lv = 0; lv < g.scale; ++lv) {
                if ((lv & 1) == 0) r = rng_at(g.seed, (uint64_t)e * (uint64_t)draws + (uint64_t)(lv >> 1));
                const uint32_t bits = (lv & 1) ? (uint32_t)r : (uint32_t)(r >> 32);
                const double x = (double)bits * (1.0 / 4294967296.0);
                uint32_t bu, bv;
                if (x < g.a) {
                    bu = 0;
                    bv = 0;
                } else if (x < g.ab) {
                    bu = 0;
                    bv = 1;
                } else if (x < g.abc) {
                    bu = 1;
                    bv = 0;
                } else {
                    bu = 1;
                    bv = 1;
                }
                u |= bu << (g.scale - 1 - lv);
                v |= bv << (g.scale - 1 - lv);
            }
        } else if (g.kind == ALBBA_GEN_ER) {
            u = (uint32_t)(rng_at(g.seed, 2 * (uint64_t)e) % (uint64_t)g.n);
            v = (uint32_t)(rng_at(g.seed, 2 * (uint64_t)e + 1) % (uint64_t)g.n);
        } else {
            if (e < g.H) {
                const int64_t i = e / (g.cols - 1), j = e % (g.cols - 1);
                u = (uint32_t)(i * g.cols + j);
                v = u + 1;
            } else {
                const int64_t k = e - g.H, i = k / g.cols, j = k % g.cols;
                u = (uint32_t)(i * g.cols + j);
                v = u + (uint32_t)g.cols;
            }
            if (u01(rng_at(g.seed, (uint64_t)e)) < g.p_remove) v = u; // removed: a self loop
        }
        const unsigned long long pu = (unsigned long long)(uint32_t)perm[u];
        const unsigned long long pv = (unsigned long long)(uint32_t)perm[v];
        if (pu == pv) {
            keys[2 * e] = keys[2 * e + 1] = ~0ull;
        } else {
            keys[2 * e] = (pu << g.bits) | pv;
            keys[2 * e + 1] = (pv << g.bits) | pu;
        }
    }
}

// columns of the unique keys, and the row pointers: row r starts at the
// first key >= r << bits
__global__ __launch_bounds__(kGB) void k_csr(const unsigned long long *keys, int64_t m, int bits, int32_t n,
                                             int64_t *rp, int32_t *col)
{
    const unsigned long long mask = (1ull << bits) - 1;
    const int64_t stride = (int64_t)gridDim.x * kGB;
    for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i < m; i += stride)
        col[i] = (int32_t)(keys[i] & mask);
    for (int64_t r = (int64_t)blockIdx.x * kGB + threadIdx.x; r <= n; r += stride) {
        const unsigned long long t = (unsigned long long)r << bits;
        int64_t lo = 0, hi = m; // first index with keys[idx] >= t
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        rp[r] = lo;
    }
}

#define ROCCHK(x)                                                          \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) return albd_set_error(e_, #x, __FILE__, __LINE__); \
    } while (0)

// device buffers freed on every return path
struct Bufs {
    albd_ctx *c;
    void *p[8] = {};
    ~Bufs()
    {
        for (void *q : p) albd_free(c, q);
    }
};

} // namespace

extern "C" int albd_gen_csr(albd_ctx *c, const albd_gen_spec *spec, const int32_t *perm_host, int64_t **rp_out,
                            int32_t **col_out, int64_t *m_out)
{
    *rp_out = nullptr;
    *col_out = nullptr;
    *m_out = 0;
    const albd_gen_spec g = *spec;
    const int64_t nkeys = 2 * g.ne;
    hipStream_t s = albd_stream(c);
    HIPCHK(hipSetDevice(albd_ctx_device(c)));
    Bufs B{c};
    int rc;
    int32_t *perm = nullptr;
    unsigned long long *keys = nullptr, *sorted = nullptr, *count = nullptr;
    if ((rc = albd_malloc(c, sizeof(int32_t) * (size_t)g.n, (void **)&perm))) return rc;
    B.p[0] = perm;
    if ((rc = albd_malloc(c, sizeof(unsigned long long) * (size_t)(nkeys > 0 ? nkeys : 1), (void **)&keys)))
        return rc;
    B.p[1] = keys;
    if ((rc = albd_malloc(c, sizeof(unsigned long long) * (size_t)(nkeys > 0 ? nkeys : 1), (void **)&sorted)))
        return rc;
    B.p[2] = sorted;
    if ((rc = albd_malloc(c, sizeof(unsigned long long), (void **)&count))) return rc;
    B.p[3] = count;
    HIPCHK(hipMemcpyAsync(perm, perm_host, sizeof(int32_t) * (size_t)g.n, hipMemcpyHostToDevice, s));
    int64_t m = 0;
    if (nkeys > 0) {
        hipLaunchKernelGGL(k_gen_keys, dim3(grid_for(g.ne, kGB)), dim3(kGB), 0, s, g, perm, keys);
        HIPCHK(hipGetLastError());
        // sort on the key bits; the all-ones self-loop keys sort last either way
        size_t bytes = 0;
        ROCCHK(rocprim::radix_sort_keys(nullptr, bytes, keys, sorted, (size_t)nkeys, 0, 2 * g.bits, s));
        void *tmp = nullptr;
        if ((rc = albd_malloc(c, bytes ? bytes : 1, &tmp))) return rc;
        B.p[4] = tmp;
        ROCCHK(rocprim::radix_sort_keys(tmp, bytes, keys, sorted, (size_t)nkeys, 0, 2 * g.bits, s));
        // unique keys back into `keys`
        size_t ubytes = 0;
        ROCCHK(rocprim::unique(nullptr, ubytes, sorted, keys, count, (size_t)nkeys,
                               rocprim::equal_to<unsigned long long>(), s));
        void *utmp = nullptr;
        if ((rc = albd_malloc(c, ubytes ? ubytes : 1, &utmp))) return rc;
        B.p[5] = utmp;
        ROCCHK(rocprim::unique(utmp, ubytes, sorted, keys, count, (size_t)nkeys,
                               rocprim::equal_to<unsigned long long>(), s));
        unsigned long long nu = 0, last = 0;
        HIPCHK(hipMemcpyAsync(&nu, count, sizeof nu, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (nu > 0) {
            HIPCHK(hipMemcpyAsync(&last, keys + nu - 1, sizeof last, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        m = (int64_t)nu - (nu > 0 && last == ~0ull ? 1 : 0);
    }
    if (m > INT32_MAX) return albd_set_error_msg(ALBBA_ERANGE, "gen: nnz beyond int32");
    int64_t *rp = nullptr;
    int32_t *col = nullptr;
    if ((rc = albd_malloc(c, sizeof(int64_t) * (size_t)(g.n + 1), (void **)&rp))) return rc;
    if ((rc = albd_malloc(c, sizeof(int32_t) * (size_t)(m > 0 ? m : 1), (void **)&col))) {
        albd_free(c, rp);
        return rc;
    }
    hipLaunchKernelGGL(k_csr, dim3(grid_for((m > g.n ? m : (int64_t)g.n) + 1, kGB)), dim3(kGB), 0, s, keys, m,
                       g.bits, g.n, rp, col);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        albd_free(c, rp);
        albd_free(c, col);
        return albd_set_error(e, "gen: row pointers", __FILE__, __LINE__);
    }
    *rp_out = rp;
    *col_out = col;
    *m_out = m;
    return 0;
}
