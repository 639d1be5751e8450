// Shared helpers of the HIP translation units (device layer only).
#pragma once
#include <hip/hip_runtime.h>

#include "../albd.h"

int albd_set_error(hipError_t e, const char *what, const char *file, int line);
int albd_set_error_msg(int code, const char *msg);
hipStream_t albd_stream(albd_ctx *c);
// cooperative launches failed once on this context: do not try again
int albd_coop_off(const albd_ctx *c);
void albd_set_coop_off(albd_ctx *c);

#define HIPCHK(x)                                                          \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) return albd_set_error(e_, #x, __FILE__, __LINE__); \
    } while (0)

// (level, parent vertex id) of one position, as one 8-byte store
__device__ __forceinline__ unsigned long long pack_lp(int32_t level, int32_t parent)
{
    return ((unsigned long long)(uint32_t)level << 32) | (unsigned long long)(uint32_t)parent;
}

// wave64 helpers (CDNA wavefronts are 64 lanes wide)
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ uint64_t lanemask_lt(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// exclusive prefix sum across the 64 lanes of a wave (Hillis-Steele on shfl)
template <typename T>
__device__ __forceinline__ T wave_exclusive_scan(T v, T *total)
{
    const int lane = lane_id();
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    *total = __shfl(x, 63, 64);
    return x - v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// ---------------------------------------------------------------------------
// cross-workgroup communication inside one launch
//
// Data that one workgroup writes and another reads in a later phase goes
// through device-coherent loads/stores (they bypass the per-CU L1 and are
// coherent across the XCDs' L2s) or atomics; each wave waits for its own
// memory operations before a barrier.  No device-scope fences: on gfx950 a
// release at agent scope writes back the whole L2.

template <class T> __device__ __forceinline__ T ld_coherent(const T *p)
{
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T> __device__ __forceinline__ void st_coherent(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// append one entry per predicated lane to an LDS list: returns the slot of
// this lane (valid where pred holds); works under any exec mask
__device__ __forceinline__ unsigned int wave_append(unsigned int *counter, bool pred)
{
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0;
    const int lane = lane_id();
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned int)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + (unsigned int)__popcll(m & lanemask_lt(lane));
}

// exclusive scan of one value per thread over a BLOCK-thread workgroup;
// lds holds BLOCK/64 entries and is free again after the next barrier
template <int BLOCK>
__device__ __forceinline__ int64_t block_scan_excl(int64_t v, int64_t *total, int64_t *lds)
{
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    int64_t wt;
    const int64_t ex = wave_exclusive_scan<int64_t>(v, &wt);
    if (lane == 0) lds[wid] = wt;
    __syncthreads();
    int64_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i) {
        if (i < wid) off += lds[i];
        tot += lds[i];
    }
    *total = tot;
    return ex + off;
}

// Barrier over `members` co-resident workgroups sharing a counter that only
// grows (zeroed before the launch).  Wave 0 of each workgroup arrives and
// polls.  Gives up after spin_limit polls or once *error is set, so every
// wave always reaches the end of the kernel; returns false then.
__device__ __forceinline__ bool team_sync(unsigned long long *counter, unsigned int *error,
                                          unsigned int members, unsigned int spin_limit, int *s_ok)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned long long arrived = atomicAdd(counter, 1ull);
        const unsigned long long target = (arrived / members + 1) * members;
        unsigned int spins = 0;
        while (ld_coherent(counter) < target) {
            if (++spins > spin_limit || ld_coherent(error) != 0) {
                atomicOr(error, 1u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

