// Shared helpers of the HIP translation units (device layer only).
#pragma once
#include <hip/hip_runtime.h>

#include "../albd.h"

int albd_set_error(hipError_t e, const char *what, const char *file, int line);
int albd_set_error_msg(int code, const char *msg);
hipStream_t albd_stream(albd_ctx *c);

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

// One DPP move of a 32-bit value across lanes (CTRL = the DPP pattern):
// lanes whose source lies outside the pattern, or whose row is not in
// ROW_MASK, read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint64_t dpp_or_zero(uint64_t v)
{
    const uint32_t lo = dpp_or_zero<CTRL, ROW_MASK>((uint32_t)v);
    const uint32_t hi = dpp_or_zero<CTRL, ROW_MASK>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Sum over the 64 lanes of a wave, returned to every lane; call with the
// whole wave active.  DPP moves inside rows of 16 lanes (quad_perm 1,0,3,2;
// quad_perm 2,3,0,1; row_shr 4; row_shr 8 leave each row's total in its lane
// 15), then row_bcast 15 into rows 1 and 3 and row_bcast 31 into rows 2 and
// 3 leave the wave's total in lane 63.  A shuffle-based tree costs a
// dependent LDS round trip per step and per 32-bit half instead.
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit integers");
    using U = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
    U x = (U)v;
    x += dpp_or_zero<0xb1>(x);
    x += dpp_or_zero<0x4e>(x);
    x += dpp_or_zero<0x114>(x);
    x += dpp_or_zero<0x118>(x);
    x += dpp_or_zero<0x142, 0xa>(x);
    x += dpp_or_zero<0x143, 0xc>(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
    if constexpr (sizeof(T) == 8) {
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), 63);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)lo;
    }
}

// ---------------------------------------------------------------------------
// cross-workgroup communication inside one launch
//
// Data that one workgroup writes and another reads in a later phase goes
// through device-coherent loads/stores (agent-scope atomic accesses: they
// bypass the per-CU L1 and are coherent across the XCDs' L2s) or atomics;
// each wave waits for its own memory operations before a barrier.  No
// device-scope fences: on gfx950 a release at agent scope writes back the
// whole L2 (measured: 13.9 -> 21.2 us per level of the small-frontier
// kernel).  A workgroup past a barrier may already write the next phase's
// data while another still reads this phase's: counters that a phase reads
// after its barrier are per phase, never shared with the next one.

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

