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
