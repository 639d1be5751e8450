// Level features and the direction choice, shared by the select kernel of
// the level loop and the batched BFS kernel.
#pragma once
#include "dev_common.h"
#include "../../../include/albba.h"

// The 12 features (all roughly in [0,1]) of the level about to run, with
// F = frontier vertices, mf = their out-edges, m = stored entries:
//   0 log2(1+F)/log2(1+n)   1 F/n   2 log2(1+mf)/log2(1+m)   3 mf/m
//   4 unvisited vertices/n  5 unvisited edges/m  6 nonzero frontier words/words
//   7 min(level,64)/64      8 log2(1+mf/F)/16    9 log2(1+m/n)/8
//   10 log2(1+F/F_prev)/16 (frontier growth over the last level; scale-free,
//      where log n would make graphs larger than the training set's an
//      extrapolation)             11 previous level pushed (1) or pulled (0)
//
// Called by all 64 lanes of one wave with the same arguments: the seven
// log2(1 + x) terms are computed one per lane (double precision, about a
// hundred instructions each), then lane 0 writes xs.
__device__ inline void level_features(float *xs, int64_t n_, int64_t nnz, int64_t nwords,
                                      unsigned long long F, unsigned long long mf,
                                      unsigned long long nzw, unsigned long long reached,
                                      unsigned long long redges, int32_t level, int32_t prev_choice,
                                      unsigned long long F_prev)
{
    const int lane = lane_id();
    const double n = (double)n_, m = nnz > 0 ? (double)nnz : 1.0;
    // x = num / den per lane (den = 1 keeps num exact):
    //   0 F   1 n   2 mf   3 m   4 mf/F   5 m/n   6 F/F_prev
    double num = 0.0, den = 1.0;
    switch (lane) {
    case 0: num = (double)F; break;
    case 1: num = n; break;
    case 2: num = (double)mf; break;
    case 3: num = m; break;
    case 4: num = (double)mf; den = (double)(F > 0 ? F : 1); break;
    case 5: num = m; den = n; break;
    case 6: num = (double)F; den = (double)(F_prev > 0 ? F_prev : 1); break;
    default: break;
    }
    const double y = log2(1.0 + num / den);
    const double lF = __shfl(y, 0), ln = __shfl(y, 1), lmf = __shfl(y, 2), lm = __shfl(y, 3),
                 ldeg = __shfl(y, 4), lmn = __shfl(y, 5), lgrow = __shfl(y, 6);
    if (lane == 0) {
        xs[0] = (float)(lF / ln);
        xs[1] = (float)((double)F / n);
        xs[2] = (float)(lmf / lm);
        xs[3] = (float)((double)mf / m);
        xs[4] = (float)((n - (double)reached) / n);
        xs[5] = (float)((m - (double)redges) / m);
        xs[6] = (float)((double)nzw / (double)(nwords > 0 ? nwords : 1));
        xs[7] = (float)((level < 64 ? level : 64) / 64.0);
        xs[8] = (float)(ldeg / 16.0);
        xs[9] = (float)(lmn / 8.0);
        xs[10] = (float)(lgrow / 16.0);
        xs[11] = prev_choice == ALBBA_CHOICE_PUSH ? 1.0f : 0.0f;
    }
}

// bfs/src/bfs_SELLCSv2.c:282-284: pull when F > (int)(n * 0.04)
__device__ inline int heuristic_choice(unsigned long long F, int64_t n)
{
    const long long thr = (long long)(int)((double)n * 0.04);
    return (long long)F > thr ? ALBBA_CHOICE_PULL_SB : ALBBA_CHOICE_PUSH;
}

// The MLP's parameters as the forward below reads them from LDS: each weight
// matrix transposed (W1T[12][32], W2T[32][32], W3T[32][4]) so that the
// threads computing neighbouring outputs read neighbouring words (the
// row-major W2[32][32] put all 32 threads on one LDS bank); biases as they
// are.  mlp_lds_index maps an index of the row-major parameter vector
// W1[32][12] b1[32] W2[32][32] b2[32] W3[4][32] b3[4] to its LDS slot.
__device__ inline int mlp_lds_index(int j)
{
    constexpr int I = ALBBA_MLP_IN, H1 = ALBBA_MLP_H1, H2 = ALBBA_MLP_H2, O = ALBBA_MLP_OUT;
    constexpr int oB1 = H1 * I, oW2 = oB1 + H1, oB2 = oW2 + H2 * H1, oW3 = oB2 + H2,
                  oB3 = oW3 + O * H2;
    if (j < oB1) return (j % I) * H1 + j / I;
    if (j >= oW2 && j < oB2) return oW2 + ((j - oW2) % H1) * H2 + (j - oW2) / H1;
    if (j >= oW3 && j < oB3) return oW3 + ((j - oW3) % H2) * O + (j - oW3) / H2;
    return j;
}

// Copy the parameters from global memory to LDS in the layout above, in
// two steps so that a kernel can issue other loads in between: mlp_load
// issues every load of this thread (NT threads take part), mlp_store writes
// them to LDS.
template <int NT> constexpr int mlp_per_thread() { return (ALBBA_MLP_NPARAM + NT - 1) / NT; }

template <int NT>
__device__ inline void mlp_load(float (&r)[mlp_per_thread<NT>()], const float *src)
{
#pragma unroll
    for (int k = 0; k < mlp_per_thread<NT>(); ++k) {
        const int j = (int)threadIdx.x + k * NT;
        r[k] = j < ALBBA_MLP_NPARAM ? src[j] : 0.f;
    }
}

template <int NT>
__device__ inline void mlp_store(float *dst, const float (&r)[mlp_per_thread<NT>()])
{
#pragma unroll
    for (int k = 0; k < mlp_per_thread<NT>(); ++k) {
        const int j = (int)threadIdx.x + k * NT;
        if (j < ALBBA_MLP_NPARAM) dst[mlp_lds_index(j)] = r[k];
    }
}

template <int NT>
__device__ inline void mlp_to_lds(float *dst, const float *src)
{
    float r[mlp_per_thread<NT>()];
    mlp_load<NT>(r, src);
    mlp_store<NT>(dst, r);
}

// The forward of one sample by all 64 lanes of one wave (no block barrier:
// LDS accesses of one wave complete in order), parameters in LDS as
// mlp_to_lds lays them out; the sums run in the same order as the training
// kernels' forward.  Returns the argmin over the choices in valid_mask to
// every lane, or -1 if none is valid (NaN outputs never win).
__device__ inline int mlp_argmin_wave(const float *params, const float *xs, float *h1, float *h2,
                                      float *ys, int valid_mask)
{
    constexpr int I = ALBBA_MLP_IN, H1 = ALBBA_MLP_H1, H2 = ALBBA_MLP_H2, O = ALBBA_MLP_OUT;
    static_assert(H1 <= 64 && H2 <= 64, "one lane per hidden unit");
    const int lane = lane_id();
    const float *W1T = params, *B1 = W1T + H1 * I, *W2T = B1 + H1, *B2 = W2T + H2 * H1,
                *W3T = B2 + H2, *B3 = W3T + O * H2;
    if (lane < H1) {
        float a = B1[lane];
#pragma unroll
        for (int i = 0; i < I; ++i) a = fmaf(W1T[i * H1 + lane], xs[i], a);
        h1[lane] = a > 0.f ? a : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < H2) {
        float a = B2[lane];
#pragma unroll
        for (int i = 0; i < H1; ++i) a = fmaf(W2T[i * H2 + lane], h1[i], a);
        h2[lane] = a > 0.f ? a : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < O) {
        float a = B3[lane];
#pragma unroll
        for (int i = 0; i < H2; ++i) a = fmaf(W3T[i * O + lane], h2[i], a);
        ys[lane] = a;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int choice = -1;
    float bv = 0.f;
#pragma unroll
    for (int k = 0; k < O; ++k) {
        const float y = ys[k];
        if (!((valid_mask >> k) & 1) || !(y == y)) continue;
        if (choice < 0 || y < bv) {
            choice = k;
            bv = y;
        }
    }
    return choice;
}

// MLP 12 -> 32 -> 32 -> 4 (ReLU) on the features in LDS (xs), computed by
// the first 32 threads of the block (every thread of the block must call:
// it contains barriers), with the parameters in LDS as mlp_to_lds lays them
// out.  The sums run in the same order as the training kernels' forward.
// Returns the argmin over the choices in valid_mask to every thread, or -1 if
// none is valid.
__device__ inline int mlp_argmin(const float *params, const float *xs, float *h1, float *h2,
                                 float *ys, int valid_mask, int *s_choice)
{
    constexpr int I = ALBBA_MLP_IN, H1 = ALBBA_MLP_H1, H2 = ALBBA_MLP_H2, O = ALBBA_MLP_OUT;
    const int tid = threadIdx.x;
    const float *W1T = params, *B1 = W1T + H1 * I, *W2T = B1 + H1, *B2 = W2T + H2 * H1,
                *W3T = B2 + H2, *B3 = W3T + O * H2;
    if (tid < H1) {
        float a = B1[tid];
#pragma unroll
        for (int i = 0; i < I; ++i) a = fmaf(W1T[i * H1 + tid], xs[i], a);
        h1[tid] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < H2) {
        float a = B2[tid];
#pragma unroll
        for (int i = 0; i < H1; ++i) a = fmaf(W2T[i * H2 + tid], h1[i], a);
        h2[tid] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < O) {
        float a = B3[tid];
#pragma unroll
        for (int i = 0; i < H2; ++i) a = fmaf(W3T[i * O + tid], h2[i], a);
        ys[tid] = a;
    }
    __syncthreads();
    if (tid == 0) {
        int choice = -1;
        float bv = 0.f;
        for (int k = 0; k < O; ++k) {
            if (!((valid_mask >> k) & 1) || !(ys[k] == ys[k])) continue; // NaN never wins
            if (choice < 0 || ys[k] < bv) {
                choice = k;
                bv = ys[k];
            }
        }
        *s_choice = choice;
    }
    __syncthreads();
    return *s_choice;
}
