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
__device__ inline void level_features(float *xs, int64_t n_, int64_t nnz, int64_t nwords,
                                      unsigned long long F, unsigned long long mf,
                                      unsigned long long nzw, unsigned long long reached,
                                      unsigned long long redges, int32_t level, int32_t prev_choice,
                                      unsigned long long F_prev)
{
    const double n = (double)n_, m = nnz > 0 ? (double)nnz : 1.0;
    xs[0] = (float)(log2(1.0 + (double)F) / log2(1.0 + n));
    xs[1] = (float)((double)F / n);
    xs[2] = (float)(log2(1.0 + (double)mf) / log2(1.0 + m));
    xs[3] = (float)((double)mf / m);
    xs[4] = (float)((n - (double)reached) / n);
    xs[5] = (float)((m - (double)redges) / m);
    xs[6] = (float)((double)nzw / (double)(nwords > 0 ? nwords : 1));
    xs[7] = (float)((level < 64 ? level : 64) / 64.0);
    xs[8] = (float)(log2(1.0 + (double)mf / (double)(F > 0 ? F : 1)) / 16.0);
    xs[9] = (float)(log2(1.0 + m / n) / 8.0);
    xs[10] = (float)(log2(1.0 + (double)F / (double)(F_prev > 0 ? F_prev : 1)) / 16.0);
    xs[11] = prev_choice == ALBBA_CHOICE_PUSH ? 1.0f : 0.0f;
}

// bfs/src/bfs_SELLCSv2.c:282-284: pull when F > (int)(n * 0.04)
__device__ inline int heuristic_choice(unsigned long long F, int64_t n)
{
    const long long thr = (long long)(int)((double)n * 0.04);
    return (long long)F > thr ? ALBBA_CHOICE_PULL_SB : ALBBA_CHOICE_PUSH;
}

// MLP 12 -> 32 -> 32 -> 4 (ReLU) on the features in LDS (xs), computed by
// the first 32 threads of the block (every thread of the block must call:
// it contains barriers).  Returns the argmin over the choices in valid_mask
// to every thread, or -1 if none is valid.  Parameter layout:
// W1[32][12] b1[32] W2[32][32] b2[32] W3[4][32] b3[4].
__device__ inline int mlp_argmin(const float *params, const float *xs, float *h1, float *h2,
                                 float *ys, int valid_mask, int *s_choice)
{
    const int tid = threadIdx.x;
    const float *W1 = params, *B1 = W1 + 32 * 12, *W2 = B1 + 32, *B2 = W2 + 32 * 32,
                *W3 = B2 + 32, *B3 = W3 + 4 * 32;
    if (tid < ALBBA_MLP_H1) {
        float a = B1[tid];
        for (int i = 0; i < ALBBA_MLP_IN; ++i) a = fmaf(W1[tid * ALBBA_MLP_IN + i], xs[i], a);
        h1[tid] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < ALBBA_MLP_H2) {
        float a = B2[tid];
        for (int i = 0; i < ALBBA_MLP_H1; ++i) a = fmaf(W2[tid * ALBBA_MLP_H1 + i], h1[i], a);
        h2[tid] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < ALBBA_MLP_OUT) {
        float a = B3[tid];
        for (int i = 0; i < ALBBA_MLP_H2; ++i) a = fmaf(W3[tid * ALBBA_MLP_H2 + i], h2[i], a);
        ys[tid] = a;
    }
    __syncthreads();
    if (tid == 0) {
        int choice = -1;
        float bv = 0.f;
        for (int k = 0; k < ALBBA_MLP_OUT; ++k) {
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
