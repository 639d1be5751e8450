/* Internal interface between the BFS driver and the learned selector. */
#ifndef ALBBA_SELECTOR_INT_H
#define ALBBA_SELECTOR_INT_H

#include "albba.h"
#include "../albd.h"

/* The 12 per-level features (all roughly in [0,1]):
 *   0 log2(1+F)/log2(1+n)      1 F/n
 *   2 log2(1+mf)/log2(1+m)     3 mf/m
 *   4 unvisited vertices / n   5 unvisited edges / m
 *   6 nonzero frontier words / words (bitmap density)
 *   7 min(level,64)/64         8 log2(1+mf/F)/16 (frontier mean degree)
 *   9 log2(1+m/n)/8            10 log2(n)/32
 *   11 previous level pushed (1) or pulled (0)
 * F = frontier vertices, mf = their out-edges, m = stored entries. */
void albba_features(int32_t n, int64_t m, int32_t nwords, int64_t frontier, int64_t frontier_edges,
                    int64_t nzwords, int64_t reached, int64_t reached_edges, int32_t level,
                    int prev_choice, float *out);

int albba_heuristic_choice(int32_t n, int64_t frontier);

/* Run the MLP on the device for one feature vector and return the argmin
 * choice (scratch: feat_dev >= 12 floats, choice_dev >= 1 int). */
int albba_selector_decide(const albba_selector *s, albd_ctx *d, const float *feats,
                          void *feat_dev, void *choice_dev, int *choice);

albd_ctx *albba_ctx_dev(albba_ctx *c);

#endif
