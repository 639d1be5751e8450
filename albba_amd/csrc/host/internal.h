/* Internal helpers shared by the host C sources of libalbba. */
#ifndef ALBBA_INTERNAL_H
#define ALBBA_INTERNAL_H

#include "albba.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ALBBA_HIDDEN __attribute__((visibility("hidden")))

/* Record a formatted message as the calling thread's last error and return
 * `code` so call sites can write `return albba_fail(ALBBA_EINVAL, ...)`. */
ALBBA_HIDDEN int albba_fail(int code, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));

/* Zero-initialised allocation that never returns NULL for n == 0. */
ALBBA_HIDDEN void *albba_calloc(size_t count, size_t size);
ALBBA_HIDDEN void *albba_malloc(size_t bytes);

/* Monotonic wall clock in milliseconds. */
ALBBA_HIDDEN double albba_now_ms(void);

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
static inline uint64_t albba_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t albba_rng_at(uint64_t seed, uint64_t idx)
{
    return albba_mix64(seed + (idx + 1) * 0x9E3779B97F4A7C15ull);
}

/* The id permutation of every generator (Fisher-Yates with perm_seed);
 * malloc'ed, NULL when out of memory. */
ALBBA_HIDDEN int32_t *albba_make_perm(int32_t n, uint64_t perm_seed);

/* Uniform double in [0,1) from the top 53 bits. */
static inline double albba_u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

/* Stable descending sort of (key,id) pairs in [lo,hi) by key; equal keys keep
 * ascending id.  Used for the sigma-window row ordering. */
ALBBA_HIDDEN void albba_sort_window_desc(int32_t *key, int32_t *id, int32_t lo, int32_t hi,
                                         int32_t *tmp_key, int32_t *tmp_id);

/* Build the sigma-sorted position -> row order for row lengths len[0..n).
 * Rows are stably sorted by length, descending, inside windows of sigma rows
 * (bfs/src/sorting_utils.c:232-253).  order[p] = row at position p and
 * sorted_len[p] = its length. */
ALBBA_HIDDEN int albba_sigma_order(const int32_t *len, int32_t n, int32_t sigma,
                                   int32_t *order, int32_t *sorted_len);

#endif
