/* The host objects that own device state, shared by the host sources that
 * drive the device (level loop, batched BFS). */
#ifndef ALBBA_DEVICE_INT_H
#define ALBBA_DEVICE_INT_H

#include "internal.h"
#include "../albd.h"

struct albba_ctx {
    albd_ctx *d;
    int device;
};

struct albba_graph {
    albba_ctx *ctx;
    albd_graph g;
    albd_state st;
    int32_t n;
    int32_t sigma;       /* the layout's sigma as requested */
    int32_t order_flags; /* ALBBA_SELL_* row-order flags of the layout */
    int64_t nnz;
    int32_t *pos_of;     /* host: vertex id -> position */
    int64_t device_bytes;
    void *scratch_visited; /* BEST-policy snapshots */
    void *scratch_ctr;
    int64_t *host_log;     /* small-frontier level sizes, 4 per level */
    int64_t *word_entries; /* host: stored entries per position word, made
                            * on the first BFS that accounts swept entries */
    int32_t *list_copy;    /* host: a pull's word list, for that accounting */
    int needs_reset;       /* a BFS failed midway: pmin/pkey may be dirty */
};

#endif
