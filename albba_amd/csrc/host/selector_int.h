/* Internal interface between the BFS driver and the learned selector. */
#ifndef ALBBA_SELECTOR_INT_H
#define ALBBA_SELECTOR_INT_H

#include "albba.h"
#include "../albd.h"

/* Device copy of the selector's parameters (read by the select kernel that
 * ends every BFS level; the features it computes are listed there). */
const float *albba_selector_dparams(const albba_selector *s);

albd_ctx *albba_ctx_dev(albba_ctx *c);

#endif
