/*
 * Learned direction selector (host side).  The reference decides push vs
 * pull with a fixed vertex-count threshold (bfs/src/bfs_SELLCSv2.c:282-290);
 * here a 12 -> 32 -> 32 -> 4 MLP predicts, per level, the relative log time
 * of each kernel choice and the level runs the argmin.  Forward, backward
 * and Adam run on the GPU (mlp.hip); this file owns the parameters, the
 * feature vector, the training loop and the weight file format.
 *
 * Weight file: a first line "albba-selector 2 1604", one parameter per line
 * printed with %.9g (exact float32 round trip), then "queue_ratio <r>" (the
 * push levels' queue-listing bound, frontier out-edges per vertex; see
 * albba_selector_set_queue_ratio).  Version 1 files end after the parameters
 * and leave the ratio unset.
 */
#include "internal.h"
#include "selector_int.h"

#include <math.h>

struct albba_selector {
    albba_ctx *ctx;
    float *dparams; /* device copy */
    float params[ALBBA_MLP_NPARAM];
    float queue_ratio; /* 0: unset */
};

static int upload(albba_selector *s)
{
    albd_ctx *d = albba_ctx_dev(s->ctx);
    int rc = albd_h2d(d, s->dparams, s->params, sizeof(s->params));
    return rc ? albba_fail(rc, "selector: %s", albd_error()) : ALBBA_OK;
}

int albba_selector_create(albba_ctx *ctx, uint64_t init_seed, albba_selector **out)
{
    if (!ctx || !out) return albba_fail(ALBBA_EINVAL, "selector_create: null");
    albba_selector *s = (albba_selector *)albba_calloc(1, sizeof(albba_selector));
    if (!s) return albba_fail(ALBBA_ENOMEM, "selector_create");
    s->ctx = ctx;
    int rc = albd_malloc(albba_ctx_dev(ctx), sizeof(s->params), (void **)&s->dparams);
    if (rc) { free(s); return albba_fail(rc, "selector_create: %s", albd_error()); }
    /* He-uniform weights (limit sqrt(6/fan_in)) for the ReLU layers, Xavier
     * limit sqrt(6/(fan_in+fan_out)) for the output layer, zero biases */
    const int fan_in[3] = {ALBBA_MLP_IN, ALBBA_MLP_H1, ALBBA_MLP_H2};
    const int fan_out[3] = {ALBBA_MLP_H1, ALBBA_MLP_H2, ALBBA_MLP_OUT};
    int at = 0;
    uint64_t k = 0;
    for (int layer = 0; layer < 3; ++layer) {
        double lim = layer < 2 ? sqrt(6.0 / fan_in[layer]) : sqrt(6.0 / (fan_in[layer] + fan_out[layer]));
        for (int i = 0; i < fan_in[layer] * fan_out[layer]; ++i)
            s->params[at++] = (float)((2.0 * albba_u01(albba_rng_at(init_seed, k++)) - 1.0) * lim);
        for (int i = 0; i < fan_out[layer]; ++i) s->params[at++] = 0.0f;
    }
    if ((rc = upload(s))) { albba_selector_free(s); return rc; }
    *out = s;
    return ALBBA_OK;
}

void albba_selector_free(albba_selector *s)
{
    if (!s) return;
    albd_free(albba_ctx_dev(s->ctx), s->dparams);
    free(s);
}

int albba_selector_get_params(const albba_selector *s, float *p)
{
    if (!s || !p) return albba_fail(ALBBA_EINVAL, "selector_get_params: null");
    memcpy(p, s->params, sizeof(s->params));
    return ALBBA_OK;
}

int albba_selector_set_params(albba_selector *s, const float *p)
{
    if (!s || !p) return albba_fail(ALBBA_EINVAL, "selector_set_params: null");
    memcpy(s->params, p, sizeof(s->params));
    return upload(s);
}

int albba_selector_save(const albba_selector *s, const char *path)
{
    if (!s || !path) return albba_fail(ALBBA_EINVAL, "selector_save: null");
    FILE *f = fopen(path, "w");
    if (!f) return albba_fail(ALBBA_EIO, "selector_save: cannot open %s", path);
    fprintf(f, "albba-selector 2 %d\n", ALBBA_MLP_NPARAM);
    for (int i = 0; i < ALBBA_MLP_NPARAM; ++i) fprintf(f, "%.9g\n", (double)s->params[i]);
    fprintf(f, "queue_ratio %.9g\n", (double)s->queue_ratio);
    int bad = ferror(f);
    fclose(f);
    return bad ? albba_fail(ALBBA_EIO, "selector_save: write error") : ALBBA_OK;
}

int albba_selector_load(albba_selector *s, const char *path)
{
    if (!s || !path) return albba_fail(ALBBA_EINVAL, "selector_load: null");
    FILE *f = fopen(path, "r");
    if (!f) return albba_fail(ALBBA_EIO, "selector_load: cannot open %s", path);
    char magic[32];
    int ver = 0, np = 0;
    float tmp[ALBBA_MLP_NPARAM];
    if (fscanf(f, "%31s %d %d", magic, &ver, &np) != 3 || strcmp(magic, "albba-selector") ||
        (ver != 1 && ver != 2) || np != ALBBA_MLP_NPARAM) {
        fclose(f);
        return albba_fail(ALBBA_EIO, "selector_load: %s is not a selector file", path);
    }
    for (int i = 0; i < np; ++i) {
        if (fscanf(f, "%f", &tmp[i]) != 1 || !isfinite(tmp[i])) {
            fclose(f);
            return albba_fail(ALBBA_EIO, "selector_load: bad parameter %d", i);
        }
    }
    float ratio = 0.0f;
    if (ver == 2) {
        char key[32];
        if (fscanf(f, "%31s %f", key, &ratio) != 2 || strcmp(key, "queue_ratio") || !isfinite(ratio) ||
            ratio < 0.0f) {
            fclose(f);
            return albba_fail(ALBBA_EIO, "selector_load: bad queue_ratio line in %s", path);
        }
    }
    fclose(f);
    memcpy(s->params, tmp, sizeof tmp);
    s->queue_ratio = ratio;
    return upload(s);
}

int albba_selector_set_queue_ratio(albba_selector *s, float ratio)
{
    if (!s || !isfinite(ratio) || ratio < 0.0f) return albba_fail(ALBBA_EINVAL, "selector_set_queue_ratio: bad argument");
    s->queue_ratio = ratio;
    return ALBBA_OK;
}

float albba_selector_queue_ratio(const albba_selector *s) { return s ? s->queue_ratio : 0.0f; }

/* device buffers for a batch of samples */
typedef struct batch_buf {
    float *x, *t, *scratch, *grad, *loss, *m, *v, *out;
} batch_buf;

static void batch_free(albd_ctx *d, batch_buf *b)
{
    albd_free(d, b->x); albd_free(d, b->t); albd_free(d, b->scratch); albd_free(d, b->grad);
    albd_free(d, b->loss); albd_free(d, b->m); albd_free(d, b->v); albd_free(d, b->out);
    memset(b, 0, sizeof(*b));
}

static int batch_alloc(albd_ctx *d, int32_t ns, int with_t, int with_train, batch_buf *b)
{
    memset(b, 0, sizeof(*b));
    int rc = albd_malloc(d, sizeof(float) * (size_t)ns * ALBBA_MLP_IN, (void **)&b->x);
    if (!rc && with_t) rc = albd_malloc(d, sizeof(float) * (size_t)ns * ALBBA_MLP_OUT, (void **)&b->t);
    if (!rc && with_t) rc = albd_malloc(d, sizeof(float) * ALBD_MLP_SCRATCH_FLOATS(ns), (void **)&b->scratch);
    if (!rc && with_t) rc = albd_malloc(d, sizeof(float) * ALBBA_MLP_NPARAM, (void **)&b->grad);
    if (!rc && with_t) rc = albd_malloc(d, sizeof(float) * 4, (void **)&b->loss);
    if (!rc && with_train) rc = albd_malloc(d, sizeof(float) * ALBBA_MLP_NPARAM, (void **)&b->m);
    if (!rc && with_train) rc = albd_malloc(d, sizeof(float) * ALBBA_MLP_NPARAM, (void **)&b->v);
    if (!rc && !with_t) rc = albd_malloc(d, sizeof(float) * (size_t)ns * ALBBA_MLP_OUT, (void **)&b->out);
    if (rc) batch_free(d, b);
    return rc;
}

int albba_selector_loss_grad(albba_selector *s, const float *x, const float *t, int32_t ns,
                             float *loss, float *grad)
{
    if (!s || !x || !t || ns <= 0) return albba_fail(ALBBA_EINVAL, "selector_loss_grad: bad arguments");
    albd_ctx *d = albba_ctx_dev(s->ctx);
    batch_buf b;
    int rc = batch_alloc(d, ns, 1, 0, &b);
    if (!rc) rc = albd_h2d(d, b.x, x, sizeof(float) * (size_t)ns * ALBBA_MLP_IN);
    if (!rc) rc = albd_h2d(d, b.t, t, sizeof(float) * (size_t)ns * ALBBA_MLP_OUT);
    if (!rc) rc = albd_mlp_loss_grad(d, s->dparams, b.x, b.t, ns, b.scratch, b.grad, b.loss);
    if (!rc && loss) rc = albd_d2h(d, loss, b.loss, sizeof(float));
    if (!rc && grad) rc = albd_d2h(d, grad, b.grad, sizeof(float) * ALBBA_MLP_NPARAM);
    batch_free(d, &b);
    return rc ? albba_fail(rc, "selector_loss_grad: %s", albd_error()) : ALBBA_OK;
}

int albba_selector_train(albba_selector *s, const float *x, const float *t, int32_t ns,
                         int32_t steps, float lr, float *final_loss)
{
    if (!s || !x || !t || ns <= 0 || steps < 0 || !(lr > 0))
        return albba_fail(ALBBA_EINVAL, "selector_train: bad arguments");
    albd_ctx *d = albba_ctx_dev(s->ctx);
    batch_buf b;
    int rc = batch_alloc(d, ns, 1, 1, &b);
    if (!rc) rc = albd_h2d(d, b.x, x, sizeof(float) * (size_t)ns * ALBBA_MLP_IN);
    if (!rc) rc = albd_h2d(d, b.t, t, sizeof(float) * (size_t)ns * ALBBA_MLP_OUT);
    if (!rc) rc = albd_fill_u32(d, b.m, 0, ALBBA_MLP_NPARAM);
    if (!rc) rc = albd_fill_u32(d, b.v, 0, ALBBA_MLP_NPARAM);
    for (int32_t it = 1; !rc && it <= steps; ++it) {
        rc = albd_mlp_loss_grad(d, s->dparams, b.x, b.t, ns, b.scratch, b.grad, b.loss);
        if (!rc) rc = albd_mlp_adam(d, s->dparams, b.grad, b.m, b.v, ALBBA_MLP_NPARAM, lr, 0.9f, 0.999f, 1e-8f, it);
    }
    /* loss of the final parameters */
    if (!rc) rc = albd_mlp_loss_grad(d, s->dparams, b.x, b.t, ns, b.scratch, b.grad, b.loss);
    if (!rc && final_loss) rc = albd_d2h(d, final_loss, b.loss, sizeof(float));
    if (!rc) rc = albd_d2h(d, s->params, s->dparams, sizeof(s->params));
    batch_free(d, &b);
    return rc ? albba_fail(rc, "selector_train: %s", albd_error()) : ALBBA_OK;
}

int albba_selector_predict(albba_selector *s, const float *x, int32_t ns, float *out)
{
    if (!s || !x || !out || ns <= 0) return albba_fail(ALBBA_EINVAL, "selector_predict: bad arguments");
    albd_ctx *d = albba_ctx_dev(s->ctx);
    batch_buf b;
    int rc = batch_alloc(d, ns, 0, 0, &b);
    if (!rc) rc = albd_h2d(d, b.x, x, sizeof(float) * (size_t)ns * ALBBA_MLP_IN);
    if (!rc) rc = albd_mlp_forward(d, s->dparams, b.x, ns, b.out);
    if (!rc) rc = albd_d2h(d, out, b.out, sizeof(float) * (size_t)ns * ALBBA_MLP_OUT);
    batch_free(d, &b);
    return rc ? albba_fail(rc, "selector_predict: %s", albd_error()) : ALBBA_OK;
}

const float *albba_selector_dparams(const albba_selector *s) { return s ? s->dparams : NULL; }
