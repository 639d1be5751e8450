// FP32 MLP 12 -> 32 -> 32 -> 4 (ReLU, ReLU, linear) of the learned direction
// selector: forward, mean-squared-error loss with its exact gradient, Adam,
// and the single-vector decision used inside the BFS level loop.
//
// Parameter layout (1604 floats): W1[32][12], b1[32], W2[32][32], b2[32],
// W3[4][32], b3[4]; W[j][i] multiplies input i into unit j.
//
// The gradient is computed in deterministic passes without atomics: one
// thread per sample stores activations and back-propagated errors; a block
// per chunk of 64 samples stages them in LDS and sums every parameter's
// contributions over its chunk in sample order; one thread per parameter then
// adds the chunks' partial sums in chunk order.  A run reproduces bit for bit
// and can be compared with a CPU restatement.
#include "dev_common.h"
#include "../../../include/albba.h"

namespace {

constexpr int NI = ALBBA_MLP_IN, H1 = ALBBA_MLP_H1, H2 = ALBBA_MLP_H2, NO = ALBBA_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + H1 * NI, OW2 = OB1 + H1, OB2 = OW2 + H2 * H1,
              OW3 = OB2 + H2, OB3 = OW3 + NO * H2, NP = OB3 + NO;
static_assert(NP == ALBBA_MLP_NPARAM, "parameter count");

__device__ __forceinline__ void load_params(const float *params, float *sp)
{
    for (int i = threadIdx.x; i < NP; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
}

__device__ __forceinline__ void forward_one(const float *sp, const float *x, float *h1, float *h2,
                                            float *y)
{
    for (int j = 0; j < H1; ++j) {
        float a = sp[OB1 + j];
        for (int i = 0; i < NI; ++i) a = fmaf(sp[OW1 + j * NI + i], x[i], a);
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < H2; ++j) {
        float a = sp[OB2 + j];
        for (int i = 0; i < H1; ++i) a = fmaf(sp[OW2 + j * H1 + i], h1[i], a);
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int k = 0; k < NO; ++k) {
        float a = sp[OB3 + k];
        for (int i = 0; i < H2; ++i) a = fmaf(sp[OW3 + k * H2 + i], h2[i], a);
        y[k] = a;
    }
}

__global__ __launch_bounds__(256) void k_forward(const float *params, const float *x, int32_t ns,
                                                 float *out)
{
    __shared__ float sp[NP];
    load_params(params, sp);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float xi[NI], h1[H1], h2[H2], y[NO];
    for (int i = 0; i < NI; ++i) xi[i] = x[(size_t)s * NI + i];
    forward_one(sp, xi, h1, h2, y);
    for (int k = 0; k < NO; ++k) out[(size_t)s * NO + k] = y[k];
}

// scratch layout per sample block: h1[ns][32] h2[ns][32] dy[ns][4] dh1[ns][32]
// dh2[ns][32] lossv[ns]
__global__ __launch_bounds__(256) void k_sample_pass(const float *params, const float *x,
                                                     const float *t, int32_t ns, float *scratch)
{
    __shared__ float sp[NP];
    load_params(params, sp);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float *H1s = scratch, *H2s = H1s + (size_t)ns * H1, *DY = H2s + (size_t)ns * H2,
          *DH1 = DY + (size_t)ns * NO, *DH2 = DH1 + (size_t)ns * H1, *LV = DH2 + (size_t)ns * H2;
    float xi[NI], h1[H1], h2[H2], y[NO], dy[NO], dh2[H2];
    for (int i = 0; i < NI; ++i) xi[i] = x[(size_t)s * NI + i];
    forward_one(sp, xi, h1, h2, y);
    const float scale = 2.0f / ((float)ns * (float)NO);
    float l = 0.f;
    for (int k = 0; k < NO; ++k) {
        const float e = y[k] - t[(size_t)s * NO + k];
        l = fmaf(e, e, l);
        dy[k] = scale * e;
        DY[(size_t)s * NO + k] = dy[k];
    }
    LV[s] = l;
    for (int i = 0; i < H2; ++i) {
        float a = 0.f;
        for (int k = 0; k < NO; ++k) a = fmaf(sp[OW3 + k * H2 + i], dy[k], a);
        dh2[i] = h2[i] > 0.f ? a : 0.f;
        DH2[(size_t)s * H2 + i] = dh2[i];
        H2s[(size_t)s * H2 + i] = h2[i];
    }
    for (int i = 0; i < H1; ++i) {
        float a = 0.f;
        for (int j = 0; j < H2; ++j) a = fmaf(sp[OW2 + j * H1 + i], dh2[j], a);
        DH1[(size_t)s * H1 + i] = h1[i] > 0.f ? a : 0.f;
        H1s[(size_t)s * H1 + i] = h1[i];
    }
}

constexpr int CH = ALBD_MLP_CHUNK;

// Partial sums of chunk blockIdx.x (samples [CH*c, CH*c + m)): partial[c][p]
// for every parameter p, partial[c][NP] the chunk's squared errors.
__global__ __launch_bounds__(256) void k_param_partial(const float *x, int32_t ns, const float *scratch,
                                                       float *partial)
{
    __shared__ float sh1[CH * H1], sh2[CH * H2], sdy[CH * NO], sdh1[CH * H1], sdh2[CH * H2], sx[CH * NI],
        slv[CH];
    const float *H1s = scratch, *H2s = H1s + (size_t)ns * H1, *DY = H2s + (size_t)ns * H2,
                *DH1 = DY + (size_t)ns * NO, *DH2 = DH1 + (size_t)ns * H1, *LV = DH2 + (size_t)ns * H2;
    const int c = blockIdx.x;
    const size_t s0 = (size_t)c * CH;
    const int m = (int)((size_t)ns - s0 < (size_t)CH ? (size_t)ns - s0 : (size_t)CH);
    for (int i = threadIdx.x; i < m * H1; i += blockDim.x) {
        sh1[i] = H1s[s0 * H1 + i];
        sdh1[i] = DH1[s0 * H1 + i];
    }
    for (int i = threadIdx.x; i < m * H2; i += blockDim.x) {
        sh2[i] = H2s[s0 * H2 + i];
        sdh2[i] = DH2[s0 * H2 + i];
    }
    for (int i = threadIdx.x; i < m * NO; i += blockDim.x) sdy[i] = DY[s0 * NO + i];
    for (int i = threadIdx.x; i < m * NI; i += blockDim.x) sx[i] = x[s0 * NI + i];
    for (int i = threadIdx.x; i < m; i += blockDim.x) slv[i] = LV[s0 + i];
    __syncthreads();
    float *out = partial + (size_t)c * (NP + 1);
    for (int p = threadIdx.x; p <= NP; p += blockDim.x) {
        float g = 0.f;
        if (p == NP) {
            for (int q = 0; q < m; ++q) g += slv[q];
        } else if (p < OB1) {
            const int j = (p - OW1) / NI, i = (p - OW1) % NI;
            for (int q = 0; q < m; ++q) g = fmaf(sdh1[q * H1 + j], sx[q * NI + i], g);
        } else if (p < OW2) {
            const int j = p - OB1;
            for (int q = 0; q < m; ++q) g += sdh1[q * H1 + j];
        } else if (p < OB2) {
            const int j = (p - OW2) / H1, i = (p - OW2) % H1;
            for (int q = 0; q < m; ++q) g = fmaf(sdh2[q * H2 + j], sh1[q * H1 + i], g);
        } else if (p < OW3) {
            const int j = p - OB2;
            for (int q = 0; q < m; ++q) g += sdh2[q * H2 + j];
        } else if (p < OB3) {
            const int k = (p - OW3) / H2, i = (p - OW3) % H2;
            for (int q = 0; q < m; ++q) g = fmaf(sdy[q * NO + k], sh2[q * H2 + i], g);
        } else {
            const int k = p - OB3;
            for (int q = 0; q < m; ++q) g += sdy[q * NO + k];
        }
        out[p] = g;
    }
}

// grad[p] = the chunks' partial sums in chunk order; loss = the mean of the
// squared errors
__global__ __launch_bounds__(256) void k_param_reduce(const float *partial, int32_t nchunks, int32_t ns,
                                                      float *grad, float *loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    float g = 0.f;
    for (int c = 0; c < nchunks; ++c) g += partial[(size_t)c * (NP + 1) + p];
    if (p == NP) *loss = g / ((float)ns * (float)NO);
    else grad[p] = g;
}

__global__ __launch_bounds__(256) void k_adam(float *params, const float *grad, float *m, float *v,
                                              int32_t np, float lr, float b1, float b2, float eps,
                                              float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

} // namespace

extern "C" {

int albd_mlp_forward(albd_ctx *c, const float *params, const float *x, int32_t ns, float *out)
{
    if (ns <= 0) return 0;
    hipLaunchKernelGGL(k_forward, dim3((ns + 255) / 256), dim3(256), 0, albd_stream(c), params, x, ns, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_mlp_loss_grad(albd_ctx *c, const float *params, const float *x, const float *t, int32_t ns,
                       float *scratch, float *grad, float *loss)
{
    if (ns <= 0) return albd_set_error_msg(ALBBA_EINVAL, "mlp: no samples");
    hipStream_t s = albd_stream(c);
    hipLaunchKernelGGL(k_sample_pass, dim3((ns + 255) / 256), dim3(256), 0, s, params, x, t, ns, scratch);
    HIPCHK(hipGetLastError());
    const int32_t nchunks = (ns + CH - 1) / CH;
    float *partial = scratch + (size_t)ns * ALBD_MLP_SCRATCH_PER_SAMPLE;
    hipLaunchKernelGGL(k_param_partial, dim3(nchunks), dim3(256), 0, s, x, ns, scratch, partial);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_param_reduce, dim3((NP + 1 + 255) / 256), dim3(256), 0, s, partial, nchunks, ns, grad,
                       loss);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_mlp_adam(albd_ctx *c, float *params, const float *grad, float *m, float *v, int32_t np,
                  float lr, float b1, float b2, float eps, int32_t step)
{
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, albd_stream(c), params, grad, m, v,
                       np, lr, b1, b2, eps, c1, c2);
    HIPCHK(hipGetLastError());
    return 0;
}

} // extern "C"
