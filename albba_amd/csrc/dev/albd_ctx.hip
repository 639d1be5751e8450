// Context, memory, events and fill kernels of the device layer.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "../albd.h"
#include "../../../include/albba.h"
#include "dev_common.h"

static thread_local char g_err[512];

const char *albd_error(void) { return g_err; }

int albd_set_error(hipError_t e, const char *what, const char *file, int line)
{
    snprintf(g_err, sizeof g_err, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return ALBBA_EHIP;
}

int albd_set_error_msg(int code, const char *msg)
{
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

struct albd_ctx {
    int device;
    int ncu;
    hipStream_t stream;
    hipEvent_t ev[ALBD_NEVENTS];
    albd_counters *hctr; // pinned
    albd_counters *mail; // pinned, host-coherent, written by the select kernel
    albd_counters *dmail; // its device address
    unsigned int mail_seq;
    unsigned long mail_hits, mail_misses; // posts seen / waits that fell back
};

int albd_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        (void)hipGetLastError();
        return albd_set_error_msg(ALBBA_ENODEV, "no HIP device");
    }
    *count = n;
    return 0;
}

int albd_ctx_create(int device, albd_ctx **out)
{
    int n = 0;
    if (albd_device_count(&n) || device < 0 || device >= n)
        return albd_set_error_msg(ALBBA_ENODEV, "requested HIP device is not present");
    HIPCHK(hipSetDevice(device));
    // the batched BFS's fixed scratch array: its address on this device
    if (int rc = albd_batch_scratch_init(device)) return rc;
    albd_ctx *c = new albd_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->ncu = prop.multiProcessorCount;
    // default priority stream: no priority games
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 0; i < ALBD_NEVENTS; ++i) HIPCHK(hipEventCreate(&c->ev[i]));
    HIPCHK(hipHostMalloc((void **)&c->hctr, sizeof(albd_counters), hipHostMallocDefault));
    memset(c->hctr, 0, sizeof(albd_counters));
    HIPCHK(hipHostMalloc((void **)&c->mail, sizeof(albd_counters),
                         hipHostMallocMapped | hipHostMallocCoherent));
    memset(c->mail, 0, sizeof(albd_counters));
    HIPCHK(hipHostGetDevicePointer((void **)&c->dmail, c->mail, 0));
    *out = c;
    return 0;
}

void albd_ctx_destroy(albd_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < ALBD_NEVENTS; ++i) (void)hipEventDestroy(c->ev[i]);
    (void)hipStreamDestroy(c->stream);
    (void)hipHostFree(c->hctr);
    (void)hipHostFree(c->mail);
    delete c;
}

int albd_ctx_device(const albd_ctx *c) { return c->device; }
int albd_num_cus(const albd_ctx *c) { return c->ncu; }
hipStream_t albd_stream(albd_ctx *c) { return c->stream; }

int albd_malloc(albd_ctx *c, size_t bytes, void **p)
{
    HIPCHK(hipSetDevice(c->device));
    hipError_t e = hipMalloc(p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        snprintf(g_err, sizeof g_err, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return ALBBA_EDEVMEM;
    }
    return 0;
}

void albd_free(albd_ctx *c, void *p)
{
    if (!p) return;
    (void)hipSetDevice(c->device);
    (void)hipFree(p);
}

int albd_h2d(albd_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int albd_d2h(albd_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int albd_d2d(albd_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

int albd_sync(albd_ctx *c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

albd_counters *albd_host_counters(albd_ctx *c) { return c->hctr; }

int albd_read_counters(albd_ctx *c, const albd_state *st)
{
    HIPCHK(hipMemcpyAsync(c->hctr, st->ctr, sizeof(albd_counters), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

albd_counters *albd_mail_next(albd_ctx *c, unsigned int *seq)
{
    *seq = ++c->mail_seq;
    if (*seq == 0) *seq = ++c->mail_seq; // 0 is the mailbox's initial value
    return c->dmail;
}

int albd_mail_wait(albd_ctx *c, const albd_state *st, unsigned int seq)
{
    volatile unsigned int *sp = &c->mail->seq;
    for (unsigned long spins = 0;; ++spins) {
        if (__atomic_load_n(sp, __ATOMIC_ACQUIRE) == seq) {
            memcpy(c->hctr, (const void *)c->mail, sizeof(albd_counters));
            ++c->mail_hits;
            return 0;
        }
        if ((spins & 1023) == 1023) {
            // the stream failed, or finished without the post becoming
            // visible: take the counters the ordinary way
            const hipError_t e = hipStreamQuery(c->stream);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return albd_set_error(e, "select kernel", __FILE__, __LINE__);
            if (__atomic_load_n(sp, __ATOMIC_ACQUIRE) == seq) continue;
            ++c->mail_misses;
            return albd_read_counters(c, st);
        }
    }
}

void albd_mail_stats(const albd_ctx *c, unsigned long *hits, unsigned long *misses)
{
    *hits = c->mail_hits;
    *misses = c->mail_misses;
}

int albd_event_record(albd_ctx *c, int idx)
{
    HIPCHK(hipEventRecord(c->ev[idx], c->stream));
    return 0;
}

int albd_event_wait(albd_ctx *c, int idx)
{
    HIPCHK(hipEventSynchronize(c->ev[idx]));
    return 0;
}

int albd_pinned_alloc(albd_ctx *c, size_t bytes, void **p)
{
    HIPCHK(hipSetDevice(c->device));
    hipError_t e = hipHostMalloc(p, bytes ? bytes : 256, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        snprintf(g_err, sizeof g_err, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return ALBBA_ENOMEM;
    }
    return 0;
}

void albd_pinned_free(albd_ctx *c, void *p)
{
    if (!p) return;
    (void)hipSetDevice(c->device);
    (void)hipHostFree(p);
}

int albd_h2d_async(albd_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return 0;
}

int albd_event_elapsed(albd_ctx *c, int a, int b, float *ms)
{
    HIPCHK(hipEventSynchronize(c->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return 0;
}

// ---------------------------------------------------------------------------
// fills

static __global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static __global__ __launch_bounds__(256) void k_fill_u64(uint64_t *p, uint64_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// plain copy, 16 bytes per lane per step: the calibration kernel for the
// device's achievable HBM bandwidth and for the memory traffic counters
static __global__ __launch_bounds__(256) void k_copy_u64(ulonglong2 *dst, const ulonglong2 *src,
                                                         size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// out[w] = the stored entries of position word w (its 64 row lengths
// summed): the column entries a full scan of the word's slices reads
static __global__ __launch_bounds__(256) void k_word_entries(const int32_t *row_len, int64_t nwords,
                                                             int64_t *out)
{
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = gt >> 6;
    if (w >= nwords) return; // whole waves: 64 lanes share w
    const unsigned long long s = wave_sum((unsigned long long)row_len[gt]);
    if ((threadIdx.x & 63) == 0) out[w] = (int64_t)s;
}

static unsigned fill_blocks(albd_ctx *c, size_t n)
{
    size_t b = (n + 255) / 256;
    size_t cap = (size_t)c->ncu * 16;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

int albd_word_entries(albd_ctx *c, const albd_graph *g, int64_t *out)
{
    const int64_t threads = (int64_t)g->nwords * 64;
    if (threads <= 0) return 0;
    hipLaunchKernelGGL(k_word_entries, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream,
                       g->row_len, (int64_t)g->nwords, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_fill_u32(albd_ctx *c, void *p, uint32_t v, size_t count)
{
    if (!count) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(c, count)), dim3(256), 0, c->stream,
                       (uint32_t *)p, v, count);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_fill_u64(albd_ctx *c, void *p, uint64_t v, size_t count)
{
    if (!count) return 0;
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(c, count)), dim3(256), 0, c->stream,
                       (uint64_t *)p, v, count);
    HIPCHK(hipGetLastError());
    return 0;
}

int albd_copy_u64(albd_ctx *c, void *dst, const void *src, size_t count)
{
    if (!count) return 0;
    if (count & 1) return albd_d2d(c, dst, src, count * 8);
    const size_t n2 = count / 2;
    size_t b = (n2 + 255) / 256, cap = (size_t)c->ncu * 32;
    hipLaunchKernelGGL(k_copy_u64, dim3((unsigned)(b < 1 ? 1 : (b > cap ? cap : b))), dim3(256), 0,
                       c->stream, (ulonglong2 *)dst, (const ulonglong2 *)src, n2);
    HIPCHK(hipGetLastError());
    return 0;
}
