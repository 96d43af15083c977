// zkg_noalloc.hip -- the batched column NTT for callers that cannot allocate.
//
// zkg_ntt_batch takes the four-step transposition buffer (height x width
// elements for heights above 2^11) from the context's allocator, which may
// call hipMalloc.  zkg_ntt_batch_noalloc runs the same kernels with that
// buffer in a fixed device array of ZKG_NOALLOC_SCRATCH_BYTES, so it allocates
// nothing: for code under stream capture, tight serving loops and harnesses
// that forbid allocation.
//
// The array exists once per GPU (each device that loads the library holds
// its own copy until the process ends).  Its device address is looked up once
// per device, when a context is created on it, and only that address is ever
// handed to kernels.  Every call on one GPU shares the array: calls on
// different streams (contexts) of one GPU that overlap in time race on it;
// calls on one context are ordered by its stream.  The transposition buffer is
// fully written before it is read, so nothing needs clearing between calls.
#include "../../include/zkg.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <mutex>

using namespace zkg;

namespace {

static __device__ __align__(256) uint8_t scratch[ZKG_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each GPU

void *scratch_on(int device)
{
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

// Called by zkg_ctx_create with `device` current.
int zkg::noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return set_err(ZKG_EINVAL, "device %d beyond %d", device, MAX_DEVICES);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return ZKG_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_err(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch symbol address");
    g_scratch_ptr[device] = p;
    return ZKG_OK;
}

extern "C" int zkg_ntt_batch_noalloc(zkg_ctx *ctx, int field, zkg_buf *buf, uint64_t height, uint32_t width,
                                     int inverse)
{
    if (!ctx || !buf) return set_err(ZKG_EINVAL, "null argument");
    if (field != FIELD_GL && field != FIELD_BB) return set_err(ZKG_EINVAL, "field %d", field);
    if (!height || (height & (height - 1)) || height > (1ull << 22))
        return set_err(ZKG_EINVAL, "height %llu not a power of two <= 2^22", (unsigned long long)height);
    if (span_bytes(height, width, field == FIELD_GL ? 8 : 4) > buf->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    if (height == 1 || width == 0) return ZKG_OK;
    const uint64_t tb = ntt_tmp_bytes(field, height, width);
    if (tb > ZKG_NOALLOC_SCRATCH_BYTES)
        return set_err(ZKG_EINVAL, "height x width needs %llu bytes of scratch, more than the %llu-byte array",
                       (unsigned long long)tb, (unsigned long long)ZKG_NOALLOC_SCRATCH_BYTES);
    void *tmp = nullptr;
    if (tb) {
        tmp = scratch_on(ctx->device);
        if (!tmp) return set_err(ZKG_ESTATE, "scratch array of device %d not set up", ctx->device);
    }
    hipError_t e = launch_ntt(field, ctx->tw[field], buf->ptr, tmp, height, width, inverse, ctx->stream);
    return e == hipSuccess ? ZKG_OK : hip_err(e, "ntt (no alloc)");
}
