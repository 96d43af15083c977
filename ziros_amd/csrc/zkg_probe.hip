// zkg_probe.hip -- micro-benchmarks that measure the VALU ceilings the
// roofline of the hashing kernels is priced against (libzkg_probe.so, used by
// bench.py; not part of the commit path).
//
// Each lane runs 8 independent multiply chains (enough ILP to keep the VALU
// issuing) for `iters` steps; the grid holds 8 waves per SIMD on every CU.
// Kinds: 0 = BabyBear Montgomery product (the exact routine the kernels use),
// 1 = Goldilocks product, 2 = raw 32-bit v_mul_lo_u32, 3 = FP64 FMA,
// 4 = BabyBear product in exact FP64 arithmetic (6 FP64 operations),
// 5 = raw v_mad_u64_u32 (multiply-with-carry step), 6 = v_mul_hi_u32 + v_add.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zkg_field.h"

using namespace zkg;

namespace {

template <int KIND>
__global__ __launch_bounds__(256) void k_chain(uint64_t *out, uint32_t iters, uint64_t seed)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (KIND == 0) {
        uint32_t x[8], y = (uint32_t)(seed % Bb::P) | 1u;
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = (t * 2654435761u + i) % Bb::P;
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = Bb::mul(x[i], y);
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc ^= x[i];
        out[t] = acc;
    } else if constexpr (KIND == 1) {
        uint64_t x[8], y = seed % Gl::P;
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = (uint64_t)t * 0x9e3779b97f4a7c15ULL % Gl::P + i;
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = Gl::mul(x[i], y);
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc ^= x[i];
        out[t] = acc;
    } else if constexpr (KIND == 3) {
        double x[8], y = 1.0000001, z = 1e-9;
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = (double)(t + i);
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = __builtin_fma(x[i], y, z);
        double acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc += x[i];
        out[t] = (uint64_t)acc;
    } else if constexpr (KIND == 4) {
        const double P = 2013265921.0, PINV = 1.0 / 2013265921.0, MAGIC = 6755399441055744.0;
        double x[8], y = (double)(seed % 2013265921ull);
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = (double)((t * 2654435761u + i) % 2013265921u);
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                double h = x[i] * y;
                double l = __builtin_fma(x[i], y, -h);
                double q = __builtin_fma(h, PINV, MAGIC) - MAGIC;
                x[i] = __builtin_fma(-q, P, h) + l;
            }
        double acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc += x[i];
        out[t] = (uint64_t)(acc + 4e9);
    } else if constexpr (KIND == 5) {
        uint64_t x[8];
        const uint32_t y = (uint32_t)seed | 1u;
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = ((uint64_t)(t + i) << 32) | (t * 2654435761u + i);
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = (uint64_t)(uint32_t)x[i] * y + (x[i] >> 32);
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc ^= x[i];
        out[t] = acc;
    } else if constexpr (KIND == 6) {
        uint32_t x[8], y = (uint32_t)seed | 0x80000001u;
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = t * 2654435761u + i;
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = __umulhi(x[i], y) + 0x9e3779b9u;
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc ^= x[i];
        out[t] = acc;
    } else {
        uint32_t x[8], y = (uint32_t)seed | 1u;
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = t + i;
        for (uint32_t k = 0; k < iters; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = x[i] * y;
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc ^= x[i];
        out[t] = acc;
    }
}

__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

// Plain copy of `bytes` (a multiple of 16) between two fresh buffers: the
// known-size reference that scales HBM traffic counters.  Returns GB/s.
extern "C" int zkg_probe_copy(uint64_t bytes, double *gbps)
{
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -3;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return -3;
    }
    (void)hipMemset(a, 1, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint64_t n = bytes / 16;
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_copy, dim3(256 * 16), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess || ms <= 0) return -2;
    *gbps = 2.0 * bytes / (ms * 1e-3) / 1e9;
    return 0;
}

// Operations per second of `kind` over the whole device (8 products per lane
// per iteration).  Returns 0 on success.
extern "C" int zkg_probe_rate(int kind, uint32_t iters, double *ops_per_s)
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -2;
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 8;  // 8 x 256 threads per CU
    const uint64_t threads = (uint64_t)blocks * 256;
    uint64_t *out = nullptr;
    if (hipMalloc(&out, threads * sizeof(uint64_t)) != hipSuccess) return -3;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    auto launch = [&](uint32_t it) {
        if (kind == 0)
            hipLaunchKernelGGL(k_chain<0>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
        else if (kind == 1)
            hipLaunchKernelGGL(k_chain<1>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
        else if (kind == 3)
            hipLaunchKernelGGL(k_chain<3>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
        else if (kind == 4)
            hipLaunchKernelGGL(k_chain<4>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
        else if (kind == 5)
            hipLaunchKernelGGL(k_chain<5>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
        else if (kind == 6)
            hipLaunchKernelGGL(k_chain<6>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
        else
            hipLaunchKernelGGL(k_chain<2>, dim3(blocks), dim3(256), 0, 0, out, it, 12345);
    };
    launch(16);  // warm-up
    (void)hipEventRecord(a, 0);
    launch(iters);
    (void)hipEventRecord(b, 0);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(out);
    if (e != hipSuccess || ms <= 0) return -2;
    *ops_per_s = (double)threads * iters * 8.0 / (ms * 1e-3);
    return 0;
}
