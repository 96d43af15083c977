// zkg_probe.hip -- micro-benchmarks that measure the VALU ceilings the
// roofline of the hashing kernels is priced against (libzkg_probe.so, used by
// bench.py; not part of the commit path).
//
// Each lane runs 8 independent multiply chains (enough ILP to keep the VALU
// issuing) for `iters` steps; the grid holds 8 waves per SIMD on every CU.
// Kinds: 0 = BabyBear Montgomery product (the exact routine the kernels use),
// 1 = Goldilocks product, 2 = raw 32-bit v_mul_lo_u32.
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

}  // namespace

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
