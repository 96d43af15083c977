// zkg_poly.hip -- PolyOpsAccelerator for Goldilocks: batch polynomial
// evaluation, evaluation on a coset of the two-adic subgroup, and the
// quotient (f(x_i) - f(z)) / (x_i - z) over the subgroup.
//
// Reference: the trait zkf-core/src/acceleration.rs:173-192 and its Metal
// implementation zkf-metal/src/poly.rs:37-246 (kernels
// shaders/poly_ops.metal: Horner per point; per-element shift^k by square and
// multiply, then the forward DFT; per-element generator^i and a Fermat
// inverse per element).  The design here is this build's own:
//   * Horner: one point per lane, the coefficients staged through LDS in
//     tiles read highest degree first, so each coefficient is one broadcast
//     LDS read for the whole block;
//   * shift powers and subgroup points: each lane walks a run of RUN
//     consecutive exponents from one square-and-multiply power, one product
//     per element;
//   * quotient: the run's denominators are inverted together (Montgomery's
//     batch inversion: one Fermat inverse and three products per element);
//     a point equal to z gives 0, as the reference does;
//   * coset evaluation: the shifted coefficients go through the same column
//     NTT as zkg_ntt_batch (natural order in and out).
#include "../../include/zkg.h"
#include "zkg_field.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <vector>

using namespace zkg;

namespace {

constexpr int POLY_THREADS = 256;
constexpr int HORNER_TILE = 2048;  // coefficients per LDS tile (16 KiB)
constexpr int RUN = 16;            // consecutive exponents per lane

__device__ __forceinline__ uint64_t gl_pow_dev(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = Gl::mul(r, a);
        a = Gl::mul(a, a);
        e >>= 1;
    }
    return r;
}

// out[i] = sum_k coeffs[k] points[i]^k by Horner, highest degree first.
__global__ __launch_bounds__(POLY_THREADS) void k_poly_eval(const uint64_t *__restrict__ coeffs, uint64_t n_coeffs,
                                                            const uint64_t *__restrict__ points, uint64_t n_points,
                                                            uint64_t *__restrict__ out)
{
    __shared__ uint64_t tile[HORNER_TILE];
    const uint64_t i = (uint64_t)blockIdx.x * POLY_THREADS + threadIdx.x;
    const uint64_t x = i < n_points ? points[i] : 0;
    uint64_t acc = 0;
    for (uint64_t hi = n_coeffs; hi > 0;) {
        const uint64_t lo = hi > (uint64_t)HORNER_TILE ? hi - HORNER_TILE : 0;
        const int len = (int)(hi - lo);
        __syncthreads();  // the previous tile is consumed
        for (int k = threadIdx.x; k < len; k += POLY_THREADS) tile[k] = coeffs[lo + k];
        __syncthreads();
        for (int k = len - 1; k >= 0; k--) acc = Gl::add(Gl::mul(acc, x), tile[k]);
        hi = lo;
    }
    if (i < n_points) out[i] = acc;
}

// c_k <- c_k shift^k, k < n (in place).
__global__ __launch_bounds__(POLY_THREADS) void k_poly_shift(uint64_t *__restrict__ c, uint64_t n, uint64_t shift)
{
    const uint64_t k0 = ((uint64_t)blockIdx.x * POLY_THREADS + threadIdx.x) * RUN;
    if (k0 >= n) return;
    uint64_t w = gl_pow_dev(shift, k0);
    const uint64_t end = k0 + RUN < n ? k0 + RUN : n;
    for (uint64_t k = k0; k < end; k++) {
        c[k] = Gl::mul(c[k], w);
        w = Gl::mul(w, shift);
    }
}

// q_i = (e_i - f_z) / (g^i - z), 0 where g^i = z; g of order n.
__global__ __launch_bounds__(POLY_THREADS) void k_poly_quotient(const uint64_t *__restrict__ e, uint64_t n, uint64_t g,
                                                                uint64_t z, uint64_t f_z, uint64_t *__restrict__ q)
{
    const uint64_t i0 = ((uint64_t)blockIdx.x * POLY_THREADS + threadIdx.x) * RUN;
    if (i0 >= n) return;
    const int len = (int)(i0 + RUN < n ? RUN : n - i0);
    uint64_t x = gl_pow_dev(g, i0);
    uint64_t d[RUN], pre[RUN];
    uint64_t acc = 1;
#pragma unroll
    for (int t = 0; t < RUN; t++) {
        d[t] = 1;
        if (t < len) {
            d[t] = Gl::sub(x, z);
            x = Gl::mul(x, g);
        }
        pre[t] = acc;                                 // product of the nonzero d before t
        if (d[t]) acc = Gl::mul(acc, d[t]);
    }
    uint64_t inv = gl_pow_dev(acc, Gl::P - 2);        // 1 / product of the nonzero d
#pragma unroll
    for (int t = RUN - 1; t >= 0; t--) {
        if (t >= len) continue;
        if (!d[t]) {
            q[i0 + t] = 0;
            continue;
        }
        const uint64_t dinv = Gl::mul(inv, pre[t]);  // 1 / d[t]
        inv = Gl::mul(inv, d[t]);
        q[i0 + t] = Gl::mul(Gl::sub(e[i0 + t], f_z), dinv);
    }
}

// p3 Goldilocks::two_adic_generator(log_n): the order-2^32 root squared down.
uint64_t gl_two_adic_generator(int log_n)
{
    uint64_t w = 0x185629dcda58878cULL;
    for (int i = log_n; i < 32; i++) w = Gl::mul(w, w);
    return w;
}

int check_canonical(const uint64_t *v, uint64_t n, const char *what)
{
    for (uint64_t i = 0; i < n; i++)
        if (v[i] >= Gl::P) return set_err(ZKG_EINVAL, "%s[%llu] is not canonical", what, (unsigned long long)i);
    return ZKG_OK;
}

// Device copies of host arrays for one call: freed at scope exit.
struct Scratch {
    zkg_ctx *ctx;
    std::vector<void *> blocks;
    explicit Scratch(zkg_ctx *c) : ctx(c) {}
    ~Scratch()
    {
        for (void *p : blocks) dev_free(ctx, p);
    }
    int alloc(uint64_t bytes, void **out)
    {
        int rc = dev_alloc(ctx, bytes ? bytes : 8, out);
        if (!rc) blocks.push_back(*out);
        return rc;
    }
};

}  // namespace

extern "C" int zkg_poly_eval_goldilocks(zkg_ctx *ctx, const uint64_t *coeffs, uint64_t n_coeffs,
                                        const uint64_t *points, uint64_t n_points, uint64_t *out)
{
    if (!ctx || (!coeffs && n_coeffs) || (!points && n_points) || (!out && n_points))
        return set_err(ZKG_EINVAL, "null argument");
    if (!n_points) return ZKG_OK;
    const uint64_t cb = span_bytes(n_coeffs, 8), pb = span_bytes(n_points, 8);
    if (cb > ctx->hbm_budget || pb > ctx->hbm_budget) return set_err(ZKG_EINVAL, "sizes exceed the HBM budget");
    if (int rc = check_canonical(coeffs, n_coeffs, "coeffs")) return rc;
    if (int rc = check_canonical(points, n_points, "points")) return rc;
    Scratch s(ctx);
    void *dc, *dp, *dout;
    int rc;
    if ((rc = s.alloc(cb, &dc)) || (rc = s.alloc(pb, &dp)) || (rc = s.alloc(pb, &dout))) return rc;
    hipStream_t st = ctx->stream;
    hipError_t e = n_coeffs ? hipMemcpyAsync(dc, coeffs, cb, hipMemcpyHostToDevice, st) : hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(dp, points, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_poly_eval, dim3((unsigned)((n_points + POLY_THREADS - 1) / POLY_THREADS)),
                           dim3(POLY_THREADS), 0, st, (const uint64_t *)dc, n_coeffs, (const uint64_t *)dp, n_points,
                           (uint64_t *)dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, pb, hipMemcpyDeviceToHost, st);
    const hipError_t done = hipStreamSynchronize(st);  // also on error: nothing in flight reads freed scratch
    if (e == hipSuccess) e = done;
    return e == hipSuccess ? ZKG_OK : hip_err(e, "poly eval");
}

extern "C" int zkg_poly_coset_eval_goldilocks(zkg_ctx *ctx, const uint64_t *coeffs, uint32_t log_n, uint64_t shift,
                                              uint64_t *out)
{
    if (!ctx || !coeffs || !out) return set_err(ZKG_EINVAL, "null argument");
    if (log_n > 22) return set_err(ZKG_EINVAL, "log_n %u above 22", log_n);
    if (shift >= Gl::P) return set_err(ZKG_EINVAL, "shift not canonical");
    const uint64_t n = 1ull << log_n, bytes = n * 8;
    if (int rc = check_canonical(coeffs, n, "coeffs")) return rc;
    Scratch s(ctx);
    void *d, *tmp = nullptr;
    int rc;
    if ((rc = s.alloc(bytes, &d))) return rc;
    const uint64_t tb = ntt_tmp_bytes(FIELD_GL, n, 1);
    if (tb && (rc = s.alloc(tb, &tmp))) return rc;
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(d, coeffs, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const uint64_t lanes = (n + RUN - 1) / RUN;
        hipLaunchKernelGGL(k_poly_shift, dim3((unsigned)((lanes + POLY_THREADS - 1) / POLY_THREADS)),
                           dim3(POLY_THREADS), 0, st, (uint64_t *)d, n, shift);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n > 1) e = launch_ntt(FIELD_GL, ctx->tw[FIELD_GL], d, tmp, n, 1, 0, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, st);
    const hipError_t done = hipStreamSynchronize(st);  // also on error: nothing in flight reads freed scratch
    if (e == hipSuccess) e = done;
    return e == hipSuccess ? ZKG_OK : hip_err(e, "poly coset eval");
}

extern "C" int zkg_poly_quotient_goldilocks(zkg_ctx *ctx, const uint64_t *evals, uint64_t n, uint64_t z, uint64_t f_z,
                                            uint64_t *out)
{
    if (!ctx || !evals || !out) return set_err(ZKG_EINVAL, "null argument");
    if (!n || (n & (n - 1)) || n > (1ull << 32)) return set_err(ZKG_EINVAL, "length %llu not a power of two <= 2^32",
                                                                 (unsigned long long)n);
    if (z >= Gl::P || f_z >= Gl::P) return set_err(ZKG_EINVAL, "z and f(z) must be canonical");
    if (span_bytes(n, 8) > ctx->hbm_budget) return set_err(ZKG_EINVAL, "length exceeds the HBM budget");
    if (int rc = check_canonical(evals, n, "evals")) return rc;
    int log_n = 0;
    while ((1ull << log_n) < n) log_n++;
    Scratch s(ctx);
    void *de, *dq;
    int rc;
    if ((rc = s.alloc(n * 8, &de)) || (rc = s.alloc(n * 8, &dq))) return rc;
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(de, evals, n * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const uint64_t lanes = (n + RUN - 1) / RUN;
        hipLaunchKernelGGL(k_poly_quotient, dim3((unsigned)((lanes + POLY_THREADS - 1) / POLY_THREADS)),
                           dim3(POLY_THREADS), 0, st, (const uint64_t *)de, n, gl_two_adic_generator(log_n), z, f_z,
                           (uint64_t *)dq);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dq, n * 8, hipMemcpyDeviceToHost, st);
    const hipError_t done = hipStreamSynchronize(st);  // also on error: nothing in flight reads freed scratch
    if (e == hipSuccess) e = done;
    return e == hipSuccess ? ZKG_OK : hip_err(e, "poly quotient");
}
