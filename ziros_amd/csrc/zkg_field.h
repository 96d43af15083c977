// zkg_field.h -- exact prime-field arithmetic for gfx950 VALU.
//
// Goldilocks p = 2^64 - 2^32 + 1 and BabyBear p = 2^31 - 2^27 + 1 (with the
// degree-4 binomial extension x^4 - 11).  Semantics follow the reference's
// field helpers zkf-metal/src/shaders/field_goldilocks.metal:1-56 and
// field_babybear.metal:1-25; the representation is this build's own:
//   * Goldilocks values are canonical u64; a 64x64 product is four
//     v_mad_u64_u32 partial products reduced with 2^64 = 2^32 - 1, 2^96 = -1.
//   * BabyBear values inside kernels are Montgomery form (R = 2^32); HBM holds
//     canonical u32 and kernels convert at load/store.  A product is
//     mad_u64_u32 + mul_lo + mad_u64_u32 (REDC), one conditional subtract.
// All ops keep outputs fully reduced, so any value written back is canonical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zkg {

// ---------------------------------------------------------------- Goldilocks
// 64-bit values are handled as 32-bit limbs with explicit carry chains
// (v_add_co/v_addc, v_sub_co/v_subb): a carry selects a correction instead of
// a 64-bit compare.  Inputs and outputs are canonical (< p).
__host__ __device__ inline uint64_t gl_join(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

// a + b mod 2^64 and the carry out
__host__ __device__ inline uint64_t add_c(uint64_t a, uint64_t b, uint32_t &c)
{
    uint32_t c0, c1;
    uint32_t lo = __builtin_addc((uint32_t)a, (uint32_t)b, 0u, &c0);
    uint32_t hi = __builtin_addc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c0, &c1);
    c = c1;
    return gl_join(lo, hi);
}
// a - b mod 2^64 and the borrow out
__host__ __device__ inline uint64_t sub_b(uint64_t a, uint64_t b, uint32_t &bw)
{
    uint32_t b0, b1;
    uint32_t lo = __builtin_subc((uint32_t)a, (uint32_t)b, 0u, &b0);
    uint32_t hi = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), b0, &b1);
    bw = b1;
    return gl_join(lo, hi);
}

#if defined(__HIP_DEVICE_COMPILE__)
// Lazily reduced Goldilocks products on gfx950 that keep their carries in SGPR
// lane masks: the carry-out of v_mad_u64_u32 and of the add/sub chains.  A
// product (also one with an addend) issues 14 VALU instructions on its common
// path, none of them a move; the same arithmetic written in C compiles to
// 22-24, with 64-bit compares standing in for the carries and moves building
// the aligned register pairs of the 64-bit addends.  (gfx950 pairs moves and
// plain 32-bit adds in one issue slot, but no instruction that reads or
// writes a carry, so the saving in issue slots is about a fifth.)  The chains
// are longer, which costs where few waves hide latency (Gl::*_lean).
// Every asm block that reads an SGPR written by a VALU instruction waits two
// cycles first (s_nop 1), the wait gfx950 needs between the two.
namespace gl_asm {

// a0 b1 + a1 b0 = x + 2^64 cx (cx as 0 / 1)
__device__ __forceinline__ uint64_t cross(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, uint32_t &cx)
{
    uint64_t x, m;
    asm("v_mad_u64_u32 %0, %1, %2, %3, 0\n\t"
        "v_mad_u64_u32 %0, %1, %4, %5, %0"
        : "=&v"(x), "=&s"(m)
        : "v"(a0), "v"(b1), "v"(a1), "v"(b0));
    asm("s_nop 1\n\tv_cndmask_b32_e64 %0, 0, 1, %1" : "=v"(cx) : "s"(m));
    return x;
}

// lo + 2^64 h0 + 2^96 h1 == lo + h0 (2^32 - 1) - h1 (mod p), as a u64
// representative.  d = lo - h1 with borrow b, r = h0 (2^32 - 1) + d with carry
// c (2^64 == 2^32 - 1), result r + (c - b)(2^32 - 1).  With c alone r is below
// (2^32 - 1)^2, so adding 2^32 - 1 cannot wrap; with b alone r >= d > 2^64 -
// 2^32, so subtracting it cannot borrow; c with b cancel.  The carry is set
// for about half the lanes and is always applied; the borrow needs lo < h1 <
// 2^32 and almost never occurs, so its correction runs only when some lane of
// the wave has it (a scalar branch on the wave's mask).
__device__ __forceinline__ uint64_t reduce(uint64_t lo, uint32_t h0, uint32_t h1)
{
    uint32_t d0, d1;
    uint64_t b;
    asm("v_sub_co_u32_e64 %0, %2, %3, %5\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %1, %2, %4, 0, %2"
        : "=&v"(d0), "=&v"(d1), "=&s"(b)
        : "v"((uint32_t)lo), "v"((uint32_t)(lo >> 32)), "v"(h1));
    uint64_t r, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(c) : "v"(h0), "v"(((uint64_t)d1 << 32) | d0));
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32), t0, t1;
    uint64_t k;
    asm("s_nop 1\n\t"
        "v_cndmask_b32_e64 %2, 0, -1, %6\n\t"  // c: + (2^32 - 1)
        "v_add_co_u32_e64 %0, %4, %0, %2\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %1, %4, %1, 0, %4\n\t"
        "s_cmp_lg_u64 %5, 0\n\t"
        "s_cbranch_scc0 .Lzkg_gl_reduce%=\n\t"
        "v_cndmask_b32_e64 %2, 0, 1, %5\n\t"  // b: + 2^64 - (2^32 - 1)
        "v_cndmask_b32_e64 %3, 0, -1, %5\n\t"
        "v_add_co_u32_e64 %0, %4, %0, %2\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %1, %4, %1, %3, %4\n"
        ".Lzkg_gl_reduce%=:"
        : "+v"(r0), "+v"(r1), "=&v"(t0), "=&v"(t1), "=&s"(k)
        : "s"(b), "s"(c)
        : "scc");
    return ((uint64_t)r1 << 32) | r0;
}

// a b for any u64 representatives.  a b = P + 2^32 (x + 2^64 cx) + 2^64 H with
// P = a0 b0, H = a1 b1; the 128-bit value is lo + 2^64 (h0 + 2^32 h1).
__device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b)
{
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint32_t cx, c1, c2, c3;
    const uint64_t x = cross(a0, a1, b0, b1, cx);
    const uint64_t p = (uint64_t)a0 * b0, h = (uint64_t)a1 * b1;
    const uint32_t lo1 = __builtin_addc((uint32_t)(p >> 32), (uint32_t)x, 0u, &c1);
    const uint32_t h0 = __builtin_addc((uint32_t)h, (uint32_t)(x >> 32), c1, &c2);
    const uint32_t h1 = __builtin_addc((uint32_t)(h >> 32), cx, c2, &c3);  // the true high word fits: c3 = 0
    return reduce(((uint64_t)lo1 << 32) | (uint32_t)p, h0, h1);
}

// a b + k for any u64 representatives: the halves of k ride in the addends of
// a0 b0 + k0 and a0 b1 + k1, neither of which can pass 2^64 - 1.
__device__ __forceinline__ uint64_t mul_add(uint64_t a, uint64_t b, uint64_t k)
{
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint32_t cx, c1, c2, c3;
    uint64_t x, m;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %6\n\t"
        "v_mad_u64_u32 %0, %1, %4, %5, %0"
        : "=&v"(x), "=&s"(m)
        : "v"(a0), "v"(b1), "v"(a1), "v"(b0), "v"((uint64_t)(uint32_t)(k >> 32)));
    asm("s_nop 1\n\tv_cndmask_b32_e64 %0, 0, 1, %1" : "=v"(cx) : "s"(m));
    const uint64_t p = (uint64_t)a0 * b0 + (uint32_t)k, h = (uint64_t)a1 * b1;
    const uint32_t lo1 = __builtin_addc((uint32_t)(p >> 32), (uint32_t)x, 0u, &c1);
    const uint32_t h0 = __builtin_addc((uint32_t)h, (uint32_t)(x >> 32), c1, &c2);
    const uint32_t h1 = __builtin_addc((uint32_t)(h >> 32), cx, c2, &c3);
    return reduce(((uint64_t)lo1 << 32) | (uint32_t)p, h0, h1);
}

}  // namespace gl_asm
#endif

struct Gl {
    using T = uint64_t;
    static constexpr uint64_t P = 0xFFFFFFFF00000001ULL;
    static constexpr uint64_t EPS = 0xFFFFFFFFULL;  // 2^64 mod p

    // s = a + b; t = s + EPS = s - p (mod 2^64).  The sum reaches p exactly
    // when it wrapped (then t is the reduced value and does not wrap again)
    // or when t wraps.
    __host__ __device__ static inline T add(T a, T b)
    {
        uint32_t c, c2;
        T s = add_c(a, b, c);
        T t = add_c(s, EPS, c2);
        return (c | c2) ? t : s;
    }
    // d = a - b; on borrow the true value d - 2^64 equals d - EPS (mod p), and
    // d > EPS then, so the correction cannot borrow again.
    __host__ __device__ static inline T sub(T a, T b)
    {
        uint32_t bw, bw2;
        T d = sub_b(a, b, bw);
        return sub_b(d, bw ? EPS : 0, bw2);
    }
    __host__ __device__ static inline T neg(T a) { return a ? P - a : 0; }
    __host__ __device__ static inline T dbl(T a) { return add(a, a); }
    __host__ __device__ static inline T sub_for_mul(T a, T b) { return sub(a, b); }

    // canonical representative of any u64: x >= p iff x + EPS wraps
    __host__ __device__ static inline T canon(T x)
    {
        uint32_t c;
        T t = add_c(x, EPS, c);
        return c ? t : x;
    }
    // x = lo + 2^64 x2 + 2^96 x3 == lo + x2 (2^32 - 1) - x3 (mod p), as some u64
    // representative (not necessarily < p).  r = x2 EPS + lo is one
    // v_mad_u64_u32; a wrap of it (+2^64) leaves r < (2^32 - 1)^2 and a borrow
    // of r - x3 (-2^64) leaves r - x3 > 2^64 - 2^32, so the net correction
    // (wrap - borrow) EPS never wraps again.
    __host__ __device__ static inline T reduce128_lazy(uint64_t lo, uint64_t hi)
    {
        const uint32_t x2 = (uint32_t)hi, x3 = (uint32_t)(hi >> 32);
        uint64_t r = (uint64_t)x2 * EPS + lo;
        const bool c = r < lo;
        const bool b = r < x3;
        r -= x3;
        if (c != b) r += c ? EPS : (0 - EPS);
        return r;
    }
    __host__ __device__ static inline T reduce128(uint64_t lo, uint64_t hi) { return canon(reduce128_lazy(lo, hi)); }
    // 128-bit product of any two u64
    __host__ __device__ static inline void mul_wide(T a, T b, uint64_t &lo, uint64_t &hi)
    {
        uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
        uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
        uint64_t p00 = (uint64_t)a0 * b0;
        uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);            // < 2^64
        uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;            // < 2^64
        hi = (uint64_t)a1 * b1 + ((t >> 32) + (u >> 32));
        lo = ((uint64_t)(uint32_t)u << 32) | (uint32_t)p00;
    }
    // 128-bit square of any u64 with three partial products:
    // a0^2 + a0 a1 2^33 + a1^2 2^64.  lo = a0^2 + (a0 a1 << 33) mod 2^64 and
    // hi = a1^2 + (a0 a1 >> 31) + carry is the exact high word (< 2^64).
    __host__ __device__ static inline void sqr_wide(T a, uint64_t &lo, uint64_t &hi)
    {
        const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
        const uint64_t p00 = (uint64_t)a0 * a0;
        const uint64_t m = (uint64_t)a0 * a1;
        uint32_t c;
        lo = add_c(p00, m << 33, c);
        hi = (uint64_t)a1 * a1 + ((m >> 31) + c);
    }
    __host__ __device__ static inline T mul(T a, T b)
    {
        uint64_t lo, hi;
        mul_wide(a, b, lo, hi);
        return reduce128(lo, hi);
    }
    // a^2 for any u64 representative, result a u64 representative
    __host__ __device__ static inline T sqr_lazy(T a)
    {
        uint64_t lo, hi;
        sqr_wide(a, lo, hi);
        return reduce128_lazy(lo, hi);
    }
    // a b for any u64 representatives, result a u64 representative
    __host__ __device__ static inline T mul_lazy(T a, T b)
    {
        uint64_t lo, hi;
        mul_wide(a, b, lo, hi);
        return reduce128_lazy(lo, hi);
    }
    // a b + k (k any u64): the 128-bit product plus k cannot pass 2^128 - 1
    __host__ __device__ static inline T mul_add_lazy(T a, T b, T k)
    {
        uint64_t lo, hi;
        mul_wide(a, b, lo, hi);
        uint32_t c;
        lo = add_c(lo, k, c);
        return reduce128_lazy(lo, hi + c);
    }
    // The same three products in the fewest issue slots (gl_asm on the
    // device): for kernels with enough waves to hide their longer dependency
    // chains (one state per lane over many rows).  Kernels bound by latency
    // (a state spread over lanes, short tree levels) keep the forms above.
    __host__ __device__ static inline T sqr_lean(T a)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return gl_asm::mul(a, a);
#else
        return sqr_lazy(a);
#endif
    }
    __host__ __device__ static inline T mul_lean(T a, T b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return gl_asm::mul(a, b);
#else
        return mul_lazy(a, b);
#endif
    }
    __host__ __device__ static inline T mul_add_lean(T a, T b, T k)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return gl_asm::mul_add(a, b, k);
#else
        return mul_add_lazy(a, b, k);
#endif
    }
    // a + k for any u64 a and canonical k: a wrap leaves a + k - 2^64 < k, so
    // adding 2^64 mod p = EPS once stays below 2^64.
    __host__ __device__ static inline T add_const_lazy(T a, T k)
    {
        T s = a + k;
        return s + (s < k ? EPS : 0);
    }
    __host__ __device__ static inline T sqr(T a) { return mul(a, a); }
    // x 2^k for canonical x and 0 <= k < 96, canonical result (2^96 = -1
    // covers the other exponents by a sign).  The small roots of unity are
    // powers of two (omega_4 = 2^48, omega_8 = 2^120, omega_16 = 2^156 for
    // the generator powers the transforms use), so a butterfly's twiddle inside
    // a register sub-DFT is a shift and one reduction instead of a product.
    // k < 64: x 2^k = (x << k) + 2^64 (x >> (64 - k)), one 128-bit reduction.
    // k >= 64: with y = x 2^(k-64) = y0 + 2^32 Y (Y < 2^(32+k-64) < p),
    // x 2^k = 2^64 y0 + 2^96 Y == y0 (2^32 - 1) - Y, and y0 (2^32 - 1) < p.
    // k is a constant wherever the callers' loops unroll.
    __host__ __device__ static inline T mul_pow2(T x, int k)
    {
        if (k == 0) return x;
        if (k < 64) return reduce128(x << k, x >> (64 - k));
        const int j = k - 64;
        const uint64_t a = (uint64_t)(uint32_t)(x << j) * EPS;
        return sub(a, x >> (32 - j));
    }
    // HBM <-> register representation (identity for Goldilocks)
    __host__ __device__ static inline T from_canon(T a) { return a; }
    __host__ __device__ static inline T to_canon(T a) { return a; }
    __host__ __device__ static inline T one() { return 1; }
};

// Sums of a few Goldilocks representatives kept as lo + top 2^64 (top small),
// folded back once with 2^64 = 2^32 - 1 (mod p).  Additions are one
// v_add_co / v_addc / v_addc chain over the three 32-bit words.
struct GlWide {
    uint32_t w0, w1, top;
};
__host__ __device__ inline GlWide glw(uint64_t x) { return GlWide{(uint32_t)x, (uint32_t)(x >> 32), 0u}; }
__host__ __device__ inline GlWide glw_add(GlWide a, uint64_t x)
{
    uint32_t c0, c1, c2;
    GlWide r;
    r.w0 = __builtin_addc(a.w0, (uint32_t)x, 0u, &c0);
    r.w1 = __builtin_addc(a.w1, (uint32_t)(x >> 32), c0, &c1);
    r.top = __builtin_addc(a.top, 0u, c1, &c2);
    (void)c2;
    return r;
}
__host__ __device__ inline GlWide glw_add(GlWide a, GlWide b)
{
    uint32_t c0, c1, c2;
    GlWide r;
    r.w0 = __builtin_addc(a.w0, b.w0, 0u, &c0);
    r.w1 = __builtin_addc(a.w1, b.w1, c0, &c1);
    r.top = __builtin_addc(a.top, b.top, c1, &c2);
    (void)c2;
    return r;
}
__host__ __device__ inline GlWide glw_dbl(uint64_t x)
{
    const uint64_t d = x << 1;
    return GlWide{(uint32_t)d, (uint32_t)(d >> 32), (uint32_t)(x >> 63)};
}
// top EPS + lo < 2^65; if it wraps the remainder is below (2^32 - 1)^2 and
// one more EPS fits.
__host__ __device__ inline uint64_t glw_fold(GlWide a)
{
    const uint64_t lo = gl_join(a.w0, a.w1);
    uint64_t r = (uint64_t)a.top * Gl::EPS + lo;
    return r + (r < lo ? Gl::EPS : 0);
}

// ---------------------------------------------------------------- BabyBear
struct Bb {
    using T = uint32_t;
    static constexpr uint32_t P = 0x78000001u;
    static constexpr uint32_t NPINV = 0x77FFFFFFu;  // -p^-1 mod 2^32
    static constexpr uint32_t R2 = 1172168163u;     // 2^64 mod p
    static constexpr uint32_t R1 = 268435454u;      // 2^32 mod p (Montgomery 1)

    // Reductions are unsigned minima: for x < 2p, x - p wraps above x exactly
    // when x < p, so min(x, x - p) is x mod p with no compare or carry mask.
    __host__ __device__ static inline T umin(T a, T b) { return __builtin_elementwise_min(a, b); }
    __host__ __device__ static inline T add(T a, T b)
    {
        T s = a + b;
        return umin(s, s - P);
    }
    // d = a - b wraps to d > 2^32 - p > p when a < b, and then d + p wraps
    // back below p; otherwise d < p <= d + p.
    __host__ __device__ static inline T sub(T a, T b)
    {
        T d = a - b;
        return umin(d, d + P);
    }
    __host__ __device__ static inline T neg(T a) { return a ? P - a : 0; }
    __host__ __device__ static inline T dbl(T a) { return add(a, a); }
    // a - b + p in [1, 2p): a valid left operand of mul (REDC needs a b < p 2^32
    // and a canonical right operand), one reduction cheaper than sub.
    __host__ __device__ static inline T sub_for_mul(T a, T b) { return a + (P - b); }
    // Montgomery REDC of t < p * 2^32:  (t + m p) / 2^32,  m = t * (-p^-1) mod 2^32
    __host__ __device__ static inline T redc(uint64_t t)
    {
        uint32_t m = (uint32_t)t * NPINV;
        uint64_t u = t + (uint64_t)m * P;
        uint32_t r = (uint32_t)(u >> 32);
        return umin(r, r - P);
    }
    __host__ __device__ static inline T mul(T a, T b) { return redc((uint64_t)a * b); }
    __host__ __device__ static inline T sqr(T a) { return mul(a, a); }
    __host__ __device__ static inline T from_canon(T a) { return mul(a, R2); }
    __host__ __device__ static inline T to_canon(T a) { return redc((uint64_t)a); }
    __host__ __device__ static inline T one() { return R1; }
    // x * 2^-k for 1 <= k <= 27, any representation (Montgomery form commutes
    // with scaling).  p = 1 (mod 2^27) so m = -x mod 2^k clears the low bits:
    // (x + m p) / 2^k = (x + m) / 2^k + m * 15 * 2^(27-k).
    template <int K>
    __host__ __device__ static inline T mul_inv_pow2(T x)
    {
        static_assert(K >= 1 && K <= 27, "shift range");
        uint32_t m = (0u - x) & ((1u << K) - 1u);
        uint32_t r = ((x + m) >> K) + m * (15u << (27 - K));
        return umin(r, r - P);
    }
};

// ------------------------------------------------- BabyBear in exact FP64
// BabyBear elements as exact integers held in doubles, for the permutation
// kernels: MI355X issues FP64 FMA at full rate, sums of field elements stay
// exact below 2^53 without any reduction, and a product needs 6 FP64
// operations.  Values are representatives (any sign, |x| < 2^41 at every
// product input); only digests leaving the kernel are canonicalised.
// Compile with -ffp-contract=off: every fused step below is an explicit fma.
struct Bf {
    using T = double;
    static constexpr double P = 2013265921.0;
    static constexpr double PINV = 1.0 / 2013265921.0;
    static constexpr double MAGIC = 6755399441055744.0;  // 1.5 * 2^52: add-round to integer

    // round(x / p) for |x| < 2^51 * p: x * PINV is exact inside the fma and
    // the sum with MAGIC lies in [2^52, 2^53), whose spacing is 1.
    __host__ __device__ static inline double quot(double x) { return __builtin_fma(x, PINV, MAGIC) - MAGIC; }
    // a * b mod p with |a|, |b| < 2^41: h + l = a b exactly (l from fma), q is
    // round(ab/p) up to +-1, and h - q p is an integer of magnitude < 2^32, so
    // the fma computing it and the final add are exact.  |result| < 1.6 p.
    __host__ __device__ static inline double mul(double a, double b)
    {
        double h = a * b;
        double l = __builtin_fma(a, b, -h);
        double q = quot(h);
        return __builtin_fma(-q, P, h) + l;
    }
    // x mod p into [-p/2 - 1, p/2 + 1] for |x| < 2^52 (exact).
    __host__ __device__ static inline double red(double x) { return __builtin_fma(-quot(x), P, x); }
    // x * (+-2^-K mod p) for an integer |x| < 2^52, 1 <= K <= 27, in 4 operations
    // instead of a product's 6.  y = +-x 2^-K is exact, f = y - floor(y) = r 2^-K
    // with r = (+-x) mod 2^K, and y - f p = floor(y) - r 15 2^(27-K) is an integer
    // (p - 1 = 15 2^27) whose 2^K-multiple is +-x - r p.  Every step is exact;
    // |result| < |x| 2^-K + 2^31.
    template <int K, bool NEG>
    __host__ __device__ static inline double mul_inv_pow2(double x)
    {
        static_assert(K >= 1 && K <= 27, "shift range");
        constexpr double c = (NEG ? -1.0 : 1.0) / (double)(1u << K);
        double y = x * c;
        // y - floor(y): at most K fraction bits, so exact, and never rounded up
        // to 1; on the device one v_fract_f64
#if defined(__HIP_DEVICE_COMPILE__)
        double f = __builtin_amdgcn_fract(y);
#else
        double f = y - __builtin_floor(y);
#endif
        return __builtin_fma(-f, P, y);
    }
    // canonical u32 in [0, p)
    __host__ __device__ static inline uint32_t canon(double x)
    {
        double r = red(x);
        r = r < 0.0 ? r + P : r;
        return (uint32_t)r;
    }
    __host__ __device__ static inline double from_u32(uint32_t v) { return (double)v; }
};

// ------------------------------------------------- BabyBear degree-4 extension
// EF4 = Bb[x] / (x^4 - 11), coefficients low-first; components in Montgomery.
struct Ef4 {
    Bb::T c[4];
};

__host__ __device__ static inline Ef4 ef4_add(const Ef4 &a, const Ef4 &b)
{
    Ef4 r;
#pragma unroll
    for (int i = 0; i < 4; i++) r.c[i] = Bb::add(a.c[i], b.c[i]);
    return r;
}
__host__ __device__ static inline Ef4 ef4_sub(const Ef4 &a, const Ef4 &b)
{
    Ef4 r;
#pragma unroll
    for (int i = 0; i < 4; i++) r.c[i] = Bb::sub(a.c[i], b.c[i]);
    return r;
}
__host__ __device__ static inline Ef4 ef4_mul_base(const Ef4 &a, Bb::T s)
{
    Ef4 r;
#pragma unroll
    for (int i = 0; i < 4; i++) r.c[i] = Bb::mul(a.c[i], s);
    return r;
}
// Schoolbook with lazy accumulation: each of the 7 convolution sums holds at
// most 4 products < p^2 < 2^62 only after REDC, so reduce per product and add.
__host__ __device__ static inline Ef4 ef4_mul(const Ef4 &a, const Ef4 &b, Bb::T w_mont)
{
    Bb::T c0 = Bb::mul(a.c[0], b.c[0]);
    Bb::T c1 = Bb::add(Bb::mul(a.c[0], b.c[1]), Bb::mul(a.c[1], b.c[0]));
    Bb::T c2 = Bb::add(Bb::add(Bb::mul(a.c[0], b.c[2]), Bb::mul(a.c[1], b.c[1])), Bb::mul(a.c[2], b.c[0]));
    Bb::T c3 = Bb::add(Bb::add(Bb::mul(a.c[0], b.c[3]), Bb::mul(a.c[1], b.c[2])),
                       Bb::add(Bb::mul(a.c[2], b.c[1]), Bb::mul(a.c[3], b.c[0])));
    Bb::T c4 = Bb::add(Bb::add(Bb::mul(a.c[1], b.c[3]), Bb::mul(a.c[2], b.c[2])), Bb::mul(a.c[3], b.c[1]));
    Bb::T c5 = Bb::add(Bb::mul(a.c[2], b.c[3]), Bb::mul(a.c[3], b.c[2]));
    Bb::T c6 = Bb::mul(a.c[3], b.c[3]);
    Ef4 r;
    r.c[0] = Bb::add(c0, Bb::mul(c4, w_mont));
    r.c[1] = Bb::add(c1, Bb::mul(c5, w_mont));
    r.c[2] = Bb::add(c2, Bb::mul(c6, w_mont));
    r.c[3] = c3;
    return r;
}

}  // namespace zkg
