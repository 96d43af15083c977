// Instruction-level model of the device's lean Goldilocks products
// (gl_asm::mul / gl_asm::mul_add / gl_asm::reduce in ziros_amd/csrc/zkg_field.h),
// checked against exact 128-bit arithmetic on the host.
//
// Every step below is one instruction of the device sequence with its
// documented semantics: v_mad_u64_u32 (32x32 + 64 -> 64 and a carry-out lane
// mask), v_sub_co / v_subb (32-bit borrow chain), v_add_co / v_addc (32-bit
// carry chain), v_cndmask (select by a lane mask).  The device code itself is
// checked bit for bit by the GPU parity tests; this file pins the case analysis
// that makes the sequence exact: which carry or borrow is applied, with which
// correction constant, and that no correction wraps.
#include <stdint.h>
#include <stdio.h>
#include <random>

typedef unsigned __int128 u128;
static const uint64_t P = 0xFFFFFFFF00000001ull;

static uint64_t mad(uint32_t a, uint32_t b, uint64_t c, uint32_t &carry)
{
    const u128 r = (u128)a * b + c;
    carry = (uint32_t)(r >> 64);
    return (uint64_t)r;
}
static uint32_t addc(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout)
{
    const uint64_t r = (uint64_t)a + b + cin;
    cout = (uint32_t)(r >> 32);
    return (uint32_t)r;
}
static uint32_t subb(uint32_t a, uint32_t b, uint32_t bin, uint32_t &bout)
{
    const uint64_t r = (uint64_t)a - b - bin;
    bout = (uint32_t)(r >> 63);
    return (uint32_t)r;
}

static long rare_path = 0, wraps = 0;

// reduce: d = lo - h1 (borrow b); r = h0 (2^32 - 1) + d (carry c);
// r += c ? 2^32 - 1 : 0; then, when some lane borrowed, r += b ? 2^64 - (2^32 - 1) : 0.
static uint64_t reduce(uint64_t lo, uint32_t h0, uint32_t h1)
{
    uint32_t b0, b, c, k;
    const uint32_t d0 = subb((uint32_t)lo, h1, 0, b0);
    const uint32_t d1 = subb((uint32_t)(lo >> 32), 0, b0, b);
    const uint64_t r = mad(h0, 0xFFFFFFFFu, ((uint64_t)d1 << 32) | d0, c);
    uint32_t r0 = addc((uint32_t)r, c ? 0xFFFFFFFFu : 0u, 0, k);
    uint32_t r1 = addc((uint32_t)(r >> 32), 0, k, k);
    if (k) wraps++;  // the carry correction must not wrap
    if (b) {
        rare_path++;
        uint32_t k2;
        r0 = addc(r0, 1u, 0, k2);
        r1 = addc(r1, 0xFFFFFFFFu, k2, k2);
        if (!c && !k2) wraps++;  // alone, the borrow correction must come back below 2^64 (a wrap of the sum)
    }
    return ((uint64_t)r1 << 32) | r0;
}
static uint64_t tail(uint64_t p, uint64_t x, uint32_t cx, uint64_t h)
{
    uint32_t c1, c2, c3;
    const uint32_t lo1 = addc((uint32_t)(p >> 32), (uint32_t)x, 0, c1);
    const uint32_t h0 = addc((uint32_t)h, (uint32_t)(x >> 32), c1, c2);
    const uint32_t h1 = addc((uint32_t)(h >> 32), cx, c2, c3);
    if (c3) wraps++;  // the high word of a 128-bit value fits
    return reduce(((uint64_t)lo1 << 32) | (uint32_t)p, h0, h1);
}
static uint64_t mul(uint64_t a, uint64_t b)
{
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint32_t t, cx;
    uint64_t x = mad(a0, b1, 0, t);
    x = mad(a1, b0, x, cx);
    return tail(mad(a0, b0, 0, t), x, cx, mad(a1, b1, 0, t));
}
static uint64_t mul_add(uint64_t a, uint64_t b, uint64_t k)
{
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint32_t t, cx;
    uint64_t x = mad(a0, b1, (uint32_t)(k >> 32), t);
    if (t) wraps++;
    x = mad(a1, b0, x, cx);
    const uint64_t p = mad(a0, b0, (uint32_t)k, t);
    if (t) wraps++;
    return tail(p, x, cx, mad(a1, b1, 0, t));
}

int main()
{
    std::mt19937_64 g(20261204);
    const uint64_t edge[] = {0, 1, 2, 3, P - 1, P, P + 1, 0xFFFFFFFFull, 0x100000000ull, 0x1FFFFFFFFull,
                             0xFFFFFFFFFFFFFFFFull, 0xFFFFFFFFFFFFFFFEull, 0xFFFFFFFF00000000ull,
                             0x8000000000000000ull, 0x7FFFFFFFFFFFFFFFull, 0xFFFFFFFEFFFFFFFFull};
    long bad = 0, n = 0;
    for (uint64_t a : edge)
        for (uint64_t b : edge) {
            bad += (u128)mul(a, b) % P != ((u128)a * b) % P;
            n++;
            for (uint64_t k : edge) {
                bad += (u128)mul_add(a, b, k) % P != ((u128)a * b + k) % P;
                n++;
            }
        }
    // the reduction alone, over its whole input range, with lo < h1 (the rare
    // borrow) and h0 at its extremes forced often
    for (long i = 0; i < 3000000; i++) {
        uint64_t lo = g();
        uint32_t h0 = (uint32_t)g(), h1 = (uint32_t)g();
        if (i % 3 == 0) lo &= 0xFFFFFFFFull;
        if (i % 5 == 0) h0 = (i & 8) ? 0u : 0xFFFFFFFFu;
        if (i % 7 == 0) lo = h1 ? h1 - 1 : 0;
        const u128 v = (u128)lo + ((u128)h0 << 64) + ((u128)h1 << 96);
        bad += (u128)reduce(lo, h0, h1) % P != v % P;
        n++;
    }
    for (long i = 0; i < 3000000; i++) {
        uint64_t a = g(), b = g(), k = g();
        if (i & 1) a |= 0xFFFFFFFF00000000ull;
        if (i & 2) b = ~(b & 0xFFFF);
        if (i & 4) k = ~(k & 0xFF);
        bad += (u128)mul_add(a, b, k) % P != ((u128)a * b + k) % P;
        bad += (u128)mul(a, b) % P != ((u128)a * b) % P;
        n += 2;
    }
    printf("checked %ld; mismatches %ld; wraps %ld; rare path %ld\n", n, bad, wraps, rare_path);
    return bad || wraps || rare_path < 100000;
}
