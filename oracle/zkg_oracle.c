/*
 * zkg_oracle.c -- CPU restatement of the ZirOS STARK trace-commitment path.
 *
 * TEST INFRASTRUCTURE ONLY: the bit-exactness checker for libzkg and the CPU
 * baseline timed beside it.  Nothing in the product path links or calls it.
 *
 * The reference computes this path with Plonky3 0.4.2 on the CPU
 * (zkf-backends/src/plonky3.rs:63-65,230-233) and with Metal shaders on macOS
 * (the .metal files under zkf-metal/src/shaders).  Plonky3 and rand are external crates that
 * are not vendored in the reference; where a function restates their
 * published semantics it names the crate, version and the reference call site.
 *
 * Parity pins: see zkg_oracle.h.  Functions marked "parity unpinned" have no
 * reference vector that could fail them; their tests check invariants only.
 */
#include "zkg_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

#define GL_EPS 0xFFFFFFFFULL /* 2^64 mod p = 2^32 - 1 */

/* ======================================================================
 * Goldilocks p = 2^64 - 2^32 + 1
 * Follows zkf-metal/src/shaders/field_goldilocks.metal:1-56 and the CPU twin
 * zkf-core/src/acceleration.rs:311-380 (reduce 2^64 = 2^32 - 1, 2^96 = -1).
 * ====================================================================== */
/* The conditional corrections below are written as masks (m = 0 or all
 * ones) rather than branches: on random data the branches mispredict and cost
 * more than the arithmetic.  The values are the same either way. */
static inline uint64_t mask_if(int c) { return (uint64_t)0 - (uint64_t)(c != 0); }

uint64_t or_gl_add(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    s += GL_EPS & mask_if(s < a);             /* wrapped past 2^64: add 2^64 mod p */
    s -= OR_GL_P & mask_if(s >= OR_GL_P);
    return s;
}

uint64_t or_gl_sub(uint64_t a, uint64_t b)
{
    uint64_t d = a - b;
    d -= GL_EPS & mask_if(a < b);             /* borrowed 2^64: subtract 2^64 mod p */
    return d;
}

/* 128-bit product x = lo + 2^64 (a + 2^32 b) == lo - b + a (2^32 - 1). */
static inline uint64_t gl_reduce128(u128 x)
{
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t hh = hi >> 32, hl = hi & GL_EPS;
    uint64_t t0 = lo - hh;
    t0 -= GL_EPS & mask_if(lo < hh);
    uint64_t t1 = hl * GL_EPS;
    uint64_t r = t0 + t1;
    r += GL_EPS & mask_if(r < t0);
    r -= OR_GL_P & mask_if(r >= OR_GL_P);
    return r;
}

uint64_t or_gl_mul(uint64_t a, uint64_t b) { return gl_reduce128((u128)a * b); }

uint64_t or_gl_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = or_gl_mul(r, a);
        a = or_gl_mul(a, a);
        e >>= 1;
    }
    return r;
}

/* ======================================================================
 * BabyBear p = 2^31 - 2^27 + 1, canonical representation.
 * Follows zkf-metal/src/shaders/field_babybear.metal:1-25 (canonical add/sub,
 * reduction of the 62-bit product).
 * ====================================================================== */
static inline uint64_t bb_add(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    return s - (OR_BB_P & mask_if(s >= OR_BB_P));
}
static inline uint64_t bb_sub(uint64_t a, uint64_t b) { return a - b + (OR_BB_P & mask_if(a < b)); }
uint64_t or_bb_mul(uint64_t a, uint64_t b) { return (a * b) % OR_BB_P; }
uint64_t or_bb_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = or_bb_mul(r, a);
        a = or_bb_mul(a, a);
        e >>= 1;
    }
    return r;
}

uint64_t or_f_add(int f, uint64_t a, uint64_t b) { return f == OR_GOLDILOCKS ? or_gl_add(a, b) : bb_add(a, b); }
uint64_t or_f_sub(int f, uint64_t a, uint64_t b) { return f == OR_GOLDILOCKS ? or_gl_sub(a, b) : bb_sub(a, b); }
uint64_t or_f_mul(int f, uint64_t a, uint64_t b) { return f == OR_GOLDILOCKS ? or_gl_mul(a, b) : or_bb_mul(a, b); }
uint64_t or_f_pow(int f, uint64_t a, uint64_t e) { return f == OR_GOLDILOCKS ? or_gl_pow(a, e) : or_bb_pow(a, e); }
static inline uint64_t f_p(int f) { return f == OR_GOLDILOCKS ? OR_GL_P : OR_BB_P; }
uint64_t or_f_inv(int f, uint64_t a) { return or_f_pow(f, a, f_p(f) - 2); }

/* Multiplicative generators: p3 `Field::GENERATOR` = 7 (Goldilocks), 31
 * (BabyBear); the LDE shift used by TwoAdicFriPcs (Val::GENERATOR). */
uint64_t or_f_generator(int f) { return f == OR_GOLDILOCKS ? 7 : 31; }

/* `TwoAdicField::two_adic_generator(bits)`: the 2^32-th (Goldilocks) or
 * 2^27-th (BabyBear) root of unity g^((p-1)/2^adicity) squared down to order
 * 2^bits.  The order-2^adicity roots are 0x185629dcda58878c and 0x1a427a41;
 * both equal GENERATOR^((p-1)/2^adicity), which tests check. */
uint64_t or_f_two_adic_generator(int f, uint32_t log_n)
{
    uint64_t w = f == OR_GOLDILOCKS ? 0x185629dcda58878cULL : 0x1a427a41ULL;
    uint32_t adicity = f == OR_GOLDILOCKS ? 32 : 27;
    for (uint32_t i = log_n; i < adicity; i++) w = or_f_mul(f, w, w);
    return w;
}

/* EF4 = BabyBear[x]/(x^4 - W), W = 11 (p3-baby-bear 0.4.2
 * BinomiallyExtendable<4>); schoolbook product with the wrap multiplied by W. */
void or_ef4_mul(const uint64_t *a, const uint64_t *b, uint64_t *out)
{
    uint64_t c[7] = {0};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) c[i + j] = bb_add(c[i + j], or_bb_mul(a[i], b[j]));
    for (int k = 0; k < 4; k++) out[k] = k < 3 ? bb_add(c[k], or_bb_mul(11, c[k + 4])) : c[k];
}
static inline void ef4_add(const uint64_t *a, const uint64_t *b, uint64_t *o)
{
    for (int i = 0; i < 4; i++) o[i] = bb_add(a[i], b[i]);
}
static inline void ef4_sub(const uint64_t *a, const uint64_t *b, uint64_t *o)
{
    for (int i = 0; i < 4; i++) o[i] = bb_sub(a[i], b[i]);
}
static inline void ef4_mul_base(const uint64_t *a, uint64_t s, uint64_t *o)
{
    for (int i = 0; i < 4; i++) o[i] = or_bb_mul(a[i], s);
}

/* ======================================================================
 * rand 0.9.2 SmallRng on 64-bit targets = Xoshiro256PlusPlus, seeded by
 * `seed_from_u64` through SplitMix64 (rand rngs/xoshiro256plusplus.rs).
 * Used by plonky3.rs:606-620 (`SmallRng::seed_from_u64(seed)`) and
 * zkf-metal/src/poseidon2/{goldilocks,babybear}.rs (flatten_round_constants).
 * Pinned through the Goldilocks Poseidon2 golden vector.
 * ====================================================================== */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void or_rng_seed(or_rng *r, uint64_t state)
{
    for (int i = 0; i < 4; i++) {
        state += 0x9e3779b97f4a7c15ULL;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        r->s[i] = z ^ (z >> 31);
    }
}

uint64_t or_rng_next_u64(or_rng *r)
{
    uint64_t *s = r->s;
    uint64_t out = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}

/* Xoshiro's low bits are weaker; rand returns the high half for u32. */
uint32_t or_rng_next_u32(or_rng *r) { return (uint32_t)(or_rng_next_u64(r) >> 32); }

/* ======================================================================
 * Synthetic trace generator (this build's own documented contract; the
 * reference defines none).  Element i of stream `seed` is the first
 * SplitMix64 finaliser output mix(seed + (i+1)*phi + a*psi), a = 0,1,...,
 * shifted to the field's bit width, that is below p.  Counter-based, so any
 * slice can be generated independently (the HIP generator uses the same rule).
 * ====================================================================== */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t or_gen_elem(int f, uint64_t seed, uint64_t i)
{
    uint64_t z0 = seed + (i + 1) * 0x9e3779b97f4a7c15ULL;
    for (uint64_t a = 0;; a++) {
        uint64_t v = mix64(z0 + a * 0x632be59bd9b4e019ULL);
        if (f == OR_BABYBEAR) v >>= 33;
        if (v < f_p(f)) return v;
    }
}

void or_gen_fill(int f, uint64_t seed, uint64_t start, uint64_t n, uint64_t *out)
{
#pragma omp parallel for schedule(static) if (n > 65536)
    for (uint64_t i = 0; i < n; i++) out[i] = or_gen_elem(f, seed, start + i);
}

/* ======================================================================
 * Poseidon2
 * Permutation schedule: zkf-metal/src/shaders/poseidon2.metal:17-264 and the
 * Rust restatement zkf-backends/src/wrapping/poseidon2_goldilocks.rs:117-339:
 *   MDS-light; rf/2 x (add RC to all, x^7 all, MDS-light);
 *   rp x (s0 += rc, s0 = s0^7, s_i = d_i s_i + sum); rf/2 x external.
 * MDS-light: circulant M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] on each
 * 4-chunk, then add the per-position sum over chunks
 * (poseidon2_goldilocks.rs:266-317).
 * ====================================================================== */

/* MATRIX_DIAG_16_GOLDILOCKS: zkf-metal/src/poseidon2/goldilocks.rs:21-38
 * (identical copy at zkf-backends/src/wrapping/poseidon2_goldilocks.rs:23-40). */
static const uint64_t GL_DIAG16[16] = {
    0xde9b91a467d6afc0ULL, 0xc5f16b9c76a9be17ULL, 0x0ab0fef2d540ac55ULL, 0x3001d27009d05773ULL,
    0xed23b1f906d3d9ebULL, 0x5ce73743cba97054ULL, 0x1c3bab944af4ba24ULL, 0x2faa105854dbafaeULL,
    0x53ffb3ae6d421a10ULL, 0xbcda9df8884ba396ULL, 0xfc1273e4a31807bbULL, 0xc77952573d5142c0ULL,
    0x56683339a819b85eULL, 0x328fcbd8f0ddc8ebULL, 0xb5101e303fce9cb7ULL, 0x774487b8c40089bbULL,
};

/* MATRIX_DIAG_8_GOLDILOCKS of p3-goldilocks 0.4.2.  Not present anywhere in
 * the reference tree: parity unpinned. */
static const uint64_t GL_DIAG8[8] = {
    0xa98811a1fed4e3a5ULL, 0x1cc48b54f377e2a0ULL, 0xe40cd4f6c5609a26ULL, 0x11de79ebca97a4a3ULL,
    0x9177c73d8b7e929cULL, 0x2a6fe8085797e791ULL, 0x3de6e93329f8d5adULL, 0x3f7af9125da962feULL,
};

/* BabyBear width-16 internal diagonal V (zkf-metal/src/poseidon2/babybear.rs:14-58):
 * [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/4, 1/8, 1/2^27, -1/2^8, -1/16, -1/2^27]. */
static void bb_diag16(uint64_t *d)
{
    const uint64_t p = OR_BB_P;
    uint64_t inv2 = or_f_inv(OR_BABYBEAR, 2);
#define INVP2(e) or_bb_pow(inv2, (e))
    uint64_t v[16] = {p - 2, 1, 2, inv2, 3, 4, p - inv2, p - 3, p - 4, INVP2(8), INVP2(2), INVP2(3),
                      INVP2(27), p - INVP2(8), p - INVP2(4), p - INVP2(27)};
#undef INVP2
    memcpy(d, v, sizeof v);
}

/* One field element from SmallRng with p3's `StandardUniform` sampling:
 *   Goldilocks (p3-goldilocks 0.4.2): next_u64, reject >= p.
 *   BabyBear (p3-monty-31 0.4.2): next_u32 >> 1, reject >= p, and the value is
 *   taken as the Montgomery form (canonical = v * 2^-32).  Parity unpinned:
 *   mode 1 keeps the canonical interpretation for comparison. */
static uint64_t sample_field(or_rng *r, int f, int bb_rc_mode)
{
    if (f == OR_GOLDILOCKS) {
        for (;;) {
            uint64_t v = or_rng_next_u64(r);
            if (v < OR_GL_P) return v;
        }
    }
    for (;;) {
        uint64_t v = or_rng_next_u32(r) >> 1;
        if (v < OR_BB_P) {
            if (bb_rc_mode == 1) return v;
            return or_bb_mul(v, or_bb_pow(or_f_inv(OR_BABYBEAR, 2), 32)); /* v * 2^-32 */
        }
    }
}

/* poseidon2_round_numbers_128 (p3-poseidon2 0.4.2) for D = 7:
 * Goldilocks width 8 and 16 -> (8, 22); BabyBear width 16 -> (8, 13). */
int or_p2_init(or_p2 *p, int field, int width, uint64_t seed, int bb_rc_mode)
{
    memset(p, 0, sizeof *p);
    if (!((field == OR_GOLDILOCKS && (width == 8 || width == 16)) ||
          (field == OR_BABYBEAR && width == 16)))
        return -1;
    p->field = field;
    p->width = width;
    p->rounds_f = 8;
    p->rounds_p = field == OR_GOLDILOCKS ? 22 : 13;
    if (field == OR_GOLDILOCKS)
        memcpy(p->diag, width == 16 ? GL_DIAG16 : GL_DIAG8, width * sizeof(uint64_t));
    else
        bb_diag16(p->diag);
    /* RNG order (poseidon2_goldilocks.rs:65-95, Poseidon2::new_from_rng):
     * initial external rf/2 x W, terminal external rf/2 x W, internal rp.
     * Flat layout (goldilocks.rs:42-83): [initial][internal][terminal]. */
    or_rng r;
    or_rng_seed(&r, seed);
    int half = p->rounds_f / 2, W = width;
    uint64_t *init = p->rc, *internal = p->rc + half * W, *term = internal + p->rounds_p;
    for (int i = 0; i < half * W; i++) init[i] = sample_field(&r, field, bb_rc_mode);
    for (int i = 0; i < half * W; i++) term[i] = sample_field(&r, field, bb_rc_mode);
    for (int i = 0; i < p->rounds_p; i++) internal[i] = sample_field(&r, field, bb_rc_mode);
    return 0;
}

size_t or_p2_flat_constants(const or_p2 *p, uint64_t *out)
{
    size_t n = (size_t)p->rounds_f * p->width + p->rounds_p;
    if (out) memcpy(out, p->rc, n * sizeof(uint64_t));
    return n;
}

static inline __attribute__((always_inline)) uint64_t sbox7(int f, uint64_t x)
{
    uint64_t x2 = or_f_mul(f, x, x);
    uint64_t x3 = or_f_mul(f, x2, x);
    uint64_t x4 = or_f_mul(f, x2, x2);
    return or_f_mul(f, x3, x4);
}

static inline __attribute__((always_inline)) void mds_light(int f, int W, uint64_t *s)
{
    for (int c = 0; c < W; c += 4) {
        uint64_t a = s[c], b = s[c + 1], cc = s[c + 2], d = s[c + 3];
        uint64_t ab = or_f_add(f, a, b), cd = or_f_add(f, cc, d);
        uint64_t t = or_f_add(f, ab, cd);
        uint64_t tb = or_f_add(f, t, b), td = or_f_add(f, t, d);
        s[c] = or_f_add(f, tb, ab);                    /* 2a+3b+c+d */
        s[c + 1] = or_f_add(f, tb, or_f_add(f, cc, cc)); /* a+2b+3c+d */
        s[c + 2] = or_f_add(f, td, cd);                /* a+b+2c+3d */
        s[c + 3] = or_f_add(f, td, or_f_add(f, a, a));   /* 3a+b+c+2d */
    }
    uint64_t sum[4];
    for (int k = 0; k < 4; k++) {
        sum[k] = s[k];
        for (int c = 4; c < W; c += 4) sum[k] = or_f_add(f, sum[k], s[c + k]);
    }
    for (int i = 0; i < W; i++) s[i] = or_f_add(f, s[i], sum[i & 3]);
}

static inline __attribute__((always_inline)) void external_round(int f, int W, uint64_t *s, const uint64_t *rc)
{
    for (int i = 0; i < W; i++) s[i] = sbox7(f, or_f_add(f, s[i], rc[i]));
    mds_light(f, W, s);
}

/* The permutation body with the field and width as arguments that the two
 * call sites below fix, so the compiler specialises the arithmetic (no
 * per-operation field dispatch); the semantics are the ones written here. */
static inline __attribute__((always_inline)) void p2_permute_body(const or_p2 *p, uint64_t *s, const int f,
                                                                   const int W)
{
    const int half = p->rounds_f / 2;
    const uint64_t *init = p->rc, *internal = p->rc + half * W, *term = internal + p->rounds_p;
    mds_light(f, W, s);
    for (int r = 0; r < half; r++) external_round(f, W, s, init + r * W);
    for (int r = 0; r < p->rounds_p; r++) {
        s[0] = sbox7(f, or_f_add(f, s[0], internal[r]));
        uint64_t sum = s[0];
        for (int i = 1; i < W; i++) sum = or_f_add(f, sum, s[i]);
        for (int i = 0; i < W; i++) s[i] = or_f_add(f, or_f_mul(f, s[i], p->diag[i]), sum);
    }
    for (int r = 0; r < half; r++) external_round(f, W, s, term + r * W);
}

void or_p2_permute(const or_p2 *p, uint64_t *s)
{
    if (p->field == OR_GOLDILOCKS && p->width == 16)
        p2_permute_body(p, s, OR_GOLDILOCKS, 16);
    else if (p->field == OR_GOLDILOCKS)
        p2_permute_body(p, s, OR_GOLDILOCKS, 8);
    else
        p2_permute_body(p, s, OR_BABYBEAR, 16);
}

/* PaddingFreeSponge<P, W, RATE=W/2, OUT=W/2>::hash_iter (p3-symmetric 0.4.2,
 * used at plonky3.rs:56; restated at poseidon2_goldilocks.rs:372-398):
 * overwrite the rate with the next chunk; a partial last chunk keeps the
 * previous rate values; no permutation after an exactly consumed input. */
void or_p2_hash(const or_p2 *p, const uint64_t *data, size_t len, uint64_t *digest)
{
    const int W = p->width, R = W / 2;
    uint64_t s[16] = {0};
    size_t pos = 0;
    while (pos < len) {
        size_t take = len - pos < (size_t)R ? len - pos : (size_t)R;
        for (size_t i = 0; i < take; i++) s[i] = data[pos + i];
        pos += take;
        or_p2_permute(p, s);
    }
    memcpy(digest, s, R * sizeof(uint64_t));
}

/* TruncatedPermutation<P, 2, W/2, W>::compress (plonky3.rs:57;
 * poseidon2_goldilocks.rs:172-188): state = l || r, permute, take W/2. */
void or_p2_compress(const or_p2 *p, const uint64_t *l, const uint64_t *r, uint64_t *out)
{
    const int W = p->width, D = W / 2;
    uint64_t s[16];
    memcpy(s, l, D * sizeof(uint64_t));
    memcpy(s + D, r, D * sizeof(uint64_t));
    or_p2_permute(p, s);
    memcpy(out, s, D * sizeof(uint64_t));
}

/* ======================================================================
 * DFT over columns of a row-major matrix (natural order in and out).
 * Semantics of p3 TwoAdicSubgroupDft::dft_batch / idft_batch, restated from
 * the reference CPU twin zkf-metal/src/ntt/p3_adapter.rs:461-572
 * (bit-reverse, radix-2 DIT with w = g^(j * n / 2^(s+1)), idft scales 1/n).
 * ====================================================================== */
static inline uint64_t rev_bits(uint64_t x, uint32_t bits)
{
    uint64_t r = 0;
    for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

static uint32_t log2_exact(uint64_t n)
{
    uint32_t l = 0;
    while ((1ULL << l) < n) l++;
    return l;
}

/* In-place DIT on one column (already bit-reversed).  tw[j] = w^j, j < n/2. */
static void dit_column(int f, uint64_t *x, uint64_t n, const uint64_t *tw)
{
    for (uint64_t half = 1; half < n; half <<= 1) {
        uint64_t step = n / (2 * half);
        for (uint64_t g = 0; g < n; g += 2 * half)
            for (uint64_t j = 0; j < half; j++) {
                uint64_t a = x[g + j], b = or_f_mul(f, x[g + j + half], tw[j * step]);
                x[g + j] = or_f_add(f, a, b);
                x[g + j + half] = or_f_sub(f, a, b);
            }
    }
}

static uint64_t *powers_table(int f, uint64_t w, uint64_t n)
{
    uint64_t *t = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t acc = 1;
    for (uint64_t i = 0; i < n; i++) { t[i] = acc; acc = or_f_mul(f, acc, w); }
    return t;
}

static void bitrev_inplace(uint64_t *x, uint64_t n)
{
    uint32_t lg = log2_exact(n);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = rev_bits(i, lg);
        if (i < j) { uint64_t t = x[i]; x[i] = x[j]; x[j] = t; }
    }
}

void or_dft_batch(int f, uint64_t *mat, uint64_t n, uint32_t width, int inverse)
{
    if (n <= 1 || width == 0) return;
    uint32_t lg = log2_exact(n);
    uint64_t w = or_f_two_adic_generator(f, lg);
    if (inverse) w = or_f_inv(f, w);
    uint64_t *tw = powers_table(f, w, n / 2);
    uint64_t ninv = or_f_inv(f, n % f_p(f));
#pragma omp parallel
    {
        uint64_t *col = (uint64_t *)malloc(n * sizeof(uint64_t));
#pragma omp for schedule(dynamic, 1)
        for (uint32_t c = 0; c < width; c++) {
            for (uint64_t r = 0; r < n; r++) col[r] = mat[r * width + c];
            bitrev_inplace(col, n);
            dit_column(f, col, n, tw);
            for (uint64_t r = 0; r < n; r++)
                mat[r * width + c] = inverse ? or_f_mul(f, col[r], ninv) : col[r];
        }
        free(col);
    }
    free(tw);
}

void or_bit_reverse_rows(uint64_t *mat, uint64_t n, uint32_t width)
{
    uint32_t lg = log2_exact(n);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = rev_bits(i, lg);
        if (i < j)
            for (uint32_t c = 0; c < width; c++) {
                uint64_t t = mat[i * width + c];
                mat[i * width + c] = mat[j * width + c];
                mat[j * width + c] = t;
            }
    }
}

/* Coset LDE as TwoAdicFriPcs::commit uses it (p3-dft/p3-fri 0.4.2 via
 * RuntimeDft, plonky3.rs:101-257): coeffs = idft(evals); coeff_k *= shift^k;
 * zero-pad to n << log_blowup; dft; then bit_reverse_rows.  Output row j holds
 * p(shift * w_{Bn}^{rev(j)}). */
void or_coset_lde_batch(int f, const uint64_t *in, uint64_t n, uint32_t width,
                        uint32_t log_blowup, uint64_t shift, uint64_t *out)
{
    uint64_t big = n << log_blowup;
    uint32_t lg = log2_exact(n), lgb = lg + log_blowup;
    uint64_t wi = or_f_inv(f, or_f_two_adic_generator(f, lg));
    uint64_t wb = or_f_two_adic_generator(f, lgb);
    uint64_t *twi = powers_table(f, wi, n / 2 ? n / 2 : 1);
    uint64_t *twb = powers_table(f, wb, big / 2);
    uint64_t ninv = or_f_inv(f, n % f_p(f));
    /* shift^k / n */
    uint64_t *wt = powers_table(f, shift, n);
    for (uint64_t k = 0; k < n; k++) wt[k] = or_f_mul(f, wt[k], ninv);
#pragma omp parallel
    {
        uint64_t *col = (uint64_t *)malloc(big * sizeof(uint64_t));
#pragma omp for schedule(dynamic, 1)
        for (uint32_t c = 0; c < width; c++) {
            for (uint64_t r = 0; r < n; r++) col[r] = in[r * width + c];
            if (n > 1) {
                bitrev_inplace(col, n);
                dit_column(f, col, n, twi);
            }
            for (uint64_t k = 0; k < n; k++) col[k] = or_f_mul(f, col[k], wt[k]);
            for (uint64_t k = n; k < big; k++) col[k] = 0;
            if (big > 1) {
                bitrev_inplace(col, big);
                dit_column(f, col, big, twb);
            }
            /* bit_reverse_rows: out row j = evaluation rev(j) */
            for (uint64_t j = 0; j < big; j++) out[j * width + c] = col[rev_bits(j, lgb)];
        }
        free(col);
    }
    free(twi);
    free(twb);
    free(wt);
}

/* ======================================================================
 * Merkle tree: p3-merkle-tree 0.4.2 as vendored in
 * vendor/p3-merkle-tree-gpu/src/merkle_tree.rs:84-403 (MerkleTree::new,
 * first_digest_layer, compress_and_inject, compress) and openings per
 * vendor/p3-merkle-tree-gpu/src/mmcs.rs:179-330 (open_batch, verify_batch).
 * ====================================================================== */
static uint64_t next_pow2(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

/* hash of the concatenated rows `row` of matrices idx[0..k) */
static void hash_rows(const or_p2 *p, int k, const int *idx, const uint64_t *const *mats,
                      const uint32_t *widths, uint64_t row, uint64_t *tmp, uint64_t *digest)
{
    size_t len = 0;
    for (int m = 0; m < k; m++) {
        memcpy(tmp + len, mats[idx[m]] + row * widths[idx[m]], widths[idx[m]] * sizeof(uint64_t));
        len += widths[idx[m]];
    }
    or_p2_hash(p, tmp, len, digest);
}

int or_merkle_build(const or_p2 *p, int nmats, const uint64_t *const *mats,
                    const uint64_t *heights, const uint32_t *widths, or_tree *out)
{
    const int D = p->width / 2;
    memset(out, 0, sizeof *out);
    if (nmats <= 0) return -1;
    /* tallest first (stable) */
    int *ord = (int *)malloc(nmats * sizeof(int));
    for (int i = 0; i < nmats; i++) ord[i] = i;
    for (int i = 1; i < nmats; i++)
        for (int j = i; j > 0 && heights[ord[j]] > heights[ord[j - 1]]; j--) {
            int t = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = t;
        }
    for (int i = 1; i < nmats; i++) {
        uint64_t a = heights[ord[i - 1]], b = heights[ord[i]];
        if (a != b && next_pow2(a) == next_pow2(b)) { free(ord); return -2; }
        if (b == 0) { free(ord); return -3; }
    }
    size_t totw = 0;
    for (int i = 0; i < nmats; i++) totw += widths[i];
    uint64_t maxh = heights[ord[0]];
    int cursor = 0;
    int ntall = 0;
    while (cursor + ntall < nmats && heights[ord[cursor + ntall]] == maxh) ntall++;

    int cap = 2;
    for (uint64_t h = maxh; h > 1; h = (h + 1) / 2) cap++;
    out->digest = D;
    out->layers = (uint64_t **)calloc(cap, sizeof(uint64_t *));
    out->layer_len = (uint64_t *)calloc(cap, sizeof(uint64_t));

    /* first_digest_layer: even length except height 1; padding = zero digest */
    uint64_t len0 = maxh == 1 ? 1 : maxh + (maxh & 1);
    uint64_t *l0 = (uint64_t *)calloc(len0 * D, sizeof(uint64_t));
#pragma omp parallel
    {
        uint64_t *tmp = (uint64_t *)malloc((totw + 1) * sizeof(uint64_t));
#pragma omp for schedule(static)
        for (uint64_t i = 0; i < maxh; i++) hash_rows(p, ntall, ord + cursor, mats, widths, i, tmp, l0 + i * D);
        free(tmp);
    }
    cursor += ntall;
    out->layers[0] = l0;
    out->layer_len[0] = len0;
    int nl = 1;
    while (out->layer_len[nl - 1] != 1) {
        const uint64_t *prev = out->layers[nl - 1];
        uint64_t plen = out->layer_len[nl - 1];
        uint64_t next_len_p2 = next_pow2(plen / 2);
        int ninj = 0;
        while (cursor + ninj < nmats && next_pow2(heights[ord[cursor + ninj]]) == next_len_p2) ninj++;
        uint64_t padded = plen == 2 ? 1 : ((plen / 2 + 1) & ~1ULL);
        uint64_t *nx = (uint64_t *)calloc(padded * D, sizeof(uint64_t));
        uint64_t half = plen / 2;
        uint64_t inj_len = ninj ? heights[ord[cursor]] : 0;
#pragma omp parallel
        {
            uint64_t *tmp = (uint64_t *)malloc((totw + 1) * sizeof(uint64_t));
            uint64_t dg[8], rd[8], zero[8] = {0};
#pragma omp for schedule(static)
            for (uint64_t i = 0; i < half; i++) {
                or_p2_compress(p, prev + 2 * i * D, prev + (2 * i + 1) * D, nx + i * D);
                if (ninj) {
                    memcpy(dg, nx + i * D, D * sizeof(uint64_t));
                    if (i < inj_len) {
                        hash_rows(p, ninj, ord + cursor, mats, widths, i, tmp, rd);
                        or_p2_compress(p, dg, rd, nx + i * D);
                    } else {
                        or_p2_compress(p, dg, zero, nx + i * D);
                    }
                }
            }
            free(tmp);
        }
        cursor += ninj;
        out->layers[nl] = nx;
        out->layer_len[nl] = padded;
        nl++;
    }
    out->nlayers = nl;
    free(ord);
    return 0;
}

void or_merkle_free(or_tree *t)
{
    for (int i = 0; i < t->nlayers; i++) free(t->layers[i]);
    free(t->layers);
    free(t->layer_len);
    memset(t, 0, sizeof *t);
}

void or_merkle_root(const or_tree *t, uint64_t *root)
{
    memcpy(root, t->layers[t->nlayers - 1], t->digest * sizeof(uint64_t));
}

/* open_batch proof: [digest_layers[i][(index >> i) ^ 1] for i < log2_ceil(max_height)] */
int or_merkle_path(const or_tree *t, uint64_t index, uint64_t *path)
{
    const int D = t->digest;
    int n = 0;
    for (int i = 0; i + 1 < t->nlayers; i++) {
        uint64_t sib = (index >> i) ^ 1;
        if (sib >= t->layer_len[i]) return -1;
        memcpy(path + (size_t)i * D, t->layers[i] + sib * D, D * sizeof(uint64_t));
        n++;
    }
    return n;
}

int or_merkle_verify(const or_p2 *p, const uint64_t *root, int nmats, const uint64_t *heights,
                     const uint32_t *widths, uint64_t index, const uint64_t *const *rows,
                     const uint64_t *path, int path_len)
{
    const int D = p->width / 2;
    int *ord = (int *)malloc(nmats * sizeof(int));
    for (int i = 0; i < nmats; i++) ord[i] = i;
    for (int i = 1; i < nmats; i++)
        for (int j = i; j > 0 && heights[ord[j]] > heights[ord[j - 1]]; j--) {
            int t = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = t;
        }
    uint64_t maxh = heights[ord[0]];
    uint64_t cur_p2 = next_pow2(maxh);
    if ((int)log2_exact(cur_p2) != path_len || index >= maxh) { free(ord); return 0; }
    size_t totw = 0;
    for (int i = 0; i < nmats; i++) totw += widths[i];
    uint64_t *tmp = (uint64_t *)malloc((totw + 1) * sizeof(uint64_t));
    uint64_t acc[8], l[8], r[8];
    int cursor = 0;
    size_t len = 0;
    while (cursor < nmats && next_pow2(heights[ord[cursor]]) == cur_p2) {
        memcpy(tmp + len, rows[ord[cursor]], widths[ord[cursor]] * sizeof(uint64_t));
        len += widths[ord[cursor]];
        cursor++;
    }
    or_p2_hash(p, tmp, len, acc);
    for (int s = 0; s < path_len; s++) {
        const uint64_t *sib = path + (size_t)s * D;
        if ((index & 1) == 0) { memcpy(l, acc, D * 8); memcpy(r, sib, D * 8); }
        else { memcpy(l, sib, D * 8); memcpy(r, acc, D * 8); }
        or_p2_compress(p, l, r, acc);
        index >>= 1;
        cur_p2 >>= 1;
        if (cursor < nmats && next_pow2(heights[ord[cursor]]) == cur_p2) {
            uint64_t h = heights[ord[cursor]];
            len = 0;
            while (cursor < nmats && heights[ord[cursor]] == h) {
                memcpy(tmp + len, rows[ord[cursor]], widths[ord[cursor]] * sizeof(uint64_t));
                len += widths[ord[cursor]];
                cursor++;
            }
            uint64_t rd[8];
            or_p2_hash(p, tmp, len, rd);
            memcpy(l, acc, D * 8);
            or_p2_compress(p, l, rd, acc);
        }
    }
    free(tmp);
    free(ord);
    return memcmp(acc, root, D * sizeof(uint64_t)) == 0;
}

int or_merkle_single(const or_p2 *p, const uint64_t *mat, uint64_t height, uint32_t width,
                     uint64_t *root, uint64_t *layers_flat, uint64_t layers_cap)
{
    or_tree t;
    const uint64_t *mats[1] = {mat};
    int rc = or_merkle_build(p, 1, mats, &height, &width, &t);
    if (rc) return rc;
    or_merkle_root(&t, root);
    if (layers_flat) {
        uint64_t off = 0;
        for (int i = 0; i < t.nlayers; i++) {
            uint64_t n = t.layer_len[i] * t.digest;
            if (off + n > layers_cap) { or_merkle_free(&t); return -4; }
            memcpy(layers_flat + off, t.layers[i], n * sizeof(uint64_t));
            off += n;
        }
    }
    int nl = t.nlayers;
    or_merkle_free(&t);
    return nl;
}

int or_commit(const or_p2 *p, const uint64_t *trace, uint64_t height, uint32_t width,
              uint32_t log_blowup, uint64_t shift, uint64_t *lde_out, uint64_t *root)
{
    uint64_t big = height << log_blowup;
    uint64_t *lde = lde_out ? lde_out : (uint64_t *)malloc(big * width * sizeof(uint64_t));
    or_coset_lde_batch(p->field, trace, height, width, log_blowup, shift, lde);
    int rc = or_merkle_single(p, lde, big, width, root, NULL, 0);
    if (!lde_out) free(lde);
    return rc < 0 ? rc : 0;
}

/* ======================================================================
 * FRI
 * ====================================================================== */

/* FriAccelerator::fold_goldilocks / fold_babybear semantics
 * (zkf-core/src/acceleration.rs:196-217, zkf-metal/src/shaders/fri.metal:14-53):
 * g[i] = (f[2i] + f[2i+1]) / 2 + alpha * (f[2i] - f[2i+1]) * inv_twiddles[i]. */
void or_fri_fold_base(int f, const uint64_t *ev, uint64_t n_out, uint64_t alpha,
                      const uint64_t *itw, uint64_t *out)
{
    uint64_t half = or_f_inv(f, 2);
#pragma omp parallel for schedule(static) if (n_out > 65536)
    for (uint64_t i = 0; i < n_out; i++) {
        uint64_t s = or_f_mul(f, or_f_add(f, ev[2 * i], ev[2 * i + 1]), half);
        uint64_t d = or_f_mul(f, or_f_mul(f, or_f_sub(f, ev[2 * i], ev[2 * i + 1]), itw[i]), alpha);
        out[i] = or_f_add(f, s, d);
    }
}

/* p3-fri 0.4.2 TwoAdicFriFolding::fold_matrix, arity 2 (called by the FRI
 * commit phase of TwoAdicFriPcs, plonky3.rs:68,596-598):
 *   g_inv = two_adic_generator(log2(n) + 1)^-1, powers[k] = (beta/2) g_inv^k,
 *   bit-reverse powers, out[i] = (1/2 + powers[i]) lo + (1/2 - powers[i]) hi
 * with (lo, hi) = row i of the width-2 matrix. */
void or_fri_fold_matrix(int ext, int f, const uint64_t *ev, uint64_t n, const uint64_t *beta,
                        uint64_t *out)
{
    const int E = ext ? 4 : 1;
    if (ext) f = OR_BABYBEAR;
    uint32_t lg = log2_exact(n);
    uint64_t ginv = or_f_inv(f, or_f_two_adic_generator(f, lg + 1));
    uint64_t half = or_f_inv(f, 2);
    uint64_t hb[4] = {0};
    for (int e = 0; e < E; e++) hb[e] = or_f_mul(f, beta[e], half);
    uint64_t *gp = powers_table(f, ginv, n);
#pragma omp parallel for schedule(static) if (n > 65536)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t g = gp[rev_bits(i, lg)];
        const uint64_t *lo = ev + 2 * i * E, *hi = ev + (2 * i + 1) * E;
        if (ext) {
            uint64_t pw[4], a[4], b[4], t0[4], t1[4];
            ef4_mul_base(hb, g, pw);
            for (int e = 0; e < 4; e++) { a[e] = pw[e]; b[e] = bb_sub(0, pw[e]); }
            a[0] = bb_add(a[0], half);
            b[0] = bb_add(b[0], half);
            or_ef4_mul(a, lo, t0);
            or_ef4_mul(b, hi, t1);
            ef4_add(t0, t1, out + i * 4);
        } else {
            uint64_t pw = or_f_mul(f, hb[0], g);
            uint64_t a = or_f_add(f, half, pw), b = or_f_sub(f, half, pw);
            out[i] = or_f_add(f, or_f_mul(f, a, lo[0]), or_f_mul(f, b, hi[0]));
        }
    }
    free(gp);
}

/* Challenges for the FRI benchmark: beta_r coefficient j is generator element
 * r*4 + j of stream `seed` (BabyBear). */
void or_gen_ef4_betas(uint64_t seed, int n, uint64_t *betas)
{
    for (int i = 0; i < 4 * n; i++) betas[i] = or_gen_elem(OR_BABYBEAR, seed, (uint64_t)i);
}

/* p3-fri 0.4.2 commit phase loop with the challenger replaced by given betas:
 * while len > 2^log_final: commit RowMajorMatrix(folded, 2) through
 * ExtensionMmcs (an EF4 row of 2 elements flattens to 8 base elements,
 * plonky3.rs:66), then fold with beta_r. */
int or_fri_commit_phase_ef4(const or_p2 *p, uint64_t *cw, uint32_t log_len, uint32_t log_final,
                            const uint64_t *betas, uint64_t *layer_roots, uint64_t *final_out,
                            or_tree *trees_out)
{
    if (p->field != OR_BABYBEAR || p->width != 16 || log_final >= log_len) return -1;
    uint64_t len = 1ULL << log_len;
    uint64_t *cur = cw;
    uint64_t *buf = (uint64_t *)malloc((len / 2) * 4 * sizeof(uint64_t));
    uint64_t *nxt = buf;
    int r = 0;
    while (len > (1ULL << log_final)) {
        uint64_t h = len / 2;
        uint32_t w = 8;
        const uint64_t *mats[1] = {cur};
        or_tree t;
        int rc = or_merkle_build(p, 1, mats, &h, &w, &t);
        if (rc) { free(buf); return rc; }
        or_merkle_root(&t, layer_roots + (size_t)r * 8);
        if (trees_out) trees_out[r] = t;
        else or_merkle_free(&t);
        or_fri_fold_matrix(1, OR_BABYBEAR, cur, h, betas + (size_t)r * 4, nxt);
        /* ping-pong between the codeword buffer and the scratch buffer */
        uint64_t *t2 = cur;
        cur = nxt;
        nxt = t2;
        len = h;
        r++;
    }
    memcpy(final_out, cur, len * 4 * sizeof(uint64_t));
    free(buf);
    return r;
}

/* Commit phase over a base field (Challenge = Val, the reference's
 * `type Challenge<F> = F` at zkf-backends/src/plonky3.rs:55, FRI parameters
 * from create_test_fri_params at :597; p3-fri 0.4.2's commit phase itself is
 * external): each layer is committed as a width-2 matrix under the field's
 * Poseidon2 MMCS (one sponge absorption of 2 elements per leaf) and folded
 * with fold_matrix, arity 2, beta = betas[r]. */
int or_fri_commit_phase_base(const or_p2 *p, uint64_t *cw, uint32_t log_len, uint32_t log_final,
                             const uint64_t *betas, uint64_t *layer_roots, uint64_t *final_out)
{
    if (p->width != 16 || log_final >= log_len) return -1;
    uint64_t len = 1ULL << log_len;
    uint64_t *cur = cw;
    uint64_t *buf = (uint64_t *)malloc((len / 2) * sizeof(uint64_t));
    uint64_t *nxt = buf;
    int r = 0;
    while (len > (1ULL << log_final)) {
        uint64_t h = len / 2;
        uint32_t w = 2;
        const uint64_t *mats[1] = {cur};
        or_tree t;
        int rc = or_merkle_build(p, 1, mats, &h, &w, &t);
        if (rc) { free(buf); return rc; }
        or_merkle_root(&t, layer_roots + (size_t)r * (p->width / 2));
        or_merkle_free(&t);
        or_fri_fold_matrix(0, p->field, cur, h, betas + r, nxt);
        uint64_t *t2 = cur;
        cur = nxt;
        nxt = t2;
        len = h;
        r++;
    }
    memcpy(final_out, cur, len * sizeof(uint64_t));
    free(buf);
    return r;
}

/* ======================================================================
 * PCS opening (DEEP quotient), restated from p3-fri 0.4.2
 * TwoAdicFriPcs::open and p3-interpolation 0.4.2 interpolate_coset; neither
 * is in the reference tree.  BabyBear base field, EF4 challenges.
 * ====================================================================== */
static void ef4_pow(const uint64_t *a, u128 e, uint64_t *out)
{
    uint64_t r[4] = {1, 0, 0, 0}, b[4] = {a[0], a[1], a[2], a[3]};
    while (e) {
        if (e & 1) or_ef4_mul(r, b, r);
        or_ef4_mul(b, b, b);
        e >>= 1;
    }
    memcpy(out, r, sizeof r);
}

/* a^(p^4 - 2): Fermat in the multiplicative group of EF4. */
void or_ef4_inv(const uint64_t *a, uint64_t *out)
{
    const u128 p = OR_BB_P, q = p * p * p * p;
    ef4_pow(a, q - 2, out);
}

/* x_i = g * w_H^rev(i) for i < height (bit-reversed coset order). */
static uint64_t *deep_points(uint64_t height, uint64_t shift)
{
    uint32_t lg = log2_exact(height);
    uint64_t w = or_f_two_adic_generator(OR_BABYBEAR, lg);
    uint64_t *x = (uint64_t *)malloc(height * sizeof(uint64_t));
    uint64_t acc = shift;
    for (uint64_t j = 0; j < height; j++) {
        x[rev_bits(j, lg)] = acc;
        acc = or_bb_mul(acc, w);
    }
    return x;
}

void or_deep_open_values(const uint64_t *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                         uint64_t shift, const uint64_t *z, uint64_t *ys)
{
    const uint64_t n = height >> log_blowup;
    uint64_t *x = deep_points(height, shift);
    memset(ys, 0, (size_t)width * 4 * sizeof(uint64_t));
    /* rows 0..n-1 are p on g*H_n (their points x_i are the coset g*H_n) */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t d[4] = {bb_sub(z[0], x[i]), z[1], z[2], z[3]}, w[4];
        or_ef4_inv(d, w);
        ef4_mul_base(w, x[i], w); /* x_i / (z - x_i) */
        for (uint32_t k = 0; k < width; k++) {
            uint64_t t[4];
            ef4_mul_base(w, lde[i * width + k], t);
            ef4_add(ys + 4 * k, t, ys + 4 * k);
        }
    }
    /* scale (z^n - g^n) / (n g^n) */
    uint64_t zn[4], gn = or_bb_pow(shift, n), s[4];
    ef4_pow(z, n, zn);
    zn[0] = bb_sub(zn[0], gn);
    ef4_mul_base(zn, or_f_inv(OR_BABYBEAR, or_bb_mul(n % OR_BB_P, gn)), s);
    for (uint32_t k = 0; k < width; k++) or_ef4_mul(ys + 4 * k, s, ys + 4 * k);
    free(x);
}

void or_deep_reduce(const uint64_t *lde, uint64_t height, uint32_t width, uint64_t shift, const uint64_t *z,
                    const uint64_t *alpha, uint64_t alpha_off, const uint64_t *ys, uint64_t *ro)
{
    uint64_t *x = deep_points(height, shift);
    uint64_t *ap = (uint64_t *)malloc(((size_t)width + 1) * 4 * sizeof(uint64_t));
    uint64_t ry[4] = {0, 0, 0, 0}, aoff[4];
    ap[0] = 1, ap[1] = ap[2] = ap[3] = 0;
    for (uint32_t k = 1; k <= width; k++) or_ef4_mul(ap + 4 * (k - 1), alpha, ap + 4 * k);
    for (uint32_t k = 0; k < width; k++) {
        uint64_t t[4];
        or_ef4_mul(ap + 4 * k, ys + 4 * k, t);
        ef4_add(ry, t, ry);
    }
    ef4_pow(alpha, alpha_off, aoff);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < height; i++) {
        uint64_t acc[4] = {0, 0, 0, 0}, t[4], d[4], inv[4];
        for (uint32_t k = 0; k < width; k++) {
            ef4_mul_base(ap + 4 * k, lde[i * width + k], t);
            ef4_add(acc, t, acc);
        }
        ef4_sub(acc, ry, acc);
        d[0] = bb_sub(x[i], z[0]);
        d[1] = bb_sub(0, z[1]), d[2] = bb_sub(0, z[2]), d[3] = bb_sub(0, z[3]);
        or_ef4_inv(d, inv);
        or_ef4_mul(acc, inv, acc);
        or_ef4_mul(acc, aoff, acc);
        ef4_add(ro + 4 * i, acc, ro + 4 * i);
    }
    free(ap);
    free(x);
}

static uint64_t *deep_points_f(int f, uint64_t height, uint64_t shift)
{
    uint32_t lg = log2_exact(height);
    uint64_t w = or_f_two_adic_generator(f, lg);
    uint64_t *x = (uint64_t *)malloc(height * sizeof(uint64_t));
    uint64_t acc = shift;
    for (uint64_t j = 0; j < height; j++) {
        x[rev_bits(j, lg)] = acc;
        acc = or_f_mul(f, acc, w);
    }
    return x;
}

void or_deep_gl_open_values(const uint64_t *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                            uint64_t shift, uint64_t z, uint64_t *ys)
{
    const int f = OR_GOLDILOCKS;
    const uint64_t n = height >> log_blowup;
    uint64_t *x = deep_points_f(f, height, shift);
    memset(ys, 0, (size_t)width * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) {
        uint64_t w = or_f_mul(f, x[i], or_f_inv(f, or_gl_sub(z, x[i]))); /* x_i / (z - x_i) */
        for (uint32_t k = 0; k < width; k++) ys[k] = or_gl_add(ys[k], or_gl_mul(w, lde[i * width + k]));
    }
    uint64_t gn = or_gl_pow(shift, n);
    uint64_t s = or_gl_mul(or_gl_sub(or_gl_pow(z, n), gn), or_f_inv(f, or_gl_mul(n % OR_GL_P, gn)));
    for (uint32_t k = 0; k < width; k++) ys[k] = or_gl_mul(ys[k], s);
    free(x);
}

void or_deep_gl_reduce(const uint64_t *lde, uint64_t height, uint32_t width, uint64_t shift, uint64_t z,
                       uint64_t alpha, uint64_t alpha_off, const uint64_t *ys, uint64_t *ro)
{
    const int f = OR_GOLDILOCKS;
    uint64_t *x = deep_points_f(f, height, shift);
    uint64_t *ap = (uint64_t *)malloc(((size_t)width + 1) * sizeof(uint64_t));
    uint64_t ry = 0, aoff = or_gl_pow(alpha, alpha_off);
    ap[0] = 1;
    for (uint32_t k = 1; k <= width; k++) ap[k] = or_gl_mul(ap[k - 1], alpha);
    for (uint32_t k = 0; k < width; k++) ry = or_gl_add(ry, or_gl_mul(ap[k], ys[k]));
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < height; i++) {
        uint64_t acc = 0;
        for (uint32_t k = 0; k < width; k++) acc = or_gl_add(acc, or_gl_mul(ap[k], lde[i * width + k]));
        acc = or_gl_mul(or_gl_sub(acc, ry), or_f_inv(f, or_gl_sub(x[i], z)));
        ro[i] = or_gl_add(ro[i], or_gl_mul(acc, aoff));
    }
    free(ap);
    free(x);
}

/* ======================================================================
 * Constraint evaluation: the stack machine of
 * zkf-metal/src/shaders/constraint_eval.metal:1-74 (opcodes :7-13, one row
 * per evaluation, 32-entry stack :29-30), Goldilocks arithmetic.
 * ====================================================================== */
static int constraint_check(const uint32_t *bc, uint32_t n_instr, uint32_t width, uint32_t n_constants,
                            uint32_t n_constraints)
{
    int sp = 0;
    for (uint32_t pc = 0; pc < n_instr; pc++) {
        uint32_t op = bc[pc] >> 24, arg = bc[pc] & 0xFFFFFFu;
        switch (op) {
        case 0: if (arg >= n_constants || sp >= 32) return -1; sp++; break;
        case 1: if (arg >= width || sp >= 32) return -1; sp++; break;
        case 2: case 3: case 4: if (sp < 2) return -1; sp--; break;
        case 5: if (sp < 1 || sp >= 32) return -1; sp++; break;
        case 6: if (sp < 1 || arg >= n_constraints) return -1; sp--; break;
        default: break;
        }
    }
    return 0;
}

int or_constraint_eval_gl(const uint64_t *trace, uint64_t n_rows, uint32_t width, const uint32_t *bytecode,
                          uint32_t n_instr, const uint64_t *constants, uint32_t n_constants,
                          uint32_t n_constraints, uint64_t *out)
{
    if (constraint_check(bytecode, n_instr, width, n_constants, n_constraints)) return -1;
    memset(out, 0, n_rows * n_constraints * sizeof(uint64_t));
#pragma omp parallel for schedule(static)
    for (uint64_t row = 0; row < n_rows; row++) {
        uint64_t st[32], a, b;
        int sp = 0;
        for (uint32_t pc = 0; pc < n_instr; pc++) {
            uint32_t op = bytecode[pc] >> 24, arg = bytecode[pc] & 0xFFFFFFu;
            switch (op) {
            case 0: st[sp++] = constants[arg]; break;
            case 1: st[sp++] = trace[row * width + arg]; break;
            case 2: b = st[--sp]; a = st[--sp]; st[sp++] = or_gl_add(a, b); break;
            case 3: b = st[--sp]; a = st[--sp]; st[sp++] = or_gl_sub(a, b); break;
            case 4: b = st[--sp]; a = st[--sp]; st[sp++] = or_gl_mul(a, b); break;
            case 5: st[sp] = st[sp - 1]; sp++; break;
            case 6: out[row * n_constraints + arg] = st[--sp]; break;
            default: break;
            }
        }
    }
    return 0;
}

int or_constraint_quotient_gl(const uint64_t *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                              uint64_t shift, const uint32_t *bytecode, uint32_t n_instr, const uint64_t *constants,
                              uint32_t n_constants, uint32_t n_constraints, const uint64_t *coef, uint64_t *q)
{
    uint64_t *c = (uint64_t *)malloc(height * n_constraints * sizeof(uint64_t));
    if (or_constraint_eval_gl(lde, height, width, bytecode, n_instr, constants, n_constants, n_constraints, c)) {
        free(c);
        return -1;
    }
    uint64_t *x = deep_points_f(OR_GOLDILOCKS, height, shift);
    const uint64_t n = height >> log_blowup;
    int rc = 0;
    for (uint64_t i = 0; i < height && !rc; i++) {
        const uint64_t v = or_gl_sub(or_gl_pow(x[i], n), 1);
        if (!v) {
            rc = -1;
            break;
        }
        uint64_t acc = 0;
        for (uint32_t k = 0; k < n_constraints; k++) acc = or_gl_add(acc, or_gl_mul(coef[k], c[i * n_constraints + k]));
        q[i] = or_gl_mul(acc, or_f_inv(OR_GOLDILOCKS, v));
    }
    free(x);
    free(c);
    return rc;
}

/* ======================================================================
 * PolyOpsAccelerator, Goldilocks.  poly_eval_goldilocks
 * (zkf-metal/src/shaders/poly_ops.metal:22-41): Horner from the top
 * coefficient.  coset_eval_goldilocks (zkf-metal/src/poly.rs:159-189):
 * poly_coset_shift_goldilocks (poly_ops.metal:85-95) then dft_batch of one
 * column.  quotient_goldilocks (poly.rs:191-246, poly_ops.metal:67-82):
 * x = generator^i, (e_i - f_z) * inverse(x - z), 0 for a zero denominator.
 * ====================================================================== */
void or_poly_eval_gl(const uint64_t *coeffs, uint64_t n_coeffs, const uint64_t *points, uint64_t n_points,
                     uint64_t *out)
{
#pragma omp parallel for schedule(static) if (n_points > 4096)
    for (uint64_t i = 0; i < n_points; i++) {
        uint64_t r = 0;
        for (uint64_t k = n_coeffs; k-- > 0;) r = or_gl_add(or_gl_mul(r, points[i]), coeffs[k]);
        out[i] = r;
    }
}

void or_poly_coset_eval_gl(const uint64_t *coeffs, uint32_t log_n, uint64_t shift, uint64_t *out)
{
    uint64_t n = 1ULL << log_n;
    for (uint64_t k = 0; k < n; k++) out[k] = or_gl_mul(coeffs[k], or_gl_pow(shift, k));
    or_dft_batch(OR_GOLDILOCKS, out, n, 1, 0);
}

void or_poly_quotient_gl(const uint64_t *evals, uint64_t n, uint64_t z, uint64_t f_z, uint64_t *out)
{
    uint64_t g = or_f_two_adic_generator(OR_GOLDILOCKS, log2_exact(n));
#pragma omp parallel for schedule(static) if (n > 4096)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t x = or_gl_pow(g, i);
        uint64_t num = or_gl_sub(evals[i], f_z), den = or_gl_sub(x, z);
        out[i] = den == 0 ? 0 : or_gl_mul(num, or_f_inv(OR_GOLDILOCKS, den));
    }
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
