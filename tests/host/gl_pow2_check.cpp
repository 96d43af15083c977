// Host build of the Goldilocks power-of-two product Gl::mul_pow2
// (ziros_amd/csrc/zkg_field.h) checked against exact 128-bit arithmetic, and the
// small roots of unity of p3's two-adic generator checked to be the powers of
// two the transforms' register sub-DFTs use.  Exit status 0 = all equal.
#include <cstdio>
#include <random>

#include "../../ziros_amd/csrc/zkg_field.h"

using namespace zkg;
typedef unsigned __int128 u128;
static const uint64_t P = 0xFFFFFFFF00000001ull;

static uint64_t mulmod(uint64_t a, uint64_t b) { return (uint64_t)((u128)a * b % P); }
static uint64_t powmod(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    for (; e; e >>= 1, a = mulmod(a, a))
        if (e & 1) r = mulmod(r, a);
    return r;
}

int main()
{
    // p3 Goldilocks two_adic_generator(32); order-2^b roots by squaring down
    const uint64_t g32 = 0x185629dcda58878cull;
    struct {
        int bits, fwd, inv;
    } roots[] = {{2, 48, 144}, {3, 120, 72}, {4, 156, 36}};
    long root_bad = 0;
    for (auto &r : roots) {
        const uint64_t w = powmod(g32, 1ull << (32 - r.bits));
        root_bad += powmod(2, r.fwd) != w;
        root_bad += mulmod(powmod(2, r.inv), w) != 1;
    }
    // edge values and random canonical inputs, every k < 96
    std::mt19937_64 rng(12345);
    uint64_t edges[] = {0, 1, 2, 0xFFFFFFFFull, 0x100000000ull, 0x100000001ull, P - 1, P - 2, P - 0xFFFFFFFFull,
                        0x7FFFFFFFFFFFFFFFull, 0x8000000000000000ull, 0xFFFFFFFE00000000ull, 0xFFFFFFFF00000000ull};
    long bad = 0, n = 0;
    for (int k = 0; k < 96; k++) {
        const uint64_t c = powmod(2, k);
        auto check = [&](uint64_t x) {
            const uint64_t got = Gl::mul_pow2(x, k);
            bad += got != mulmod(x, c) || got >= P;
            n++;
        };
        for (uint64_t x : edges) check(x);
        for (int i = 0; i < 200000; i++) {
            uint64_t x = rng();
            if (x >= P) x -= P;
            check(x);
        }
        // inputs with only high or only low bits set stress the shifted halves
        for (int i = 0; i < 20000; i++) {
            check(rng() & 0xFFFFFFFF00000000ull & (P - 1));
            check(rng() & 0xFFFFFFFFull);
        }
    }
    printf("roots mismatches %ld; pow2 products %ld mismatches %ld\n", root_bad, n, bad);
    return root_bad || bad ? 1 : 0;
}
