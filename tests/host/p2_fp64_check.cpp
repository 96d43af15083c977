// Host build of the device BabyBear FP64 arithmetic (ziros_amd/csrc/zkg_field.h
// Bf, zkg_poseidon2.h p2_permute_bf) checked against exact integer arithmetic
// and the oracle's Poseidon2 (oracle/zkg_oracle.c).  Exit status 0 = all equal.
#include <cstdio>
#include <random>

#include "../../oracle/zkg_oracle.h"
#include "../../ziros_amd/csrc/zkg_poseidon2.h"

using namespace zkg;
static const int64_t P = 2013265921;

static int64_t modp(__int128 v)
{
    int64_t r = (int64_t)(v % P);
    return r < 0 ? r + P : r;
}

// 2^-k mod p by exponentiation of (p + 1) / 2
static int64_t inv_pow2(int k)
{
    __int128 r = 1;
    for (int i = 0; i < k; i++) r = r * ((P + 1) / 2) % P;
    return (int64_t)r;
}

template <int K, bool NEG>
static long check_inv_pow2(std::mt19937_64 &g, int n)
{
    const int64_t c = NEG ? (P - inv_pow2(K)) % P : inv_pow2(K);
    long bad = 0;
    auto one = [&](int64_t x) {
        double r = Bf::mul_inv_pow2<K, NEG>((double)x);
        // the result must be an integer congruent to x c and small
        bool ok = r == (double)(int64_t)r && modp((__int128)(int64_t)r) == modp((__int128)x * c) &&
                  __builtin_fabs(r) < __builtin_fabs((double)x) / (double)(1u << K) + 2147483648.0;
        bad += !ok;
    };
    const int64_t lim = 1ll << 41;
    const int64_t edges[] = {0, 1, -1, 2, -2, (1ll << K) - 1, -(1ll << K) + 1, 1ll << K, -(1ll << K),
                             P - 1, -(P - 1), P, -P, lim - 1, -(lim - 1), (lim >> K) << K, -((lim >> K) << K)};
    for (int64_t x : edges) one(x);
    for (int i = 0; i < n; i++) one((int64_t)(g() % (uint64_t)(2 * lim - 1)) - (lim - 1));
    return bad;
}

int main()
{
    std::mt19937_64 g(20260101);
    long bad_pow2 = 0;
    bad_pow2 += check_inv_pow2<1, false>(g, 200000);
    bad_pow2 += check_inv_pow2<1, true>(g, 200000);
    bad_pow2 += check_inv_pow2<2, false>(g, 200000);
    bad_pow2 += check_inv_pow2<3, false>(g, 200000);
    bad_pow2 += check_inv_pow2<4, true>(g, 200000);
    bad_pow2 += check_inv_pow2<8, false>(g, 200000);
    bad_pow2 += check_inv_pow2<8, true>(g, 200000);
    bad_pow2 += check_inv_pow2<27, false>(g, 200000);
    bad_pow2 += check_inv_pow2<27, true>(g, 200000);

    long bad_mul = 0;
    for (int t = 0; t < 1000000; t++) {
        int64_t a = (int64_t)(g() % (1ull << 38)) - (1ll << 37);
        int64_t b = (int64_t)(g() % (1ull << 38)) - (1ll << 37);
        bad_mul += Bf::canon(Bf::mul((double)a, (double)b)) != (uint32_t)modp((__int128)a * b);
    }

    or_p2 p;
    if (or_p2_init(&p, OR_BABYBEAR, 16, 42, 0) != 0) return 2;
    double rc[141];
    for (int i = 0; i < 141; i++) rc[i] = (double)p.rc[i];
    long bad_perm = 0, n = 0;
    for (int t = 0; t < 100000; t++) {
        uint64_t st[16];
        double s[16];
        const int mode = t % 6;
        for (int i = 0; i < 16; i++) {
            uint64_t v = g() % (uint64_t)P;
            if (mode == 1) v = (uint64_t)P - 1 - (g() % 3);
            if (mode == 2) v = (g() & 1) ? (uint64_t)P - 1 : 0;
            if (mode == 5) v = 0xFFFFFFFFull - (g() % 3);  // fresh but not canonical
            st[i] = v % (uint64_t)P;
            s[i] = (double)v;
        }
        or_p2_permute(&p, st);
        if (mode == 4) p2_permute_bf(s, rc);  // every lane reduced on entry
        else p2_permute_bf<16>(s, rc);       // fresh lanes, no reduction
        for (int i = 0; i < 16; i++) n++, bad_perm += Bf::canon(s[i]) != st[i];
        if (mode == 3 || mode == 5) {  // second absorptions over the unreduced capacity, as in a sponge
            for (int i = 0; i < 8; i++) {
                uint64_t v = g() % (uint64_t)P;
                st[i] = v;
                s[i] = (double)v;
            }
            or_p2_permute(&p, st);
            p2_permute_bf<8>(s, rc);  // full-rate absorption: capacity reduced
            for (int i = 0; i < 16; i++) n++, bad_perm += Bf::canon(s[i]) != st[i];
            for (int i = 0; i < 3; i++) {  // partial absorption: lanes 3..7 keep earlier outputs
                uint64_t v = g() % (uint64_t)P;
                st[i] = v;
                s[i] = (double)v;
            }
            or_p2_permute(&p, st);
            p2_permute_bf(s, rc);
            for (int i = 0; i < 16; i++) n++, bad_perm += Bf::canon(s[i]) != st[i];
        }
    }
    printf("inv_pow2 mismatches %ld; mul mismatches %ld; permutation mismatches %ld of %ld\n", bad_pow2, bad_mul,
           bad_perm, n);
    return (bad_pow2 || bad_mul || bad_perm) ? 1 : 0;
}
