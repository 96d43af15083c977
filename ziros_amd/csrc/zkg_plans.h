// zkg_plans.h -- the tile plans of the column transforms (NTT and coset LDE).
//
// A plan fixes, for 2^L-point in-LDS tiles: the largest radix round per lane
// (BabyBear 2^rmax32, Goldilocks 2^rmax64, and 2^rmax64_small for Goldilocks
// tiles of at most 2^small_l64 points), whether a two-round plan takes the
// full radix first (first_heavy: wider tiles), how far the tile store loops
// unroll, the waves per SIMD LDE pass B is compiled for (0: the
// compiler's choice), and whether the Goldilocks products of the tile passes
// take the fewest-slot lazy form (lean64) or the canonical one.  The radix fixes the tile shape (columns per tile =
// 256 * radix / 2^L), the registers and LDS per block and so the blocks per
// CU.  zkg_ntt.hip is compiled once per plan; which plan a transform runs is
// chosen per shape at run time (zkg_tuning.hip).
#pragma once

namespace zkg {

struct LdePlan {
    const char *name;
    int rmax32, rmax64, rmax64_small, small_l64;
    int unroll;
    int waves32, waves64;
    int first_heavy;
    int lean64;
};

constexpr int LDE_PLANS = 4;
constexpr LdePlan LDE_PLAN_TABLE[LDE_PLANS] = {
    {"r6_r4s3_u8", 6, 4, 3, 8, 8, 0, 0, 0, 0},
    {"r6_r4s3_u4", 6, 4, 3, 8, 4, 0, 0, 0, 0},
    {"r6_r4s3_u4_o33", 6, 4, 3, 8, 4, 3, 3, 0, 0},
    {"r6h_r4s3_u8_lean", 6, 4, 3, 8, 8, 0, 0, 1, 1},
};

}  // namespace zkg
