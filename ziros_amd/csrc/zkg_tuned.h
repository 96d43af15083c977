// zkg_tuned.h -- LDE tile plan chosen by ziros_amd/tune.py (python -m ziros_amd.tune --measure)
// on gfx950.  Coset LDE milliseconds per shape (field: 1 BabyBear, 0 Goldilocks;
// log2 height, width, log_blowup), median of 3; the candidate with the least total wins.
//
// candidate  1/20/64/1 1/22/16/1 1/16/128/2 1/13/256/1 1/18/48/1 0/20/64/1 0/18/64/2 0/16/32/1 0/21/16/1 0/14/96/2  total
// r6_r4_u8   1.242 1.589 0.178 0.052 0.185 2.681 1.056 0.092 1.506 0.111  8.691
// r5_r5_u8   1.270 1.587 0.179 0.053 0.185 2.731 1.067 0.092 1.506 0.111  8.781
// r6_r4_u4   1.267 1.594 0.178 0.052 0.183 2.740 1.070 0.092 1.510 0.111  8.798
// r6_r4_u16  1.277 1.584 0.181 0.053 0.185 2.719 1.079 0.092 1.518 0.111  8.800
// r6_r5_u8   1.230 1.581 0.179 0.055 0.185 2.772 1.080 0.094 1.520 0.112  8.807
// r5_r4_u8   1.255 1.592 0.179 0.053 0.186 2.768 1.091 0.094 1.526 0.112  8.858
// r6_r4_u4_o3 1.301 1.586 0.179 0.052 0.207 2.797 1.093 0.094 1.531 0.113  8.953
// r6_r4_u4_o33 1.319 1.584 0.178 0.052 0.207 3.211 0.996 0.096 1.519 0.112  9.274
#pragma once
#define ZKG_PLAN_RMAX32 6
#define ZKG_PLAN_RMAX64 4
#define ZKG_TILE_STORE_UNROLL 8
