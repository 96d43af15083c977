// zkg_tuned.h -- LDE tile plan chosen by ziros_amd/tune.py (python -m ziros_amd.tune --measure)
// on gfx950.  Coset LDE milliseconds per shape (field: 1 BabyBear, 0 Goldilocks;
// log2 height, width, log_blowup), median of 3; the candidate with the least total wins.
//
// candidate  1/20/64/1 1/22/16/1 1/16/128/2 1/13/256/1 1/18/48/1 0/20/64/1 0/18/64/2 0/16/32/1 0/21/16/1 0/14/96/2  total
// r6_r4_u4   1.273 1.592 0.180 0.053 0.186 2.747 1.085 0.094 1.518 0.112  8.840
// r6_r4_u8   1.287 1.616 0.181 0.052 0.187 2.778 1.064 0.093 1.515 0.110  8.883
// r5_r4_u8   1.264 2.378 0.179 0.052 0.186 2.740 1.068 0.093 1.518 0.110  9.587
// r6_r5_u8   1.284 1.616 0.183 0.052 0.187 3.792 1.453 0.092 1.587 0.110  10.357
// r6_r4_u16  1.294 1.666 0.176 0.052 0.195 3.764 1.078 0.092 2.135 0.110  10.563
// r5_r5_u8   1.287 2.394 0.176 0.051 0.182 3.818 1.456 0.092 1.593 0.110  11.160
#pragma once
#define ZKG_PLAN_RMAX32 6
#define ZKG_PLAN_RMAX64 4
#define ZKG_TILE_STORE_UNROLL 4
