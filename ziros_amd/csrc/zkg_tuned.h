// zkg_tuned.h -- LDE tile plan chosen by ziros_amd/tune.py (python -m ziros_amd.tune --measure)
// on gfx950.  Coset LDE milliseconds per shape (field: 1 BabyBear, 0 Goldilocks;
// log2 height, width, log_blowup), median of 3.  Score = sum over shapes of the time relative to
// the best candidate on that shape; the least score wins.
//
// candidate    1/20/64/1 1/22/16/1 1/16/128/2 1/13/256/1 1/18/48/1 0/20/64/1 0/18/64/2 0/16/32/1 0/21/16/1 0/14/96/2 0/14/8/1  score
// r6_r4s3_u8   1.290 1.612 0.182 0.052 0.187 2.771 1.058 0.090 1.530 0.092 0.041  11.169
// r6_r4_u8     1.291 1.623 0.181 0.052 0.184 2.757 1.054 0.093 1.523 0.110 0.058  11.772
// r6_r4_u16    1.278 1.623 0.181 0.052 0.185 2.768 1.059 0.093 1.522 0.111 0.058  11.775
// r5_r4_u8     1.273 1.623 0.182 0.052 0.186 2.787 1.055 0.092 1.520 0.111 0.058  11.781
// r5_r5_u8     1.306 1.615 0.180 0.053 0.187 2.740 1.056 0.092 1.524 0.111 0.058  11.791
// r6_r4_u4     1.263 1.620 0.181 0.052 0.186 2.762 1.066 0.093 1.533 0.111 0.058  11.794
// r6_r5_u8     1.297 1.613 0.183 0.053 0.187 2.769 1.064 0.094 1.519 0.112 0.058  11.874
// r6_r4_u4_o3  1.282 1.616 0.182 0.052 0.211 2.749 1.051 0.092 1.528 0.110 0.058  11.906
// r6_r4_u4_o33 1.298 1.632 0.180 0.052 0.208 3.444 0.971 0.099 1.530 0.114 0.060  12.223
#pragma once
#define ZKG_PLAN_RMAX32 6
#define ZKG_PLAN_RMAX64 4
#define ZKG_PLAN_RMAX64_SMALL 3
#define ZKG_PLAN_SMALL_L64 8
#define ZKG_TILE_STORE_UNROLL 8
