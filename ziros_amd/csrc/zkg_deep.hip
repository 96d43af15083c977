// zkg_deep.hip -- PCS opening: the DEEP quotient over a committed BabyBear LDE
// with EF4 (x^4 - 11) challenges.
//
// Semantics: p3-fri 0.4.2 TwoAdicFriPcs::open (reduced openings) and
// p3-interpolation 0.4.2 interpolate_coset, external to the reference (its
// graph node for this step is ProverOp::FriQueryOpen, zkf-runtime/src/
// graph.rs:105-108).  An LDE of height H holds p on the coset g*H_H in
// bit-reversed row order: row i = p(x_i), x_i = g w_H^rev(i).
//
// Design (this build's own):
//   * k_deep_inv: inv_i = 1 / (z - x_i) for every row.  A lane takes a run of
//     RUN consecutive rows, multiplies the denominators up (Montgomery's
//     batch trick) and inverts once, through the tower u = X^2, u^2 = 11:
//     a = A(u) + X B(u), a^-1 = (A - X B) / (A^2 - u B^2), the Bb[u] norm
//     inverted by a Fermat power.
//   * k_deep_colsum / k_deep_colsum_final: p_k(z) = (z^N - g^N) / (N g^N)
//     sum_{i<N} e_ik x_i inv_i (the first N rows are p on g*H_N).  A wave
//     reads whole 64-column row segments, per-block partial sums meet in LDS,
//     then one block per column sums the partials.
//   * k_deep_reduce: ro_i += a^off (sum_k a^k y_k - sum_k a^k e_ik) inv_i,
//     i.e. a^off sum_k a^k (p_k(x_i) - y_k) / (x_i - z).  One lane per row
//     over an LDS-staged tile; the base x EF4 dot product runs in exact FP64
//     on the 16-bit halves of the coefficients (see the kernel).
// HBM holds canonical values.  Constants (twiddles, powers of alpha, inverse
// denominators) are Montgomery form, so data x constant products come out
// canonical without conversions.
#include "zkg_field.h"
#include "zkg_internal.h"

namespace zkg {

namespace {

constexpr int DEEP_THREADS = 256;
constexpr int INV_RUN = 8;          // rows per lane in the batch inversion
constexpr int COLSUM_ROWS = 1024;   // rows per block in the column sums

__device__ __forceinline__ uint32_t brev(uint32_t x, int bits)
{
    return bits ? (__brev(x) >> (32 - bits)) : 0u;
}

__device__ __forceinline__ uint32_t root_pow(const uint32_t *__restrict__ lo, const uint32_t *__restrict__ hi,
                                             uint32_t e, int logN)
{
    uint32_t ep = (e << (LOG_TW_TOP - logN)) & ((1u << LOG_TW_TOP) - 1u);
    return Bb::mul(lo[ep & ((1u << TW_SPLIT) - 1u)], hi[ep >> TW_SPLIT]);
}

// Montgomery a^(p-2)
__device__ __forceinline__ uint32_t bb_inv(uint32_t a)
{
    uint32_t r = Bb::one(), b = a;
    uint32_t e = Bb::P - 2;
#pragma unroll 1
    while (e) {
        if (e & 1u) r = Bb::mul(r, b);
        b = Bb::sqr(b);
        e >>= 1;
    }
    return r;
}

// (x0 + x1 u)(y0 + y1 u), u^2 = 11
__device__ __forceinline__ void ef2_mul(uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, uint32_t w11, uint32_t &r0,
                                        uint32_t &r1)
{
    r0 = Bb::add(Bb::mul(x0, y0), Bb::mul(Bb::mul(x1, y1), w11));
    r1 = Bb::add(Bb::mul(x0, y1), Bb::mul(x1, y0));
}

__device__ Ef4 ef4_inv(const Ef4 &a, uint32_t w11)
{
    const uint32_t A0 = a.c[0], A1 = a.c[2], B0 = a.c[1], B1 = a.c[3];
    uint32_t a20, a21, b20, b21;
    ef2_mul(A0, A1, A0, A1, w11, a20, a21);
    ef2_mul(B0, B1, B0, B1, w11, b20, b21);
    const uint32_t n0 = Bb::sub(a20, Bb::mul(b21, w11)), n1 = Bb::sub(a21, b20);  // A^2 - u B^2
    const uint32_t mi = bb_inv(Bb::sub(Bb::sqr(n0), Bb::mul(Bb::sqr(n1), w11)));
    const uint32_t ni0 = Bb::mul(n0, mi), ni1 = Bb::neg(Bb::mul(n1, mi));
    Ef4 r;
    ef2_mul(A0, A1, ni0, ni1, w11, r.c[0], r.c[2]);
    ef2_mul(Bb::neg(B0), Bb::neg(B1), ni0, ni1, w11, r.c[1], r.c[3]);
    return r;
}

__device__ __forceinline__ Ef4 ef4_ld(const uint32_t *__restrict__ p)
{
    const uint4 v = *(const uint4 *)p;
    return Ef4{{v.x, v.y, v.z, v.w}};
}

__device__ __forceinline__ void ef4_st(uint32_t *__restrict__ p, const Ef4 &a)
{
    *(uint4 *)p = make_uint4(a.c[0], a.c[1], a.c[2], a.c[3]);
}

struct DeepConsts {
    Ef4 z;        // Montgomery
    uint32_t g;   // coset shift, Montgomery
    uint32_t w11; // 11, Montgomery
};

// inv[i] = 1 / (z - x_i), Montgomery EF4, rows [RUN t, RUN t + RUN) per lane.
__global__ __launch_bounds__(DEEP_THREADS) void k_deep_inv(uint32_t *__restrict__ inv, uint64_t height, int log_h,
                                                           DeepConsts k, const uint32_t *__restrict__ lo,
                                                           const uint32_t *__restrict__ hi)
{
    const uint64_t r0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * INV_RUN;
    if (r0 >= height) return;
    const int run = height - r0 < (uint64_t)INV_RUN ? (int)(height - r0) : INV_RUN;
    Ef4 d[INV_RUN], pre[INV_RUN];
#pragma unroll
    for (int j = 0; j < INV_RUN; j++) {
        if (j < run) {
            const uint32_t x = Bb::mul(k.g, root_pow(lo, hi, brev((uint32_t)(r0 + j), log_h), log_h));
            d[j] = k.z;
            d[j].c[0] = Bb::sub(k.z.c[0], x);
        } else {
            d[j] = Ef4{{Bb::one(), 0u, 0u, 0u}};
        }
        pre[j] = j ? ef4_mul(pre[j - 1], d[j], k.w11) : d[j];
    }
    Ef4 acc = ef4_inv(pre[INV_RUN - 1], k.w11);
#pragma unroll
    for (int j = INV_RUN - 1; j >= 0; j--) {
        const Ef4 v = j ? ef4_mul(acc, pre[j - 1], k.w11) : acc;
        if (j < run) ef4_st(inv + (r0 + j) * 4, v);
        if (j) acc = ef4_mul(acc, d[j], k.w11);
    }
}

// partial[blk][col] = sum over the block's rows i of e_ik * x_i * inv_i
// (canonical EF4).  Rows [COLSUM_ROWS blk, ...) of the first n rows; grid.y
// takes 64-column chunks.  Each wave walks every fourth row, lane = column, so
// a load is 64 consecutive columns of one row; the four waves' sums meet in
// LDS.  The row weights w_i = x_i inv_i are wave-uniform: the block writes
// them to LDS, in halves of COLSUM_HALF rows, as the 16-bit halves of their
// canonical coefficients in doubles, and every element adds e_ik times them
// with 8 exact FP64 fmas (each product < 2^47; the sums are reduced with
// Bf::red after every COLSUM_BATCH rows, so they stay below 2^53).
constexpr int COLSUM_HALF = COLSUM_ROWS / 2;
constexpr int COLSUM_BATCH = 16;

__global__ __launch_bounds__(DEEP_THREADS) void k_deep_colsum(const uint32_t *__restrict__ lde, uint32_t width,
                                                              uint64_t n, int log_h, const uint32_t *__restrict__ inv,
                                                              DeepConsts k, const uint32_t *__restrict__ lo,
                                                              const uint32_t *__restrict__ hi,
                                                              uint32_t *__restrict__ partial)
{
    __shared__ double w[COLSUM_HALF][8];  // [lo0..lo3, hi0..hi3] of canonical x_i inv_i
    __shared__ uint4 red[DEEP_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * COLSUM_ROWS;
    const int rows = n - base < (uint64_t)COLSUM_ROWS ? (int)(n - base) : COLSUM_ROWS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = blockIdx.y * 64 + lane;
    constexpr int STEP = DEEP_THREADS / 64;
    double al[4] = {0, 0, 0, 0}, ah[4] = {0, 0, 0, 0};
    for (int h0 = 0; h0 < rows; h0 += COLSUM_HALF) {
        const int hr = rows - h0 < COLSUM_HALF ? rows - h0 : COLSUM_HALF;
        __syncthreads();  // the previous half's reads are done
        for (int r = threadIdx.x; r < hr; r += DEEP_THREADS) {
            const uint64_t i = base + h0 + r;
            const uint32_t x = Bb::mul(k.g, root_pow(lo, hi, brev((uint32_t)i, log_h), log_h));
            const Ef4 v = ef4_mul_base(ef4_ld(inv + i * 4), x);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t c = Bb::to_canon(v.c[j]);
                w[r][j] = (double)(c & 0xFFFFu);
                w[r][4 + j] = (double)(c >> 16);
            }
        }
        __syncthreads();
        if (col < width) {
            const uint32_t *e = lde + (base + h0) * width + col;
            int r = wave;
            // full batches: all COLSUM_BATCH loads issue before the first use
            for (; r + STEP * (COLSUM_BATCH - 1) < hr; r += STEP * COLSUM_BATCH) {
                uint32_t v[COLSUM_BATCH];
#pragma unroll
                for (int u = 0; u < COLSUM_BATCH; u++) v[u] = e[(uint64_t)(r + STEP * u) * width];
#pragma unroll
                for (int u = 0; u < COLSUM_BATCH; u++) {
                    const double ev = (double)v[u];
                    const double *wr = w[r + STEP * u];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        al[j] = __builtin_fma(ev, wr[j], al[j]);
                        ah[j] = __builtin_fma(ev, wr[4 + j], ah[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    al[j] = Bf::red(al[j]);
                    ah[j] = Bf::red(ah[j]);
                }
            }
            // the tail: fewer than COLSUM_BATCH rows
            for (; r < hr; r += STEP) {
                const double ev = (double)e[(uint64_t)r * width];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    al[j] = __builtin_fma(ev, w[r][j], al[j]);
                    ah[j] = __builtin_fma(ev, w[r][4 + j], ah[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                al[j] = Bf::red(al[j]);
                ah[j] = Bf::red(ah[j]);
            }
        }
    }
    Ef4 acc{{0u, 0u, 0u, 0u}};
    if (col < width) {
#pragma unroll
        for (int j = 0; j < 4; j++) acc.c[j] = Bf::canon(__builtin_fma(ah[j], 65536.0, al[j]));
    }
    red[threadIdx.x] = make_uint4(acc.c[0], acc.c[1], acc.c[2], acc.c[3]);
    __syncthreads();
    if (wave == 0 && col < width) {
#pragma unroll
        for (int q = 1; q < DEEP_THREADS / 64; q++) {
            const uint4 v = red[q * 64 + lane];
            acc = ef4_add(acc, Ef4{{v.x, v.y, v.z, v.w}});
        }
        ef4_st(partial + ((uint64_t)blockIdx.x * width + col) * 4, acc);
    }
}

// ys[col] = scale * sum over blocks of partial[blk][col]; scale Montgomery.
// One block per column: a strided sum per lane, then a tree in LDS.
__global__ __launch_bounds__(DEEP_THREADS) void k_deep_colsum_final(const uint32_t *__restrict__ partial,
                                                                    uint32_t nblk, uint32_t width, Ef4 scale,
                                                                    uint32_t w11, uint32_t *__restrict__ ys)
{
    __shared__ uint4 red[DEEP_THREADS];
    const uint32_t col = blockIdx.x;
    Ef4 acc{{0u, 0u, 0u, 0u}};
    for (uint32_t b = threadIdx.x; b < nblk; b += DEEP_THREADS)
        acc = ef4_add(acc, ef4_ld(partial + ((uint64_t)b * width + col) * 4));
    red[threadIdx.x] = make_uint4(acc.c[0], acc.c[1], acc.c[2], acc.c[3]);
    __syncthreads();
    for (int h = DEEP_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const uint4 a = red[threadIdx.x], b = red[threadIdx.x + h];
            const Ef4 s2 = ef4_add(Ef4{{a.x, a.y, a.z, a.w}}, Ef4{{b.x, b.y, b.z, b.w}});
            red[threadIdx.x] = make_uint4(s2.c[0], s2.c[1], s2.c[2], s2.c[3]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint4 v = red[0];
        ef4_st(ys + (uint64_t)col * 4, ef4_mul(Ef4{{v.x, v.y, v.z, v.w}}, scale, w11));
    }
}

// ro_i += (ry - sum_k c_k e_ik) * inv_i with c_k = alpha^off alpha^k and
// ry = alpha^off sum_k alpha^k y_k (canonical).  One lane per row.  c_k is
// given as 8 doubles per column, the 16-bit halves of its 4 coefficients
// [lo0..lo3, hi0..hi3]; every lane reads the same column at the same time, so
// they are wave-uniform (scalar) loads.  sum_k c_k e_ik accumulates exactly in
// FP64 (each product < 2^47, reduced with Bf::red after every REDUCE_COLS
// columns, sums stay below 2^52).  A wave stages 64 rows x REDUCE_COLS
// columns through LDS with row-contiguous loads (row stride padded by one
// word, so lane l's column-k read hits bank l + k).
constexpr int REDUCE_COLS = 32;
constexpr int REDUCE_WAVES = DEEP_THREADS / 64;

__global__ __launch_bounds__(DEEP_THREADS) void k_deep_reduce(const uint32_t *__restrict__ lde, uint64_t height,
                                                              uint32_t width, const uint32_t *__restrict__ inv,
                                                              const double *__restrict__ coef, Ef4 ry, uint32_t w11,
                                                              uint32_t *__restrict__ ro)
{
    __shared__ uint32_t tile[REDUCE_WAVES][64][REDUCE_COLS + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t(*t)[REDUCE_COLS + 1] = tile[wave];
    // the block takes REDUCE_WAVES consecutive 64-row tiles per step; every wave
    // runs the same number of steps, so the barriers are uniform
    const uint64_t step = (uint64_t)gridDim.x * REDUCE_WAVES * 64;
    for (uint64_t r0 = ((uint64_t)blockIdx.x * REDUCE_WAVES + wave) * 64; r0 - wave * 64 < height; r0 += step) {
        const uint64_t row = r0 + lane;
        const int nrows = r0 >= height ? 0 : (height - r0 < 64 ? (int)(height - r0) : 64);
        double lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
        for (uint32_t c0 = 0; c0 < width; c0 += REDUCE_COLS) {
            const int kc = width - c0 < (uint32_t)REDUCE_COLS ? (int)(width - c0) : REDUCE_COLS;
            // stage rows r0..r0+63, columns c0..c0+kc-1: lane = column, 2 rows per
            // load; rows past the end re-read the last row (never used), so the
            // 32 loads are unconditional and issue before the first LDS write
            if (nrows && lane < 2 * kc) {
                const int half = lane >= kc, c = lane - half * kc;
                uint32_t v[32];
#pragma unroll
                for (int i = 0; i < 32; i++) {
                    const int r = 2 * i + half < nrows ? 2 * i + half : nrows - 1;
                    v[i] = lde[(r0 + r) * width + c0 + c];
                }
#pragma unroll
                for (int i = 0; i < 32; i++) t[2 * i + half][c] = v[i];
            }
            __syncthreads();
            if (lane < nrows) {
#pragma unroll 8
                for (int k = 0; k < kc; k++) {
                    const double e = (double)t[lane][k];
                    const double *ck = coef + (uint64_t)(c0 + k) * 8;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        lo[j] = __builtin_fma(e, ck[j], lo[j]);
                        hi[j] = __builtin_fma(e, ck[4 + j], hi[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    lo[j] = Bf::red(lo[j]);
                    hi[j] = Bf::red(hi[j]);
                }
            }
            __syncthreads();
        }
        if (lane < nrows) {
            Ef4 acc;
#pragma unroll
            for (int j = 0; j < 4; j++) acc.c[j] = Bf::canon(__builtin_fma(hi[j], 65536.0, lo[j]));
            const Ef4 v = ef4_mul(ef4_sub(ry, acc), ef4_ld(inv + row * 4), w11);
            ef4_st(ro + row * 4, ef4_add(ef4_ld(ro + row * 4), v));
        }
    }
}

// ---------------------------------------------------------------------------
// Goldilocks with base-field challenges (z, alpha in Goldilocks): the same
// three passes in canonical u64 arithmetic.
constexpr int INV_RUN_GL = 16;

__device__ __forceinline__ uint64_t gl_root_pow(const uint64_t *__restrict__ lo, const uint64_t *__restrict__ hi,
                                                uint32_t e, int logN)
{
    uint32_t ep = (e << (LOG_TW_TOP - logN)) & ((1u << LOG_TW_TOP) - 1u);
    return Gl::mul(lo[ep & ((1u << TW_SPLIT) - 1u)], hi[ep >> TW_SPLIT]);
}

__device__ uint64_t gl_inv(uint64_t a)
{
    uint64_t r = 1, b = a, e = Gl::P - 2;
#pragma unroll 1
    while (e) {
        if (e & 1) r = Gl::mul(r, b);
        b = Gl::mul(b, b);
        e >>= 1;
    }
    return r;
}

__global__ __launch_bounds__(DEEP_THREADS) void k_deep_inv_gl(uint64_t *__restrict__ inv, uint64_t height, int log_h,
                                                              uint64_t z, uint64_t g, const uint64_t *__restrict__ lo,
                                                              const uint64_t *__restrict__ hi)
{
    const uint64_t r0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * INV_RUN_GL;
    if (r0 >= height) return;
    const int run = height - r0 < (uint64_t)INV_RUN_GL ? (int)(height - r0) : INV_RUN_GL;
    uint64_t d[INV_RUN_GL], pre[INV_RUN_GL];
#pragma unroll
    for (int j = 0; j < INV_RUN_GL; j++) {
        d[j] = j < run ? Gl::sub(z, Gl::mul(g, gl_root_pow(lo, hi, brev((uint32_t)(r0 + j), log_h), log_h))) : 1;
        pre[j] = j ? Gl::mul(pre[j - 1], d[j]) : d[j];
    }
    uint64_t acc = gl_inv(pre[INV_RUN_GL - 1]);
#pragma unroll
    for (int j = INV_RUN_GL - 1; j >= 0; j--) {
        const uint64_t v = j ? Gl::mul(acc, pre[j - 1]) : acc;
        if (j < run) inv[r0 + j] = v;
        if (j) acc = Gl::mul(acc, d[j]);
    }
}

__global__ __launch_bounds__(DEEP_THREADS) void k_deep_colsum_gl(const uint64_t *__restrict__ lde, uint32_t width,
                                                                 uint64_t n, int log_h,
                                                                 const uint64_t *__restrict__ inv, uint64_t g,
                                                                 const uint64_t *__restrict__ lo,
                                                                 const uint64_t *__restrict__ hi,
                                                                 uint64_t *__restrict__ partial)
{
    __shared__ uint64_t w[COLSUM_ROWS];  // x_i inv_i
    __shared__ uint64_t red[DEEP_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * COLSUM_ROWS;
    const int rows = n - base < (uint64_t)COLSUM_ROWS ? (int)(n - base) : COLSUM_ROWS;
    for (int r = threadIdx.x; r < rows; r += DEEP_THREADS)
        w[r] = Gl::mul(Gl::mul(g, gl_root_pow(lo, hi, brev((uint32_t)(base + r), log_h), log_h)), inv[base + r]);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = blockIdx.y * 64 + lane;
    uint64_t acc = 0;
    if (col < width) {
        const uint64_t *e = lde + base * width + col;
        constexpr int STEP = DEEP_THREADS / 64, BATCH = 16;
        int r = wave;
        for (; r + STEP * (BATCH - 1) < rows; r += STEP * BATCH) {
            uint64_t v[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; u++) v[u] = e[(uint64_t)(r + STEP * u) * width];
#pragma unroll
            for (int u = 0; u < BATCH; u++) acc = Gl::mul_add_lazy(v[u], w[r + STEP * u], acc);
        }
        for (; r < rows; r += STEP) acc = Gl::mul_add_lazy(e[(uint64_t)r * width], w[r], acc);
    }
    red[threadIdx.x] = Gl::canon(acc);
    __syncthreads();
    if (wave == 0 && col < width) {
        acc = red[lane];
#pragma unroll
        for (int q = 1; q < DEEP_THREADS / 64; q++) acc = Gl::add(acc, red[q * 64 + lane]);
        partial[(uint64_t)blockIdx.x * width + col] = acc;
    }
}

__global__ __launch_bounds__(DEEP_THREADS) void k_deep_colsum_final_gl(const uint64_t *__restrict__ partial,
                                                                       uint32_t nblk, uint32_t width, uint64_t scale,
                                                                       uint64_t *__restrict__ ys)
{
    __shared__ uint64_t red[DEEP_THREADS];
    const uint32_t col = blockIdx.x;
    uint64_t acc = 0;
    for (uint32_t b = threadIdx.x; b < nblk; b += DEEP_THREADS) acc = Gl::add(acc, partial[(uint64_t)b * width + col]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = DEEP_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = Gl::add(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) ys[col] = Gl::mul(red[0], scale);
}

// ro_i += (ry - sum_k c_k e_ik) * inv_i, c_k = alpha^off alpha^k (wave-uniform
// scalar loads); one lane per row over an LDS tile of 64 rows x REDUCE_COLS.
__global__ __launch_bounds__(DEEP_THREADS) void k_deep_reduce_gl(const uint64_t *__restrict__ lde, uint64_t height,
                                                                 uint32_t width, const uint64_t *__restrict__ inv,
                                                                 const uint64_t *__restrict__ coef, uint64_t ry,
                                                                 uint64_t *__restrict__ ro)
{
    __shared__ uint64_t tile[REDUCE_WAVES][64][REDUCE_COLS + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t(*t)[REDUCE_COLS + 1] = tile[wave];
    const uint64_t step = (uint64_t)gridDim.x * REDUCE_WAVES * 64;
    for (uint64_t r0 = ((uint64_t)blockIdx.x * REDUCE_WAVES + wave) * 64; r0 - wave * 64 < height; r0 += step) {
        const uint64_t row = r0 + lane;
        const int nrows = r0 >= height ? 0 : (height - r0 < 64 ? (int)(height - r0) : 64);
        uint64_t acc = 0;
        for (uint32_t c0 = 0; c0 < width; c0 += REDUCE_COLS) {
            const int kc = width - c0 < (uint32_t)REDUCE_COLS ? (int)(width - c0) : REDUCE_COLS;
            if (nrows && lane < 2 * kc) {
                const int half = lane >= kc, c = lane - half * kc;
                uint64_t v[32];
#pragma unroll
                for (int i = 0; i < 32; i++) {
                    const int r = 2 * i + half < nrows ? 2 * i + half : nrows - 1;
                    v[i] = lde[(r0 + r) * width + c0 + c];
                }
#pragma unroll
                for (int i = 0; i < 32; i++) t[2 * i + half][c] = v[i];
            }
            __syncthreads();
            if (lane < nrows) {
#pragma unroll 4
                for (int k = 0; k < kc; k++) acc = Gl::mul_add_lazy(t[lane][k], coef[c0 + k], acc);
            }
            __syncthreads();
        }
        if (lane < nrows) ro[row] = Gl::add(ro[row], Gl::mul(Gl::sub(ry, Gl::canon(acc)), inv[row]));
    }
}

}  // namespace

hipError_t launch_deep_inv(const TwiddleSet &tw, void *inv, uint64_t height, const uint32_t *z_mont,
                           uint32_t g_mont, uint32_t w11, hipStream_t st)
{
    DeepConsts k{{{z_mont[0], z_mont[1], z_mont[2], z_mont[3]}}, g_mont, w11};
    int log_h = 0;
    while ((1ull << log_h) < height) log_h++;
    const uint64_t lanes = (height + INV_RUN - 1) / INV_RUN;
    hipLaunchKernelGGL(k_deep_inv, dim3((unsigned)((lanes + DEEP_THREADS - 1) / DEEP_THREADS)), dim3(DEEP_THREADS),
                       0, st, (uint32_t *)inv, height, log_h, k, (const uint32_t *)tw.top_lo_fwd,
                       (const uint32_t *)tw.top_hi_fwd);
    return hipGetLastError();
}

uint64_t deep_partial_blocks(uint64_t n) { return (n + COLSUM_ROWS - 1) / COLSUM_ROWS; }

hipError_t launch_deep_open_values(const TwiddleSet &tw, const void *lde, uint64_t height, uint32_t width,
                                   uint64_t n, const void *inv, const uint32_t *z_mont, uint32_t g_mont, uint32_t w11,
                                   const uint32_t *scale_mont, void *partial, void *ys, hipStream_t st)
{
    DeepConsts k{{{z_mont[0], z_mont[1], z_mont[2], z_mont[3]}}, g_mont, w11};
    int log_h = 0;
    while ((1ull << log_h) < height) log_h++;
    const uint64_t nblk = deep_partial_blocks(n);
    hipLaunchKernelGGL(k_deep_colsum, dim3((unsigned)nblk, (width + 63) / 64), dim3(DEEP_THREADS), 0, st,
                       (const uint32_t *)lde, width,
                       n, log_h, (const uint32_t *)inv, k, (const uint32_t *)tw.top_lo_fwd,
                       (const uint32_t *)tw.top_hi_fwd, (uint32_t *)partial);
    hipError_t e = hipGetLastError();
    if (e) return e;
    Ef4 s{{scale_mont[0], scale_mont[1], scale_mont[2], scale_mont[3]}};
    hipLaunchKernelGGL(k_deep_colsum_final, dim3(width), dim3(DEEP_THREADS), 0, st, (const uint32_t *)partial,
                       (uint32_t)nblk, width, s, w11, (uint32_t *)ys);
    return hipGetLastError();
}

hipError_t launch_deep_reduce(const void *lde, uint64_t height, uint32_t width, const void *inv, const void *coef,
                              const uint32_t *ry, uint32_t w11, void *ro, hipStream_t st)
{
    Ef4 r{{ry[0], ry[1], ry[2], ry[3]}};
    const uint64_t rows_per_block = REDUCE_WAVES * 64;
    uint64_t blocks = (height + rows_per_block - 1) / rows_per_block;
    if (blocks > 4096) blocks = 4096;  // grid-stride beyond 2^20 rows
    hipLaunchKernelGGL(k_deep_reduce, dim3((unsigned)blocks), dim3(DEEP_THREADS), 0, st, (const uint32_t *)lde, height,
                       width, (const uint32_t *)inv, (const double *)coef, r, w11, (uint32_t *)ro);
    return hipGetLastError();
}

}  // namespace zkg

namespace zkg {

static int log2_of(uint64_t height)
{
    int log_h = 0;
    while ((1ull << log_h) < height) log_h++;
    return log_h;
}

hipError_t launch_deep_gl_inv(const TwiddleSet &tw, void *inv, uint64_t height, uint64_t z, uint64_t g,
                              hipStream_t st)
{
    const uint64_t lanes = (height + INV_RUN_GL - 1) / INV_RUN_GL;
    hipLaunchKernelGGL(k_deep_inv_gl, dim3((unsigned)((lanes + DEEP_THREADS - 1) / DEEP_THREADS)), dim3(DEEP_THREADS),
                       0, st, (uint64_t *)inv, height, log2_of(height), z, g, (const uint64_t *)tw.top_lo_fwd,
                       (const uint64_t *)tw.top_hi_fwd);
    return hipGetLastError();
}

hipError_t launch_deep_gl_open_values(const TwiddleSet &tw, const void *lde, uint64_t height, uint32_t width,
                                      uint64_t n, const void *inv, uint64_t g, uint64_t scale, void *partial, void *ys,
                                      hipStream_t st)
{
    const uint64_t nblk = deep_partial_blocks(n);
    hipLaunchKernelGGL(k_deep_colsum_gl, dim3((unsigned)nblk, (width + 63) / 64), dim3(DEEP_THREADS), 0, st,
                       (const uint64_t *)lde, width, n, log2_of(height), (const uint64_t *)inv, g,
                       (const uint64_t *)tw.top_lo_fwd, (const uint64_t *)tw.top_hi_fwd, (uint64_t *)partial);
    hipError_t e = hipGetLastError();
    if (e) return e;
    hipLaunchKernelGGL(k_deep_colsum_final_gl, dim3(width), dim3(DEEP_THREADS), 0, st, (const uint64_t *)partial,
                       (uint32_t)nblk, width, scale, (uint64_t *)ys);
    return hipGetLastError();
}

hipError_t launch_deep_gl_reduce(const void *lde, uint64_t height, uint32_t width, const void *inv, const void *coef,
                                 uint64_t ry, void *ro, hipStream_t st)
{
    const uint64_t rows_per_block = REDUCE_WAVES * 64;
    uint64_t blocks = (height + rows_per_block - 1) / rows_per_block;
    if (blocks > 4096) blocks = 4096;  // grid-stride beyond 2^20 rows
    hipLaunchKernelGGL(k_deep_reduce_gl, dim3((unsigned)blocks), dim3(DEEP_THREADS), 0, st, (const uint64_t *)lde,
                       height, width, (const uint64_t *)inv, (const uint64_t *)coef, ry, (uint64_t *)ro);
    return hipGetLastError();
}

}  // namespace zkg
