// zkg_ntt.hip -- batched column NTTs and the coset LDE for gfx950.
//
// Semantics: p3 TwoAdicSubgroupDft::dft_batch / idft_batch (natural order in
// and out, idft scaled by 1/N; restated from the reference CPU twin
// zkf-metal/src/ntt/p3_adapter.rs:461-572) and the coset LDE that
// TwoAdicFriPcs::commit runs through RuntimeDft (zkf-backends/src/plonky3.rs:
// 101-257): coeffs = idft(evals), coeff_k *= shift^k, zero-pad to B N, dft,
// bit_reverse_rows.
//
// Design (this build's own, not the reference's log2(N) global radix-2 passes
// of zkf-metal/src/ntt/radix2.rs):
//   * A column transform of size N = N1 * N2 (N1, N2 <= 2^11) is a four-step:
//     size-N1 transforms over rows strided by N2, a twiddle omega_N^(a b), and
//     size-N2 transforms over contiguous rows.
//   * Each sub-transform of M = 2^L points runs on a block tile of M points x C
//     columns in 1-3 radix-2^r rounds.  In a round every lane holds one
//     2^r-point sub-DFT in VGPRs (radix-2 DIF fully unrolled; its twiddles are
//     compile-time indices into the resident table, i.e. wave-uniform scalar
//     loads, or for Goldilocks shifts: the roots of order <= 16 are powers of
//     two), applies the inter-round twiddle and writes to the LDS tile in
//     place, so a round needs one barrier instead of one per radix-2 stage.
//     Round 0 loads its inputs straight from HBM (a wave reads C-column row
//     segments); the last round hands its outputs to the caller, which stores
//     them to HBM, keeps them in VGPRs (LDE coefficients) or stages them
//     through the tile (stores with a twiddle product, bit-reversed rows).  The
//     LDS tile is padded by one row every 2^r_last rows so the last round's
//     strided accesses spread over the banks.
//   * The LDE output rows are in bit-reversed order.  Row b*N + q holds
//     p(s_b omega_N^rev(q)) with s_b = shift * omega_{BN}^{rev_B(b)}, so the
//     LDE is one inverse transform followed by B forward size-N transforms of
//     the coefficients weighted by s_b^k.  Passes: A (inverse step 1), B
//     (inverse step 2 fused with the B coset weightings and forward step 1,
//     coefficients kept in VGPRs between cosets), C (forward step 2, canonical
//     store of contiguous output row blocks).  Three HBM round trips for
//     N > 2^11, one for N <= 2^11.
//   * Block -> tile mapping is XCD-aware: tiles that share rows (adjacent
//     column chunks) are given to blocks that land on the same XCD.
#include <vector>

#include "zkg_field.h"
#include "zkg_internal.h"

// This file is compiled once per tile plan (zkg_plans.h): -DZKG_LDE_VARIANT=k
// puts plan k's kernels and entry points into namespace zkg::lde_plan<k>.  The
// plan-independent parts (the reference-structured plain path, scratch sizes,
// the per-shape dispatch) are compiled with plan 0 only.
#ifndef ZKG_LDE_VARIANT
#define ZKG_LDE_VARIANT 0
#endif
#define ZKG_PLAN_NS_(k) lde_plan##k
#define ZKG_PLAN_NS(k) ZKG_PLAN_NS_(k)

namespace zkg {
namespace ZKG_PLAN_NS(ZKG_LDE_VARIANT) {

namespace {

// This plan's parameters.
constexpr int P_RMAX32 = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].rmax32;
constexpr int P_RMAX64 = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].rmax64;
constexpr int P_RMAX64_SMALL = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].rmax64_small;
constexpr int P_SMALL_L64 = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].small_l64;
constexpr int P_UNROLL = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].unroll;
constexpr int P_WAVES32 = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].waves32;
constexpr int P_WAVES64 = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].waves64;
constexpr int P_FIRST_HEAVY = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].first_heavy;
constexpr int P_LEAN64 = LDE_PLAN_TABLE[ZKG_LDE_VARIANT].lean64;

// Radix split of a 2^L-point in-LDS transform: NR rounds of at most 2^rmax
// points per lane, larger first (or, first-heavy, the full radix first when
// the second round would keep at least 2^(rmax-2) points).
__host__ __device__ constexpr int plan_rmax(int elem_bytes, int L)
{
    return elem_bytes == 4 ? P_RMAX32 : (L <= P_SMALL_L64 ? P_RMAX64_SMALL : P_RMAX64);
}
__host__ __device__ constexpr int plan_nr(int rmax, int L)
{
    return ((L + rmax - 1) / rmax < 2 && L >= 4) ? 2 : ((L + rmax - 1) / rmax < 1 ? 1 : (L + rmax - 1) / rmax);
}
__host__ __device__ constexpr bool plan_first_heavy(int rmax, int L)
{
    return P_FIRST_HEAVY && plan_nr(rmax, L) == 2 && L - rmax >= rmax - 2;
}
__host__ __device__ constexpr int plan_logr(int rmax, int L, int i)
{
    return plan_first_heavy(rmax, L) ? (i == 0 ? rmax : L - rmax)
                                     : L / plan_nr(rmax, L) + (i < L % plan_nr(rmax, L) ? 1 : 0);
}

constexpr int NTT_THREADS = 256;

__device__ __forceinline__ uint32_t brev(uint32_t x, int bits)
{
    return bits ? (__brev(x) >> (32 - bits)) : 0u;
}

// Blocks are dispatched round-robin over the 8 XCDs.  Give XCD x the
// contiguous tile range [x*q + min(x, r), ...) so neighbouring tiles share L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks)
{
    uint32_t q = nblocks >> 3, r = nblocks & 7u;
    uint32_t x = b & 7u, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// A product inside the tile passes.  With lean64, Goldilocks takes the
// fewest-slot lazy product and one canonicalisation: fewer issue slots, a
// longer dependency chain; which wins depends on the shape.
template <class F>
__device__ __forceinline__ typename F::T tmul(typename F::T a, typename F::T b)
{
    if constexpr (sizeof(typename F::T) == 8 && P_LEAN64)
        return Gl::canon(Gl::mul_lean(a, b));
    else
        return F::mul(a, b);
}

template <class F>
__device__ __forceinline__ typename F::T top_pow(const typename F::T *__restrict__ lo,
                                                 const typename F::T *__restrict__ hi, uint64_t e, int logN)
{
    uint32_t ep = (uint32_t)(e << (LOG_TW_TOP - logN)) & ((1u << LOG_TW_TOP) - 1u);
    return tmul<F>(lo[ep & ((1u << TW_SPLIT) - 1u)], hi[ep >> TW_SPLIT]);
}

// ---------------------------------------------------------------------------
// Tile plan: M = 2^L points split into NR radix rounds (most significant
// digit first), at most 2^6 (BabyBear) / 2^4 (Goldilocks) points per lane.
template <class F, int L>
struct Plan {
    static constexpr int RMAXLOG = plan_rmax(sizeof(typename F::T), L);
    static constexpr int NR = plan_nr(RMAXLOG, L);
    static constexpr int logr(int i) { return plan_logr(RMAXLOG, L, i); }
    static constexpr int LOGRMAX = logr(0);
    static constexpr int LOGRLAST = logr(NR - 1);
    static constexpr int LOG_TC = L - LOGRMAX;  // lanes per column
    static constexpr int LOG_C = 8 - LOG_TC;    // columns per block tile
    static constexpr int C = 1 << LOG_C;
    static constexpr int M = 1 << L;
    static constexpr int RSTEP = NTT_THREADS >> LOG_C;               // tile rows per sweep of the block
    static constexpr int ITERS = (M << LOG_C) / NTT_THREADS;         // tile elements per lane
    static constexpr int PADROWS = M + (M >> LOGRLAST);
    static constexpr int LDS_ELEMS = PADROWS * C;
    static_assert(LOG_TC <= 8, "plan needs <= 256 lanes per column");
    static_assert(NR <= 3, "inter-round twiddle table holds two rounds per plan");
    // LDS slot of tile row `pos`, column `c`
    static __device__ __forceinline__ int slot(int pos, int c) { return (pos + (pos >> LOGRLAST)) * C + c; }
    // natural transform index held at row `pos` after the rounds: the digit
    // groups of pos (q1 most significant) read back least significant first
    static __device__ __forceinline__ uint32_t index_of(uint32_t pos)
    {
        uint32_t idx = 0;
        int below = L, shift = 0;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            below -= logr(i);
            idx |= ((pos >> below) & ((1u << logr(i)) - 1u)) << shift;
            shift += logr(i);
        }
        return idx;
    }
    // tile row that holds transform index `idx` (inverse of index_of)
    static __device__ __forceinline__ uint32_t pos_of(uint32_t idx)
    {
        uint32_t pos = 0;
        int below = L, shift = 0;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            below -= logr(i);
            pos |= ((idx >> shift) & ((1u << logr(i)) - 1u)) << below;
            shift += logr(i);
        }
        return pos;
    }
};

// Per-block twiddle vectors in LDS are indexed k + k / 32: the rows a wave
// reads at once differ by multiples of 32 and land on distinct banks.
__device__ __forceinline__ int vpad(int k) { return k + (k >> 5); }
template <int M>
constexpr int vpad_len()
{
    return M + (M >> 5) + 1;
}

// Where the twiddles inside a register sub-DFT come from: the tile table
// passed in (TW_TABLE), or, for Goldilocks, the powers of two that the roots of
// order <= 16 are (TW_FWD / TW_INV: the forward or the inverse transform).
enum { TW_TABLE = 0, TW_FWD = 1, TW_INV = 2 };
template <class F>
__host__ __device__ constexpr int tw_mode(bool inverse)
{
    return sizeof(typename F::T) == 8 ? (inverse ? TW_INV : TW_FWD) : TW_TABLE;
}
// omega_2048^e for e a multiple of 128 (a root of order <= 16) as 2^k, k < 192:
// p3's omega_16 is 2^156 (its inverse 2^36), so k = 156 e/128 or 36 e/128 mod 192.
template <int DIR>
__device__ constexpr int gl_root_log2(int e)
{
    return ((DIR == TW_FWD ? 156 : 36) * (e / 128)) % 192;
}

// R-point DFT in registers: radix-2 DIF, natural order in, x[p] = X[rev(p)]
// out.  omega_{2h}^t = omega_2048^(t*1024/h) = tw[t * (1024/h)]; with a
// power-of-two mode the product by 2^k is Gl::mul_pow2, and 2^96 = -1 turns
// the exponents k >= 96 into a swapped difference.
template <class F, int R, int DIR>
__device__ __forceinline__ void reg_dft(typename F::T (&x)[R], const typename F::T *__restrict__ tw)
{
    using T = typename F::T;
    constexpr bool POW2 = DIR != TW_TABLE;
    static_assert(!POW2 || (sizeof(T) == 8 && R <= 16), "power-of-two twiddles: Goldilocks roots of order <= 16");
#pragma unroll
    for (int h = R / 2; h >= 1; h >>= 1) {
#pragma unroll
        for (int i = 0; i < R; i++) {
            if ((i & h) == 0) {
                T a = x[i], b = x[i + h];
                x[i] = F::add(a, b);
                const int t = i & (h - 1);
                if constexpr (POW2) {
                    const int k = gl_root_log2<DIR>(t * (1024 / h));
                    x[i + h] = k < 96 ? F::mul_pow2(F::sub(a, b), k) : F::mul_pow2(F::sub(b, a), k - 96);
                } else {
                    x[i + h] = t == 0 ? F::sub(a, b) : F::mul(F::sub_for_mul(a, b), tw[t * (1024 / h)]);
                }
            }
        }
    }
}

template <int R>
__device__ constexpr int crev(int q)
{
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (q & 1);
        q >>= 1;
    }
    return r;
}

// One radix round of the tile transform (round I of the plan).  Sub-DFT s of
// a column covers rows b*Mi + a + Mi1*j (j < R) and writes output q, times
// omega_Mi^(a q) unless this is the last round, to row b*Mi + q*Mi1 + a: the
// same rows, so rounds in LDS work in place and need no barrier inside.  The
// inter-round twiddles come from the resident table rtw ([q][a] per round).
// The round reads input j of sub-DFT u of the lane through src(u, j, row)
// and hands output q to sink(u, q, row, value); u, j and q are constants once
// the loops unroll, so a source or sink may index a register array by them.
template <class F, int L, int I, int SHIFT, int DIR, class Src, class Sink>
__device__ __forceinline__ void tile_round(const typename F::T *__restrict__ tw, const typename F::T *__restrict__ rtw,
                                           Src src, Sink sink)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    constexpr int LOGR = P::logr(I);
    constexpr int R = 1 << LOGR;
    constexpr int LOG_MI = L - SHIFT;          // current sub-problem size
    constexpr int LOG_MI1 = LOG_MI - LOGR;     // stride between its points
    constexpr int NSUB = (P::M >> LOGR) >> P::LOG_TC;
    const int t = threadIdx.x >> P::LOG_C;
#pragma unroll
    for (int u = 0; u < NSUB; u++) {
        const int s = t + (u << P::LOG_TC);
        const int a = s & ((1 << LOG_MI1) - 1);
        const int base = (s >> LOG_MI1) << LOG_MI;
        T x[R];
#pragma unroll
        for (int j = 0; j < R; j++) x[j] = src(u, j, base + a + (j << LOG_MI1));
        reg_dft<F, R, DIR>(x, tw);
        if constexpr (I + 1 < P::NR) {
            const T *r = rtw + rtw_offset(L, I) + a;
#pragma unroll
            for (int q = 1; q < R; q++) x[crev<R>(q)] = tmul<F>(x[crev<R>(q)], r[q << LOG_MI1]);
        }
#pragma unroll
        for (int q = 0; q < R; q++) sink(u, q, base + (q << LOG_MI1) + a, x[crev<R>(q)]);
    }
}

// Full tile transform of the lane's column c = tid % C.  Round 0 reads its
// inputs from src0 (tile row `row` of the column: HBM, registers, or weighted
// coefficients), the last round hands transform index Plan::index_of(pos) to
// sinkN(u, q, pos, value); only the exchanges between rounds go through the
// LDS tile, one barrier each.  There is no barrier at either end: a caller
// that wrote LDS data a round reads (twiddle vectors) or that reuses the tile
// afterwards places its own (plans of one round have none inside).
template <class F, int L, int DIR, class Src, class Sink>
__device__ __forceinline__ void tile_dft(typename F::T *lds, const typename F::T *__restrict__ tw,
                                         const typename F::T *__restrict__ rtw, Src src0, Sink sinkN)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    const int c = threadIdx.x & (P::C - 1);
    auto from_tile = [&](int, int, int row) { return lds[P::slot(row, c)]; };
    auto to_tile = [&](int, int, int pos, T v) { lds[P::slot(pos, c)] = v; };
    if constexpr (P::NR == 1) {
        tile_round<F, L, 0, 0, DIR>(tw, rtw, src0, sinkN);
    } else {
        tile_round<F, L, 0, 0, DIR>(tw, rtw, src0, to_tile);
        __syncthreads();
        if constexpr (P::NR == 2) {
            tile_round<F, L, 1, P::logr(0), DIR>(tw, rtw, from_tile, sinkN);
        } else {
            tile_round<F, L, 1, P::logr(0), DIR>(tw, rtw, from_tile, to_tile);
            __syncthreads();
            tile_round<F, L, 2, P::logr(0) + P::logr(1), DIR>(tw, rtw, from_tile, sinkN);
        }
    }
}

struct Geo {
    uint32_t width;
    int logN, logN1, logN2, logB;
    uint32_t nchunks;
};

// A lane of a tile pass handles column c = tid % C of its tile.  Whether that
// column exists (the last chunk of a matrix may be partial) is known once per
// lane; a missing column loads column 0 (its transform touches no other
// column) and stores nothing.
#define TILE_STORE_LOOP(P, row) \
    _Pragma("unroll P_UNROLL") \
    for (int row = (int)(threadIdx.x >> P::LOG_C); row < P::M; row += P::RSTEP)

// The last round's outputs through the LDS tile: the caller stores them
// afterwards in TILE_STORE_LOOP order (row `pos` holds transform index
// Plan::index_of(pos)), a few at a time.  Stores that also take a twiddle
// product go this way: issued straight from the last round, with every
// product's operands in flight at once, they cost more VGPRs than the
// exchange saves.
template <class F, int L>
struct SinkToTile {
    typename F::T *lds;
    int c;
    __device__ void operator()(int, int, int pos, typename F::T v) const { lds[Plan<F, L>::slot(pos, c)] = v; }
};

// Round 0 straight from HBM: tile row `row` of the lane's column is matrix row
// row0 + row * row_stride.  A wave reads C columns of 64 / C rows per load,
// the same row segments an LDS staging pass would read, and every load of a
// lane is issued before the first is waited for.
template <class F>
struct TileFromHbm {
    const typename F::T *p;
    uint64_t step;
    __device__ TileFromHbm(const typename F::T *src, uint64_t row0, uint64_t row_stride, uint32_t width, uint32_t col)
        : p(src + row0 * width + (col < width ? col : 0u)), step(row_stride * width)
    {
    }
    __device__ typename F::T operator()(int, int, int row) const { return p[(uint64_t)row * step]; }
};

// Products inside the transforms are always data x constant, and the
// constants (twiddles, coset weights, 1/N) are Montgomery form, so a product
// is the true product whatever scale the data carries: the kernels below
// transform canonical values directly and their outputs are canonical.

// ---------------------------------------------------------------------------
// Four-step pass 1 (also LDE pass A): group n2 reads rows n2 + N2 n1,
// transforms over n1, multiplies by omega_N^(+-n2 k1) and stores element k1 at
// row n2 + N2 k1.
template <class F, int L, int DIR>
__global__ __launch_bounds__(NTT_THREADS) void k_fs_pass1(const typename F::T *__restrict__ src,
                                                          typename F::T *__restrict__ dst, Geo g,
                                                          const typename F::T *__restrict__ tile_tw,
                                                          const typename F::T *__restrict__ round_tw,
                                                          const typename F::T *__restrict__ top_lo,
                                                          const typename F::T *__restrict__ top_hi)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    __shared__ T lds[P::LDS_ELEMS];
    __shared__ T wk[vpad_len<P::M>()];  // omega_N^(n2 k1) by k1
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t n2 = tile / g.nchunks;
    const uint32_t c0 = (tile % g.nchunks) << P::LOG_C;
    const int c = threadIdx.x & (P::C - 1);
    const bool col_ok = c0 + c < g.width;
    for (int k = threadIdx.x; k < P::M; k += NTT_THREADS)
        wk[vpad(k)] = top_pow<F>(top_lo, top_hi, (uint64_t)n2 * k, g.logN);
    tile_dft<F, L, DIR>(lds, tile_tw, round_tw, TileFromHbm<F>(src, n2, 1ull << g.logN2, g.width, c0 + c),
                        SinkToTile<F, L>{lds, c});
    __syncthreads();
    if (!col_ok) return;
    T *d = dst + (uint64_t)n2 * g.width + c0 + c;
    const uint64_t rs = (uint64_t)g.width << g.logN2;
    TILE_STORE_LOOP(P, pos)
    {
        const uint32_t k1 = P::index_of(pos);
        d[k1 * rs] = tmul<F>(lds[P::slot(pos, c)], wk[vpad(k1)]);
    }
}

// Plain four-step pass 2: group k1 reads contiguous rows N2 k1 + n2,
// transforms over n2 and stores element k2 at row k1 + N1 k2, times `scale`
// (1/N for the inverse).
template <class F, int L, int DIR>
__global__ __launch_bounds__(NTT_THREADS) void k_fs_pass2(const typename F::T *__restrict__ src,
                                                          typename F::T *__restrict__ dst, Geo g,
                                                          typename F::T scale,
                                                          const typename F::T *__restrict__ tile_tw,
                                                          const typename F::T *__restrict__ round_tw)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    __shared__ T lds[P::LDS_ELEMS];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t k1 = tile / g.nchunks;
    const uint32_t c0 = (tile % g.nchunks) << P::LOG_C;
    const int c = threadIdx.x & (P::C - 1);
    const bool col_ok = c0 + c < g.width;
    T *d = dst + (uint64_t)k1 * g.width + c0 + c;
    const uint64_t rs = (uint64_t)g.width << g.logN1;
    tile_dft<F, L, DIR>(lds, tile_tw, round_tw, TileFromHbm<F>(src, (uint64_t)k1 << g.logN2, 1, g.width, c0 + c),
                        [&](int, int, int pos, T v) {
                            if (col_ok) d[P::index_of(pos) * rs] = tmul<F>(v, scale);
                        });
}

// Single-pass transform (N <= 2^11): natural in/out, in place.  Every lane
// has loaded its inputs before the round-0 barrier, and a one-round plan reads
// and writes only the lane's own rows, so no store overtakes a load.
template <class F, int L, int DIR>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_small(typename F::T *__restrict__ data, Geo g,
                                                           typename F::T scale,
                                                           const typename F::T *__restrict__ tile_tw,
                                                           const typename F::T *__restrict__ round_tw)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    __shared__ T lds[P::LDS_ELEMS];
    const uint32_t c0 = xcd_tile(blockIdx.x, gridDim.x) << P::LOG_C;
    const int c = threadIdx.x & (P::C - 1);
    const bool col_ok = c0 + c < g.width;
    T *d = data + c0 + c;
    tile_dft<F, L, DIR>(lds, tile_tw, round_tw, TileFromHbm<F>(data, 0, 1, g.width, c0 + c),
                        [&](int, int, int pos, T v) {
                            if (col_ok) d[(uint64_t)P::index_of(pos) * g.width] = tmul<F>(v, scale);
                        });
}

// ---------------------------------------------------------------------------
// The inverse transform's coefficients feed every coset's first forward round.
// A lane's share of them (2^LOGRMAX values) stays in VGPRs when it is at most
// 128 bytes; larger shares (BabyBear 2^11-point tiles) go to an LDS array,
// since 64 more live registers would cost more occupancy than the LDS does.
template <class F, int L>
constexpr bool lde_coef_in_regs()
{
    return (1 << Plan<F, L>::LOGRMAX) * sizeof(typename F::T) <= 128;
}

// Where the inverse transform leaves the coefficients.  The forward round 0
// (one sub-DFT per lane) takes input j, coefficient t + j 2^LOG_TC with
// t = tid / C, from the lane's registers or from the LDS array.  In a
// two-round plan the inverse's last round ends with exactly those
// coefficients in the lane: sub-DFT u (s = t + u 2^LOG_TC) output q is
// coefficient s + q 2^logr(0), i.e. j = u + q 2^(logr(0) - LOG_TC), so they go
// straight to the registers.  Otherwise the inverse ends in the LDS tile and
// the lane gathers them (load_coefficients).
template <class F, int L>
constexpr bool lde_coef_direct()
{
    return lde_coef_in_regs<F, L>() && Plan<F, L>::NR == 2;
}

// After an inverse transform that ended in the LDS tile (row pos holds
// coefficient index_of(pos)): registers get the coefficients of rows
// tid / C + j 2^LOG_TC, or the LDS array gets every coefficient of the tile in
// natural order.
template <class F, int L, bool CREG>
__device__ __forceinline__ void load_coefficients(const typename F::T *lds, typename F::T *creg,
                                                  typename F::T *coef)
{
    using P = Plan<F, L>;
    const int c = threadIdx.x & (P::C - 1);
    if constexpr (CREG) {
#pragma unroll
        for (int j = 0; j < (1 << P::LOGRMAX); j++)
            creg[j] = lds[P::slot(P::pos_of((threadIdx.x >> P::LOG_C) + (j << P::LOG_TC)), c)];
    } else {
        TILE_STORE_LOOP(P, pos)
        {
            coef[(P::index_of(pos) << P::LOG_C) + c] = lds[P::slot(pos, c)];
        }
    }
}

// The inverse transform of an LDE tile, leaving its coefficients in creg
// (lde_coef_in_regs) or in the LDS array coef.  Ends with a barrier when the
// coefficients passed through LDS.
template <class F, int L, class Src>
__device__ __forceinline__ void lde_inverse(typename F::T *lds, const typename F::T *__restrict__ tile_inv,
                                            const typename F::T *__restrict__ round_inv, Src src,
                                            typename F::T *creg, typename F::T *coef)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    constexpr bool CREG = lde_coef_in_regs<F, L>();
    const int c = threadIdx.x & (P::C - 1);
    if constexpr (lde_coef_direct<F, L>()) {
        constexpr int QS = P::logr(0) - P::LOG_TC;
        static_assert(QS >= 0, "two-round plans put the larger radix first");
        tile_dft<F, L, tw_mode<F>(true)>(lds, tile_inv, round_inv, src,
                                         [&](int u, int q, int, T v) { creg[u + (q << QS)] = v; });
    } else {
        tile_dft<F, L, tw_mode<F>(true)>(lds, tile_inv, round_inv, src,
                                         [&](int, int, int pos, T v) { lds[P::slot(pos, c)] = v; });
        __syncthreads();
        load_coefficients<F, L, CREG>(lds, creg, coef);
    }
}

// Round-0 input of a coset's forward transform: coefficient `row` times its
// coset weight wt[vpad(row)], from the lane's registers or the LDS array.
template <class F, int L, bool CREG>
struct WeightedCoef {
    const typename F::T *creg, *coef, *wt;
    int c;
    __device__ typename F::T operator()(int, int j, int row) const
    {
        using P = Plan<F, L>;
        if constexpr (CREG)
            return tmul<F>(creg[j], wt[vpad(row)]);
        else
            return tmul<F>(coef[(row << P::LOG_C) + c], wt[vpad(row)]);
    }
};

// ---------------------------------------------------------------------------
// LDE pass B: group k1 reads contiguous rows N2 k1 + n2 of pass A's output,
// finishes the inverse over n2 (c[k1 + N1 k2]), keeps the coefficients,
// and for each coset b applies s_b^k / N, the forward transform over k2 and
// omega_N^(j2 k1); the result goes to row b N + N2 k1 + j2 of the scratch
// matrix.  The coset weights (by k2) and omega_N^(j2 k1) (by j2) are vectors
// in LDS shared by the tile's columns; the weighting is fused into the first
// forward round.
// Occupancy of pass B: plans with waves32 / waves64 set compile it for at
// least that many waves per SIMD (BabyBear / Goldilocks) for tiles of up to
// 2^10 points; 0 leaves the register budget to the compiler.  2^11-point tiles
// hold too much LDS for more than one block per CU.
template <class F, int L>
constexpr int lde_mid_waves()
{
    const int w = sizeof(typename F::T) == 4 ? P_WAVES32 : P_WAVES64;
    return L <= 10 && w > 0 ? w : 1;
}
// the occupancy plan (zkg_plans.h row 2) is the one that sets waves
static_assert((P_WAVES32 > 0 || P_WAVES64 > 0) == (ZKG_LDE_VARIANT == 2), "plan 2 carries the occupancy attribute");
#if ZKG_LDE_VARIANT == 2
#define ZKG_LDE_MID_OCC __attribute__((amdgpu_waves_per_eu(lde_mid_waves<F, L>())))
#else
#define ZKG_LDE_MID_OCC
#endif

template <class F, int L>
__global__ __launch_bounds__(NTT_THREADS) ZKG_LDE_MID_OCC void k_lde_mid(const typename F::T *__restrict__ src,
                                                         typename F::T *__restrict__ dst, Geo g,
                                                         const typename F::T *__restrict__ tile_inv,
                                                         const typename F::T *__restrict__ tile_fwd,
                                                         const typename F::T *__restrict__ round_inv,
                                                         const typename F::T *__restrict__ round_fwd,
                                                         const typename F::T *__restrict__ top_lo,
                                                         const typename F::T *__restrict__ top_hi,
                                                         const typename F::T *__restrict__ cos_lo,
                                                         const typename F::T *__restrict__ cos_hi)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    __shared__ T lds[P::LDS_ELEMS];
    __shared__ T wj[vpad_len<P::M>()];  // omega_N^(j2 k1) by j2
    __shared__ T wb[vpad_len<P::M>()];  // s_b^(k1 + N1 k2) / N by k2, per coset
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t k1 = tile / g.nchunks;
    const uint32_t c0 = (tile % g.nchunks) << P::LOG_C;
    const int c = threadIdx.x & (P::C - 1);
    const bool col_ok = c0 + c < g.width;
    for (int j = threadIdx.x; j < P::M; j += NTT_THREADS)
        wj[vpad(j)] = top_pow<F>(top_lo, top_hi, (uint64_t)j * k1, g.logN);
    // the coefficients c[k1 + N1 k2] this lane feeds into every coset's first
    // forward round (k2 = tid / C + j 2^LOG_TC), kept in VGPRs across cosets
    constexpr bool CREG = lde_coef_in_regs<F, L>();
    __shared__ T coef[CREG ? 1 : (P::M << P::LOG_C)];  // c[k1 + N1 k2] at [k2][c] otherwise
    T creg[CREG ? (1 << P::LOGRMAX) : 1];
    lde_inverse<F, L>(lds, tile_inv, round_inv, TileFromHbm<F>(src, (uint64_t)k1 << g.logN2, 1, g.width, c0 + c),
                      creg, coef);
    for (int b = 0; b < (1 << g.logB); b++) {
        const T *clo = cos_lo + (size_t)b * (1u << TW_SPLIT);
        const T *chi = cos_hi + (size_t)b * (1u << (LOG_TW_TOP - TW_SPLIT - 2));
        for (int k2 = threadIdx.x; k2 < P::M; k2 += NTT_THREADS) {
            const uint32_t k = k1 + ((uint32_t)k2 << g.logN1);
            wb[vpad(k2)] = tmul<F>(clo[k & ((1u << TW_SPLIT) - 1u)], chi[k >> TW_SPLIT]);
        }
        // wj and wb complete; every lane is done with the tile (the inverse,
        // or the previous coset's stores) and with the previous weights
        __syncthreads();
        tile_dft<F, L, tw_mode<F>(false)>(lds, tile_fwd, round_fwd, WeightedCoef<F, L, CREG>{creg, coef, wb, c},
                                          SinkToTile<F, L>{lds, c});
        __syncthreads();
        if (col_ok) {
            T *d = dst + (((uint64_t)b << g.logN) + ((uint64_t)k1 << g.logN2)) * g.width + c0 + c;
            TILE_STORE_LOOP(P, pos)
            {
                const uint32_t j2 = P::index_of(pos);
                d[(uint64_t)j2 * g.width] = tmul<F>(lds[P::slot(pos, c)], wj[vpad(j2)]);
            }
        }
    }
}

// LDE pass C: coset b, group j2 reads rows b N + N2 k1 + j2, transforms over
// k1 and writes output row b N + rev(j2) N1 + rev(j1) (one contiguous block
// of N1 rows per group).
template <class F, int L>
__global__ __launch_bounds__(NTT_THREADS) void k_lde_final(const typename F::T *__restrict__ src,
                                                           typename F::T *__restrict__ out, Geo g,
                                                           const typename F::T *__restrict__ tile_fwd,
                                                           const typename F::T *__restrict__ round_fwd)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    __shared__ T lds[P::LDS_ELEMS];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t chunk = tile % g.nchunks;
    const uint32_t grp = tile / g.nchunks;  // b * N2 + j2
    const uint32_t j2 = grp & ((1u << g.logN2) - 1u), b = grp >> g.logN2;
    const uint32_t c0 = chunk << P::LOG_C;
    const int c = threadIdx.x & (P::C - 1);
    const bool col_ok = c0 + c < g.width;
    const uint64_t base = (uint64_t)b << g.logN;
    tile_dft<F, L, tw_mode<F>(false)>(lds, tile_fwd, round_fwd,
                                      TileFromHbm<F>(src, base + j2, 1ull << g.logN2, g.width, c0 + c),
                                      [&](int, int, int pos, T v) { lds[P::slot(pos, c)] = v; });
    __syncthreads();
    if (!col_ok) return;
    // Transform index j1 = index_of(pos) goes to output row obase + rev_L(j1).
    // The stores go through the tile: walking pos in order, a wave writes 8
    // adjacent output rows per store, where the last round's lanes hold rows
    // 2^(L - r0) apart.  Stored straight from the round, the same 32-byte
    // segments moved 4% more HBM bytes (fewer partial lines merge in L2) in
    // the same time.
    T *d = out + (base + ((uint64_t)brev(j2, g.logN2) << L)) * g.width + c0 + c;
    TILE_STORE_LOOP(P, pos)
    {
        d[(uint64_t)brev(P::index_of(pos), L) * g.width] = lds[P::slot(pos, c)];
    }
}

// Whole LDE in one pass for N <= 2^11: one block per column chunk.
template <class F, int L>
__global__ __launch_bounds__(NTT_THREADS) void k_lde_small(const typename F::T *__restrict__ in,
                                                           typename F::T *__restrict__ out, Geo g,
                                                           const typename F::T *__restrict__ tile_inv,
                                                           const typename F::T *__restrict__ tile_fwd,
                                                           const typename F::T *__restrict__ round_inv,
                                                           const typename F::T *__restrict__ round_fwd,
                                                           const typename F::T *__restrict__ cos_lo)
{
    using T = typename F::T;
    using P = Plan<F, L>;
    __shared__ T lds[P::LDS_ELEMS];
    __shared__ T wb[vpad_len<P::M>()];  // s_b^k / N by k, per coset
    const uint32_t c0 = xcd_tile(blockIdx.x, gridDim.x) << P::LOG_C;
    const int c = threadIdx.x & (P::C - 1);
    const bool col_ok = c0 + c < g.width;
    constexpr bool CREG = lde_coef_in_regs<F, L>();
    __shared__ T coef[CREG ? 1 : (P::M << P::LOG_C)];  // coefficient k at [k][c] otherwise
    T creg[CREG ? (1 << P::LOGRMAX) : 1];
    lde_inverse<F, L>(lds, tile_inv, round_inv, TileFromHbm<F>(in, 0, 1, g.width, c0 + c), creg, coef);
    for (int b = 0; b < (1 << g.logB); b++) {
        const T *clo = cos_lo + (size_t)b * (1u << TW_SPLIT);
        for (int k = threadIdx.x; k < P::M; k += NTT_THREADS) wb[vpad(k)] = clo[k];
        __syncthreads();
        tile_dft<F, L, tw_mode<F>(false)>(lds, tile_fwd, round_fwd, WeightedCoef<F, L, CREG>{creg, coef, wb, c},
                                          SinkToTile<F, L>{lds, c});
        __syncthreads();
        if (col_ok) {
            T *d = out + ((uint64_t)b << L) * g.width + c0 + c;
            TILE_STORE_LOOP(P, pos)
            {
                d[(uint64_t)brev(P::index_of(pos), L) * g.width] = lds[P::slot(pos, c)];
            }
        }
    }
}

// Height-1 LDE: the constant polynomial, B copies of the row.
template <class F>
__global__ void k_lde_one(const typename F::T *__restrict__ in, typename F::T *__restrict__ out, uint32_t width,
                          int logB)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (width << logB)) out[t] = in[t % width];
}

// ---------------------------------------------------------------------------
// Plain (reference-structured) path: one global radix-2 pass per stage.
template <class F>
__global__ void k_bitrev_rows(const typename F::T *__restrict__ src, typename F::T *__restrict__ dst,
                              uint64_t n, int logn, uint32_t width, uint64_t src_rows, int to_reg)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * width) return;
    uint64_t r = t / width, c = t - r * width;
    uint64_t sr = logn ? (uint64_t)(__brevll(r) >> (64 - logn)) : 0;
    typename F::T v = sr < src_rows ? src[sr * width + c] : typename F::T(0);
    dst[t] = to_reg ? F::from_canon(v) : v;
}

template <class F>
__global__ void k_dit_stage(typename F::T *__restrict__ data, uint64_t n, int logn, int s, uint32_t width,
                            const typename F::T *__restrict__ top_lo, const typename F::T *__restrict__ top_hi)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (n >> 1) * width) return;
    uint64_t bf = t / width, c = t - bf * width;
    uint64_t half = 1ull << s, j = bf & (half - 1), grp = bf >> s;
    uint64_t i0 = (grp << (s + 1)) + j, i1 = i0 + half;
    typename F::T w = top_pow<F>(top_lo, top_hi, j, s + 1);
    typename F::T a = data[i0 * width + c], b = F::mul(data[i1 * width + c], w);
    data[i0 * width + c] = F::add(a, b);
    data[i1 * width + c] = F::sub(a, b);
}

// coefficients (register form, n rows) -> c_k * shift^k / n
template <class F>
__global__ void k_scale_shift(typename F::T *__restrict__ data, uint64_t n, uint32_t width, typename F::T shift,
                              typename F::T ninv)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * width) return;
    uint64_t k = t / width;
    typename F::T w = ninv, base = shift;
    for (uint64_t e = k; e; e >>= 1) {  // shift^k by square-and-multiply per element
        if (e & 1) w = F::mul(w, base);
        base = F::sqr(base);
    }
    data[t] = F::mul(data[t], w);
}

template <class F>
__global__ void k_bitrev_swap_canon(typename F::T *__restrict__ data, uint64_t n, int logn, uint32_t width)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * width) return;
    uint64_t r = t / width, c = t - r * width;
    uint64_t j = logn ? (uint64_t)(__brevll(r) >> (64 - logn)) : 0;
    if (r > j) return;
    typename F::T a = data[r * width + c], b = data[j * width + c];
    data[r * width + c] = F::to_canon(b);
    if (r != j) data[j * width + c] = F::to_canon(a);
}

// ---------------------------------------------------------------------------
template <class F>
struct Tabs {
    const typename F::T *tile_fwd, *tile_inv, *round_fwd, *round_inv, *lo_fwd, *hi_fwd, *lo_inv, *hi_inv;
    explicit Tabs(const TwiddleSet &t)
        : tile_fwd((const typename F::T *)t.tile_fwd), tile_inv((const typename F::T *)t.tile_inv),
          round_fwd((const typename F::T *)t.round_fwd[ZKG_LDE_VARIANT]),
          round_inv((const typename F::T *)t.round_inv[ZKG_LDE_VARIANT]),
          lo_fwd((const typename F::T *)t.top_lo_fwd), hi_fwd((const typename F::T *)t.top_hi_fwd),
          lo_inv((const typename F::T *)t.top_lo_inv), hi_inv((const typename F::T *)t.top_hi_inv)
    {
    }
};

template <class F>
void fill_round_twiddles_t(typename F::T w2048, typename F::T *tab)
{
    using T = typename F::T;
    for (int i = 0; i < RTW_ENTRIES; i++) tab[i] = 0;
    for (int L = 1; L <= LOG_TW_TILE; L++) {
        const int RM = plan_rmax(sizeof(T), L);
        const int nr = plan_nr(RM, L);
        int shift = 0;
        for (int I = 0; I + 1 < nr; I++) {
            const int logr = plan_logr(RM, L, I), log_mi = L - shift, log_mi1 = log_mi - logr;
            T wm = F::one();  // omega_Mi = omega_2048^(2048 / Mi)
            for (int k = 0; k < (1 << (LOG_TW_TILE - log_mi)); k++) wm = F::mul(wm, w2048);
            std::vector<T> pw((size_t)1 << log_mi);
            T acc = F::one();
            for (auto &v : pw) {
                v = acc;
                acc = F::mul(acc, wm);
            }
            for (int q = 0; q < (1 << logr); q++)
                for (int a = 0; a < (1 << log_mi1); a++)
                    tab[rtw_offset(L, I) + (q << log_mi1) + a] = pw[(size_t)(a * q) & ((1u << log_mi) - 1u)];
            shift += logr;
        }
    }
}

int ilog2(uint64_t n)
{
    int l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

Geo make_geo(uint64_t height, uint32_t width, int logB)
{
    Geo g;
    g.width = width;
    g.logN = ilog2(height);
    g.logN1 = g.logN / 2;
    g.logN2 = g.logN - g.logN1;
    g.logB = logB;
    g.nchunks = 0;
    return g;
}

template <class F, int L>
uint32_t chunks(uint32_t width)
{
    return (width + Plan<F, L>::C - 1) >> Plan<F, L>::LOG_C;
}

template <class F>
typename F::T host_ninv(uint64_t n)
{
    using T = typename F::T;
    T inv = F::from_canon(1);
    T half;
    if constexpr (sizeof(T) == 8)
        half = (F::P + 1) / 2;
    else
        half = F::from_canon((F::P + 1) / 2);
    for (uint64_t m = n; m > 1; m >>= 1) inv = F::mul(inv, half);
    return inv;
}

// Runtime L -> template L dispatch for L = 1 .. 11.
template <template <class, int> class K, class F, class... A>
hipError_t dispatch_l(int L, A &&...a)
{
    switch (L) {
    case 1: return K<F, 1>::run(a...);
    case 2: return K<F, 2>::run(a...);
    case 3: return K<F, 3>::run(a...);
    case 4: return K<F, 4>::run(a...);
    case 5: return K<F, 5>::run(a...);
    case 6: return K<F, 6>::run(a...);
    case 7: return K<F, 7>::run(a...);
    case 8: return K<F, 8>::run(a...);
    case 9: return K<F, 9>::run(a...);
    case 10: return K<F, 10>::run(a...);
    case 11: return K<F, 11>::run(a...);
    default: return hipErrorInvalidValue;
    }
}

// The plain transform passes run in either direction: `inverse` picks the
// kernel instance (the same one for BabyBear, whose twiddles all come from the
// tables passed in).
template <class F, int L>
struct RunPass1 {
    static hipError_t run(const void *src, void *dst, Geo g, const void *tile, const void *round, const void *lo,
                          const void *hi, bool inverse, hipStream_t st)
    {
        using T = typename F::T;
        g.nchunks = chunks<F, L>(g.width);
        auto k = inverse ? k_fs_pass1<F, L, tw_mode<F>(true)> : k_fs_pass1<F, L, tw_mode<F>(false)>;
        hipLaunchKernelGGL(k, dim3(g.nchunks << g.logN2), dim3(NTT_THREADS), 0, st, (const T *)src, (T *)dst, g,
                           (const T *)tile, (const T *)round, (const T *)lo, (const T *)hi);
        return hipGetLastError();
    }
};

template <class F, int L>
struct RunPass2 {
    static hipError_t run(const void *src, void *dst, Geo g, typename F::T scale, const void *tile, const void *round,
                          bool inverse, hipStream_t st)
    {
        using T = typename F::T;
        g.nchunks = chunks<F, L>(g.width);
        auto k = inverse ? k_fs_pass2<F, L, tw_mode<F>(true)> : k_fs_pass2<F, L, tw_mode<F>(false)>;
        hipLaunchKernelGGL(k, dim3(g.nchunks << g.logN1), dim3(NTT_THREADS), 0, st, (const T *)src, (T *)dst, g,
                           scale, (const T *)tile, (const T *)round);
        return hipGetLastError();
    }
};

template <class F, int L>
struct RunSmall {
    static hipError_t run(void *data, Geo g, typename F::T scale, const void *tile, const void *round, bool inverse,
                          hipStream_t st)
    {
        using T = typename F::T;
        g.nchunks = chunks<F, L>(g.width);
        auto k = inverse ? k_ntt_small<F, L, tw_mode<F>(true)> : k_ntt_small<F, L, tw_mode<F>(false)>;
        hipLaunchKernelGGL(k, dim3(g.nchunks), dim3(NTT_THREADS), 0, st, (T *)data, g, scale, (const T *)tile,
                           (const T *)round);
        return hipGetLastError();
    }
};

template <class F, int L>
struct RunMid {
    static hipError_t run(const void *src, void *dst, Geo g, const Tabs<F> &t, const CosetSet &cs, hipStream_t st)
    {
        using T = typename F::T;
        g.nchunks = chunks<F, L>(g.width);
        hipLaunchKernelGGL((k_lde_mid<F, L>), dim3(g.nchunks << g.logN1), dim3(NTT_THREADS), 0, st, (const T *)src,
                           (T *)dst, g, t.tile_inv, t.tile_fwd, t.round_inv, t.round_fwd, t.lo_fwd, t.hi_fwd,
                           (const T *)cs.lo, (const T *)cs.hi);
        return hipGetLastError();
    }
};

template <class F, int L>
struct RunFinal {
    static hipError_t run(const void *src, void *out, Geo g, const Tabs<F> &t, hipStream_t st)
    {
        using T = typename F::T;
        g.nchunks = chunks<F, L>(g.width);
        hipLaunchKernelGGL((k_lde_final<F, L>), dim3(g.nchunks << (g.logN2 + g.logB)), dim3(NTT_THREADS), 0, st,
                           (const T *)src, (T *)out, g, t.tile_fwd, t.round_fwd);
        return hipGetLastError();
    }
};

template <class F, int L>
struct RunLdeSmall {
    static hipError_t run(const void *in, void *out, Geo g, const Tabs<F> &t, const CosetSet &cs, hipStream_t st)
    {
        using T = typename F::T;
        g.nchunks = chunks<F, L>(g.width);
        hipLaunchKernelGGL((k_lde_small<F, L>), dim3(g.nchunks), dim3(NTT_THREADS), 0, st, (const T *)in, (T *)out, g,
                           t.tile_inv, t.tile_fwd, t.round_inv, t.round_fwd, (const T *)cs.lo);
        return hipGetLastError();
    }
};

template <class F>
hipError_t coset_lde_impl(const TwiddleSet &tws, const CosetSet &cs, const void *in, void *out, void *tmp,
                          uint64_t height, uint32_t width, uint32_t log_blowup, hipStream_t st)
{
    using T = typename F::T;
    Tabs<F> t(tws);
    Geo g = make_geo(height, width, log_blowup);
    if (g.logN == 0) {
        uint32_t n = width << log_blowup;
        hipLaunchKernelGGL(k_lde_one<F>, dim3((n + 255) / 256), dim3(256), 0, st, (const T *)in, (T *)out, width,
                           (int)log_blowup);
        return hipGetLastError();
    }
    if (g.logN <= LOG_TW_TILE) return dispatch_l<RunLdeSmall, F>(g.logN, in, out, g, t, cs, st);
    // pass A: in -> out[0 .. N*W) (scratch), pass B: -> tmp (B*N*W), pass C: -> out
    hipError_t e = dispatch_l<RunPass1, F>(g.logN1, in, out, g, (const void *)t.tile_inv, (const void *)t.round_inv,
                                           (const void *)t.lo_inv, (const void *)t.hi_inv, true, st);
    if (e) return e;
    e = dispatch_l<RunMid, F>(g.logN2, (const void *)out, tmp, g, t, cs, st);
    if (e) return e;
    return dispatch_l<RunFinal, F>(g.logN1, (const void *)tmp, out, g, t, st);
}

template <class F>
hipError_t ntt_impl(const TwiddleSet &tws, void *data, void *tmp, uint64_t height, uint32_t width, int inverse,
                    hipStream_t st)
{
    using T = typename F::T;
    Tabs<F> t(tws);
    Geo g = make_geo(height, width, 0);
    if (g.logN == 0) return hipSuccess;
    T scale = inverse ? host_ninv<F>(height) : F::from_canon(1);
    const void *tile = inverse ? t.tile_inv : t.tile_fwd;
    const void *round = inverse ? t.round_inv : t.round_fwd;
    const bool inv = inverse != 0;
    if (g.logN <= LOG_TW_TILE) return dispatch_l<RunSmall, F>(g.logN, data, g, scale, tile, round, inv, st);
    hipError_t e = dispatch_l<RunPass1, F>(g.logN1, (const void *)data, tmp, g, tile, round,
                                           (const void *)(inv ? t.lo_inv : t.lo_fwd),
                                           (const void *)(inv ? t.hi_inv : t.hi_fwd), inv, st);
    if (e) return e;
    return dispatch_l<RunPass2, F>(g.logN2, (const void *)tmp, data, g, scale, tile, round, inv, st);
}

template <class F>
hipError_t lde_plain_impl(const TwiddleSet &tws, const void *in, void *out, void *tmp, uint64_t height,
                          uint32_t width, uint32_t log_blowup, uint64_t shift, hipStream_t st, bool sync_each)
{
    using T = typename F::T;
    Tabs<F> t(tws);
    const int logN = ilog2(height), logB = logN + (int)log_blowup;
    const uint64_t big = height << log_blowup;
    T *o = (T *)out, *c = (T *)tmp;
    const unsigned TB = 256;
    auto grid = [&](uint64_t n) { return dim3((unsigned)((n + TB - 1) / TB)); };
    auto step = [&]() -> hipError_t {
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && sync_each) e = hipStreamSynchronize(st);
        return e;
    };
    hipError_t e;
    // iNTT: bit-reverse permute, then log N DIT stages with inverse twiddles
    hipLaunchKernelGGL(k_bitrev_rows<F>, grid(height * width), dim3(TB), 0, st, (const T *)in, c, height, logN, width,
                       height, 1);
    if ((e = step())) return e;
    for (int s = 0; s < logN; s++) {
        hipLaunchKernelGGL(k_dit_stage<F>, grid((height >> 1) * width), dim3(TB), 0, st, c, height, logN, s, width,
                           t.lo_inv, t.hi_inv);
        if ((e = step())) return e;
    }
    // coefficients * shift^k / N (one pow per element)
    hipLaunchKernelGGL(k_scale_shift<F>, grid(height * width), dim3(TB), 0, st, c, height, width,
                       F::from_canon((T)shift), host_ninv<F>(height));
    if ((e = step())) return e;
    // zero-pad to B N while bit-reversing, then log(B N) DIT stages
    hipLaunchKernelGGL(k_bitrev_rows<F>, grid(big * width), dim3(TB), 0, st, (const T *)c, o, big, logB, width,
                       height, 0);
    if ((e = step())) return e;
    for (int s = 0; s < logB; s++) {
        hipLaunchKernelGGL(k_dit_stage<F>, grid((big >> 1) * width), dim3(TB), 0, st, o, big, logB, s, width,
                           t.lo_fwd, t.hi_fwd);
        if ((e = step())) return e;
    }
    // bit_reverse_rows of the evaluations, canonical store
    hipLaunchKernelGGL(k_bitrev_swap_canon<F>, grid(big * width), dim3(TB), 0, st, o, big, logB, width);
    return step();
}

}  // namespace

hipError_t coset_lde(int field, const TwiddleSet &tw, const CosetSet &cs, const void *in, void *out, void *tmp,
                     uint64_t height, uint32_t width, uint32_t log_blowup, hipStream_t st)
{
    if (field == FIELD_GL) return coset_lde_impl<Gl>(tw, cs, in, out, tmp, height, width, log_blowup, st);
    return coset_lde_impl<Bb>(tw, cs, in, out, tmp, height, width, log_blowup, st);
}

hipError_t ntt(int field, const TwiddleSet &tw, void *data, void *tmp, uint64_t height, uint32_t width, int inverse,
               hipStream_t st)
{
    if (field == FIELD_GL) return ntt_impl<Gl>(tw, data, tmp, height, width, inverse, st);
    return ntt_impl<Bb>(tw, data, tmp, height, width, inverse, st);
}

void fill_round_twiddles(int field, uint64_t w2048, void *table)
{
    if (field == FIELD_GL)
        fill_round_twiddles_t<Gl>(w2048, (uint64_t *)table);
    else
        fill_round_twiddles_t<Bb>((uint32_t)w2048, (uint32_t *)table);
}

#if ZKG_LDE_VARIANT == 0
hipError_t lde_plain(int field, const TwiddleSet &tw, const void *in, void *out, void *tmp, uint64_t height,
                     uint32_t width, uint32_t log_blowup, uint64_t shift, hipStream_t st, bool sync_each)
{
    if (field == FIELD_GL) return lde_plain_impl<Gl>(tw, in, out, tmp, height, width, log_blowup, shift, st, sync_each);
    return lde_plain_impl<Bb>(tw, in, out, tmp, height, width, log_blowup, shift, st, sync_each);
}
#endif

}  // namespace lde_plan<k>

#if ZKG_LDE_VARIANT == 0
uint64_t coset_lde_tmp_bytes(int field, uint64_t height, uint32_t width, uint32_t log_blowup)
{
    if (lde_plan0::ilog2(height) <= LOG_TW_TILE) return 0;
    uint64_t es = field == FIELD_GL ? 8 : 4;
    return (height << log_blowup) * width * es;
}

uint64_t ntt_tmp_bytes(int field, uint64_t height, uint32_t width)
{
    if (lde_plan0::ilog2(height) <= LOG_TW_TILE) return 0;
    return height * width * (field == FIELD_GL ? 8 : 4);
}

void fill_round_twiddles(int plan, int field, uint64_t w2048, void *table)
{
    switch (plan) {
    case 1: return lde_plan1::fill_round_twiddles(field, w2048, table);
    case 2: return lde_plan2::fill_round_twiddles(field, w2048, table);
    case 3: return lde_plan3::fill_round_twiddles(field, w2048, table);
    default: return lde_plan0::fill_round_twiddles(field, w2048, table);
    }
}

hipError_t launch_coset_lde_plan(int plan, int field, const TwiddleSet &tw, const CosetSet &cs, const void *in,
                                 void *out, void *tmp, uint64_t height, uint32_t width, uint32_t log_blowup,
                                 hipStream_t st)
{
    switch (plan) {
    case 1: return lde_plan1::coset_lde(field, tw, cs, in, out, tmp, height, width, log_blowup, st);
    case 2: return lde_plan2::coset_lde(field, tw, cs, in, out, tmp, height, width, log_blowup, st);
    case 3: return lde_plan3::coset_lde(field, tw, cs, in, out, tmp, height, width, log_blowup, st);
    default: return lde_plan0::coset_lde(field, tw, cs, in, out, tmp, height, width, log_blowup, st);
    }
}

// The plan of each transform is chosen by its shape (lde_plan_for); every
// plan computes the same values.
hipError_t launch_coset_lde(int field, const TwiddleSet &tw, const CosetSet &cs, const void *in, void *out, void *tmp,
                            uint64_t height, uint32_t width, uint32_t log_blowup, hipStream_t st)
{
    return launch_coset_lde_plan(lde_plan_for(field, lde_plan0::ilog2(height), width, log_blowup), field, tw, cs, in,
                                 out, tmp, height, width, log_blowup, st);
}

hipError_t launch_ntt(int field, const TwiddleSet &tw, void *data, void *tmp, uint64_t height, uint32_t width,
                      int inverse, hipStream_t st)
{
    switch (lde_plan_for(field, lde_plan0::ilog2(height), width, 0)) {
    case 1: return lde_plan1::ntt(field, tw, data, tmp, height, width, inverse, st);
    case 2: return lde_plan2::ntt(field, tw, data, tmp, height, width, inverse, st);
    case 3: return lde_plan3::ntt(field, tw, data, tmp, height, width, inverse, st);
    default: return lde_plan0::ntt(field, tw, data, tmp, height, width, inverse, st);
    }
}

hipError_t launch_lde_plain(int field, const TwiddleSet &tw, const void *in, void *out, void *tmp, uint64_t height,
                            uint32_t width, uint32_t log_blowup, uint64_t shift, hipStream_t st, bool sync_each)
{
    return lde_plan0::lde_plain(field, tw, in, out, tmp, height, width, log_blowup, shift, st, sync_each);
}
#endif

}  // namespace zkg
