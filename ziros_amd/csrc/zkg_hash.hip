// zkg_hash.hip -- Poseidon2 row hashing, Merkle trees and FRI folding for gfx950.
//
// Semantics (reference file:line):
//   * permutation: zkf-metal/src/shaders/poseidon2.metal:17-264;
//   * row hash = PaddingFreeSponge<P, W, W/2, W/2> and node = TruncatedPermutation
//     <P, 2, W/2, W> (zkf-backends/src/plonky3.rs:56-57; restated at
//     zkf-backends/src/wrapping/poseidon2_goldilocks.rs:143-225,372-398);
//   * tree layers: vendor/p3-merkle-tree-gpu/src/merkle_tree.rs:84-403
//     (power-of-two heights: layer i+1 = compress of pairs of layer i);
//   * openings: vendor/p3-merkle-tree-gpu/src/mmcs.rs:179-219;
//   * FRI fold: p3-fri 0.4.2 TwoAdicFriFolding::fold_matrix (arity 2) and the
//     reference FriAccelerator formula zkf-metal/src/shaders/fri.metal:14-53.
//
// Design: one Poseidon2 state per lane, held in VGPRs for the whole
// permutation (constants are wave-uniform scalar loads).  Merkle levels are
// fused: a block compresses its inputs 3 levels down, passing each level
// through LDS.  The small levels at the top of a tree spread each compression
// over W lanes, 5 (W = 16) or 6 (W = 8) levels per launch.  Goldilocks work
// of middle size (short matrices, the first levels below the fused ones) runs
// with one 4-chunk of the state per lane.  The FRI fold of a
// pair of outputs is fused with the Poseidon2 hash of the next layer's leaf
// (one leaf = two folded EF4 values = 8 base elements = one absorption).
#include "zkg_field.h"
#include "zkg_internal.h"
#include "zkg_poseidon2.h"
#include "zkg_poseidon2_lanes.h"

namespace zkg {

namespace {

constexpr int HASH_THREADS = 256;
constexpr int MERKLE_THREADS = 256;
constexpr int MAX_FUSE = 9;  // 2 * MERKLE_THREADS leaves -> 1 root

// A tuning override from the environment: 2^k for NAME=<k>, 0 <= k <= 30,
// else `dflt`.
uint64_t env_pow2(const char *name, uint64_t dflt)
{
    const char *e = getenv(name);
    int k = -1;
    if (!e || sscanf(e, "%d", &k) != 1 || k < 0 || k > 30) return dflt;
    return 1ull << k;
}

// Merkle levels of at most lanes_split_max digests run with W lanes per
// compression (one element each); Goldilocks levels above that and of at most
// chunks_split_max digests with W/4 lanes per compression (one 4-chunk each);
// larger levels with one compression per lane.  Row hashes of Goldilocks
// matrices of at most hash_chunks_max rows run with W/4 lanes per row.  Below
// these sizes one compression (or row) per lane leaves the chip mostly idle,
// while the W-lane form repeats the internal S-box in every lane and has the
// most work per compression.  The sizes come from the tuning layers
// (zkg_tuning.hip: override > learned > static table > built-in).
template <class F, int W>
uint64_t lanes_split_max()
{
    return 1ull << split_value(std::is_same<F, Bb>::value ? FIELD_BB : FIELD_GL, W, SPLIT_LANES);
}
template <int W>
uint64_t chunks_split_max()
{
    return 1ull << split_value(FIELD_GL, W, SPLIT_CHUNKS);
}
template <int W>
uint64_t hash_chunks_max()
{
    return 1ull << split_value(FIELD_GL, W, SPLIT_HASH_CHUNKS);
}

// BabyBear width-16 permutations run in exact FP64 (Bf); the Montgomery form
// serves the 16-lane tree-top kernel and the FRI fold.
template <class F, int W>
constexpr bool IS_BB16 = std::is_same<F, Bb>::value && W == 16;

__device__ __forceinline__ uint32_t brev(uint32_t x, int bits)
{
    return bits ? (__brev(x) >> (32 - bits)) : 0u;
}

// ---------------------------------------------------------------- permute
template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_p2_permute(typename F::T *__restrict__ states, uint64_t n,
                                                             const typename F::T *__restrict__ rc,
                                                             const typename F::T *__restrict__ diag,
                                                             const double *__restrict__ rcf)
{
    using T = typename F::T;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (IS_BB16<F, W>) {
        double s[16];
#pragma unroll
        for (int k = 0; k < 16; k++) s[k] = Bf::from_u32(states[i * 16 + k]);
        p2_permute_bf<16>(s, rcf);
#pragma unroll
        for (int k = 0; k < 16; k++) states[i * 16 + k] = Bf::canon(s[k]);
        return;
    }
    T s[W];
#pragma unroll
    for (int k = 0; k < W; k++) s[k] = F::from_canon(states[i * W + k]);
    p2_permute<F, W>(s, rc, diag);
#pragma unroll
    for (int k = 0; k < W; k++) states[i * W + k] = F::to_canon(s[k]);
}

// ---------------------------------------------------------------- row hash
// Sponge over one row per lane: overwrite the rate with the next <= W/2
// elements, permute; a partial last chunk leaves the other rate lanes as they
// were.  A zero-width row hashes to the zero digest.
template <class F, int W>
__device__ __forceinline__ void hash_one_row(const typename F::T *__restrict__ mat, uint32_t width, uint64_t row,
                                             typename F::T *__restrict__ digests, const typename F::T *__restrict__ rc,
                                             const typename F::T *__restrict__ diag, const double *__restrict__ rcf)
{
    using T = typename F::T;
    constexpr int R = W / 2;
    if constexpr (IS_BB16<F, W>) {
        double s[16];
#pragma unroll
        for (int k = 0; k < 16; k++) s[k] = 0.0;
        const uint32_t *src = mat + row * width;
        uint32_t pos = 0;
        for (; pos + 8 <= width; pos += 8) {
#pragma unroll
            for (int k = 0; k < 8; k++) s[k] = Bf::from_u32(src[pos + k]);
            p2_permute_bf<8>(s, rcf);
        }
        if (pos < width) {
            uint32_t take = width - pos;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if ((uint32_t)k < take) s[k] = Bf::from_u32(src[pos + k]);
            p2_permute_bf(s, rcf);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) digests[row * 8 + k] = Bf::canon(s[k]);
        return;
    }
    T s[W];
#pragma unroll
    for (int k = 0; k < W; k++) s[k] = 0;
    const T *src = mat + row * width;
    uint32_t pos = 0;
    for (; pos + R <= width; pos += R) {
#pragma unroll
        for (int k = 0; k < R; k++) s[k] = F::from_canon(src[pos + k]);
        p2_permute<F, W>(s, rc, diag);
    }
    if (pos < width) {
        uint32_t take = width - pos;
#pragma unroll
        for (int k = 0; k < R; k++)
            if ((uint32_t)k < take) s[k] = F::from_canon(src[pos + k]);
        p2_permute<F, W>(s, rc, diag);
    }
#pragma unroll
    for (int k = 0; k < R; k++) digests[row * R + k] = F::to_canon(s[k]);
}

template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_hash_rows(const typename F::T *__restrict__ mat, uint64_t height,
                                                            uint32_t width, typename F::T *__restrict__ digests,
                                                            const typename F::T *__restrict__ rc,
                                                            const typename F::T *__restrict__ diag,
                                                            const double *__restrict__ rcf)
{
    uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= height) return;
    hash_one_row<F, W>(mat, width, row, digests, rc, diag, rcf);
}

// The same sponge for Goldilocks with one row per L = W/4 lanes (one 4-chunk
// of the state per lane, zkg_poseidon2_lanes.h): for matrices too short to
// give every SIMD several waves at one row per lane.  Lanes j < R/4 hold the
// rate.
template <int W>
__global__ __launch_bounds__(HASH_THREADS) void k_hash_rows_chunks(const uint64_t *__restrict__ mat, uint64_t height,
                                                                   uint32_t width, uint64_t *__restrict__ digests,
                                                                   const uint64_t *__restrict__ rc,
                                                                   const uint64_t *__restrict__ diag)
{
    constexpr int R = W / 2, L = ChunkMap<W>::L;
    const int j = threadIdx.x & (L - 1);
    const uint64_t row = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / L;
    if (row >= height) return;  // a whole lane group leaves together
    uint64_t d[4], s[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        d[k] = diag[4 * j + k];
        s[k] = 0;
    }
    const uint64_t *src = mat + row * width + 4 * j;
    const bool rate_lane = 4 * j < R;
    uint32_t pos = 0;
    for (; pos + R <= width; pos += R) {
        if (rate_lane) {
#pragma unroll
            for (int k = 0; k < 4; k++) s[k] = src[pos + k];
        }
        p2_permute_gl_chunk<W>(s, j, d, rc);
    }
    if (pos < width) {
        const uint32_t take = width - pos;
        if (rate_lane) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if ((uint32_t)(4 * j + k) < take) s[k] = src[pos + k];
        }
        p2_permute_gl_chunk<W>(s, j, d, rc);
    }
    if (rate_lane) {
#pragma unroll
        for (int k = 0; k < 4; k++) digests[row * R + 4 * j + k] = s[k];
    }
}

// ---------------------------------------------------------------- Merkle
template <class T>
struct LayerPtrs {
    T *p[MAX_FUSE];
};

// Block b compresses input digests [2 b nb1, 2 (b+1) nb1) of `in` into k
// levels (nb1 >= 2^(k-1) nodes at the first, halving per level); level l
// node i goes to out.p[l-1][i].
// Block `block` of one tree's launch; lds holds W/2 x 256 elements.
template <class F, int W>
__device__ __forceinline__ void merkle_fused_block(const typename F::T *__restrict__ in,
                                                   const LayerPtrs<typename F::T> &out, int k, int nb1,
                                                   uint32_t block, const typename F::T *__restrict__ rc,
                                                   const typename F::T *__restrict__ diag,
                                                   const double *__restrict__ rcf, typename F::T *lds)
{
    using T = typename F::T;
    constexpr int D = W / 2;
    const int t = threadIdx.x;
    int nodes = nb1;
    if constexpr (IS_BB16<F, W>) {
        // canonical digests in HBM and LDS, FP64 state in registers
        double s[16];
        if (t < nodes) {
            uint64_t i = (uint64_t)block * nodes + t;
#pragma unroll
            for (int e = 0; e < 16; e++) s[e] = Bf::from_u32(in[2 * i * 8 + e]);
            p2_permute_bf<16>(s, rcf);
#pragma unroll
            for (int e = 0; e < 8; e++) {
                uint32_t v = Bf::canon(s[e]);
                out.p[0][i * 8 + e] = v;
                lds[e * MERKLE_THREADS + t] = v;
            }
        }
        for (int lvl = 1; lvl < k; lvl++) {
            __syncthreads();
            nodes >>= 1;
            if (t < nodes) {
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    s[e] = Bf::from_u32(lds[e * MERKLE_THREADS + 2 * t]);
                    s[8 + e] = Bf::from_u32(lds[e * MERKLE_THREADS + 2 * t + 1]);
                }
            }
            __syncthreads();
            if (t < nodes) {
                p2_permute_bf<16>(s, rcf);
                uint64_t i = (uint64_t)block * nodes + t;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    uint32_t v = Bf::canon(s[e]);
                    out.p[lvl][i * 8 + e] = v;
                    lds[e * MERKLE_THREADS + t] = v;
                }
            }
        }
        return;
    }
    T s[W];
    if (t < nodes) {
        uint64_t i = (uint64_t)block * nodes + t;
#pragma unroll
        for (int e = 0; e < W; e++) s[e] = F::from_canon(in[2 * i * D + e]);
        p2_permute<F, W>(s, rc, diag);
#pragma unroll
        for (int e = 0; e < D; e++) {
            out.p[0][i * D + e] = F::to_canon(s[e]);
            lds[e * MERKLE_THREADS + t] = s[e];
        }
    }
    for (int lvl = 1; lvl < k; lvl++) {
        __syncthreads();
        nodes >>= 1;
        if (t < nodes) {
#pragma unroll
            for (int e = 0; e < D; e++) {
                s[e] = lds[e * MERKLE_THREADS + 2 * t];
                s[D + e] = lds[e * MERKLE_THREADS + 2 * t + 1];
            }
        }
        __syncthreads();
        if (t < nodes) {
            p2_permute<F, W>(s, rc, diag);
            uint64_t i = (uint64_t)block * nodes + t;
#pragma unroll
            for (int e = 0; e < D; e++) {
                out.p[lvl][i * D + e] = F::to_canon(s[e]);
                lds[e * MERKLE_THREADS + t] = s[e];
            }
        }
    }
}

template <class F, int W>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_fused(const typename F::T *__restrict__ in,
                                                                 LayerPtrs<typename F::T> out, int k, int nb1,
                                                                 const typename F::T *__restrict__ rc,
                                                                 const typename F::T *__restrict__ diag,
                                                                 const double *__restrict__ rcf)
{
    __shared__ typename F::T lds[(W / 2) * MERKLE_THREADS];  // element-major: lds[e * 256 + node]
    merkle_fused_block<F, W>(in, out, k, nb1, blockIdx.x, rc, diag, rcf, lds);
}

// Tree top: the same level fusion with each compression spread over W lanes
// (one state element per lane).  A block holds 256 / W compressions and
// compresses 2^k input digests, k <= log2(256 / W) + 1, into levels 1..k.
template <int W>
constexpr int lanes_fuse()
{
    return W == 16 ? 5 : 6;
}

// Block `block` of one tree's launch; lds holds 256 / W digests.
template <class F, int W>
__device__ __forceinline__ void merkle_lanes_block(const typename F::T *__restrict__ in,
                                                   const LayerPtrs<typename F::T> &out, int k, uint32_t block,
                                                   const typename F::T *__restrict__ rc,
                                                   const typename F::T *__restrict__ diag, typename F::T *lds)
{
    using T = typename F::T;
    constexpr int D = W / 2;
    using M = LaneMap<W>;
    static_assert(M::PER_BLOCK * W == MERKLE_THREADS && (2 << (lanes_fuse<W>() - 1)) == 2 * M::PER_BLOCK,
                  "a block's first level holds 256 / W compressions");
    const int g = M::slot(threadIdx.x), e = M::elem(threadIdx.x);
    LaneConsts<F> kc;
    load_lane_consts<F, W>(kc, e, rc, diag);
    int nodes = 1 << (k - 1);
    T x = 0;
    if (g < nodes) {
        const uint64_t i = (uint64_t)block * nodes + g;
        x = p2_permute_lanes<F, W>(F::from_canon(in[2 * i * D + e]), e, kc, rc);
        if (e < D) {
            out.p[0][i * D + e] = F::to_canon(x);
            lds[g * D + e] = x;
        }
    }
    for (int lvl = 1; lvl < k; lvl++) {
        __syncthreads();
        nodes >>= 1;
        if (g < nodes) x = lds[2 * g * D + e];  // elements D..W-1 read the right child
        __syncthreads();
        if (g < nodes) {
            const uint64_t i = (uint64_t)block * nodes + g;
            x = p2_permute_lanes<F, W>(x, e, kc, rc);
            if (e < D) {
                out.p[lvl][i * D + e] = F::to_canon(x);
                lds[g * D + e] = x;
            }
        }
    }
}

template <class F, int W>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_lanes(const typename F::T *__restrict__ in,
                                                                 LayerPtrs<typename F::T> out, int k,
                                                                 const typename F::T *__restrict__ rc,
                                                                 const typename F::T *__restrict__ diag)
{
    __shared__ typename F::T lds[LaneMap<W>::PER_BLOCK * (W / 2)];
    merkle_lanes_block<F, W>(in, out, k, blockIdx.x, rc, diag, lds);
}

// The tops of several trees in one launch: tree t takes blocks [first[t],
// first[t + 1]) and compresses its level `in` into k[t] levels, as one
// k_merkle_lanes launch would.  Small trees that are ready at the same time
// then cost one launch per lanes_fuse<W>() levels together, not each.
constexpr int MAX_TOPS = 16;
template <class T>
struct TopJobs {
    const T *in[MAX_TOPS];
    LayerPtrs<T> out[MAX_TOPS];
    int k[MAX_TOPS];
    uint32_t first[MAX_TOPS + 1];
    int n;
};

template <class F, int W>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_lanes_multi(TopJobs<typename F::T> jobs,
                                                                       const typename F::T *__restrict__ rc,
                                                                       const typename F::T *__restrict__ diag)
{
    __shared__ typename F::T lds[LaneMap<W>::PER_BLOCK * (W / 2)];
    int t = 0;
    while (t + 1 < jobs.n && blockIdx.x >= jobs.first[t + 1]) t++;
    merkle_lanes_block<F, W>(jobs.in[t], jobs.out[t], jobs.k[t], blockIdx.x - jobs.first[t], rc, diag, lds);
}

// The first levels of several larger trees in one launch, one compression per
// lane as in k_merkle_fused: tree t takes blocks [first[t], first[t + 1]) of
// nb1[t] first-level compressions each and k[t] levels.
template <class T>
struct FusedJobs {
    const T *in[MAX_TOPS];
    LayerPtrs<T> out[MAX_TOPS];
    int k[MAX_TOPS];
    int nb1[MAX_TOPS];
    uint32_t first[MAX_TOPS + 1];
    int n;
};

template <class F, int W>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_fused_multi(FusedJobs<typename F::T> jobs,
                                                                       const typename F::T *__restrict__ rc,
                                                                       const typename F::T *__restrict__ diag,
                                                                       const double *__restrict__ rcf)
{
    __shared__ typename F::T lds[(W / 2) * MERKLE_THREADS];
    int t = 0;
    while (t + 1 < jobs.n && blockIdx.x >= jobs.first[t + 1]) t++;
    merkle_fused_block<F, W>(jobs.in[t], jobs.out[t], jobs.k[t], jobs.nb1[t], blockIdx.x - jobs.first[t], rc, diag,
                             rcf, lds);
}

// Middle levels of a Goldilocks tree: level fusion as in k_merkle_fused with
// one 4-chunk of the state per lane (L = W/4 lanes per compression).  Block b
// compresses input digests [2 b nb1, 2 (b+1) nb1) into k levels, nb1 <=
// 256 / L first-level compressions and nb1 >= 2^(k-1); level l node i goes to
// out.p[l-1][i].
template <int W>
constexpr int chunks_fuse()
{
    return W == 16 ? 7 : 8;
}

template <int W>
__device__ __forceinline__ void merkle_chunks_block(const uint64_t *__restrict__ in, const LayerPtrs<uint64_t> &out,
                                                    int k, int nb1, uint32_t block, const uint64_t *__restrict__ rc,
                                                    const uint64_t *__restrict__ diag, uint64_t *lds)
{
    constexpr int D = W / 2, L = ChunkMap<W>::L;
    static_assert((2 << (chunks_fuse<W>() - 1)) == 2 * ChunkMap<W>::PER_BLOCK && chunks_fuse<W>() <= MAX_FUSE,
                  "a block's first level holds 256 / L compressions");
    const int j = threadIdx.x & (L - 1), g = threadIdx.x / L;
    const bool digest_lane = 4 * j < D;
    uint64_t d[4], s[4];
#pragma unroll
    for (int e = 0; e < 4; e++) d[e] = diag[4 * j + e];
    int nodes = nb1;
    if (g < nodes) {
        const uint64_t i = (uint64_t)block * nodes + g;
#pragma unroll
        for (int e = 0; e < 4; e++) s[e] = in[2 * i * D + 4 * j + e];
        p2_permute_gl_chunk<W>(s, j, d, rc);
        if (digest_lane) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                out.p[0][i * D + 4 * j + e] = s[e];
                lds[g * D + 4 * j + e] = s[e];
            }
        }
    }
    for (int lvl = 1; lvl < k; lvl++) {
        __syncthreads();
        nodes >>= 1;
        if (g < nodes) {
#pragma unroll
            for (int e = 0; e < 4; e++) s[e] = lds[2 * g * D + 4 * j + e];  // chunks D/4.. read the right child
        }
        __syncthreads();
        if (g < nodes) {
            const uint64_t i = (uint64_t)block * nodes + g;
            p2_permute_gl_chunk<W>(s, j, d, rc);
            if (digest_lane) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    out.p[lvl][i * D + 4 * j + e] = s[e];
                    lds[g * D + 4 * j + e] = s[e];
                }
            }
        }
    }
}

template <int W>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_chunks(const uint64_t *__restrict__ in,
                                                                  LayerPtrs<uint64_t> out, int k, int nb1,
                                                                  const uint64_t *__restrict__ rc,
                                                                  const uint64_t *__restrict__ diag)
{
    __shared__ uint64_t lds[ChunkMap<W>::PER_BLOCK * (W / 2)];
    merkle_chunks_block<W>(in, out, k, nb1, blockIdx.x, rc, diag, lds);
}

// One level, one compression per lane (the unfused per-level path).
template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_merkle_level(const typename F::T *__restrict__ in,
                                                               typename F::T *__restrict__ out, uint64_t nodes,
                                                               const typename F::T *__restrict__ rc,
                                                               const typename F::T *__restrict__ diag)
{
    using T = typename F::T;
    constexpr int D = W / 2;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodes) return;
    T s[W];
#pragma unroll
    for (int e = 0; e < W; e++) s[e] = F::from_canon(in[2 * i * D + e]);
    p2_permute<F, W, false>(s, rc, diag);  // latency-bound short levels: the compiler's products
#pragma unroll
    for (int e = 0; e < D; e++) out[i * D + e] = F::to_canon(s[e]);
}

// ---------------------------------------------------------------- shared launches
// Row hashes of every job of a table, one persistent grid: each block takes the
// next tile of HASH_TILE_ROWS rows from ctr[0] until none is left, finds its job
// by binary search over the tiles' first indices, and hashes one row per lane.
// The last block to finish (ctr[1]) puts both counters back to zero.
static_assert(HASH_TILE_ROWS == HASH_THREADS, "a tile is one row per lane of a block");
static_assert(SEG_MAX_LEVELS == MAX_FUSE, "a segment holds at most MAX_FUSE levels");

template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_hash_jobs(const HashJob *__restrict__ jobs, int njobs,
                                                            uint32_t ntiles, uint32_t *__restrict__ ctr,
                                                            const typename F::T *__restrict__ rc,
                                                            const typename F::T *__restrict__ diag,
                                                            const double *__restrict__ rcf)
{
    using T = typename F::T;
    __shared__ uint32_t s_tile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&ctr[0], 1u);
        __syncthreads();
        const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();  // every lane has read s_tile before it is written again
        if (tile >= ntiles) break;
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].tile0 <= tile)
                lo = mid;
            else
                hi = mid - 1;
        }
        const HashJob &j = jobs[lo];
        const uint64_t row = (uint64_t)(tile - j.tile0) * HASH_TILE_ROWS + threadIdx.x;
        if (row < j.rows) hash_one_row<F, W>((const T *)j.mat, j.width, row, (T *)j.digests, rc, diag, rcf);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctr[1], 1u) == gridDim.x - 1) {  // every other block is past its last ctr[0] access
            atomicExch(&ctr[0], 0u);
            atomicExch(&ctr[1], 0u);
        }
    }
}

// One launch over the segments of several trees: block b runs block
// b - first of the segment that holds it, in that segment's form.
template <class F, int W>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_segs(const MerkleSeg *__restrict__ segs, int nsegs,
                                                                const typename F::T *__restrict__ rc,
                                                                const typename F::T *__restrict__ diag,
                                                                const double *__restrict__ rcf)
{
    using T = typename F::T;
    __shared__ T lds[(W / 2) * MERKLE_THREADS];  // the largest form's (fused) LDS
    static_assert(LaneMap<W>::PER_BLOCK * (W / 2) <= (W / 2) * MERKLE_THREADS, "lanes form fits");
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].first <= blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const MerkleSeg &sg = segs[lo];
    LayerPtrs<T> lp;
#pragma unroll
    for (int i = 0; i < MAX_FUSE; i++) lp.p[i] = (T *)sg.out[i];
    const uint32_t block = blockIdx.x - sg.first;
    if (sg.form == SEG_LANES) {
        merkle_lanes_block<F, W>((const T *)sg.in, lp, sg.k, block, rc, diag, lds);
        return;
    }
    if constexpr (std::is_same<F, Gl>::value) {
        static_assert(ChunkMap<W>::PER_BLOCK * (W / 2) <= (W / 2) * MERKLE_THREADS, "chunks form fits");
        if (sg.form == SEG_CHUNKS) {
            merkle_chunks_block<W>((const uint64_t *)sg.in, lp, sg.k, sg.nb1, block, rc, diag, lds);
            return;
        }
    }
    merkle_fused_block<F, W>((const T *)sg.in, lp, sg.k, sg.nb1, block, rc, diag, rcf, lds);
}

// ---------------------------------------------------------------- generator
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Element i of stream `seed`: first mix(seed + (i+1) phi + a psi), a = 0,1,..,
// shifted to the field's bit width, that is below p.
template <class F>
__global__ void k_gen(typename F::T *__restrict__ out, uint64_t seed, uint64_t start, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z0 = seed + (start + i + 1) * 0x9e3779b97f4a7c15ULL;
    for (uint64_t a = 0;; a++) {
        uint64_t v = mix64(z0 + a * 0x632be59bd9b4e019ULL);
        if (sizeof(typename F::T) == 4) v >>= 33;
        if (v < F::P) {
            out[i] = (typename F::T)v;
            return;
        }
    }
}

// ---------------------------------------------------------------- FRI
template <class F>
__device__ __forceinline__ typename F::T inv_root_pow(const typename F::T *__restrict__ lo,
                                                      const typename F::T *__restrict__ hi, uint32_t e, int logN)
{
    uint32_t ep = (e << (LOG_TW_TOP - logN)) & ((1u << LOG_TW_TOP) - 1u);
    return F::mul(lo[ep & ((1u << TW_SPLIT) - 1u)], hi[ep >> TW_SPLIT]);
}

__device__ __forceinline__ Ef4 ef4_load(const uint32_t *__restrict__ p)
{
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    Ef4 r;
    r.c[0] = Bb::from_canon(v.x);
    r.c[1] = Bb::from_canon(v.y);
    r.c[2] = Bb::from_canon(v.z);
    r.c[3] = Bb::from_canon(v.w);
    return r;
}

__device__ __forceinline__ void ef4_store(uint32_t *__restrict__ p, const Ef4 &a)
{
    uint4 v;
    v.x = Bb::to_canon(a.c[0]);
    v.y = Bb::to_canon(a.c[1]);
    v.z = Bb::to_canon(a.c[2]);
    v.w = Bb::to_canon(a.c[3]);
    *reinterpret_cast<uint4 *>(p) = v;
}

struct FoldArgs {
    uint32_t beta[4];  // Montgomery
    uint32_t half;     // 1/2, Montgomery
    uint32_t w11;      // 11, Montgomery
};

// out[i] = (lo + hi)/2 + beta (lo - hi) t_i,  t_i = (1/2) g^-rev(i),
// g = two_adic_generator(log n + 1): the arity-2 fold of p3-fri's
// fold_matrix, (1/2 + beta t_i) lo + (1/2 - beta t_i) hi, regrouped.
__device__ __forceinline__ Ef4 fold_one(const Ef4 &lo, const Ef4 &hi, uint32_t t, const FoldArgs &a)
{
    Ef4 beta;
#pragma unroll
    for (int e = 0; e < 4; e++) beta.c[e] = a.beta[e];
    Ef4 sum = ef4_mul_base(ef4_add(lo, hi), a.half);
    Ef4 diff = ef4_mul_base(ef4_sub(lo, hi), t);
    return ef4_add(sum, ef4_mul(beta, diff, a.w11));
}

// Each lane folds output pair (2j, 2j+1) and, if `digests` is set, hashes the
// pair as one leaf (8 base elements) of the next layer's tree.
__global__ __launch_bounds__(HASH_THREADS) void k_fri_fold_ef4(const uint32_t *__restrict__ in, uint64_t n_out,
                                                               int log_n, uint32_t *__restrict__ out,
                                                               uint32_t *__restrict__ digests, FoldArgs a,
                                                               const uint32_t *__restrict__ lo_inv,
                                                               const uint32_t *__restrict__ hi_inv,
                                                               const double *__restrict__ rcf)
{
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t npairs = (n_out + 1) / 2;
    if (j >= npairs) return;
    Ef4 r[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint64_t i = 2 * j + h;
        if (i < n_out) {
            Ef4 lo = ef4_load(in + 8 * i), hi = ef4_load(in + 8 * i + 4);
            uint32_t g = inv_root_pow<Bb>(lo_inv, hi_inv, brev((uint32_t)i, log_n), log_n + 1);
            r[h] = fold_one(lo, hi, Bb::mul(g, a.half), a);
            ef4_store(out + 4 * i, r[h]);
        }
    }
    if (digests && 2 * j + 1 < n_out) {
        double s[16];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            s[e] = Bf::from_u32(Bb::to_canon(r[0].c[e]));
            s[4 + e] = Bf::from_u32(Bb::to_canon(r[1].c[e]));
        }
#pragma unroll
        for (int e = 8; e < 16; e++) s[e] = 0.0;
        p2_permute_bf<16>(s, rcf);
#pragma unroll
        for (int e = 0; e < 8; e++) digests[j * 8 + e] = Bf::canon(s[e]);
    }
}

// Base-field version of the same fold (beta in the base field).
template <class F>
__global__ __launch_bounds__(HASH_THREADS) void k_fri_fold_base(const typename F::T *__restrict__ in, uint64_t n_out,
                                                                int log_n, typename F::T *__restrict__ out,
                                                                typename F::T beta, typename F::T half,
                                                                const typename F::T *__restrict__ lo_inv,
                                                                const typename F::T *__restrict__ hi_inv)
{
    using T = typename F::T;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    T lo = F::from_canon(in[2 * i]), hi = F::from_canon(in[2 * i + 1]);
    T g = inv_root_pow<F>(lo_inv, hi_inv, brev((uint32_t)i, log_n), log_n + 1);
    T t = F::mul(g, half);
    T sum = F::mul(F::add(lo, hi), half);
    out[i] = F::to_canon(F::add(sum, F::mul(beta, F::mul(F::sub(lo, hi), t))));
}

// The base-field fold fused with the next layer's leaf hash: each lane folds
// output pair (2j, 2j+1) and hashes it as one 2-element row of the next
// layer's tree (one sponge absorption, Poseidon2-16: FP64 for BabyBear).
template <class F>
__global__ __launch_bounds__(HASH_THREADS) void k_fri_fold_base_leaf(
    const typename F::T *__restrict__ in, uint64_t n_out, int log_n, typename F::T *__restrict__ out,
    typename F::T *__restrict__ digests, typename F::T beta, typename F::T half,
    const typename F::T *__restrict__ lo_inv, const typename F::T *__restrict__ hi_inv,
    const typename F::T *__restrict__ rc, const typename F::T *__restrict__ diag, const double *__restrict__ rcf)
{
    using T = typename F::T;
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_out / 2) return;
    T r[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint64_t i = 2 * j + h;
        T lo = F::from_canon(in[2 * i]), hi = F::from_canon(in[2 * i + 1]);
        T g = inv_root_pow<F>(lo_inv, hi_inv, brev((uint32_t)i, log_n), log_n + 1);
        T t = F::mul(g, half);
        T sum = F::mul(F::add(lo, hi), half);
        r[h] = F::to_canon(F::add(sum, F::mul(beta, F::mul(F::sub(lo, hi), t))));
        out[i] = r[h];
    }
    if constexpr (std::is_same<F, Bb>::value) {
        double s[16];
        s[0] = Bf::from_u32(r[0]);
        s[1] = Bf::from_u32(r[1]);
#pragma unroll
        for (int e = 2; e < 16; e++) s[e] = 0.0;
        p2_permute_bf<16>(s, rcf);
#pragma unroll
        for (int e = 0; e < 8; e++) digests[j * 8 + e] = Bf::canon(s[e]);
    } else {
        T s[16];
        s[0] = F::from_canon(r[0]);
        s[1] = F::from_canon(r[1]);
#pragma unroll
        for (int e = 2; e < 16; e++) s[e] = 0;
        p2_permute<F, 16>(s, rc, diag);
#pragma unroll
        for (int e = 0; e < 8; e++) digests[j * 8 + e] = F::to_canon(s[e]);
    }
}

// FriAccelerator::fold_*: g[i] = (f[2i] + f[2i+1]) / 2 + alpha (f[2i] - f[2i+1]) itw[i]
template <class F>
__global__ void k_fri_fold_explicit(const typename F::T *__restrict__ in, uint64_t n_out,
                                    typename F::T *__restrict__ out, typename F::T alpha, typename F::T half,
                                    const typename F::T *__restrict__ itw)
{
    using T = typename F::T;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    T a = F::from_canon(in[2 * i]), b = F::from_canon(in[2 * i + 1]);
    T s = F::mul(F::add(a, b), half);
    T d = F::mul(F::mul(F::sub(a, b), F::from_canon(itw[i])), alpha);
    out[i] = F::to_canon(F::add(s, d));
}

// ---------------------------------------------------------------- openings
template <class T>
struct OpenLayers {
    T *p[32];
};

// Block q: opened row index[q] of the leaf matrix and its sibling path
// [layer_l[(index >> l) ^ 1] for l < nlayers - 1].
template <class T>
__global__ void k_gather_open(const T *__restrict__ mat, uint32_t mat_width, OpenLayers<T> layers, int nlayers,
                              int D, const uint64_t *__restrict__ idx, T *__restrict__ rows_out,
                              T *__restrict__ paths_out)
{
    const uint32_t q = blockIdx.x;
    const uint64_t index = idx[q];
    for (uint32_t c = threadIdx.x; c < mat_width; c += blockDim.x)
        rows_out[(uint64_t)q * mat_width + c] = mat[index * mat_width + c];
    const int plen = nlayers - 1;
    for (int t = threadIdx.x; t < plen * D; t += blockDim.x) {
        int l = t / D, e = t - l * D;
        uint64_t sib = (index >> l) ^ 1ull;
        paths_out[((uint64_t)q * plen + l) * D + e] = layers.p[l][sib * D + e];
    }
}

// Block s copies piece s of an opening.
__global__ void k_gather_segs(const GatherSeg *__restrict__ segs, uint32_t n)
{
    const GatherSeg g = segs[blockIdx.x];
    for (uint32_t w = threadIdx.x; w < g.words; w += blockDim.x) g.dst[w] = g.src[w];
}

// Levels per k_merkle_fused launch (tuning layers; built-in 3).
template <class F, int W>
int fused_levels()
{
    return split_value(std::is_same<F, Bb>::value ? FIELD_BB : FIELD_GL, W, SPLIT_FUSED);
}

// The launches of one tree after its level 0.  Large levels: one compression
// per lane, fused_levels() levels per launch (every level keeps full waves).
// Goldilocks levels of at most chunks_split_max digests: W/4 lanes per
// compression, down to the lanes split.  Levels of at most lanes_split_max
// digests are latency-bound and finish with W-lane compressions,
// lanes_fuse<W>() levels per launch.  No launch takes more than `fuse` levels.
template <class F, int W>
void plan_impl(void *const *layers, int nlayers, uint64_t n, int fuse, std::vector<MerkleSeg> &segs)
{
    const uint64_t split = lanes_split_max<F, W>();
    constexpr bool GL = std::is_same<F, Gl>::value;
    const uint64_t csplit = GL ? chunks_split_max<W>() : 0;
    int lvl = 0;
    while (lvl < nlayers - 1) {
        const int left = nlayers - 1 - lvl;
        const uint64_t level_n = n >> lvl;
        const bool lanes = level_n <= split;
        const bool chunks = !lanes && level_n <= csplit;
        int k = lanes ? lanes_fuse<W>() : fused_levels<F, W>();
        if (chunks) {
            // stop where the W-lane form takes over: a fused level with few
            // compressions costs the chunk form's longer chain
            k = chunks_fuse<W>();
            int to_split = 0;
            while ((level_n >> to_split) > split) to_split++;
            if (k > to_split) k = to_split;
        }
        if (k > left) k = left;
        if (k > fuse) k = fuse;
        MerkleSeg sg{};
        sg.in = layers[lvl];
        for (int i = 0; i < k; i++) sg.out[i] = layers[lvl + 1 + i];
        sg.k = k;
        if (lanes) {
            // a block = 256 / W compressions = up to 2^k inputs
            sg.form = SEG_LANES;
            sg.blocks = (uint32_t)(level_n >> k);
        } else {
            // a block = nb1 first-level compressions (up to one per lane, or
            // 256 / (W/4) in the chunk form)
            const uint64_t per = chunks ? (uint64_t)ChunkMap<W>::PER_BLOCK : (uint64_t)MERKLE_THREADS;
            const uint64_t nb1 = level_n / 2 < per ? level_n / 2 : per;
            sg.form = chunks ? SEG_CHUNKS : SEG_FUSED;
            sg.nb1 = (int)nb1;
            sg.blocks = (uint32_t)(level_n / 2 / nb1);
        }
        segs.push_back(sg);
        lvl += k;
    }
}

template <class F, int W>
hipError_t merkle_impl(const P2Set &p2, void *const *layers, int nlayers, uint64_t n, int fuse, hipStream_t st,
                       bool sync_each)
{
    using T = typename F::T;
    int lvl = 0;
    if (fuse == 1) {
        for (; lvl < nlayers - 1; lvl++) {
            uint64_t nodes = n >> (lvl + 1);
            hipLaunchKernelGGL((k_merkle_level<F, W>), dim3((unsigned)((nodes + HASH_THREADS - 1) / HASH_THREADS)),
                               dim3(HASH_THREADS), 0, st, (const T *)layers[lvl], (T *)layers[lvl + 1], nodes,
                               (const T *)p2.rc, (const T *)p2.diag);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess && sync_each) e = hipStreamSynchronize(st);
            if (e) return e;
        }
        return hipSuccess;
    }
    std::vector<MerkleSeg> segs;
    plan_impl<F, W>(layers, nlayers, n, fuse, segs);
    for (const MerkleSeg &sg : segs) {
        LayerPtrs<T> lp;
        for (int i = 0; i < MAX_FUSE; i++) lp.p[i] = (T *)sg.out[i];
        if (sg.form == SEG_LANES) {
            hipLaunchKernelGGL((k_merkle_lanes<F, W>), dim3(sg.blocks), dim3(MERKLE_THREADS), 0, st, (const T *)sg.in,
                               lp, sg.k, (const T *)p2.rc, (const T *)p2.diag);
        } else if (sg.form == SEG_CHUNKS) {
            if constexpr (std::is_same<F, Gl>::value)
                hipLaunchKernelGGL((k_merkle_chunks<W>), dim3(sg.blocks), dim3(MERKLE_THREADS), 0, st,
                                   (const uint64_t *)sg.in, lp, sg.k, sg.nb1, (const uint64_t *)p2.rc,
                                   (const uint64_t *)p2.diag);
        } else {
            hipLaunchKernelGGL((k_merkle_fused<F, W>), dim3(sg.blocks), dim3(MERKLE_THREADS), 0, st, (const T *)sg.in,
                               lp, sg.k, sg.nb1, (const T *)p2.rc, (const T *)p2.diag, (const double *)p2.rc_f64);
        }
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && sync_each) e = hipStreamSynchronize(st);
        if (e) return e;
    }
    return hipSuccess;
}

// Up to MAX_TOPS trees per launch.  Levels of more than tops_fused_above()
// digests first go through k_merkle_fused_multi (one compression per lane, 3
// levels per launch), where the W-lane form would repeat every internal S-box
// 16 times over enough compressions to fill the chip; then each launch takes
// the next lanes_fuse<W>() levels of every tree that has levels left.
uint64_t tops_fused_above()
{
    static const uint64_t v = env_pow2("ZKG_TOPS_FUSED_LOG2", 1ull << 14);
    return v;
}

template <class F, int W>
hipError_t merkle_tops_impl(const P2Set &p2, const MerkleTop *trees, int ntrees, hipStream_t st)
{
    using T = typename F::T;
    for (int i = 0; i < ntrees; i++)
        if (trees[i].nlayers < 1 || trees[i].nlayers > 40 || (1ull << (trees[i].nlayers - 1)) != trees[i].n_leaves)
            return hipErrorInvalidValue;
    for (int g0 = 0; g0 < ntrees; g0 += MAX_TOPS) {
        const int gn = ntrees - g0 < MAX_TOPS ? ntrees - g0 : MAX_TOPS;
        int lvl[MAX_TOPS] = {};
        for (;;) {
            FusedJobs<T> j{};
            uint32_t blocks = 0;
            for (int i = 0; i < gn; i++) {
                const MerkleTop &tr = trees[g0 + i];
                const uint64_t level_n = tr.n_leaves >> lvl[i];
                const int left = tr.nlayers - 1 - lvl[i];
                if (left <= 0 || level_n <= tops_fused_above()) continue;
                const int k = left < 3 ? left : 3;
                const uint64_t nb1 = level_n / 2 < MERKLE_THREADS ? level_n / 2 : MERKLE_THREADS;
                j.in[j.n] = (const T *)tr.layers[lvl[i]];
                for (int m = 0; m < MAX_FUSE; m++) j.out[j.n].p[m] = m < k ? (T *)tr.layers[lvl[i] + 1 + m] : nullptr;
                j.k[j.n] = k;
                j.nb1[j.n] = (int)nb1;
                j.first[j.n] = blocks;
                blocks += (uint32_t)(level_n / 2 / nb1);
                j.n++;
                lvl[i] += k;
            }
            if (!j.n) break;
            j.first[j.n] = blocks;
            hipLaunchKernelGGL((k_merkle_fused_multi<F, W>), dim3(blocks), dim3(MERKLE_THREADS), 0, st, j,
                               (const T *)p2.rc, (const T *)p2.diag, (const double *)p2.rc_f64);
            hipError_t e = hipGetLastError();
            if (e) return e;
        }
        for (;;) {
            TopJobs<T> j{};
            uint32_t blocks = 0;
            for (int i = 0; i < gn; i++) {
                const MerkleTop &tr = trees[g0 + i];
                const int left = tr.nlayers - 1 - lvl[i];
                if (left <= 0) continue;
                const int k = left < lanes_fuse<W>() ? left : lanes_fuse<W>();
                j.in[j.n] = (const T *)tr.layers[lvl[i]];
                for (int m = 0; m < MAX_FUSE; m++) j.out[j.n].p[m] = m < k ? (T *)tr.layers[lvl[i] + 1 + m] : nullptr;
                j.k[j.n] = k;
                j.first[j.n] = blocks;
                blocks += (uint32_t)((tr.n_leaves >> lvl[i]) >> k);
                j.n++;
                lvl[i] += k;
            }
            if (!j.n) break;
            j.first[j.n] = blocks;
            hipLaunchKernelGGL((k_merkle_lanes_multi<F, W>), dim3(blocks), dim3(MERKLE_THREADS), 0, st, j,
                               (const T *)p2.rc, (const T *)p2.diag);
            hipError_t e = hipGetLastError();
            if (e) return e;
        }
    }
    return hipSuccess;
}

// ---------------------------------------------------------------- mixed-height MMCS
// PaddingFreeSponge over row `row` of every matrix of a group, concatenated in
// commit order (p3 first_digest_layer / compress_and_inject hash of
// `matrices.flat_map(|m| m.row(i))`): stream element k overwrites rate slot
// k % RATE, each full or final partial chunk permutes; an empty stream leaves
// the zero state.  State in register form.
template <class F, int W>
__device__ void group_row_hash(const MatDesc *__restrict__ g, int n, uint64_t row, typename F::T (&s)[W],
                               const typename F::T *__restrict__ rc, const typename F::T *__restrict__ diag)
{
    using T = typename F::T;
    constexpr int RATE = W / 2;
#pragma unroll
    for (int e = 0; e < W; e++) s[e] = 0;
    int m = 0;
    uint64_t k = 0;
    auto next_elem = [&](T &v) -> bool {
        while (m < n && k >= g[m].width) {
            m++;
            k = 0;
        }
        if (m >= n) return false;
        v = F::from_canon(((const T *)g[m].ptr)[row * g[m].width + k]);
        k++;
        return true;
    };
    bool more = true;
    while (more) {
        int filled = 0;
#pragma unroll
        for (int j = 0; j < RATE; j++) {
            T v;
            if (more && next_elem(v)) {
                s[j] = v;
                filled++;
            } else {
                more = false;
            }
        }
        if (filled) p2_permute<F, W>(s, rc, diag);
    }
}

// The same sponge for BabyBear width 16 in exact FP64 (p2_permute_bf reduces
// the carried state on entry).
__device__ void group_row_hash_bf(const MatDesc *__restrict__ g, int n, uint64_t row, double (&s)[16],
                                  const double *__restrict__ rcf)
{
#pragma unroll
    for (int e = 0; e < 16; e++) s[e] = 0.0;
    int m = 0;
    uint64_t k = 0;
    auto next_elem = [&](double &v) -> bool {
        while (m < n && k >= g[m].width) {
            m++;
            k = 0;
        }
        if (m >= n) return false;
        v = Bf::from_u32(((const uint32_t *)g[m].ptr)[row * g[m].width + k]);
        k++;
        return true;
    };
    bool more = true;
    while (more) {
        int filled = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            double v;
            if (more && next_elem(v)) {
                s[j] = v;
                filled++;
            } else {
                more = false;
            }
        }
        if (filled) p2_permute_bf(s, rcf);
    }
}

template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_mmcs_leaves(const MatDesc *__restrict__ g, int n, uint64_t rows,
                                                              typename F::T *__restrict__ digests,
                                                              const typename F::T *__restrict__ rc,
                                                              const typename F::T *__restrict__ diag,
                                                              const double *__restrict__ rcf)
{
    using T = typename F::T;
    constexpr int D = W / 2;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    if constexpr (IS_BB16<F, W>) {
        double s[16];
        group_row_hash_bf(g, n, i, s, rcf);
#pragma unroll
        for (int e = 0; e < 8; e++) digests[i * 8 + e] = Bf::canon(s[e]);
        return;
    }
    T s[W];
    group_row_hash<F, W>(g, n, i, s, rc, diag);
#pragma unroll
    for (int e = 0; e < D; e++) digests[i * D + e] = F::to_canon(s[e]);
}

template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_mmcs_layer(const typename F::T *__restrict__ prev, uint64_t half,
                                                             typename F::T *__restrict__ next,
                                                             const MatDesc *__restrict__ inj, int n_inj,
                                                             uint64_t inj_len, const typename F::T *__restrict__ rc,
                                                             const typename F::T *__restrict__ diag,
                                                             const double *__restrict__ rcf)
{
    using T = typename F::T;
    constexpr int D = W / 2;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    if constexpr (IS_BB16<F, W>) {
        double s[16];
#pragma unroll
        for (int e = 0; e < 16; e++) s[e] = Bf::from_u32(prev[2 * i * 8 + e]);
        p2_permute_bf<16>(s, rcf);
        if (n_inj) {
            double r[16];
            if (i < inj_len) {
                group_row_hash_bf(inj, n_inj, i, r, rcf);
            } else {
#pragma unroll
                for (int e = 0; e < 16; e++) r[e] = 0.0;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) s[8 + e] = r[e];
            p2_permute_bf(s, rcf);
        }
#pragma unroll
        for (int e = 0; e < 8; e++) next[i * 8 + e] = Bf::canon(s[e]);
        return;
    }
    T s[W];
#pragma unroll
    for (int e = 0; e < W; e++) s[e] = F::from_canon(prev[2 * i * D + e]);
    p2_permute<F, W>(s, rc, diag);
    if (n_inj) {
        T r[W];
        if (i < inj_len) {
            group_row_hash<F, W>(inj, n_inj, i, r, rc, diag);
        } else {
#pragma unroll
            for (int e = 0; e < W; e++) r[e] = 0;
        }
#pragma unroll
        for (int e = 0; e < D; e++) s[D + e] = r[e];
        p2_permute<F, W>(s, rc, diag);
    }
#pragma unroll
    for (int e = 0; e < D; e++) next[i * D + e] = F::to_canon(s[e]);
}

// Mmcs::verify_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:210-330), one
// lane per opening q.  Matrices are sorted tallest first; group g (descriptors
// gstart[g] .. gstart[g+1]) holds the rows of the matrices whose heights round
// up to the same power of two, each laid out [opening][width], and is hashed
// in after glevel[g] compressions (group 0 gives the leaf digest).  ok[q] = 1
// if the recomputed root equals `root`.
template <class F, int W>
__global__ __launch_bounds__(HASH_THREADS) void k_mmcs_verify(
    const MatDesc *__restrict__ desc, const int *__restrict__ gstart, const int *__restrict__ glevel, int n_groups,
    uint32_t n_q, const uint64_t *__restrict__ idx, const typename F::T *__restrict__ proof, int log_h,
    const typename F::T *__restrict__ root, uint32_t *__restrict__ ok, const typename F::T *__restrict__ rc,
    const typename F::T *__restrict__ diag)
{
    using T = typename F::T;
    constexpr int D = W / 2;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_q) return;
    T s[W];
    group_row_hash<F, W>(desc + gstart[0], gstart[1] - gstart[0], q, s, rc, diag);
    uint64_t index = idx[q];
    int g = 1;
    for (int l = 0; l < log_h; l++) {
        const T *sib = proof + ((uint64_t)q * log_h + l) * D;
        const bool right = index & 1;  // the current node is the right child
#pragma unroll
        for (int e = 0; e < D; e++) {
            const T node = s[e], other = F::from_canon(sib[e]);
            s[e] = right ? other : node;
            s[D + e] = right ? node : other;
        }
        p2_permute<F, W>(s, rc, diag);
        index >>= 1;
        if (g < n_groups && glevel[g] == l + 1) {  // compress the injected rows' digest in
            T r[W];
            group_row_hash<F, W>(desc + gstart[g], gstart[g + 1] - gstart[g], q, r, rc, diag);
#pragma unroll
            for (int e = 0; e < D; e++) s[D + e] = r[e];
            p2_permute<F, W>(s, rc, diag);
            g++;
        }
    }
    uint32_t good = 1;
#pragma unroll
    for (int e = 0; e < D; e++) good &= F::to_canon(s[e]) == root[e] ? 1u : 0u;
    ok[q] = good;
}

}  // namespace

hipError_t launch_mmcs_verify(int field, int width, const P2Set &p2, const MatDesc *desc, const int *gstart,
                              const int *glevel, int n_groups, uint32_t n_q, const uint64_t *idx, const void *proof,
                              int log_h, const void *root, uint32_t *ok, hipStream_t st)
{
    if (!n_q) return hipSuccess;
    dim3 g((n_q + HASH_THREADS - 1) / HASH_THREADS), b(HASH_THREADS);
    if (field == FIELD_BB && width == 16)
        hipLaunchKernelGGL((k_mmcs_verify<Bb, 16>), g, b, 0, st, desc, gstart, glevel, n_groups, n_q, idx,
                           (const uint32_t *)proof, log_h, (const uint32_t *)root, ok, (const uint32_t *)p2.rc,
                           (const uint32_t *)p2.diag);
    else if (field == FIELD_GL && width == 16)
        hipLaunchKernelGGL((k_mmcs_verify<Gl, 16>), g, b, 0, st, desc, gstart, glevel, n_groups, n_q, idx,
                           (const uint64_t *)proof, log_h, (const uint64_t *)root, ok, (const uint64_t *)p2.rc,
                           (const uint64_t *)p2.diag);
    else if (field == FIELD_GL && width == 8)
        hipLaunchKernelGGL((k_mmcs_verify<Gl, 8>), g, b, 0, st, desc, gstart, glevel, n_groups, n_q, idx,
                           (const uint64_t *)proof, log_h, (const uint64_t *)root, ok, (const uint64_t *)p2.rc,
                           (const uint64_t *)p2.diag);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_mmcs_leaves(int field, int width, const P2Set &p2, const MatDesc *group, int n, uint64_t rows,
                              void *digests, hipStream_t st)
{
    if (!rows) return hipSuccess;
    dim3 g((unsigned)((rows + HASH_THREADS - 1) / HASH_THREADS)), b(HASH_THREADS);
    if (field == FIELD_BB && width == 16)
        hipLaunchKernelGGL((k_mmcs_leaves<Bb, 16>), g, b, 0, st, group, n, rows, (uint32_t *)digests,
                           (const uint32_t *)p2.rc, (const uint32_t *)p2.diag, (const double *)p2.rc_f64);
    else if (field == FIELD_GL && width == 16)
        hipLaunchKernelGGL((k_mmcs_leaves<Gl, 16>), g, b, 0, st, group, n, rows, (uint64_t *)digests,
                           (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
    else if (field == FIELD_GL && width == 8)
        hipLaunchKernelGGL((k_mmcs_leaves<Gl, 8>), g, b, 0, st, group, n, rows, (uint64_t *)digests,
                           (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_mmcs_layer(int field, int width, const P2Set &p2, const void *prev, uint64_t half, void *next,
                             const MatDesc *inj, int n_inj, uint64_t inj_len, hipStream_t st)
{
    if (!half) return hipSuccess;
    dim3 g((unsigned)((half + HASH_THREADS - 1) / HASH_THREADS)), b(HASH_THREADS);
    if (field == FIELD_BB && width == 16)
        hipLaunchKernelGGL((k_mmcs_layer<Bb, 16>), g, b, 0, st, (const uint32_t *)prev, half, (uint32_t *)next, inj,
                           n_inj, inj_len, (const uint32_t *)p2.rc, (const uint32_t *)p2.diag,
                           (const double *)p2.rc_f64);
    else if (field == FIELD_GL && width == 16)
        hipLaunchKernelGGL((k_mmcs_layer<Gl, 16>), g, b, 0, st, (const uint64_t *)prev, half, (uint64_t *)next, inj,
                           n_inj, inj_len, (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
    else if (field == FIELD_GL && width == 8)
        hipLaunchKernelGGL((k_mmcs_layer<Gl, 8>), g, b, 0, st, (const uint64_t *)prev, half, (uint64_t *)next, inj,
                           n_inj, inj_len, (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_p2_permute(int field, int width, const P2Set &p2, void *states, uint64_t n, hipStream_t st)
{
    if (!n) return hipSuccess;
    dim3 g((unsigned)((n + HASH_THREADS - 1) / HASH_THREADS)), b(HASH_THREADS);
    if (field == FIELD_BB && width == 16)
        hipLaunchKernelGGL((k_p2_permute<Bb, 16>), g, b, 0, st, (uint32_t *)states, n, (const uint32_t *)p2.rc,
                           (const uint32_t *)p2.diag, (const double *)p2.rc_f64);
    else if (field == FIELD_GL && width == 16)
        hipLaunchKernelGGL((k_p2_permute<Gl, 16>), g, b, 0, st, (uint64_t *)states, n, (const uint64_t *)p2.rc,
                           (const uint64_t *)p2.diag, nullptr);
    else if (field == FIELD_GL && width == 8)
        hipLaunchKernelGGL((k_p2_permute<Gl, 8>), g, b, 0, st, (uint64_t *)states, n, (const uint64_t *)p2.rc,
                           (const uint64_t *)p2.diag, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_hash_rows(int field, int width, const P2Set &p2, const void *mat, uint64_t height,
                            uint32_t mat_width, void *digests, hipStream_t st)
{
    if (!height) return hipSuccess;
    dim3 g((unsigned)((height + HASH_THREADS - 1) / HASH_THREADS)), b(HASH_THREADS);
    if (field == FIELD_BB && width == 16)
        hipLaunchKernelGGL((k_hash_rows<Bb, 16>), g, b, 0, st, (const uint32_t *)mat, height, mat_width,
                           (uint32_t *)digests, (const uint32_t *)p2.rc, (const uint32_t *)p2.diag,
                           (const double *)p2.rc_f64);
    else if (field == FIELD_GL && width == 16 && height <= hash_chunks_max<16>())
        hipLaunchKernelGGL((k_hash_rows_chunks<16>), dim3((unsigned)((height * 4 + HASH_THREADS - 1) / HASH_THREADS)),
                           b, 0, st, (const uint64_t *)mat, height, mat_width, (uint64_t *)digests,
                           (const uint64_t *)p2.rc, (const uint64_t *)p2.diag);
    else if (field == FIELD_GL && width == 8 && height <= hash_chunks_max<8>())
        hipLaunchKernelGGL((k_hash_rows_chunks<8>), dim3((unsigned)((height * 2 + HASH_THREADS - 1) / HASH_THREADS)),
                           b, 0, st, (const uint64_t *)mat, height, mat_width, (uint64_t *)digests,
                           (const uint64_t *)p2.rc, (const uint64_t *)p2.diag);
    else if (field == FIELD_GL && width == 16)
        hipLaunchKernelGGL((k_hash_rows<Gl, 16>), g, b, 0, st, (const uint64_t *)mat, height, mat_width,
                           (uint64_t *)digests, (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
    else if (field == FIELD_GL && width == 8)
        hipLaunchKernelGGL((k_hash_rows<Gl, 8>), g, b, 0, st, (const uint64_t *)mat, height, mat_width,
                           (uint64_t *)digests, (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_merkle_tree(int field, int width, const P2Set &p2, void *const *layers, int nlayers,
                              uint64_t n_leaves, int fuse_levels, hipStream_t st, bool sync_each)
{
    if (fuse_levels < 1) fuse_levels = 1;
    if (field == FIELD_BB && width == 16)
        return merkle_impl<Bb, 16>(p2, layers, nlayers, n_leaves, fuse_levels, st, sync_each);
    if (field == FIELD_GL && width == 16)
        return merkle_impl<Gl, 16>(p2, layers, nlayers, n_leaves, fuse_levels, st, sync_each);
    if (field == FIELD_GL && width == 8)
        return merkle_impl<Gl, 8>(p2, layers, nlayers, n_leaves, fuse_levels, st, sync_each);
    return hipErrorInvalidValue;
}

hipError_t launch_merkle_tops(int field, int width, const P2Set &p2, const MerkleTop *trees, int ntrees,
                              hipStream_t st)
{
    if (field == FIELD_BB && width == 16) return merkle_tops_impl<Bb, 16>(p2, trees, ntrees, st);
    if (field == FIELD_GL && width == 16) return merkle_tops_impl<Gl, 16>(p2, trees, ntrees, st);
    return hipErrorInvalidValue;
}

int merkle_plan(int field, int width, void *const *layers, int nlayers, uint64_t n_leaves, std::vector<MerkleSeg> &segs)
{
    if (field == FIELD_BB && width == 16)
        plan_impl<Bb, 16>(layers, nlayers, n_leaves, MAX_FUSE, segs);
    else if (field == FIELD_GL && width == 16)
        plan_impl<Gl, 16>(layers, nlayers, n_leaves, MAX_FUSE, segs);
    else if (field == FIELD_GL && width == 8)
        plan_impl<Gl, 8>(layers, nlayers, n_leaves, MAX_FUSE, segs);
    else
        return -1;
    return 0;
}

// Blocks of a persistent launch: as many as stay resident on the whole chip.
template <class K>
static uint32_t resident_blocks(K kernel, int threads)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                  hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || cus <= 0 || per <= 0) {
        (void)hipGetLastError();
        return 1024;
    }
    return (uint32_t)cus * (uint32_t)per;
}

template <class F, int W>
static hipError_t hash_jobs_impl(const P2Set &p2, const HashJob *jobs, int njobs, uint32_t ntiles, uint32_t *ctr,
                                 hipStream_t st)
{
    using T = typename F::T;
    static const uint32_t resident = resident_blocks(k_hash_jobs<F, W>, HASH_THREADS);
    const uint32_t grid = ntiles < resident ? ntiles : resident;
    hipLaunchKernelGGL((k_hash_jobs<F, W>), dim3(grid), dim3(HASH_THREADS), 0, st, jobs, njobs, ntiles, ctr,
                       (const T *)p2.rc, (const T *)p2.diag, (const double *)p2.rc_f64);
    return hipGetLastError();
}

hipError_t launch_hash_jobs(int field, int width, const P2Set &p2, const HashJob *jobs, int njobs, uint32_t ntiles,
                            uint32_t *ctr, hipStream_t st)
{
    if (!ntiles || njobs <= 0) return hipSuccess;
    if (field == FIELD_BB && width == 16) return hash_jobs_impl<Bb, 16>(p2, jobs, njobs, ntiles, ctr, st);
    if (field == FIELD_GL && width == 16) return hash_jobs_impl<Gl, 16>(p2, jobs, njobs, ntiles, ctr, st);
    if (field == FIELD_GL && width == 8) return hash_jobs_impl<Gl, 8>(p2, jobs, njobs, ntiles, ctr, st);
    return hipErrorInvalidValue;
}

hipError_t launch_merkle_segs(int field, int width, const P2Set &p2, const MerkleSeg *segs, int nsegs,
                              uint32_t nblocks, hipStream_t st)
{
    if (!nblocks || nsegs <= 0) return hipSuccess;
    dim3 g(nblocks), b(MERKLE_THREADS);
    if (field == FIELD_BB && width == 16)
        hipLaunchKernelGGL((k_merkle_segs<Bb, 16>), g, b, 0, st, segs, nsegs, (const uint32_t *)p2.rc,
                           (const uint32_t *)p2.diag, (const double *)p2.rc_f64);
    else if (field == FIELD_GL && width == 16)
        hipLaunchKernelGGL((k_merkle_segs<Gl, 16>), g, b, 0, st, segs, nsegs, (const uint64_t *)p2.rc,
                           (const uint64_t *)p2.diag, nullptr);
    else if (field == FIELD_GL && width == 8)
        hipLaunchKernelGGL((k_merkle_segs<Gl, 8>), g, b, 0, st, segs, nsegs, (const uint64_t *)p2.rc,
                           (const uint64_t *)p2.diag, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// 2^16: measured against 2^15, 2^17 and 2^18 on the 2^22 -> 2^8 commit phase.
uint64_t fri_tops_max_leaves()
{
    static const uint64_t v = env_pow2("ZKG_FRI_TOPS_LOG2", 1ull << 16);
    return v;
}

hipError_t launch_gen(int field, void *out, uint64_t seed, uint64_t start, uint64_t n, hipStream_t st)
{
    if (!n) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (field == FIELD_GL)
        hipLaunchKernelGGL(k_gen<Gl>, g, b, 0, st, (uint64_t *)out, seed, start, n);
    else
        hipLaunchKernelGGL(k_gen<Bb>, g, b, 0, st, (uint32_t *)out, seed, start, n);
    return hipGetLastError();
}

hipError_t launch_fri_fold(int ext, int field, const TwiddleSet &tw, const P2Set &p2, const void *in,
                           uint64_t n_out, const uint32_t *beta_reg, void *out, void *leaf_digests, hipStream_t st)
{
    if (!n_out) return hipSuccess;
    int log_n = 0;
    while ((1ull << log_n) < n_out) log_n++;
    if (ext) {
        FoldArgs a;
        for (int e = 0; e < 4; e++) a.beta[e] = beta_reg[e];
        a.half = bb_to_mont_host((Bb::P + 1) / 2);
        a.w11 = bb_to_mont_host(11);
        uint64_t npairs = (n_out + 1) / 2;
        hipLaunchKernelGGL(k_fri_fold_ef4, dim3((unsigned)((npairs + HASH_THREADS - 1) / HASH_THREADS)),
                           dim3(HASH_THREADS), 0, st, (const uint32_t *)in, n_out, log_n, (uint32_t *)out,
                           (uint32_t *)leaf_digests, a, (const uint32_t *)tw.top_lo_inv,
                           (const uint32_t *)tw.top_hi_inv, (const double *)p2.rc_f64);
        return hipGetLastError();
    }
    // next-layer leaves are pairs of outputs
    if (leaf_digests && (n_out & 1)) return hipErrorInvalidValue;
    if (leaf_digests && field == FIELD_GL && n_out / 2 <= hash_chunks_max<16>()) {
        // short Goldilocks layers: the fold, then the leaves with 4 lanes per
        // permutation, whose chain is a third of one lane's
        hipError_t e = launch_fri_fold(0, field, tw, p2, in, n_out, beta_reg, out, nullptr, st);
        if (e == hipSuccess) e = launch_hash_rows(field, 16, p2, out, n_out / 2, 2, leaf_digests, st);
        return e;
    }
    if (leaf_digests) {
        dim3 g((unsigned)((n_out / 2 + HASH_THREADS - 1) / HASH_THREADS)), b(HASH_THREADS);
        if (field == FIELD_GL)
            hipLaunchKernelGGL(k_fri_fold_base_leaf<Gl>, g, b, 0, st, (const uint64_t *)in, n_out, log_n,
                               (uint64_t *)out, (uint64_t *)leaf_digests,
                               (uint64_t)beta_reg[0] | ((uint64_t)beta_reg[1] << 32), (uint64_t)(Gl::P + 1) / 2,
                               (const uint64_t *)tw.top_lo_inv, (const uint64_t *)tw.top_hi_inv,
                               (const uint64_t *)p2.rc, (const uint64_t *)p2.diag, nullptr);
        else
            hipLaunchKernelGGL(k_fri_fold_base_leaf<Bb>, g, b, 0, st, (const uint32_t *)in, n_out, log_n,
                               (uint32_t *)out, (uint32_t *)leaf_digests, beta_reg[0],
                               bb_to_mont_host((Bb::P + 1) / 2), (const uint32_t *)tw.top_lo_inv,
                               (const uint32_t *)tw.top_hi_inv, (const uint32_t *)p2.rc, (const uint32_t *)p2.diag,
                               (const double *)p2.rc_f64);
        return hipGetLastError();
    }
    dim3 g((unsigned)((n_out + HASH_THREADS - 1) / HASH_THREADS)), b(HASH_THREADS);
    if (field == FIELD_GL)
        hipLaunchKernelGGL(k_fri_fold_base<Gl>, g, b, 0, st, (const uint64_t *)in, n_out, log_n, (uint64_t *)out,
                           (uint64_t)beta_reg[0] | ((uint64_t)beta_reg[1] << 32), (uint64_t)(Gl::P + 1) / 2,
                           (const uint64_t *)tw.top_lo_inv, (const uint64_t *)tw.top_hi_inv);
    else
        hipLaunchKernelGGL(k_fri_fold_base<Bb>, g, b, 0, st, (const uint32_t *)in, n_out, log_n, (uint32_t *)out,
                           beta_reg[0], bb_to_mont_host((Bb::P + 1) / 2), (const uint32_t *)tw.top_lo_inv,
                           (const uint32_t *)tw.top_hi_inv);
    return hipGetLastError();
}

hipError_t launch_fri_fold_explicit(int field, const void *in, uint64_t n_out, uint64_t alpha, const void *inv_tw,
                                    void *out, hipStream_t st)
{
    if (!n_out) return hipSuccess;
    dim3 g((unsigned)((n_out + 255) / 256)), b(256);
    if (field == FIELD_GL)
        hipLaunchKernelGGL(k_fri_fold_explicit<Gl>, g, b, 0, st, (const uint64_t *)in, n_out, (uint64_t *)out, alpha,
                           (uint64_t)(Gl::P + 1) / 2, (const uint64_t *)inv_tw);
    else
        hipLaunchKernelGGL(k_fri_fold_explicit<Bb>, g, b, 0, st, (const uint32_t *)in, n_out, (uint32_t *)out,
                           bb_to_mont_host((uint32_t)alpha), bb_to_mont_host((Bb::P + 1) / 2),
                           (const uint32_t *)inv_tw);
    return hipGetLastError();
}

hipError_t launch_gather_open(int field, int digest, const void *mat, uint32_t mat_width, void *const *layers,
                              int nlayers, const uint64_t *d_indices, uint32_t n, void *rows_out, void *paths_out,
                              hipStream_t st)
{
    if (!n) return hipSuccess;
    if (nlayers > 32) return hipErrorInvalidValue;
    if (field == FIELD_GL) {
        OpenLayers<uint64_t> ol;
        for (int i = 0; i < 32; i++) ol.p[i] = i < nlayers ? (uint64_t *)layers[i] : nullptr;
        hipLaunchKernelGGL(k_gather_open<uint64_t>, dim3(n), dim3(64), 0, st, (const uint64_t *)mat, mat_width, ol,
                           nlayers, digest, d_indices, (uint64_t *)rows_out, (uint64_t *)paths_out);
    } else {
        OpenLayers<uint32_t> ol;
        for (int i = 0; i < 32; i++) ol.p[i] = i < nlayers ? (uint32_t *)layers[i] : nullptr;
        hipLaunchKernelGGL(k_gather_open<uint32_t>, dim3(n), dim3(64), 0, st, (const uint32_t *)mat, mat_width, ol,
                           nlayers, digest, d_indices, (uint32_t *)rows_out, (uint32_t *)paths_out);
    }
    return hipGetLastError();
}

hipError_t launch_gather_segs(const GatherSeg *segs, uint32_t n, hipStream_t st)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_gather_segs, dim3(n), dim3(64), 0, st, segs, n);
    return hipGetLastError();
}

uint32_t bb_to_mont_host(uint32_t x) { return Bb::from_canon(x); }
uint32_t bb_from_mont_host(uint32_t x) { return Bb::to_canon(x); }

}  // namespace zkg
