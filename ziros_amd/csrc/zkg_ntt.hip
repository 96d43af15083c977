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
//     size-N2 transforms over contiguous rows.  Every sub-transform is staged
//     in LDS as a tile of M points x C columns (C adjacent columns of a row are
//     one coalesced segment).
//   * The LDE output rows are in bit-reversed order.  Row b*N + q holds
//     p(s_b omega_N^rev(q)) with s_b = shift * omega_{BN}^{rev_B(b)}, so the
//     LDE is one inverse transform followed by B forward size-N transforms of
//     the coefficients weighted by s_b^k, and the last forward pass (a DIF,
//     bit-reversed out) writes whole contiguous blocks of output rows.
//   * Passes: A (inverse step 1), B (inverse step 2 fused with the B coset
//     weightings and forward step 1, coefficients kept in registers between
//     cosets), C (forward step 2, canonical store).  Three HBM round trips for
//     N > 2^11, one for N <= 2^11.
//   * Twiddles are resident in HBM: the 2^11-point tile table is copied to LDS
//     per block; omega_N^e for the four-step comes from two 4096-entry tables of
//     omega_{2^24} powers (one product per lookup).
//   * Block -> tile mapping is XCD-aware: tiles that share rows (adjacent
//     column chunks) are given to blocks that land on the same XCD, so their
//     row segments share an L2.
#include "zkg_field.h"
#include "zkg_internal.h"

namespace zkg {

namespace {

constexpr int NTT_THREADS = 256;
constexpr int LDS_DATA_BYTES = 65536;

template <class F>
struct Lds {
    static constexpr int ELEMS = LDS_DATA_BYTES / (int)sizeof(typename F::T);
    static constexpr int KEEP = ELEMS / NTT_THREADS;  // register slots per thread
};

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

template <class F>
__device__ __forceinline__ typename F::T top_pow(const typename F::T *__restrict__ lo,
                                                 const typename F::T *__restrict__ hi, uint64_t e,
                                                 int logN)
{
    uint32_t ep = (uint32_t)(e << (LOG_TW_TOP - logN)) & ((1u << LOG_TW_TOP) - 1u);
    return F::mul(lo[ep & ((1u << TW_SPLIT) - 1u)], hi[ep >> TW_SPLIT]);
}

template <class F>
__device__ __forceinline__ void load_tile_twiddles(typename F::T *tw, const typename F::T *__restrict__ g)
{
    for (int i = threadIdx.x; i < (1 << (LOG_TW_TILE - 1)); i += blockDim.x) tw[i] = g[i];
}

// DIF on an LDS tile [M][C]: natural order in, bit-reversed order out.
template <class F>
__device__ void lds_dif(typename F::T *buf, int logM, int logC, const typename F::T *tw)
{
    using T = typename F::T;
    const int nbf = (1 << (logM - 1)) << logC;
    for (int s = logM - 1; s >= 0; s--) {
        for (int idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
            int c = idx & ((1 << logC) - 1), bf = idx >> logC;
            int j = bf & ((1 << s) - 1), g = bf >> s;
            int i0 = (g << (s + 1)) + j, i1 = i0 + (1 << s);
            T a = buf[(i0 << logC) + c], b = buf[(i1 << logC) + c];
            buf[(i0 << logC) + c] = F::add(a, b);
            buf[(i1 << logC) + c] = F::mul(F::sub(a, b), tw[j << (LOG_TW_TILE - 1 - s)]);
        }
        __syncthreads();
    }
}

// DIT on an LDS tile: bit-reversed order in, natural order out.
template <class F>
__device__ void lds_dit(typename F::T *buf, int logM, int logC, const typename F::T *tw)
{
    using T = typename F::T;
    const int nbf = (1 << (logM - 1)) << logC;
    for (int s = 0; s < logM; s++) {
        for (int idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
            int c = idx & ((1 << logC) - 1), bf = idx >> logC;
            int j = bf & ((1 << s) - 1), g = bf >> s;
            int i0 = (g << (s + 1)) + j, i1 = i0 + (1 << s);
            T a = buf[(i0 << logC) + c];
            T b = F::mul(buf[(i1 << logC) + c], tw[j << (LOG_TW_TILE - 1 - s)]);
            buf[(i0 << logC) + c] = F::add(a, b);
            buf[(i1 << logC) + c] = F::sub(a, b);
        }
        __syncthreads();
    }
}

struct Geo {
    uint32_t width;
    int logN, logN1, logN2, logC, logB;
    uint32_t nchunks;
};

// ---------------------------------------------------------------------------
// Pass 1 of a four-step transform (also LDE pass A): group n2 reads rows
// n2 + N2 n1, transforms over n1 (DIF), multiplies by omega_N^(+-n2 k1) and
// stores element k1 at row n2 + N2 k1.  Output stays in register form.
template <class F>
__global__ __launch_bounds__(NTT_THREADS) void k_fs_pass1(const typename F::T *__restrict__ src,
                                                          typename F::T *__restrict__ dst, Geo g,
                                                          int in_canon,
                                                          const typename F::T *__restrict__ tile_tw,
                                                          const typename F::T *__restrict__ top_lo,
                                                          const typename F::T *__restrict__ top_hi)
{
    using T = typename F::T;
    __shared__ T buf[Lds<F>::ELEMS];
    __shared__ T tw[1 << (LOG_TW_TILE - 1)];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t n2 = tile / g.nchunks;
    const uint32_t c0 = (tile % g.nchunks) << g.logC;
    const int logM = g.logN1, C = 1 << g.logC;
    load_tile_twiddles<F>(tw, tile_tw);
    for (int idx = threadIdx.x; idx < (1 << (logM + g.logC)); idx += blockDim.x) {
        uint32_t m = idx >> g.logC, col = c0 + (idx & (C - 1));
        T v = 0;
        if (col < g.width) v = src[((uint64_t)n2 + ((uint64_t)m << g.logN2)) * g.width + col];
        buf[idx] = in_canon ? F::from_canon(v) : v;
    }
    __syncthreads();
    lds_dif<F>(buf, logM, g.logC, tw);
    for (int idx = threadIdx.x; idx < (1 << (logM + g.logC)); idx += blockDim.x) {
        uint32_t p = idx >> g.logC, col = c0 + (idx & (C - 1));
        if (col >= g.width) continue;
        uint32_t k1 = brev(p, logM);
        T w = top_pow<F>(top_lo, top_hi, (uint64_t)n2 * k1, g.logN);
        dst[((uint64_t)n2 + ((uint64_t)k1 << g.logN2)) * g.width + col] = F::mul(buf[idx], w);
    }
}

// Pass 2 of a plain four-step transform: group k1 reads contiguous rows
// N2 k1 + n2, transforms over n2 (DIF) and stores element k2 at row k1 + N1 k2
// in canonical form, times `scale` (1/N for the inverse, register form).
template <class F>
__global__ __launch_bounds__(NTT_THREADS) void k_fs_pass2(const typename F::T *__restrict__ src,
                                                          typename F::T *__restrict__ dst, Geo g,
                                                          typename F::T scale,
                                                          const typename F::T *__restrict__ tile_tw)
{
    using T = typename F::T;
    __shared__ T buf[Lds<F>::ELEMS];
    __shared__ T tw[1 << (LOG_TW_TILE - 1)];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t k1 = tile / g.nchunks;
    const uint32_t c0 = (tile % g.nchunks) << g.logC;
    const int logM = g.logN2, C = 1 << g.logC;
    load_tile_twiddles<F>(tw, tile_tw);
    for (int idx = threadIdx.x; idx < (1 << (logM + g.logC)); idx += blockDim.x) {
        uint32_t m = idx >> g.logC, col = c0 + (idx & (C - 1));
        T v = 0;
        if (col < g.width) v = src[(((uint64_t)k1 << g.logN2) + m) * g.width + col];
        buf[idx] = v;
    }
    __syncthreads();
    lds_dif<F>(buf, logM, g.logC, tw);
    for (int idx = threadIdx.x; idx < (1 << (logM + g.logC)); idx += blockDim.x) {
        uint32_t p = idx >> g.logC, col = c0 + (idx & (C - 1));
        if (col >= g.width) continue;
        uint32_t k2 = brev(p, logM);
        dst[((uint64_t)k1 + ((uint64_t)k2 << g.logN1)) * g.width + col] = F::to_canon(F::mul(buf[idx], scale));
    }
}

// Single-pass transform (N <= 2^11): natural in/out, canonical in/out.
template <class F>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_small(typename F::T *__restrict__ data, Geo g,
                                                           typename F::T scale,
                                                           const typename F::T *__restrict__ tile_tw)
{
    using T = typename F::T;
    __shared__ T buf[Lds<F>::ELEMS];
    __shared__ T tw[1 << (LOG_TW_TILE - 1)];
    const uint32_t c0 = xcd_tile(blockIdx.x, gridDim.x) << g.logC;
    const int logM = g.logN, C = 1 << g.logC;
    load_tile_twiddles<F>(tw, tile_tw);
    for (int idx = threadIdx.x; idx < (1 << (logM + g.logC)); idx += blockDim.x) {
        uint32_t m = idx >> g.logC, col = c0 + (idx & (C - 1));
        T v = 0;
        if (col < g.width) v = F::from_canon(data[(uint64_t)m * g.width + col]);
        buf[idx] = v;
    }
    __syncthreads();
    lds_dif<F>(buf, logM, g.logC, tw);
    for (int idx = threadIdx.x; idx < (1 << (logM + g.logC)); idx += blockDim.x) {
        uint32_t p = idx >> g.logC, col = c0 + (idx & (C - 1));
        if (col >= g.width) continue;
        data[(uint64_t)brev(p, logM) * g.width + col] = F::to_canon(F::mul(buf[idx], scale));
    }
}

// ---------------------------------------------------------------------------
// LDE pass B: group k1 reads contiguous rows N2 k1 + n2 of pass A's output,
// finishes the inverse (DIF over n2 -> c[k1 + N1 k2] at slot rev(k2)), keeps
// the coefficients in registers, and for each coset b applies s_b^k / N, the
// forward step 1 (DIT over k2 -> j2 natural) and omega_N^(j2 k1); the result
// goes to row b N + N2 k1 + j2 of the scratch matrix.
template <class F>
__global__ __launch_bounds__(NTT_THREADS) void k_lde_mid(const typename F::T *__restrict__ src,
                                                         typename F::T *__restrict__ dst, Geo g,
                                                         const typename F::T *__restrict__ tile_inv,
                                                         const typename F::T *__restrict__ tile_fwd,
                                                         const typename F::T *__restrict__ top_lo,
                                                         const typename F::T *__restrict__ top_hi,
                                                         const typename F::T *__restrict__ cos_lo,
                                                         const typename F::T *__restrict__ cos_hi)
{
    using T = typename F::T;
    __shared__ T buf[Lds<F>::ELEMS];
    __shared__ T tw[1 << (LOG_TW_TILE - 1)];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t k1 = tile / g.nchunks;
    const uint32_t c0 = (tile % g.nchunks) << g.logC;
    const int logM = g.logN2, C = 1 << g.logC;
    const int total = 1 << (logM + g.logC);
    load_tile_twiddles<F>(tw, tile_inv);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        uint32_t m = idx >> g.logC, col = c0 + (idx & (C - 1));
        T v = 0;
        if (col < g.width) v = src[(((uint64_t)k1 << g.logN2) + m) * g.width + col];
        buf[idx] = v;
    }
    __syncthreads();
    lds_dif<F>(buf, logM, g.logC, tw);
    T keep[Lds<F>::KEEP];
#pragma unroll
    for (int i = 0; i < Lds<F>::KEEP; i++) {
        int idx = threadIdx.x + i * NTT_THREADS;
        keep[i] = idx < total ? buf[idx] : T(0);
    }
    __syncthreads();
    load_tile_twiddles<F>(tw, tile_fwd);
    for (int b = 0; b < (1 << g.logB); b++) {
        const T *clo = cos_lo + (size_t)b * (1u << TW_SPLIT);
        const T *chi = cos_hi + (size_t)b * (1u << (LOG_TW_TOP - TW_SPLIT - 2));
#pragma unroll
        for (int i = 0; i < Lds<F>::KEEP; i++) {
            int idx = threadIdx.x + i * NTT_THREADS;
            if (idx < total) {
                uint32_t p = idx >> g.logC;
                uint32_t k = k1 + (brev(p, logM) << g.logN1);
                T w = F::mul(clo[k & ((1u << TW_SPLIT) - 1u)], chi[k >> TW_SPLIT]);
                buf[idx] = F::mul(keep[i], w);
            }
        }
        __syncthreads();
        lds_dit<F>(buf, logM, g.logC, tw);
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            uint32_t j2 = idx >> g.logC, col = c0 + (idx & (C - 1));
            if (col >= g.width) continue;
            T w = top_pow<F>(top_lo, top_hi, (uint64_t)j2 * k1, g.logN);
            dst[(((uint64_t)b << g.logN) + ((uint64_t)k1 << g.logN2) + j2) * g.width + col] = F::mul(buf[idx], w);
        }
        __syncthreads();
    }
}

// LDE pass C: coset b, group j2 reads rows b N + N2 k1 + j2, DIF over k1
// (slot p holds j1 = rev(p)) and writes the contiguous block of output rows
// b N + rev(j2) N1 + p in canonical form.
template <class F>
__global__ __launch_bounds__(NTT_THREADS) void k_lde_final(const typename F::T *__restrict__ src,
                                                           typename F::T *__restrict__ out, Geo g,
                                                           const typename F::T *__restrict__ tile_fwd)
{
    using T = typename F::T;
    __shared__ T buf[Lds<F>::ELEMS];
    __shared__ T tw[1 << (LOG_TW_TILE - 1)];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t chunk = tile % g.nchunks;
    const uint32_t grp = tile / g.nchunks;  // b * N2 + j2
    const uint32_t j2 = grp & ((1u << g.logN2) - 1u), b = grp >> g.logN2;
    const uint32_t c0 = chunk << g.logC;
    const int logM = g.logN1, C = 1 << g.logC;
    const int total = 1 << (logM + g.logC);
    load_tile_twiddles<F>(tw, tile_fwd);
    const uint64_t base = (uint64_t)b << g.logN;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        uint32_t k1 = idx >> g.logC, col = c0 + (idx & (C - 1));
        T v = 0;
        if (col < g.width) v = src[(base + ((uint64_t)k1 << g.logN2) + j2) * g.width + col];
        buf[idx] = v;
    }
    __syncthreads();
    lds_dif<F>(buf, logM, g.logC, tw);
    const uint64_t obase = base + ((uint64_t)brev(j2, g.logN2) << g.logN1);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        uint32_t p = idx >> g.logC, col = c0 + (idx & (C - 1));
        if (col >= g.width) continue;
        out[(obase + p) * g.width + col] = F::to_canon(buf[idx]);
    }
}

// Whole LDE in one pass for N <= 2^11: one block per column chunk.
template <class F>
__global__ __launch_bounds__(NTT_THREADS) void k_lde_small(const typename F::T *__restrict__ in,
                                                           typename F::T *__restrict__ out, Geo g,
                                                           const typename F::T *__restrict__ tile_inv,
                                                           const typename F::T *__restrict__ tile_fwd,
                                                           const typename F::T *__restrict__ cos_lo)
{
    using T = typename F::T;
    __shared__ T buf[Lds<F>::ELEMS];
    __shared__ T tw[1 << (LOG_TW_TILE - 1)];
    const uint32_t c0 = xcd_tile(blockIdx.x, gridDim.x) << g.logC;
    const int logM = g.logN, C = 1 << g.logC;
    const int total = 1 << (logM + g.logC);
    load_tile_twiddles<F>(tw, tile_inv);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        uint32_t m = idx >> g.logC, col = c0 + (idx & (C - 1));
        T v = 0;
        if (col < g.width) v = F::from_canon(in[(uint64_t)m * g.width + col]);
        buf[idx] = v;
    }
    __syncthreads();
    if (logM > 0) lds_dif<F>(buf, logM, g.logC, tw);
    T keep[Lds<F>::KEEP];
#pragma unroll
    for (int i = 0; i < Lds<F>::KEEP; i++) {
        int idx = threadIdx.x + i * NTT_THREADS;
        keep[i] = idx < total ? buf[idx] : T(0);
    }
    __syncthreads();
    load_tile_twiddles<F>(tw, tile_fwd);
    for (int b = 0; b < (1 << g.logB); b++) {
        const T *clo = cos_lo + (size_t)b * (1u << TW_SPLIT);
#pragma unroll
        for (int i = 0; i < Lds<F>::KEEP; i++) {
            int idx = threadIdx.x + i * NTT_THREADS;
            if (idx < total) buf[idx] = F::mul(keep[i], clo[brev(idx >> g.logC, logM)]);
        }
        __syncthreads();
        if (logM > 0) lds_dit<F>(buf, logM, g.logC, tw);
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            uint32_t q = idx >> g.logC, col = c0 + (idx & (C - 1));
            if (col >= g.width) continue;
            out[(((uint64_t)b << g.logN) + brev(q, logM)) * g.width + col] = F::to_canon(buf[idx]);
        }
        __syncthreads();
    }
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

// coefficients (register form, n rows) -> c_k * shift^k / n, rows >= n are zero
template <class F>
__global__ void k_scale_shift(typename F::T *__restrict__ data, uint64_t n, uint32_t width,
                              typename F::T shift, typename F::T ninv)
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
__global__ void k_to_canon(typename F::T *__restrict__ data, uint64_t n)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) data[t] = F::to_canon(data[t]);
}

// ---------------------------------------------------------------------------
template <class F>
struct Tabs {
    const typename F::T *tile_fwd, *tile_inv, *lo_fwd, *hi_fwd, *lo_inv, *hi_inv;
    explicit Tabs(const TwiddleSet &t)
        : tile_fwd((const typename F::T *)t.tile_fwd), tile_inv((const typename F::T *)t.tile_inv),
          lo_fwd((const typename F::T *)t.top_lo_fwd), hi_fwd((const typename F::T *)t.top_hi_fwd),
          lo_inv((const typename F::T *)t.top_lo_inv), hi_inv((const typename F::T *)t.top_hi_inv)
    {
    }
};

int ilog2(uint64_t n)
{
    int l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

// Column chunk for a tile of 2^logM points: as many columns as fit the LDS
// tile, capped at the (power-of-two rounded) width.
template <class F>
int pick_logc(int logM, uint32_t width)
{
    int cap = ilog2(Lds<F>::ELEMS) - logM;
    int lw = ilog2(width);
    int lc = cap < lw ? cap : lw;
    return lc < 0 ? 0 : lc;
}

template <class F>
Geo make_geo(uint64_t height, uint32_t width, int logB, int logM)
{
    Geo g;
    g.width = width;
    g.logN = ilog2(height);
    g.logN1 = g.logN / 2;
    g.logN2 = g.logN - g.logN1;
    g.logB = logB;
    g.logC = pick_logc<F>(logM, width);
    g.nchunks = (width + (1u << g.logC) - 1) >> g.logC;
    return g;
}

template <class F>
typename F::T host_ninv(uint64_t n)
{
    // 1/n in register form
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

template <class F>
hipError_t coset_lde_impl(const TwiddleSet &tws, const CosetSet &cs, const void *in, void *out, void *tmp,
                          uint64_t height, uint32_t width, uint32_t log_blowup, hipStream_t st)
{
    using T = typename F::T;
    Tabs<F> t(tws);
    const int logN = ilog2(height);
    if (logN <= LOG_TW_TILE) {
        Geo g = make_geo<F>(height, width, log_blowup, logN);
        hipLaunchKernelGGL(k_lde_small<F>, dim3(g.nchunks), dim3(NTT_THREADS), 0, st, (const T *)in, (T *)out, g,
                           t.tile_inv, t.tile_fwd, (const T *)cs.lo);
        return hipGetLastError();
    }
    Geo g = make_geo<F>(height, width, log_blowup, (logN + 1) / 2);
    // pass A: in -> out[0 .. N*W) (scratch), pass B: -> tmp (B*N*W), pass C: -> out
    hipLaunchKernelGGL(k_fs_pass1<F>, dim3(g.nchunks << g.logN2), dim3(NTT_THREADS), 0, st, (const T *)in, (T *)out,
                       g, 1, t.tile_inv, t.lo_inv, t.hi_inv);
    hipLaunchKernelGGL(k_lde_mid<F>, dim3(g.nchunks << g.logN1), dim3(NTT_THREADS), 0, st, (const T *)out, (T *)tmp,
                       g, t.tile_inv, t.tile_fwd, t.lo_fwd, t.hi_fwd, (const T *)cs.lo, (const T *)cs.hi);
    hipLaunchKernelGGL(k_lde_final<F>, dim3(g.nchunks << (g.logN2 + g.logB)), dim3(NTT_THREADS), 0, st,
                       (const T *)tmp, (T *)out, g, t.tile_fwd);
    return hipGetLastError();
}

template <class F>
hipError_t ntt_impl(const TwiddleSet &tws, void *data, void *tmp, uint64_t height, uint32_t width, int inverse,
                    hipStream_t st)
{
    using T = typename F::T;
    Tabs<F> t(tws);
    const int logN = ilog2(height);
    T scale = inverse ? host_ninv<F>(height) : F::from_canon(1);
    const T *tile = inverse ? t.tile_inv : t.tile_fwd;
    if (logN <= LOG_TW_TILE) {
        Geo g = make_geo<F>(height, width, 0, logN);
        hipLaunchKernelGGL(k_ntt_small<F>, dim3(g.nchunks), dim3(NTT_THREADS), 0, st, (T *)data, g, scale, tile);
        return hipGetLastError();
    }
    Geo g = make_geo<F>(height, width, 0, (logN + 1) / 2);
    hipLaunchKernelGGL(k_fs_pass1<F>, dim3(g.nchunks << g.logN2), dim3(NTT_THREADS), 0, st, (const T *)data,
                       (T *)tmp, g, 1, tile, inverse ? t.lo_inv : t.lo_fwd, inverse ? t.hi_inv : t.hi_fwd);
    hipLaunchKernelGGL(k_fs_pass2<F>, dim3(g.nchunks << g.logN1), dim3(NTT_THREADS), 0, st, (const T *)tmp,
                       (T *)data, g, scale, tile);
    return hipGetLastError();
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

uint64_t coset_lde_tmp_bytes(int field, uint64_t height, uint32_t width, uint32_t log_blowup)
{
    if (ilog2(height) <= LOG_TW_TILE) return 0;
    uint64_t es = field == FIELD_GL ? 8 : 4;
    return (height << log_blowup) * width * es;
}

uint64_t ntt_tmp_bytes(int field, uint64_t height, uint32_t width)
{
    if (ilog2(height) <= LOG_TW_TILE) return 0;
    return height * width * (field == FIELD_GL ? 8 : 4);
}

hipError_t launch_coset_lde(int field, const TwiddleSet &tw, const CosetSet &cs, const void *in, void *out,
                            void *tmp, uint64_t height, uint32_t width, uint32_t log_blowup, hipStream_t st)
{
    if (field == FIELD_GL) return coset_lde_impl<Gl>(tw, cs, in, out, tmp, height, width, log_blowup, st);
    return coset_lde_impl<Bb>(tw, cs, in, out, tmp, height, width, log_blowup, st);
}

hipError_t launch_ntt(int field, const TwiddleSet &tw, void *data, void *tmp, uint64_t height, uint32_t width,
                      int inverse, hipStream_t st)
{
    if (field == FIELD_GL) return ntt_impl<Gl>(tw, data, tmp, height, width, inverse, st);
    return ntt_impl<Bb>(tw, data, tmp, height, width, inverse, st);
}

hipError_t launch_lde_plain(int field, const TwiddleSet &tw, const void *in, void *out, void *tmp, uint64_t height,
                            uint32_t width, uint32_t log_blowup, uint64_t shift, hipStream_t st, bool sync_each)
{
    if (field == FIELD_GL) return lde_plain_impl<Gl>(tw, in, out, tmp, height, width, log_blowup, shift, st, sync_each);
    return lde_plain_impl<Bb>(tw, in, out, tmp, height, width, log_blowup, shift, st, sync_each);
}

}  // namespace zkg
