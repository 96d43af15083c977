// zkg_poseidon2_lanes.h -- Poseidon2 with the state spread over lanes.
//
// Same permutation as zkg_poseidon2.h (schedule of zkf-metal/src/shaders/
// poseidon2.metal:17-264) in two forms that trade throughput for latency,
// for work with few independent permutations:
//   * one state element per lane (widths 8 and 16; BabyBear in Montgomery
//     form, Goldilocks in the lazy form of p2_permute_gl): the top of a
//     Merkle tree;
//   * Goldilocks with one 4-chunk of the state per lane (further below): row
//     hashes of short matrices and the Merkle levels between.
//
// Lane layout inside a 16-lane DPP row (lane bits b3 b2 b1 b0):
//   * width 16: one permutation per row, element e on lane e;
//   * width 8: two permutations per row, element e = (b3 << 2) | (b1 b0) of
//     permutation b2.  Each 4-chunk is one quad, and the two chunks of a
//     permutation sit 8 lanes apart.
// The cross-lane steps are DPP moves inside the row:
//   * M4 on each 4-chunk: new_k = (x0+x1+x2+x3) + x_k + 2 x_{k+1} (indices mod
//     4), the circulant [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]], with quad
//     permutes fetching x_{k+1}, x_{k+2}, x_{k+3};
//   * the sum over chunks of each position: row rotations by 8 and 4 (width
//     16), or by 8 (width 8);
//   * the internal-round sum of a permutation's lanes: row rotations by 8, 4,
//     2, 1 (width 16), or two quad butterflies and a rotation by 8 (width 8).
// Per-lane constants (the lane's external round constants and internal
// diagonal entry) sit in VGPRs; internal round constants are wave-uniform and
// loaded one round ahead.
#pragma once
#include "zkg_field.h"
#include "zkg_poseidon2.h"

namespace zkg {

// DPP controls: quad_perm encodes the source lane of each quad position.
constexpr int DPP_QUAD_ROT1 = 0x39;  // [1,2,3,0]
constexpr int DPP_QUAD_ROT2 = 0x4E;  // [2,3,0,1]
constexpr int DPP_QUAD_ROT3 = 0x93;  // [3,0,1,2]
constexpr int DPP_QUAD_SWAP1 = 0xB1;  // [1,0,3,2]
constexpr int DPP_QUAD_0TO1 = 0xE0;   // [0,0,2,3]: position 1 reads position 0
constexpr int DPP_QUAD_1TO0 = 0xE5;   // [1,1,2,3]: position 0 reads position 1
constexpr int DPP_PAIR_0TO1 = 0xA0;   // [0,0,2,2]: odd positions read the even one before
constexpr int DPP_PAIR_1TO0 = 0xF5;   // [1,1,3,3]: even positions read the odd one after
constexpr int DPP_ROW_ROR1 = 0x121;
constexpr int DPP_ROW_ROR2 = 0x122;
constexpr int DPP_ROW_ROR4 = 0x124;
constexpr int DPP_ROW_ROR8 = 0x128;

// Every control above reads a lane of the same row, so no lane takes the
// bound_ctrl zero; setting it spares the copy of an "old" value.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v)
{
    uint32_t lo = dpp32<CTRL>((uint32_t)v), hi = dpp32<CTRL>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

template <class F, int CTRL>
__device__ __forceinline__ typename F::T dpp(typename F::T v)
{
    if constexpr (sizeof(typename F::T) == 4)
        return dpp32<CTRL>(v);
    else
        return dpp64<CTRL>(v);
}

template <int CTRL>
__device__ __forceinline__ GlWide glw_dpp(GlWide v)
{
    return GlWide{dpp32<CTRL>(v.w0), dpp32<CTRL>(v.w1), dpp32<CTRL>(v.top)};
}

// Element index and permutation slot of a lane (thread index t) within a
// 256-thread block: 256 / W permutations per block.
template <int W>
struct LaneMap {
    static_assert(W == 8 || W == 16, "lanes form: widths 8 and 16");
    static constexpr int PER_BLOCK = 256 / W;
    __device__ static int elem(int t)
    {
        if constexpr (W == 16) return t & 15;
        return (t & 3) | ((t >> 1) & 4);
    }
    __device__ static int slot(int t)
    {
        if constexpr (W == 16) return t >> 4;
        return ((t >> 4) << 1) | ((t >> 2) & 1);
    }
};

template <class F, int W>
__device__ __forceinline__ typename F::T mds_light_lanes(typename F::T x)
{
    using T = typename F::T;
    T r1 = dpp<F, DPP_QUAD_ROT1>(x), r2 = dpp<F, DPP_QUAD_ROT2>(x), r3 = dpp<F, DPP_QUAD_ROT3>(x);
    T quad = F::add(F::add(x, r1), F::add(r2, r3));
    T y = F::add(F::add(quad, x), F::dbl(r1));
    T col = F::add(y, dpp<F, DPP_ROW_ROR8>(y));
    if constexpr (W == 16) col = F::add(col, dpp<F, DPP_ROW_ROR4>(col));
    return F::add(y, col);
}

template <class F, int W>
__device__ __forceinline__ typename F::T lanes_sum(typename F::T x)
{
    if constexpr (W == 16) {
        x = F::add(x, dpp<F, DPP_ROW_ROR8>(x));
        x = F::add(x, dpp<F, DPP_ROW_ROR4>(x));
        x = F::add(x, dpp<F, DPP_ROW_ROR2>(x));
        return F::add(x, dpp<F, DPP_ROW_ROR1>(x));
    } else {
        x = F::add(x, dpp<F, DPP_QUAD_ROT2>(x));
        x = F::add(x, dpp<F, DPP_QUAD_SWAP1>(x));
        return F::add(x, dpp<F, DPP_ROW_ROR8>(x));
    }
}

// Per-lane constants of one permutation.
template <class F>
struct LaneConsts {
    typename F::T ext[8];  // initial rounds 0..3, terminal rounds 0..3 (this lane)
    typename F::T diag;    // internal diagonal entry of this lane
};

template <class F, int W>
__device__ __forceinline__ void load_lane_consts(LaneConsts<F> &k, int e, const typename F::T *__restrict__ rc,
                                                 const typename F::T *__restrict__ diag)
{
    constexpr int RP = P2Spec<F, W>::RP;
    static_assert(P2Spec<F, W>::RF == 8, "4 + 4 external rounds");
#pragma unroll
    for (int r = 0; r < 4; r++) {
        k.ext[r] = rc[r * W + e];
        k.ext[4 + r] = rc[4 * W + RP + r * W + e];
    }
    if constexpr (std::is_same<F, Bb>::value) {
        static_assert(W == 16, "BabyBear: width 16");
        // diag in register form: V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 2^-8,
        // 1/4, 1/8, 2^-27, -2^-8, -1/16, -2^-27] (zkf-metal/src/poseidon2/babybear.rs:14-58)
        const uint32_t one = Bb::one();
        uint32_t v;
        switch (e) {
        case 0: v = Bb::neg(Bb::dbl(one)); break;
        case 1: v = one; break;
        case 2: v = Bb::dbl(one); break;
        case 3: v = Bb::mul_inv_pow2<1>(one); break;
        case 4: v = Bb::add(Bb::dbl(one), one); break;
        case 5: v = Bb::dbl(Bb::dbl(one)); break;
        case 6: v = Bb::neg(Bb::mul_inv_pow2<1>(one)); break;
        case 7: v = Bb::neg(Bb::add(Bb::dbl(one), one)); break;
        case 8: v = Bb::neg(Bb::dbl(Bb::dbl(one))); break;
        case 9: v = Bb::mul_inv_pow2<8>(one); break;
        case 10: v = Bb::mul_inv_pow2<2>(one); break;
        case 11: v = Bb::mul_inv_pow2<3>(one); break;
        case 12: v = Bb::mul_inv_pow2<27>(one); break;
        case 13: v = Bb::neg(Bb::mul_inv_pow2<8>(one)); break;
        case 14: v = Bb::neg(Bb::mul_inv_pow2<4>(one)); break;
        default: v = Bb::neg(Bb::mul_inv_pow2<27>(one)); break;
        }
        k.diag = v;
    } else {
        k.diag = diag[e];
    }
}

// Goldilocks in the lazy representation of p2_permute_gl (any u64
// representative; one reduction per product; sums as three-word GlWide, folded
// once), with the cross-lane steps of mds_light_lanes / lanes_sum on all three
// words.  MDS-light outputs sum at most 35 terms, as in mds_light_gl.
template <int W>
__device__ __forceinline__ uint64_t mds_light_lanes_gl(uint64_t x)
{
    const uint64_t r1 = dpp64<DPP_QUAD_ROT1>(x), r2 = dpp64<DPP_QUAD_ROT2>(x), r3 = dpp64<DPP_QUAD_ROT3>(x);
    const GlWide quad = glw_add(glw_add(glw(x), r1), glw_add(glw(r2), r3));
    const GlWide y = glw_add(glw_add(quad, x), glw_dbl(r1));
    GlWide col = glw_add(y, glw_dpp<DPP_ROW_ROR8>(y));
    if constexpr (W == 16) col = glw_add(col, glw_dpp<DPP_ROW_ROR4>(col));
    return glw_fold(glw_add(y, col));
}

template <int W>
__device__ __forceinline__ uint64_t lanes_sum_gl(uint64_t x)
{
    GlWide v = glw(x);
    if constexpr (W == 16) {
        v = glw_add(v, glw_dpp<DPP_ROW_ROR8>(v));
        v = glw_add(v, glw_dpp<DPP_ROW_ROR4>(v));
        v = glw_add(v, glw_dpp<DPP_ROW_ROR2>(v));
        v = glw_add(v, glw_dpp<DPP_ROW_ROR1>(v));
    } else {
        v = glw_add(v, glw_dpp<DPP_QUAD_ROT2>(v));
        v = glw_add(v, glw_dpp<DPP_QUAD_SWAP1>(v));
        v = glw_add(v, glw_dpp<DPP_ROW_ROR8>(v));
    }
    return glw_fold(v);
}

template <int W>
__device__ __forceinline__ uint64_t p2_permute_lanes_gl(uint64_t x, int e, const LaneConsts<Gl> &k,
                                                        const uint64_t *__restrict__ rc)
{
    constexpr int RP = P2Spec<Gl, W>::RP;
    x = mds_light_lanes_gl<W>(x);
#pragma unroll
    for (int r = 0; r < 4; r++) x = mds_light_lanes_gl<W>(sbox7_gl<false>(Gl::add_const_lazy(x, k.ext[r])));
    // ir[RP] is the first terminal constant, so the look-ahead stays in bounds
    const uint64_t *ir = rc + 4 * W;
    uint64_t c = ir[0];
    // Only element 0 takes the S-box.  Its chain a^2 -> (a^3, a^4) -> a^7 is
    // shared with element 1's lane, in the same quad: one product forms a^3
    // there and a^4 here, which saves the second squaring of every lane.
#pragma unroll 1
    for (int r = 0; r < RP; r++) {
        const uint64_t next = ir[r + 1];
        const uint64_t a = Gl::add_const_lazy(x, c);
        const uint64_t a2 = dpp64<DPP_QUAD_0TO1>(Gl::sqr_lazy(a));  // lanes of elements 0 and 1: a^2
        const uint64_t p = Gl::mul_lazy(a2, e == 0 ? a : a2);        // element 0: a^3, element 1: a^4
        const uint64_t y = Gl::mul_lazy(p, dpp64<DPP_QUAD_1TO0>(p));  // element 0: a^7
        x = e == 0 ? y : x;
        x = Gl::mul_add_lazy(x, k.diag, lanes_sum_gl<W>(x));
        c = next;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) x = mds_light_lanes_gl<W>(sbox7_gl<false>(Gl::add_const_lazy(x, k.ext[4 + r])));
    return Gl::canon(x);
}

// x is this lane's state element; e = its element index.
template <class F, int W>
__device__ __forceinline__ typename F::T p2_permute_lanes(typename F::T x, int e, const LaneConsts<F> &k,
                                                          const typename F::T *__restrict__ rc)
{
    if constexpr (std::is_same<F, Gl>::value) {
        return p2_permute_lanes_gl<W>(x, e, k, rc);
    } else {
        using T = typename F::T;
        constexpr int RP = P2Spec<F, W>::RP;
        x = mds_light_lanes<F, W>(x);
#pragma unroll
        for (int r = 0; r < 4; r++) x = mds_light_lanes<F, W>(sbox7<F>(F::add(x, k.ext[r])));
        // ir[RP] is the first terminal constant, so the look-ahead stays in bounds
        const T *ir = rc + 4 * W;
        T c = ir[0];
#pragma unroll 1
        for (int r = 0; r < RP; r++) {
            const T next = ir[r + 1];
            T y = sbox7<F>(F::add(x, c));
            x = e == 0 ? y : x;
            T sum = lanes_sum<F, W>(x);
            x = F::add(F::mul(x, k.diag), sum);
            c = next;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) x = mds_light_lanes<F, W>(sbox7<F>(F::add(x, k.ext[4 + r])));
        return x;
    }
}

// ---------------------------------------------------------------------------
// Goldilocks Poseidon2 (widths 8 and 16) with one 4-chunk of the state per
// lane: L = W / 4 adjacent lanes of a quad hold one permutation, lane j
// elements 4j..4j+3.  Between one state per lane (least work, longest chain)
// and one element per lane (shortest chain, each lane repeating the internal
// S-box): about 1.3x the first's work at a quarter of its chain for W = 16.
// M4 is local to a lane; the sums over chunks (MDS-light) and over the state
// (internal rounds) are quad butterflies on the lazy three-word sums (GlWide),
// with the same lazy reduction as p2_permute_gl.
template <int W>
struct ChunkMap {
    static_assert(W == 8 || W == 16, "chunk form: widths 8 and 16");
    static constexpr int L = W / 4;  // lanes per permutation
    static constexpr int PER_BLOCK = 256 / L;
};

// sum of v over the L lanes of a permutation
template <int W>
__device__ __forceinline__ GlWide chunk_sum(GlWide v)
{
    v = glw_add(v, glw_dpp<DPP_QUAD_SWAP1>(v));
    if constexpr (W == 16) v = glw_add(v, glw_dpp<DPP_QUAD_ROT2>(v));
    return v;
}

template <int W>
__device__ __forceinline__ void mds_light_gl_chunk(uint64_t (&s)[4])
{
    const uint64_t a = s[0], b = s[1], cc = s[2], d = s[3];
    const GlWide t = glw_add(glw_add(glw(a), b), glw_add(glw(cc), d));
    GlWide y[4];
    y[0] = glw_add(glw_add(t, a), glw_dbl(b));   // 2a + 3b + c + d
    y[1] = glw_add(glw_add(t, b), glw_dbl(cc));  // a + 2b + 3c + d
    y[2] = glw_add(glw_add(t, cc), glw_dbl(d));  // a + b + 2c + 3d
    y[3] = glw_add(glw_add(t, d), glw_dbl(a));   // 3a + b + c + 2d
#pragma unroll
    for (int k = 0; k < 4; k++) s[k] = glw_fold(glw_add(y[k], chunk_sum<W>(y[k])));
}

// External rounds r0..r0+3 over this lane's chunk; er = the round block's
// constants (4 rounds x W) offset to the chunk.  Each round's constants are
// loaded one round ahead.
template <int W>
__device__ __forceinline__ void gl_chunk_external(uint64_t (&s)[4], const uint64_t *__restrict__ er)
{
    uint64_t c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = er[k];
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
        const int rn = r + 1 < 4 ? r + 1 : r;
        uint64_t cn[4];
#pragma unroll
        for (int k = 0; k < 4; k++) cn[k] = er[rn * W + k];
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = sbox7_gl<false>(Gl::add_const_lazy(s[k], c[k]));
        mds_light_gl_chunk<W>(s);
#pragma unroll
        for (int k = 0; k < 4; k++) c[k] = cn[k];
    }
}

// s = this lane's chunk (j = chunk index), d = diag[4j..4j+3]; rc flat
// [initial 4W][internal 22][terminal 4W], canonical.  Canonical out.
template <int W>
__device__ __forceinline__ void p2_permute_gl_chunk(uint64_t (&s)[4], int j, const uint64_t (&d)[4],
                                                    const uint64_t *__restrict__ rc)
{
    constexpr int RP = P2Spec<Gl, W>::RP;
    mds_light_gl_chunk<W>(s);
    gl_chunk_external<W>(s, rc + 4 * j);
    // ir[RP] is the first terminal constant, so the look-ahead stays in bounds
    const uint64_t *ir = rc + 4 * W;
    uint64_t ci = ir[0];
    // element 0's S-box chain shared with chunk 1's lane (the next lane):
    // one product forms a^3 there and a^4 here
#pragma unroll 1
    for (int r = 0; r < RP; r++) {
        const uint64_t next = ir[r + 1];
        const uint64_t a = Gl::add_const_lazy(s[0], ci);
        const uint64_t a2 = dpp64<DPP_PAIR_0TO1>(Gl::sqr_lazy(a));  // chunks 0 and 1: a^2
        const uint64_t p = Gl::mul_lazy(a2, j == 0 ? a : a2);        // chunk 0: a^3, chunk 1: a^4
        const uint64_t y = Gl::mul_lazy(p, dpp64<DPP_PAIR_1TO0>(p));  // chunk 0: a^7
        s[0] = j == 0 ? y : s[0];
        const GlWide acc = glw_add(glw_add(glw(s[0]), s[1]), glw_add(glw(s[2]), s[3]));
        const uint64_t sum = glw_fold(chunk_sum<W>(acc));
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = Gl::mul_add_lazy(s[k], d[k], sum);
        ci = next;
    }
    gl_chunk_external<W>(s, ir + RP + 4 * j);
#pragma unroll
    for (int k = 0; k < 4; k++) s[k] = Gl::canon(s[k]);
}

}  // namespace zkg
