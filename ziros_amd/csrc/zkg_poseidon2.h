// zkg_poseidon2.h -- register-resident Poseidon2 permutation, one state per lane.
//
// Round schedule of zkf-metal/src/shaders/poseidon2.metal:17-264 (restated in
// zkf-backends/src/wrapping/poseidon2_goldilocks.rs:117-339):
//   MDS-light; RF/2 external rounds (add RC to every lane of the state, x^7,
//   MDS-light); RP internal rounds (s0 += rc, s0 = s0^7, s_i = d_i s_i + sum);
//   RF/2 external rounds.
// MDS-light applies the circulant M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]]
// to every 4-chunk and then adds the per-position sum over chunks.
//
// Constants live in device memory in register form (BabyBear: Montgomery) in
// the reference's flat GPU layout [initial RF/2*W][internal RP][terminal RF/2*W]
// (zkf-metal/src/poseidon2/goldilocks.rs:42-83, babybear.rs:76-111); every
// index is wave-uniform, so the loads are scalar.
#pragma once
#include "zkg_field.h"

namespace zkg {

template <class F>
__host__ __device__ __forceinline__ typename F::T sbox7(typename F::T x)
{
    typename F::T x2 = F::sqr(x);
    typename F::T x3 = F::mul(x2, x);
    typename F::T x4 = F::sqr(x2);
    return F::mul(x3, x4);
}

template <class F, int W>
__host__ __device__ __forceinline__ void mds_light(typename F::T (&s)[W])
{
    using T = typename F::T;
#pragma unroll
    for (int c = 0; c < W; c += 4) {
        T a = s[c], b = s[c + 1], cc = s[c + 2], d = s[c + 3];
        T ab = F::add(a, b), cd = F::add(cc, d);
        T t = F::add(ab, cd);
        T tb = F::add(t, b), td = F::add(t, d);
        s[c] = F::add(tb, ab);               // 2a + 3b + c + d
        s[c + 1] = F::add(tb, F::dbl(cc));   // a + 2b + 3c + d
        s[c + 2] = F::add(td, cd);           // a + b + 2c + 3d
        s[c + 3] = F::add(td, F::dbl(a));    // 3a + b + c + 2d
    }
    if (W > 4) {
        T sum[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            sum[k] = s[k];
#pragma unroll
            for (int c = 4; c < W; c += 4) sum[k] = F::add(sum[k], s[c + k]);
        }
#pragma unroll
        for (int i = 0; i < W; i++) s[i] = F::add(s[i], sum[i & 3]);
    }
}

// Generic internal diffusion: s_i = d_i * s_i + sum (d in register form).
template <class F, int W>
__host__ __device__ __forceinline__ void internal_diag_generic(typename F::T (&s)[W], typename F::T sum,
                                                      const typename F::T *__restrict__ diag)
{
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = F::add(F::mul(s[i], diag[i]), sum);
}

// BabyBear width 16: V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 2^-8, 1/4, 1/8,
// 2^-27, -2^-8, -1/16, -2^-27] (zkf-metal/src/poseidon2/babybear.rs:14-58),
// applied with doublings and exact power-of-two divisions instead of products.
__host__ __device__ __forceinline__ void internal_diag_bb16(uint32_t (&s)[16], uint32_t sum)
{
    using F = Bb;
    s[0] = F::sub(sum, F::dbl(s[0]));
    s[1] = F::add(sum, s[1]);
    s[2] = F::add(sum, F::dbl(s[2]));
    s[3] = F::add(sum, F::mul_inv_pow2<1>(s[3]));
    s[4] = F::add(sum, F::add(F::dbl(s[4]), s[4]));
    s[5] = F::add(sum, F::dbl(F::dbl(s[5])));
    s[6] = F::sub(sum, F::mul_inv_pow2<1>(s[6]));
    s[7] = F::sub(sum, F::add(F::dbl(s[7]), s[7]));
    s[8] = F::sub(sum, F::dbl(F::dbl(s[8])));
    s[9] = F::add(sum, F::mul_inv_pow2<8>(s[9]));
    s[10] = F::add(sum, F::mul_inv_pow2<2>(s[10]));
    s[11] = F::add(sum, F::mul_inv_pow2<3>(s[11]));
    s[12] = F::add(sum, F::mul_inv_pow2<27>(s[12]));
    s[13] = F::sub(sum, F::mul_inv_pow2<8>(s[13]));
    s[14] = F::sub(sum, F::mul_inv_pow2<4>(s[14]));
    s[15] = F::sub(sum, F::mul_inv_pow2<27>(s[15]));
}

template <class F, int W>
struct P2Spec;
template <>
struct P2Spec<Gl, 16> {
    static constexpr int RF = 8, RP = 22;
};
template <>
struct P2Spec<Gl, 8> {
    static constexpr int RF = 8, RP = 22;
};
template <>
struct P2Spec<Bb, 16> {
    static constexpr int RF = 8, RP = 13;
};

template <class F, int W>
__host__ __device__ __forceinline__ void external_round(typename F::T (&s)[W], const typename F::T *__restrict__ rc)
{
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = sbox7<F>(F::add(s[i], rc[i]));
    mds_light<F, W>(s);
}

// ---------------------------------------------------------------------------
// Goldilocks permutation with lazy reduction (width 8 or 16).  The state holds
// u64 representatives (any value, congruent mod p); every product reduces its
// 128-bit value once (Gl::reduce128_lazy), sums of several terms accumulate as
// lo + top 2^64 and fold once, and the state is made canonical only at the end.
// Bounds: MDS-light outputs are sums of at most 35 u64 terms (top <= 35) and
// an internal-round sum has W terms, so every top stays far below 2^32.
template <int W>
__host__ __device__ __forceinline__ void mds_light_gl(uint64_t (&s)[W])
{
    GlWide y[W];
#pragma unroll
    for (int c = 0; c < W; c += 4) {
        const uint64_t a = s[c], b = s[c + 1], cc = s[c + 2], d = s[c + 3];
        const GlWide ab = glw_add(glw(a), b), cd = glw_add(glw(cc), d);
        const GlWide t = glw_add(ab, cd);
        const GlWide tb = glw_add(t, b), td = glw_add(t, d);
        y[c] = glw_add(tb, ab);                 // 2a + 3b + c + d
        y[c + 1] = glw_add(tb, glw_dbl(cc));    // a + 2b + 3c + d
        y[c + 2] = glw_add(td, cd);             // a + b + 2c + 3d
        y[c + 3] = glw_add(td, glw_dbl(a));     // 3a + b + c + 2d
    }
    GlWide sum[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        sum[k] = y[k];
#pragma unroll
        for (int c = 4; c < W; c += 4) sum[k] = glw_add(sum[k], y[c + k]);
    }
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = glw_fold(glw_add(y[i], sum[i & 3]));
}

// LEAN picks the products with the fewest issue slots (Gl::*_lean), for
// kernels with enough waves to hide their latency; the others use Gl::*_lazy.
template <bool LEAN>
__host__ __device__ __forceinline__ uint64_t sbox7_gl(uint64_t x)
{
    if constexpr (LEAN) {
        const uint64_t x2 = Gl::sqr_lean(x);
        const uint64_t x3 = Gl::mul_lean(x2, x);
        const uint64_t x4 = Gl::sqr_lean(x2);
        return Gl::mul_lean(x3, x4);
    } else {
        const uint64_t x2 = Gl::sqr_lazy(x);
        const uint64_t x3 = Gl::mul_lazy(x2, x);
        const uint64_t x4 = Gl::sqr_lazy(x2);
        return Gl::mul_lazy(x3, x4);
    }
}

// rc: flat [initial 4W][internal RP][terminal 4W] and diag[W], canonical.
template <int W, bool LEAN>
__host__ __device__ __forceinline__ void p2_permute_gl(uint64_t (&s)[W], const uint64_t *__restrict__ rc,
                                                       const uint64_t *__restrict__ diag)
{
    constexpr int RP = 22;
    mds_light_gl<W>(s);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < W; i++) s[i] = sbox7_gl<LEAN>(Gl::add_const_lazy(s[i], rc[r * W + i]));
        mds_light_gl<W>(s);
    }
    const uint64_t *ir = rc + 4 * W;
#pragma unroll 1
    for (int r = 0; r < RP; r++) {
        s[0] = sbox7_gl<LEAN>(Gl::add_const_lazy(s[0], ir[r]));
        GlWide acc = glw(s[0]);
#pragma unroll
        for (int i = 1; i < W; i++) acc = glw_add(acc, s[i]);
        const uint64_t sum = glw_fold(acc);
#pragma unroll
        for (int i = 0; i < W; i++) s[i] = LEAN ? Gl::mul_add_lean(s[i], diag[i], sum) : Gl::mul_add_lazy(s[i], diag[i], sum);
    }
    const uint64_t *tr = ir + RP;
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < W; i++) s[i] = sbox7_gl<LEAN>(Gl::add_const_lazy(s[i], tr[r * W + i]));
        mds_light_gl<W>(s);
    }
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = Gl::canon(s[i]);
}

template <class F, int W, bool LEAN = true>
__host__ __device__ __forceinline__ void p2_permute(typename F::T (&s)[W], const typename F::T *__restrict__ rc,
                                           const typename F::T *__restrict__ diag)
{
    if constexpr (std::is_same<F, Gl>::value) {
        static_assert(P2Spec<Gl, W>::RF == 8 && P2Spec<Gl, W>::RP == 22, "Goldilocks schedule");
        p2_permute_gl<W, LEAN>(s, rc, diag);
        return;
    }
    using T = typename F::T;
    constexpr int HALF = P2Spec<F, W>::RF / 2, RP = P2Spec<F, W>::RP;
    mds_light<F, W>(s);
#pragma unroll 1
    for (int r = 0; r < HALF; r++) external_round<F, W>(s, rc + r * W);
    const T *ir = rc + HALF * W;
#pragma unroll 1
    for (int r = 0; r < RP; r++) {
        s[0] = sbox7<F>(F::add(s[0], ir[r]));
        T sum = s[0];
#pragma unroll
        for (int i = 1; i < W; i++) sum = F::add(sum, s[i]);
        if constexpr (std::is_same<F, Bb>::value && W == 16)
            internal_diag_bb16(s, sum);
        else
            internal_diag_generic<F, W>(s, sum, diag);
    }
    const T *tr = ir + RP;
#pragma unroll 1
    for (int r = 0; r < HALF; r++) external_round<F, W>(s, tr + r * W);
}

// ---------------------------------------------------------------------------
// BabyBear width-16 permutation in exact FP64 (see Bf in zkg_field.h).
// Bounds (p < 2^31): input lanes are fresh u32 values (< 2^32) or reduced to
// |s| <= p/2 + 1 (capacity lanes carry unreduced values between sponge
// absorptions); MDS-light multiplies a bound by at most 35, so the first
// S-box inputs are < 35 2^32 + p < 2^37.2; S-box outputs are < 1.6 p, so the
// later external rounds see |s| < 57 p < 2^37; internal rounds reduce the sum every round and
// the lanes with integer |V_i| >= 2 every fourth round (they stay below 171 p <
// 2^39, see internal_round_bf), scale the V = +-2^-k lanes exactly
// (Bf::mul_inv_pow2: |x| 2^-k + 2^31, so they shrink) and the V = 1 lane grows
// by < p/2 per round.  The last internal round reduces, so the terminal
// external rounds start from the same bounds; every product input stays below
// 2^37 (the product routine needs < 2^41) and every sum below 2^42 (exactness
// needs < 2^53).
__host__ __device__ __forceinline__ void mds_light_bf(double (&s)[16])
{
#pragma unroll
    for (int c = 0; c < 16; c += 4) {
        double a = s[c], b = s[c + 1], cc = s[c + 2], d = s[c + 3];
        double ab = a + b, cd = cc + d;
        double t = ab + cd;
        double tb = t + b, td = t + d;
        s[c] = tb + ab;                        // 2a + 3b + c + d
        s[c + 1] = __builtin_fma(2.0, cc, tb);  // a + 2b + 3c + d
        s[c + 2] = td + cd;                    // a + b + 2c + 3d
        s[c + 3] = __builtin_fma(2.0, a, td);   // 3a + b + c + 2d
    }
    double sum[4];
#pragma unroll
    for (int k = 0; k < 4; k++) sum[k] = (s[k] + s[4 + k]) + (s[8 + k] + s[12 + k]);
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = s[i] + sum[i & 3];
}

__host__ __device__ __forceinline__ double sbox7_bf(double x)
{
    double x2 = Bf::mul(x, x);
    double x3 = Bf::mul(x2, x);
    double x4 = Bf::mul(x2, x2);
    return Bf::mul(x3, x4);
}

// One internal round: s_i = V_i s_i + sum, V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3,
// -4, 2^-8, 1/4, 1/8, 2^-27, -2^-8, -1/16, -2^-27].  With REDUCE the lanes with
// integer |V_i| >= 2 are reduced before they are scaled (to 2.5 p at most after
// the round); without it they grow to at most 4 |s| + p/2, so three rounds
// without a reduction after one with it leave them below 171 p < 2^39.
template <bool REDUCE>
__host__ __device__ __forceinline__ void internal_round_bf(double (&s)[16], double c)
{
    s[0] = sbox7_bf(s[0] + c);
    double sum = s[0];
#pragma unroll
    for (int i = 1; i < 16; i++) sum += s[i];
    sum = Bf::red(sum);
    auto lazy = [](double x) { return REDUCE ? Bf::red(x) : x; };
    s[0] = __builtin_fma(-2.0, s[0], sum);
    s[1] = s[1] + sum;
    s[2] = __builtin_fma(2.0, lazy(s[2]), sum);
    s[3] = Bf::mul_inv_pow2<1, false>(s[3]) + sum;
    s[4] = __builtin_fma(3.0, lazy(s[4]), sum);
    s[5] = __builtin_fma(4.0, lazy(s[5]), sum);
    s[6] = Bf::mul_inv_pow2<1, true>(s[6]) + sum;
    s[7] = __builtin_fma(-3.0, lazy(s[7]), sum);
    s[8] = __builtin_fma(-4.0, lazy(s[8]), sum);
    s[9] = Bf::mul_inv_pow2<8, false>(s[9]) + sum;
    s[10] = Bf::mul_inv_pow2<2, false>(s[10]) + sum;
    s[11] = Bf::mul_inv_pow2<3, false>(s[11]) + sum;
    s[12] = Bf::mul_inv_pow2<27, false>(s[12]) + sum;
    s[13] = Bf::mul_inv_pow2<8, true>(s[13]) + sum;
    s[14] = Bf::mul_inv_pow2<4, true>(s[14]) + sum;
    s[15] = Bf::mul_inv_pow2<27, true>(s[15]) + sum;
}

// rc: flat [initial 4x16][internal 13][terminal 4x16], canonical, as doubles.
// Lanes [RED_FROM, 16) are reduced on entry, as they may carry an unreduced
// output of an earlier permutation (a sponge's capacity); the lanes below
// RED_FROM must hold fresh values in [0, 2^32), canonical or not.  16: every
// lane is fresh (Merkle nodes, single permutations); 8: a full-rate absorption.
template <int RED_FROM = 0>
__host__ __device__ __forceinline__ void p2_permute_bf(double (&s)[16], const double *__restrict__ rc)
{
    static_assert(RED_FROM >= 0 && RED_FROM <= 16, "lane range");
#pragma unroll
    for (int i = RED_FROM; i < 16; i++) s[i] = Bf::red(s[i]);
    mds_light_bf(s);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = sbox7_bf(s[i] + rc[r * 16 + i]);
        mds_light_bf(s);
    }
    const double *ir = rc + 64;
    // 13 internal rounds; the integer-V lanes are reduced in rounds 0, 4, 8, 12
#pragma unroll 1
    for (int r = 0; r < 12; r += 4) {
        internal_round_bf<true>(s, ir[r]);
        internal_round_bf<false>(s, ir[r + 1]);
        internal_round_bf<false>(s, ir[r + 2]);
        internal_round_bf<false>(s, ir[r + 3]);
    }
    internal_round_bf<true>(s, ir[12]);
    const double *tr = ir + 13;
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = sbox7_bf(s[i] + tr[r * 16 + i]);
        mds_light_bf(s);
    }
}

}  // namespace zkg
