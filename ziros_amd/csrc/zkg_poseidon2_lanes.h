// zkg_poseidon2_lanes.h -- width-16 Poseidon2 with the state spread over 16
// lanes (lane l of a 16-lane row holds s_l).
//
// Same permutation as zkg_poseidon2.h (schedule of zkf-metal/src/shaders/
// poseidon2.metal:17-264); this form trades throughput for latency and is used
// where few permutations are independent, at the top of a Merkle tree.  The
// cross-lane steps are DPP moves inside a 16-lane row:
//   * M4 on each 4-chunk: new_k = (x0+x1+x2+x3) + x_k + 2 x_{k+1} (indices mod
//     4), the circulant [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]], with quad
//     permutes fetching x_{k+1}, x_{k+2}, x_{k+3};
//   * the sum over chunks of each position: two row rotations by 8 and 4;
//   * the internal-round sum of all 16 lanes: row rotations by 8, 4, 2, 1.
// Per-lane constants (the lane's external round constants and internal
// diagonal entry) sit in VGPRs; internal round constants are wave-uniform.
#pragma once
#include "zkg_field.h"
#include "zkg_poseidon2.h"

namespace zkg {

// DPP controls: quad_perm encodes the source lane of each quad position.
constexpr int DPP_QUAD_ROT1 = 0x39;  // [1,2,3,0]
constexpr int DPP_QUAD_ROT2 = 0x4E;  // [2,3,0,1]
constexpr int DPP_QUAD_ROT3 = 0x93;  // [3,0,1,2]
constexpr int DPP_ROW_ROR1 = 0x121;
constexpr int DPP_ROW_ROR2 = 0x122;
constexpr int DPP_ROW_ROR4 = 0x124;
constexpr int DPP_ROW_ROR8 = 0x128;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
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

template <class F>
__device__ __forceinline__ typename F::T mds_light_lanes(typename F::T x)
{
    using T = typename F::T;
    T r1 = dpp<F, DPP_QUAD_ROT1>(x), r2 = dpp<F, DPP_QUAD_ROT2>(x), r3 = dpp<F, DPP_QUAD_ROT3>(x);
    T quad = F::add(F::add(x, r1), F::add(r2, r3));
    T y = F::add(F::add(quad, x), F::dbl(r1));
    T t = F::add(y, dpp<F, DPP_ROW_ROR8>(y));
    T col = F::add(t, dpp<F, DPP_ROW_ROR4>(t));
    return F::add(y, col);
}

template <class F>
__device__ __forceinline__ typename F::T row_sum16(typename F::T x)
{
    x = F::add(x, dpp<F, DPP_ROW_ROR8>(x));
    x = F::add(x, dpp<F, DPP_ROW_ROR4>(x));
    x = F::add(x, dpp<F, DPP_ROW_ROR2>(x));
    return F::add(x, dpp<F, DPP_ROW_ROR1>(x));
}

// Per-lane constants of one width-16 permutation.
template <class F>
struct LaneConsts {
    typename F::T ext[8];  // initial rounds 0..3, terminal rounds 0..3 (this lane)
    typename F::T diag;    // internal diagonal entry of this lane
};

template <class F>
__device__ __forceinline__ void load_lane_consts(LaneConsts<F> &k, int lane, const typename F::T *__restrict__ rc,
                                                 const typename F::T *__restrict__ diag)
{
    constexpr int RP = P2Spec<F, 16>::RP;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        k.ext[r] = rc[r * 16 + lane];
        k.ext[4 + r] = rc[4 * 16 + RP + r * 16 + lane];
    }
    if constexpr (std::is_same<F, Bb>::value) {
        // diag in register form: V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 2^-8,
        // 1/4, 1/8, 2^-27, -2^-8, -1/16, -2^-27] (zkf-metal/src/poseidon2/babybear.rs:14-58)
        const uint32_t one = Bb::one();
        uint32_t v;
        switch (lane) {
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
        k.diag = diag[lane];
    }
}

// x is this lane's state element; lane = lane index within its 16-lane row.
template <class F>
__device__ __forceinline__ typename F::T p2_permute_lanes(typename F::T x, int lane, const LaneConsts<F> &k,
                                                          const typename F::T *__restrict__ rc)
{
    using T = typename F::T;
    constexpr int RP = P2Spec<F, 16>::RP;
    x = mds_light_lanes<F>(x);
#pragma unroll
    for (int r = 0; r < 4; r++) x = mds_light_lanes<F>(sbox7<F>(F::add(x, k.ext[r])));
    const T *ir = rc + 4 * 16;
#pragma unroll 1
    for (int r = 0; r < RP; r++) {
        T y = sbox7<F>(F::add(x, ir[r]));
        x = lane == 0 ? y : x;
        T sum = row_sum16<F>(x);
        x = F::add(F::mul(x, k.diag), sum);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) x = mds_light_lanes<F>(sbox7<F>(F::add(x, k.ext[4 + r])));
    return x;
}

}  // namespace zkg
