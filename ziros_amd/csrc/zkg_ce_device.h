// zkg_ce_device.h -- device helpers of the constraint evaluator shared by
// the interpreter kernel (zkg_constraints.hip) and the straight-line kernels
// compiled at run time for one program (their source embeds this file).
//
// A wave evaluates 64 R rows (lane l holds rows l + 64 j).  LDS panels hold PC
// columns of those rows column-major with a padded column stride of
// 64 R * 8 + 8 bytes, so lane l reads row l + 64 j of column c at
// c * stride + (l + 64 j) * 8 without bank conflicts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zkg {

__device__ __forceinline__ uint64_t lds_u64(const unsigned char *p) { return *(const uint64_t *)p; }

// Stage panel `panel` (columns panel*PC .. +PC-1) of this wave's 64 R rows
// into the LDS slot at `dst`: each group of PC lanes reads one row's
// PC * 8-byte segment, so a wave instruction covers 64 / PC whole segments.
template <int R, int PC>
__device__ __forceinline__ void stage_panel(const uint64_t *__restrict__ trace, uint64_t n_rows, uint32_t width,
                                            uint64_t row0, uint32_t panel, uint32_t lane, unsigned char *dst)
{
    constexpr uint32_t COLS = 64 * R * 8 + 8, STEP = 64 / PC;
    constexpr int PASSES = R * PC, BATCH = PASSES < 16 ? PASSES : 16;
    const uint32_t c = lane % PC, r_off = lane / PC;
    const uint32_t col = min(panel * PC + c, width - 1);  // columns past the width: clamped, never read
#pragma unroll
    for (int b0 = 0; b0 < PASSES; b0 += BATCH) {
        uint64_t v[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const uint64_t r = min(row0 + r_off + (uint64_t)(b0 + k) * STEP, n_rows - 1);
            v[k] = trace[r * width + col];
        }
        if (b0 == 0) {
            // Lanes of this wave may still read the slot's previous panel.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) *(uint64_t *)(dst + c * COLS + (r_off + (b0 + k) * STEP) * 8) = v[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Write the output panel (constraints panel*PC .. +PC-1 of the wave's rows,
// those set in `mask`) from LDS at `src` to the row-major output: each group
// of PC lanes writes one row's PC * 8-byte segment.
template <int R, int PC>
__device__ __forceinline__ void flush_panel(uint64_t *__restrict__ out, uint64_t n_rows, uint32_t n_out,
                                            uint64_t row0, uint32_t panel, uint32_t mask, uint32_t lane,
                                            const unsigned char *src)
{
    constexpr uint32_t COLS = 64 * R * 8 + 8, STEP = 64 / PC;
    const uint32_t c = lane % PC, r_off = lane / PC;
    // EMITs of other lanes must have landed.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if ((mask >> c) & 1) {
        uint64_t *const col = out + panel * PC + c;
#pragma unroll
        for (int k = 0; k < R * PC; k++) {
            const uint32_t r = r_off + k * STEP;
            // rows past the end hold the last row's values (see the kernel)
            col[min(row0 + r, n_rows - 1) * n_out] = *(const uint64_t *)(src + c * COLS + r * 8);
        }
    }
    // The next EMITs overwrite the panel.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

}  // namespace zkg
