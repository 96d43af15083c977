// zkg_constraints.hip -- constraint evaluation over a Goldilocks trace: the
// stack-machine bytecode of ConstraintEvalAccelerator::eval_trace_goldilocks
// (zkf-core/src/acceleration.rs:219-242; instruction set and semantics of
// zkf-metal/src/shaders/constraint_eval.metal:1-74 and
// zkf-metal/src/constraint_eval.rs:15-27), evaluated for every trace row.
//
// The reference interprets the bytecode in every thread with a private
// 32-entry stack and reads trace[row * width + col] straight from memory.
// Here the host first translates the bytecode into a device program:
//
//  * the program is wave-uniform, so its words are read through the scalar
//    unit and dispatch is a scalar branch; the stack pointer is an SGPR;
//  * the top of the stack lives in a VGPR pair, the entries below it in LDS
//    ([depth][lane], conflict-free), sized by the program's real depth;
//  * "push x; op" pairs become one instruction with x as the operand (a
//    column, a constant, or the top itself for "dup; op"), which removes the
//    LDS push and pop of most expression leaves;
//  * trace columns are staged per wave into LDS panels of 16 columns x 64
//    rows with coalesced 128-byte row segments (the trace is row-major, so a
//    direct per-lane load would touch 64 cache lines for 8 useful bytes
//    each).  The host plans which panel sits in which of K slots with
//    Belady's rule (evict the panel used furthest ahead), so a program that
//    walks its columns in order stages each panel once;
//  * constraint indices the program never emits are written as zero in the
//    same kernel (the reference returns a zero-filled output buffer).
//
// One lane evaluates one row; a 256-thread block holds four independent
// waves with private LDS regions and no block-level barriers.
#include "../../include/zkg.h"
#include "zkg_field.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace zkg;

namespace {

// Device program opcodes (word = op << 24 | arg).
enum : uint32_t {
    CE_PUSH_L = 0,  // push; top = lds[arg]                 (LOAD col)
    CE_PUSH_K,      // push; top = consts[arg]              (CONST k)
    CE_ADD_S,       // top = pop() + top                    (ADD)
    CE_SUB_S,       // top = pop() - top                    (SUB)
    CE_MUL_S,       // top = pop() * top                    (MUL)
    CE_DUP,         // push (top stays)                     (DUP)
    CE_EMIT,        // out[row][arg] = top; top = pop()     (EMIT j)
    CE_ADD_L,       // top = top + lds[arg]                 (LOAD col; ADD)
    CE_SUB_L,       // top = top - lds[arg]                 (LOAD col; SUB)
    CE_MUL_L,       // top = top * lds[arg]                 (LOAD col; MUL)
    CE_ADD_K,       // top = top + consts[arg]              (CONST k; ADD)
    CE_SUB_K,       // top = top - consts[arg]              (CONST k; SUB)
    CE_MUL_K,       // top = top * consts[arg]              (CONST k; MUL)
    CE_DBL,         // top = top + top                      (DUP; ADD)
    CE_ZERO_T,      // top = 0                              (DUP; SUB)
    CE_SQR,         // top = top * top                      (DUP; MUL)
    CE_STAGE,       // slot arg >> 20 <- panel arg & 0xFFFFF
    CE_ZERO,        // out[row][arg] = 0 (never emitted)
    CE_END,
};

constexpr int CE_PANEL = 16;                             // columns per panel
constexpr int CE_ROWS = 64;                              // rows per wave
constexpr int CE_WAVES = 4;                              // waves per block
constexpr uint32_t CE_COL_STRIDE = CE_ROWS * 8 + 8;      // LDS bytes per panel column (padded)
constexpr uint32_t CE_SLOT_BYTES = CE_PANEL * CE_COL_STRIDE;
constexpr uint32_t CE_STACK_ROW = CE_ROWS * 8;           // LDS bytes per stack depth
constexpr int CE_MAX_STACK = 32;                         // the reference's stack depth
constexpr int CE_MAX_SLOTS = 4;
constexpr uint32_t CE_MAX_PANELS = 1u << 20;
constexpr uint32_t CE_LDS_BYTES = 160u * 1024;            // LDS per CU on gfx950
static_assert(CE_WAVES * (CE_SLOT_BYTES + (CE_MAX_STACK - 1) * CE_STACK_ROW) <= CE_LDS_BYTES,
              "one panel slot and a full stack must fit");

__device__ __forceinline__ uint64_t lds_u64(const unsigned char *p) { return *(const uint64_t *)p; }

// Stage panel `panel` (columns panel*16 .. +15) of this wave's 64 rows into
// the LDS slot at `dst`, column-major with a padded column stride: each
// 16-lane quarter reads one row's 128-byte segment.
__device__ __forceinline__ void stage_panel(const uint64_t *__restrict__ trace, uint64_t n_rows, uint32_t width,
                                            uint64_t row0, uint32_t panel, uint32_t lane, unsigned char *dst)
{
    const uint32_t c = lane % CE_PANEL;
    const uint32_t col = min(panel * CE_PANEL + c, width - 1);  // columns past the width: clamped, never read
    constexpr int PASSES = CE_ROWS * CE_PANEL / 64;
    uint64_t v[PASSES];
#pragma unroll
    for (int k = 0; k < PASSES; k++) {
        const uint64_t r = min(row0 + lane / CE_PANEL + (uint64_t)k * (64 / CE_PANEL), n_rows - 1);
        v[k] = trace[r * width + col];
    }
    // Lanes of this wave may still read the slot's previous panel.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < PASSES; k++)
        *(uint64_t *)(dst + c * CE_COL_STRIDE + (lane / CE_PANEL + k * (64 / CE_PANEL)) * 8) = v[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(CE_WAVES * 64) void k_constraint_eval_gl(
    const uint64_t *__restrict__ trace, uint64_t n_rows, uint32_t width, const uint32_t *__restrict__ prog,
    const uint64_t *__restrict__ consts, uint32_t n_out, uint32_t wave_bytes, uint32_t stack_off,
    uint64_t *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char ce_lds[];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t row0 = ((uint64_t)blockIdx.x * CE_WAVES + wave) * CE_ROWS;
    if (row0 >= n_rows) return;
    unsigned char *const base = ce_lds + wave * wave_bytes;
    unsigned char *const mine = base + lane * 8;  // this lane's column in panels and stack
    unsigned char *const stk = mine + stack_off;
    const uint64_t row = row0 + lane;
    const bool live = row < n_rows;
    uint64_t *const orow = out + row * n_out;
    uint64_t top = 0;
    uint32_t d = 0;  // stack depth; the top is in `top`, entries 0..d-2 in LDS

    for (uint32_t pc = 0;; pc += 4) {
        const uint4 w4 = *(const uint4 *)(prog + pc);
        const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t op = ws[i] >> 24, arg = ws[i] & 0xFFFFFFu;
            switch (op) {
            case CE_PUSH_L:
                if (d) *(uint64_t *)(stk + (d - 1) * CE_STACK_ROW) = top;
                d++;
                top = lds_u64(mine + arg);
                break;
            case CE_PUSH_K:
                if (d) *(uint64_t *)(stk + (d - 1) * CE_STACK_ROW) = top;
                d++;
                top = consts[arg];
                break;
            case CE_ADD_S: top = Gl::add(lds_u64(stk + (--d - 1) * CE_STACK_ROW), top); break;
            case CE_SUB_S: top = Gl::sub(lds_u64(stk + (--d - 1) * CE_STACK_ROW), top); break;
            case CE_MUL_S: top = Gl::mul(lds_u64(stk + (--d - 1) * CE_STACK_ROW), top); break;
            case CE_DUP: *(uint64_t *)(stk + (d++ - 1) * CE_STACK_ROW) = top; break;
            case CE_EMIT:
                if (live) orow[arg] = top;
                if (--d) top = lds_u64(stk + (d - 1) * CE_STACK_ROW);
                break;
            case CE_ADD_L: top = Gl::add(top, lds_u64(mine + arg)); break;
            case CE_SUB_L: top = Gl::sub(top, lds_u64(mine + arg)); break;
            case CE_MUL_L: top = Gl::mul(top, lds_u64(mine + arg)); break;
            case CE_ADD_K: top = Gl::add(top, consts[arg]); break;
            case CE_SUB_K: top = Gl::sub(top, consts[arg]); break;
            case CE_MUL_K: top = Gl::mul(top, consts[arg]); break;
            case CE_DBL: top = Gl::add(top, top); break;
            case CE_ZERO_T: top = 0; break;
            case CE_SQR: top = Gl::mul(top, top); break;
            case CE_STAGE:
                stage_panel(trace, n_rows, width, row0, arg & (CE_MAX_PANELS - 1), lane,
                            base + (arg >> 20) * CE_SLOT_BYTES);
                break;
            case CE_ZERO:
                if (live) orow[arg] = 0;
                break;
            default: return;  // CE_END
            }
        }
    }
}

// ------------------------------------------------------------------ host translation
struct SrcIns {
    uint32_t op, arg;
};

struct MidIns {
    uint32_t op, arg;
    int64_t col;  // trace column read by the instruction, -1 if none
};

}  // namespace

struct zkg_constraint_program {
    zkg_ctx *ctx = nullptr;
    uint32_t width = 0, n_constraints = 0;
    uint32_t slots = 0, stack_slots = 0, stages = 0, n_words = 0;
    uint32_t wave_bytes = 0, stack_off = 0, lds_budget = 0;
    void *d_prog = nullptr, *d_consts = nullptr;
};

namespace {

// Validate the reference bytecode and translate it.  Returns ZKG_OK and fills
// `p` (host side) and `words`.
int ce_translate(const uint32_t *bc, uint32_t n, uint32_t width, uint32_t n_consts, uint32_t n_out,
                 uint32_t lds_budget, zkg_constraint_program &p, std::vector<uint32_t> &words)
{
    // 1. validate: reference opcodes 0..6, anything else is a no-op
    //    (constraint_eval.metal:38-73 has no default case)
    std::vector<SrcIns> src;
    std::vector<int> depth_before;
    int sp = 0, maxd = 0;
    std::vector<bool> emitted(n_out, false);
    for (uint32_t pc = 0; pc < n; pc++) {
        const uint32_t op = bc[pc] >> 24, arg = bc[pc] & 0xFFFFFFu;
        if (op > 6) continue;
        depth_before.push_back(sp);
        switch (op) {
        case 0:
            if (arg >= n_consts) return set_err(ZKG_EINVAL, "instruction %u: constant %u of %u", pc, arg, n_consts);
            sp++;
            break;
        case 1:
            if (arg >= width) return set_err(ZKG_EINVAL, "instruction %u: column %u of width %u", pc, arg, width);
            sp++;
            break;
        case 2: case 3: case 4:
            if (sp < 2) return set_err(ZKG_EINVAL, "instruction %u: stack underflow", pc);
            sp--;
            break;
        case 5:
            if (sp < 1) return set_err(ZKG_EINVAL, "instruction %u: stack underflow", pc);
            sp++;
            break;
        case 6:
            if (sp < 1) return set_err(ZKG_EINVAL, "instruction %u: stack underflow", pc);
            if (arg >= n_out) return set_err(ZKG_EINVAL, "instruction %u: constraint %u of %u", pc, arg, n_out);
            emitted[arg] = true;
            sp--;
            break;
        }
        if (sp > CE_MAX_STACK) return set_err(ZKG_EINVAL, "instruction %u: stack deeper than %d", pc, CE_MAX_STACK);
        maxd = std::max(maxd, sp);
        src.push_back({op, arg});
    }

    // 2. fuse "push x; binop" (the push must sit on a non-empty stack, which a
    //    following binop guarantees) and map reference ops to device ops
    std::vector<MidIns> mid;
    for (size_t i = 0; i < src.size(); i++) {
        const SrcIns s = src[i];
        const bool next_bin = i + 1 < src.size() && src[i + 1].op >= 2 && src[i + 1].op <= 4;
        const uint32_t bin = next_bin ? src[i + 1].op - 2 : 0;  // 0 add, 1 sub, 2 mul
        if (next_bin && (s.op == 0 || s.op == 1 || s.op == 5)) {
            if (s.op == 1) mid.push_back({CE_ADD_L + bin, 0, (int64_t)s.arg});
            else if (s.op == 0) mid.push_back({CE_ADD_K + bin, s.arg, -1});
            else mid.push_back({CE_DBL + bin, 0, -1});
            i++;
            continue;
        }
        switch (s.op) {
        case 0: mid.push_back({CE_PUSH_K, s.arg, -1}); break;
        case 1: mid.push_back({CE_PUSH_L, 0, (int64_t)s.arg}); break;
        case 2: mid.push_back({CE_ADD_S, 0, -1}); break;
        case 3: mid.push_back({CE_SUB_S, 0, -1}); break;
        case 4: mid.push_back({CE_MUL_S, 0, -1}); break;
        case 5: mid.push_back({CE_DUP, 0, -1}); break;
        case 6: mid.push_back({CE_EMIT, s.arg, -1}); break;
        }
    }

    // 3. panel slots: simulate Belady replacement for K = 1..4 slots and take
    //    the smallest K whose stage count is within 25% of the best
    std::vector<uint32_t> uses;  // panel of each column read, in order
    for (const MidIns &m : mid)
        if (m.col >= 0) uses.push_back((uint32_t)(m.col / CE_PANEL));
    std::vector<size_t> next_use(uses.size());
    {
        std::vector<size_t> last((width + CE_PANEL - 1) / CE_PANEL, SIZE_MAX);
        for (size_t i = uses.size(); i-- > 0;) {
            next_use[i] = last[uses[i]];
            last[uses[i]] = i;
        }
    }
    auto simulate = [&](int K, std::vector<int> *slot_of_use, std::vector<bool> *stage_at) {
        std::vector<int64_t> res(K, -1);      // panel in each slot
        std::vector<size_t> nxt(K, SIZE_MAX);  // next use of that panel
        uint32_t stages = 0;
        for (size_t i = 0; i < uses.size(); i++) {
            int s = -1;
            for (int k = 0; k < K; k++)
                if (res[k] == (int64_t)uses[i]) s = k;
            const bool miss = s < 0;
            if (miss) {
                for (int k = 0; k < K && s < 0; k++)
                    if (res[k] < 0) s = k;
                if (s < 0) {
                    s = 0;
                    for (int k = 1; k < K; k++)
                        if (nxt[k] > nxt[s]) s = k;
                }
                res[s] = uses[i];
                stages++;
            }
            nxt[s] = next_use[i];
            if (slot_of_use) (*slot_of_use)[i] = s;
            if (stage_at) (*stage_at)[i] = miss;
        }
        return stages;
    };
    const uint32_t n_panels = (width + CE_PANEL - 1) / CE_PANEL;
    const int kmax = (int)std::max<uint32_t>(1, std::min<uint32_t>(CE_MAX_SLOTS, n_panels));
    const uint32_t best = simulate(kmax, nullptr, nullptr);
    int K = 1;
    while (K < kmax && (uint64_t)simulate(K, nullptr, nullptr) * 4 > (uint64_t)best * 5) K++;
    const uint32_t stack_bytes = (uint32_t)std::max(0, maxd - 1) * CE_STACK_ROW;
    while (K > 1 && CE_WAVES * (K * CE_SLOT_BYTES + stack_bytes) > lds_budget) K--;
    if (CE_WAVES * ((uses.empty() ? 0 : CE_SLOT_BYTES) + stack_bytes) > lds_budget)
        return set_err(ZKG_EINVAL, "program needs more than %u bytes of LDS per block", lds_budget);
    std::vector<int> slot_of_use(uses.size());
    std::vector<bool> stage_at(uses.size());
    p.stages = simulate(K, &slot_of_use, &stage_at);
    p.slots = uses.empty() ? 0 : K;

    // 4. emit device words
    p.stack_off = p.slots * CE_SLOT_BYTES;
    p.stack_slots = (uint32_t)std::max(0, maxd - 1);
    p.wave_bytes = (p.stack_off + stack_bytes + 15) & ~15u;
    words.clear();
    size_t u = 0;
    for (const MidIns &m : mid) {
        uint32_t arg = m.arg;
        if (m.col >= 0) {
            const uint32_t panel = (uint32_t)(m.col / CE_PANEL), slot = (uint32_t)slot_of_use[u];
            if (stage_at[u]) words.push_back(CE_STAGE << 24 | slot << 20 | panel);
            arg = slot * CE_SLOT_BYTES + (uint32_t)(m.col % CE_PANEL) * CE_COL_STRIDE;
            u++;
        }
        words.push_back(m.op << 24 | arg);
    }
    for (uint32_t j = 0; j < n_out; j++)
        if (!emitted[j]) words.push_back(CE_ZERO << 24 | j);
    do words.push_back(CE_END << 24); while (words.size() % 4);
    p.n_words = (uint32_t)words.size();
    return ZKG_OK;
}

}  // namespace

static int ce_check_args(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions, uint32_t n_constraints)
{
    if (!bytecode && n_instructions) return set_err(ZKG_EINVAL, "null argument");
    if (!width || width > (1u << 24)) return set_err(ZKG_EINVAL, "width %u outside 1..2^24", width);
    if (!n_instructions) return set_err(ZKG_EINVAL, "empty bytecode");
    if (!n_constraints || n_constraints > (1u << 24))
        return set_err(ZKG_EINVAL, "n_constraints %u outside 1..2^24", n_constraints);
    return ZKG_OK;
}

static void ce_info(const zkg_constraint_program &p, uint32_t info[6])
{
    info[0] = p.n_words;
    info[1] = p.slots;
    info[2] = p.stages;
    info[3] = p.stack_slots;
    info[4] = p.wave_bytes;
    info[5] = CE_PANEL;
}

extern "C" int zkg_constraint_translate(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                                        uint32_t n_constants, uint32_t n_constraints, uint32_t *words, uint32_t cap,
                                        uint32_t info[6])
{
    if (!info || (!words && cap)) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = ce_check_args(width, bytecode, n_instructions, n_constraints)) return rc;
    zkg_constraint_program p;
    std::vector<uint32_t> w;
    if (int rc = ce_translate(bytecode, n_instructions, width, n_constants, n_constraints, CE_LDS_BYTES, p, w))
        return rc;
    ce_info(p, info);
    std::memcpy(words, w.data(), std::min<size_t>(cap, w.size()) * 4);
    if (cap < w.size()) return set_err(ZKG_EINVAL, "%zu words needed, room for %u", w.size(), cap);
    return ZKG_OK;
}

extern "C" int zkg_constraint_program_create(zkg_ctx *ctx, uint32_t width, const uint32_t *bytecode,
                                             uint32_t n_instructions, const uint64_t *constants,
                                             uint32_t n_constants, uint32_t n_constraints,
                                             zkg_constraint_program **out)
{
    if (!ctx || !out || (!constants && n_constants)) return set_err(ZKG_EINVAL, "null argument");
    *out = nullptr;
    if (int rc = ce_check_args(width, bytecode, n_instructions, n_constraints)) return rc;
    for (uint32_t i = 0; i < n_constants; i++)
        if (constants[i] >= Gl::P) return set_err(ZKG_EINVAL, "constant %u is not canonical", i);
    auto *p = new zkg_constraint_program;
    p->ctx = ctx;
    p->width = width;
    p->n_constraints = n_constraints;
    std::vector<uint32_t> words;
    int lds_max = 0;
    int rc = ZKG_OK;
    hipError_t e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    if (e != hipSuccess) rc = hip_err(e, "device LDS size");
    p->lds_budget = std::min<uint32_t>(CE_LDS_BYTES, (uint32_t)std::max(lds_max, 0));
    if (!rc) rc = ce_translate(bytecode, n_instructions, width, n_constants, n_constraints, p->lds_budget, *p, words);
    if (!rc) rc = dev_alloc(ctx, words.size() * 4, &p->d_prog);
    if (!rc) rc = dev_alloc(ctx, std::max<uint32_t>(n_constants, 1) * 8, &p->d_consts);
    if (!rc) e = hipMemcpyAsync(p->d_prog, words.data(), words.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess && n_constants)
        e = hipMemcpyAsync(p->d_consts, constants, (uint64_t)n_constants * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // host vectors go out of scope
    if (!rc && e != hipSuccess) rc = hip_err(e, "constraint program upload");
    if (rc) {
        zkg_constraint_program_free(p);
        return rc;
    }
    *out = p;
    return ZKG_OK;
}

extern "C" int zkg_constraint_program_info(const zkg_constraint_program *p, uint32_t info[6])
{
    if (!p || !info) return set_err(ZKG_EINVAL, "null argument");
    ce_info(*p, info);
    return ZKG_OK;
}

extern "C" int zkg_constraint_program_eval(zkg_constraint_program *p, const zkg_buf *trace, uint64_t n_rows,
                                           zkg_buf *out)
{
    if (!p || !trace || !out) return set_err(ZKG_EINVAL, "null argument");
    if (!n_rows) return set_err(ZKG_EINVAL, "no rows");
    if (n_rows > (UINT64_MAX / 8) / p->width || n_rows * p->width * 8 > trace->bytes)
        return set_err(ZKG_EINVAL, "trace buffer too small for %llu rows of width %u", (unsigned long long)n_rows,
                       p->width);
    if (n_rows > (UINT64_MAX / 8) / p->n_constraints || n_rows * p->n_constraints * 8 > out->bytes)
        return set_err(ZKG_EINVAL, "output buffer too small");
    if (trace->ctx != p->ctx || out->ctx != p->ctx) return set_err(ZKG_EINVAL, "buffers of another context");
    const uint32_t lds = CE_WAVES * p->wave_bytes;
    if (lds > 64u * 1024) {
        // above the default dynamic-LDS limit; gfx950 has 160 KiB per CU
        hipError_t e = hipFuncSetAttribute((const void *)k_constraint_eval_gl,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_budget);
        if (e != hipSuccess) return hip_err(e, "constraint eval LDS limit");
    }
    const uint64_t waves = (n_rows + CE_ROWS - 1) / CE_ROWS;
    const uint64_t blocks = (waves + CE_WAVES - 1) / CE_WAVES;
    if (blocks > 0x7FFFFFFFull) return set_err(ZKG_EINVAL, "too many rows");
    hipLaunchKernelGGL(k_constraint_eval_gl, dim3((uint32_t)blocks), dim3(CE_WAVES * 64), lds,
                       p->ctx->stream, (const uint64_t *)trace->ptr, n_rows, p->width, (const uint32_t *)p->d_prog,
                       (const uint64_t *)p->d_consts, p->n_constraints, p->wave_bytes, p->stack_off,
                       (uint64_t *)out->ptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_err(e, "constraint eval launch");
    return ZKG_OK;
}

extern "C" void zkg_constraint_program_free(zkg_constraint_program *p)
{
    if (!p) return;
    if (p->d_prog || p->d_consts) (void)hipStreamSynchronize(p->ctx->stream);
    if (p->d_prog) dev_free(p->ctx, p->d_prog);
    if (p->d_consts) dev_free(p->ctx, p->d_consts);
    delete p;
}

extern "C" int zkg_constraint_eval_gl(zkg_ctx *ctx, const zkg_buf *trace, uint64_t n_rows, uint32_t width,
                                      const uint32_t *bytecode, uint32_t n_instructions, const uint64_t *constants,
                                      uint32_t n_constants, uint32_t n_constraints, zkg_buf *out)
{
    zkg_constraint_program *p = nullptr;
    int rc = zkg_constraint_program_create(ctx, width, bytecode, n_instructions, constants, n_constants,
                                           n_constraints, &p);
    if (!rc) rc = zkg_constraint_program_eval(p, trace, n_rows, out);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_err(e, "constraint eval");
    }
    zkg_constraint_program_free(p);
    return rc;
}
