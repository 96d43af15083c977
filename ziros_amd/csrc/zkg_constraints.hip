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
#include "zkg_ce_device.h"
#include "zkg_field.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace zkg;

namespace {

// Device program: one word per instruction, op << 27 | slot << 22 | arg.
// `slot` is the LDS stack entry the instruction spills to or reloads from
// (the host knows the stack depth at every instruction), or the panel slot of
// CE_STAGE; `arg` is an LDS byte offset, a constant index, a constraint index
// or a panel index.
enum : uint32_t {
    CE_PUSH_L = 0,  // stk[slot] = top; top = lds[arg]      (LOAD col)
    CE_PUSH_L0,     // top = lds[arg] (empty stack)         (LOAD col)
    CE_PUSH_K,      // stk[slot] = top; top = consts[arg]   (CONST k)
    CE_PUSH_K0,     // top = consts[arg] (empty stack)      (CONST k)
    CE_ADD_S,       // top = stk[slot] + top                (ADD)
    CE_SUB_S,       // top = stk[slot] - top                (SUB)
    CE_MUL_S,       // top = stk[slot] * top                (MUL)
    CE_DUP,         // stk[slot] = top                      (DUP)
    CE_EMIT,        // lds[arg] = top; top = stk[slot]      (EMIT j, into the output panel)
    CE_EMIT0,       // lds[arg] = top (stack empties)       (EMIT j)
    CE_ADD_L,       // top = top + lds[arg]                 (LOAD col; ADD)
    CE_SUB_L,       // top = top - lds[arg]                 (LOAD col; SUB)
    CE_MUL_L,       // top = top * lds[arg]                 (LOAD col; MUL)
    CE_ADD_K,       // top = top + consts[arg]              (CONST k; ADD)
    CE_SUB_K,       // top = top - consts[arg]              (CONST k; SUB)
    CE_MUL_K,       // top = top * consts[arg]              (CONST k; MUL)
    CE_DBL,         // top = top + top                      (DUP; ADD)
    CE_ZERO_T,      // top = 0                              (DUP; SUB)
    CE_SQR,         // top = top * top                      (DUP; MUL)
    CE_STAGE,       // panel slot `slot` <- panel `arg`
    CE_OCLEAR,      // zero the output panel at LDS offset arg
    CE_OFLUSH,      // output panel -> constraints arg*PC.., columns in the next word's mask
    CE_ACC,         // acc += coef[arg] * top; top = stk[slot]  (quotient: last EMIT of constraint arg)
    CE_ACC0,        // acc += coef[arg] * top (stack empties)
    CE_DROP,        // top = stk[slot]                           (quotient: an EMIT overwritten later)
    CE_DROP0,       // stack empties
    CE_END,
};

__host__ __device__ constexpr uint32_t ce_word(uint32_t op, uint32_t slot, uint32_t arg)
{
    return op << 27 | slot << 22 | arg;
}


constexpr int CE_WAVES = 4;                   // waves per block
constexpr int CE_MAX_STACK = 32;              // the reference's stack depth
constexpr int CE_MAX_SLOTS = 4;
constexpr uint32_t CE_ARG_LIMIT = 1u << 22;
constexpr uint32_t CE_LDS_BYTES = 160u * 1024;  // LDS per CU on gfx950

// Work shape of a wave: each lane evaluates R rows (rows lane + 64 j), so one
// decoded instruction drives R independent field operations; panels hold PC
// columns of the wave's 64 R rows, column-major with a padded column stride.
struct CeShape {
    uint32_t rows_per_lane, panel;
    uint32_t wave_rows() const { return 64 * rows_per_lane; }
    uint32_t col_stride() const { return wave_rows() * 8 + 8; }
    uint32_t slot_bytes() const { return panel * col_stride(); }
    uint32_t stack_row() const { return wave_rows() * 8; }
};

#ifndef ZKG_CE_ROWS_PER_LANE
#define ZKG_CE_ROWS_PER_LANE 2
#endif
#ifndef ZKG_CE_PANEL
#define ZKG_CE_PANEL 8
#endif

#define CE_EACH for (int j = 0; j < R; j++)

template <int R, int PC>
__global__ __launch_bounds__(CE_WAVES * 64) void k_constraint_eval_gl(
    const uint64_t *__restrict__ trace, uint64_t n_rows, uint32_t width, const uint32_t *__restrict__ prog,
    const uint64_t *__restrict__ consts, uint32_t n_out, uint32_t wave_bytes, uint32_t stack_off,
    uint32_t out_off, uint64_t *__restrict__ out, const uint64_t *__restrict__ coef,
    const uint64_t *__restrict__ invz, uint32_t log_block, uint64_t *__restrict__ qout)
{
    constexpr uint32_t WR = 64 * R, COLS = WR * 8 + 8, SLOT = PC * COLS, SROW = WR * 8;
    extern __shared__ __align__(16) unsigned char ce_lds[];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t row0 = ((uint64_t)blockIdx.x * CE_WAVES + wave) * WR;
    if (row0 >= n_rows) return;
    unsigned char *const base = ce_lds + wave * wave_bytes;
    unsigned char *const mine = base + lane * 8;  // row lane + 64 j sits at mine + 512 j
    unsigned char *const stk = mine + stack_off;
    // Lanes past the last row evaluate the last row again (staging clamps
    // their loads to it); flushes store those copies over the same values.
    uint64_t top[R], acc[R];
    CE_EACH top[j] = acc[j] = 0;

    uint32_t pc = 0, w = prog[0];
    for (;;) {
        const uint32_t op = w >> 27, arg = w & (CE_ARG_LIMIT - 1);
        if (op == CE_END) break;
        const uint32_t slot = (w >> 22) & 31;
        unsigned char *const sp = stk + slot * SROW;
        w = prog[++pc];  // the next word loads while this one executes
        switch (op) {
        case CE_PUSH_L: CE_EACH *(uint64_t *)(sp + 512 * j) = top[j]; [[fallthrough]];
        case CE_PUSH_L0: CE_EACH top[j] = lds_u64(mine + arg + 512 * j); break;
        case CE_PUSH_K: CE_EACH *(uint64_t *)(sp + 512 * j) = top[j]; [[fallthrough]];
        case CE_PUSH_K0: {
            const uint64_t k = consts[arg];
            CE_EACH top[j] = k;
            break;
        }
        case CE_ADD_S: CE_EACH top[j] = Gl::add(lds_u64(sp + 512 * j), top[j]); break;
        case CE_SUB_S: CE_EACH top[j] = Gl::sub(lds_u64(sp + 512 * j), top[j]); break;
        case CE_MUL_S: CE_EACH top[j] = Gl::mul(lds_u64(sp + 512 * j), top[j]); break;
        case CE_DUP: CE_EACH *(uint64_t *)(sp + 512 * j) = top[j]; break;
        case CE_EMIT:
            CE_EACH *(uint64_t *)(mine + arg + 512 * j) = top[j];
            CE_EACH top[j] = lds_u64(sp + 512 * j);
            break;
        case CE_EMIT0: CE_EACH *(uint64_t *)(mine + arg + 512 * j) = top[j]; break;
        case CE_ADD_L: CE_EACH top[j] = Gl::add(top[j], lds_u64(mine + arg + 512 * j)); break;
        case CE_SUB_L: CE_EACH top[j] = Gl::sub(top[j], lds_u64(mine + arg + 512 * j)); break;
        case CE_MUL_L: CE_EACH top[j] = Gl::mul(top[j], lds_u64(mine + arg + 512 * j)); break;
        case CE_ADD_K: { const uint64_t k = consts[arg]; CE_EACH top[j] = Gl::add(top[j], k); break; }
        case CE_SUB_K: { const uint64_t k = consts[arg]; CE_EACH top[j] = Gl::sub(top[j], k); break; }
        case CE_MUL_K: { const uint64_t k = consts[arg]; CE_EACH top[j] = Gl::mul(top[j], k); break; }
        case CE_DBL: CE_EACH top[j] = Gl::add(top[j], top[j]); break;
        case CE_ZERO_T: CE_EACH top[j] = 0; break;
        case CE_SQR: CE_EACH top[j] = Gl::mul(top[j], top[j]); break;
        case CE_STAGE: stage_panel<R, PC>(trace, n_rows, width, row0, arg, lane, base + slot * SLOT); break;
        case CE_OCLEAR:
#pragma unroll
            for (int c = 0; c < PC; c++) CE_EACH *(uint64_t *)(mine + arg + c * COLS + 512 * j) = 0;
            break;
        case CE_OFLUSH: {
            const uint32_t mask = w;  // the mask word
            w = prog[++pc];
            flush_panel<R, PC>(out, n_rows, n_out, row0, arg, mask, lane, base + out_off);
            break;
        }
        case CE_ACC: {
            const uint64_t k = coef[arg];
            CE_EACH acc[j] = Gl::add(acc[j], Gl::mul(top[j], k));
            CE_EACH top[j] = lds_u64(sp + 512 * j);
            break;
        }
        case CE_ACC0: {
            const uint64_t k = coef[arg];
            CE_EACH acc[j] = Gl::add(acc[j], Gl::mul(top[j], k));
            break;
        }
        case CE_DROP: CE_EACH top[j] = lds_u64(sp + 512 * j); break;
        case CE_DROP0: break;
        default: break;
        }
    }
    if (qout) {  // quotient: one value per row, divided by the vanishing polynomial
        CE_EACH {
            const uint64_t r = min(row0 + lane + 64 * j, n_rows - 1);
            qout[r] = Gl::mul(acc[j], invz[r >> log_block]);
        }
    }
}

#undef CE_EACH

static_assert(CE_WAVES * (2 * 16 * (64 * 8 + 8) + (CE_MAX_STACK - 1) * 64 * 8) <= CE_LDS_BYTES,
              "at one row per lane, an input and an output panel of 16 columns and a full stack must fit");

// Every work shape the library is built with.
#define CE_SHAPES(X) X(1, 16) X(2, 4) X(2, 8) X(2, 16) X(4, 4) X(4, 8)

bool ce_shape_built(CeShape s)
{
#define X(R, PC) if (s.rows_per_lane == R && s.panel == PC) return true;
    CE_SHAPES(X)
#undef X
    return false;
}

struct CeQuot {  // quotient mode operands (all null for plain evaluation)
    const uint64_t *coef = nullptr, *invz = nullptr;
    uint32_t log_block = 0;
    uint64_t *qout = nullptr;
};

hipError_t ce_launch(CeShape s, dim3 grid, uint32_t lds, hipStream_t st, const uint64_t *trace, uint64_t n_rows,
                     uint32_t width, const uint32_t *prog, const uint64_t *consts, uint32_t n_out, uint32_t wave_bytes,
                     uint32_t stack_off, uint32_t out_off, uint64_t *out, uint32_t lds_limit, const CeQuot &q)
{
#define X(R, PC)                                                                                                  \
    if (s.rows_per_lane == R && s.panel == PC) {                                                                  \
        if (lds > 64u * 1024) {                                                                                   \
            /* above the default dynamic-LDS limit */                                                             \
            hipError_t e = hipFuncSetAttribute((const void *)k_constraint_eval_gl<R, PC>,                         \
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit);       \
            if (e != hipSuccess) return e;                                                                        \
        }                                                                                                         \
        hipLaunchKernelGGL((k_constraint_eval_gl<R, PC>), grid, dim3(CE_WAVES * 64), lds, st, trace, n_rows,     \
                           width, prog, consts, n_out, wave_bytes, stack_off, out_off, out, q.coef, q.invz,     \
                           q.log_block, q.qout);                                                                  \
        return hipGetLastError();                                                                                 \
    }
    CE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

// The work shape: `dflt`, or ZKG_CE_SHAPE="R,PC" (a tuning override that
// applies to interpreted and compiled programs; it must name a built shape).
int ce_pick_shape(CeShape dflt, CeShape *out)
{
    CeShape s = dflt;
    if (const char *e = getenv("ZKG_CE_SHAPE")) {
        unsigned r = 0, pc = 0;
        if (sscanf(e, "%u,%u", &r, &pc) != 2 || !ce_shape_built(CeShape{r, pc}))
            return set_err(ZKG_EINVAL, "ZKG_CE_SHAPE=%s is not a built work shape", e);
        s = CeShape{r, pc};
    }
    *out = s;
    return ZKG_OK;
}

constexpr CeShape CE_INTERP_SHAPE{ZKG_CE_ROWS_PER_LANE, ZKG_CE_PANEL};
constexpr CeShape CE_COMPILED_SHAPE{1, 16};

// ------------------------------------------------------------------ host translation
struct SrcIns {
    uint32_t op, arg;
};

struct MidIns {
    uint32_t op, arg;
    int64_t col;  // trace column read by the instruction, -1 if none
};

// A translated program: the device words and the per-wave LDS layout they
// address (input panel slots, then the stack below the top, then the output
// panel).
struct CeLayout {
    CeShape shape{1, 16};
    uint32_t slots = 0, stages = 0, stack_slots = 0;
    uint32_t stack_off = 0, out_off = 0, wave_bytes = 0;
    std::vector<uint32_t> words;
};

// Validate the reference bytecode and translate it for `shape`.  The stack
// below the top lives in LDS for the interpreter; compiled kernels keep it in
// registers (stack_in_lds false).  A block holds `block_waves` waves and may
// use `lds_budget` bytes.  In quotient mode the last EMIT of each constraint
// accumulates coef[k] * value instead (earlier EMITs of it only pop) and
// there is no output panel.  ZKG_ENOMEM: does not fit at this shape.
int ce_translate(const uint32_t *bc, uint32_t n, uint32_t width, uint32_t n_consts, uint32_t n_out, CeShape shape,
                 bool stack_in_lds, uint32_t block_waves, uint32_t lds_budget, bool quotient, CeLayout &L)
{
    const uint32_t PC = shape.panel;
    L = CeLayout{};
    L.shape = shape;
    // 1. validate: reference opcodes 0..6, anything else is a no-op
    //    (constraint_eval.metal:38-73 has no default case)
    std::vector<SrcIns> src;
    int sp = 0, maxd = 0;
    std::vector<bool> emitted(n_out, false);
    for (uint32_t pc = 0; pc < n; pc++) {
        const uint32_t op = bc[pc] >> 24, arg = bc[pc] & 0xFFFFFFu;
        if (op > 6) continue;
        switch (op) {
        case 0:
            if (arg >= n_consts) return set_err(ZKG_EINVAL, "instruction %u: constant %u of %u", pc, arg, n_consts);
            if (arg >= CE_ARG_LIMIT) return set_err(ZKG_EINVAL, "instruction %u: constant index beyond 2^22", pc);
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
        if (m.col >= 0) uses.push_back((uint32_t)(m.col / PC));
    std::vector<size_t> next_use(uses.size());
    {
        std::vector<size_t> last((width + PC - 1) / PC, SIZE_MAX);
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
    const uint32_t n_panels = (width + PC - 1) / PC;
    const int kmax = (int)std::max<uint32_t>(1, std::min<uint32_t>(CE_MAX_SLOTS, n_panels));
    const uint32_t best = simulate(kmax, nullptr, nullptr);
    int K = 1;
    while (K < kmax && (uint64_t)simulate(K, nullptr, nullptr) * 4 > (uint64_t)best * 5) K++;
    // LDS per wave: K input panel slots, the stack below the top, one output panel
    const uint32_t stack_bytes = stack_in_lds ? (uint32_t)std::max(0, maxd - 1) * shape.stack_row() : 0;
    const uint32_t fixed = stack_bytes + (quotient ? 0 : shape.slot_bytes());
    while (K > 1 && block_waves * (K * shape.slot_bytes() + fixed) > lds_budget) K--;
    if (block_waves * ((uses.empty() ? 0 : shape.slot_bytes()) + fixed) > lds_budget)
        return ZKG_ENOMEM;  // the caller tries a smaller shape
    std::vector<int> slot_of_use(uses.size());
    std::vector<bool> stage_at(uses.size());
    L.stages = simulate(K, &slot_of_use, &stage_at);
    L.slots = uses.empty() ? 0 : K;

    // 4. emit device words
    L.stack_off = L.slots * shape.slot_bytes();
    L.stack_slots = stack_bytes / shape.stack_row();
    L.out_off = L.stack_off + stack_bytes;
    L.wave_bytes = (L.out_off + (quotient ? 0 : shape.slot_bytes()) + 15) & ~15u;
    std::vector<uint32_t> &words = L.words;
    size_t u = 0;
    uint32_t d = 0;  // stack depth before each instruction
    // Output panel sessions: EMITs land in the LDS output panel of constraint
    // panel `cur`; switching panels flushes it.  A panel's first flush also
    // writes zeros for its constraints that are never emitted.
    const uint32_t n_opanels = (n_out + PC - 1) / PC;
    std::vector<uint32_t> zero_mask(n_opanels, 0);
    for (uint32_t j = 0; j < n_out; j++)
        if (!emitted[j]) zero_mask[j / PC] |= 1u << (j % PC);
    std::vector<bool> flushed(n_opanels, false);
    int64_t cur = -1;
    uint32_t cur_mask = 0;
    auto flush = [&]() {
        uint32_t mask = cur_mask;
        if (!flushed[cur]) mask |= zero_mask[cur];
        flushed[cur] = true;
        words.push_back(ce_word(CE_OFLUSH, 0, (uint32_t)cur));
        words.push_back(mask);
        cur = -1;
        cur_mask = 0;
    };
    auto open = [&](uint32_t q) {
        if (!flushed[q] && zero_mask[q]) words.push_back(ce_word(CE_OCLEAR, 0, L.out_off));
        cur = q;
    };
    std::vector<size_t> last_emit(n_out, SIZE_MAX);
    for (size_t i = 0; i < mid.size(); i++)
        if (mid[i].op == CE_EMIT) last_emit[mid[i].arg] = i;
    for (size_t mi = 0; mi < mid.size(); mi++) {
        const MidIns &m = mid[mi];
        uint32_t op = m.op, arg = m.arg, slot = 0;
        if (m.col >= 0) {
            const uint32_t panel = (uint32_t)(m.col / PC), ps = (uint32_t)slot_of_use[u];
            if (stage_at[u]) words.push_back(ce_word(CE_STAGE, ps, panel));
            arg = ps * shape.slot_bytes() + (uint32_t)(m.col % PC) * shape.col_stride();
            u++;
        }
        switch (op) {
        case CE_PUSH_L: case CE_PUSH_K:
            if (d) slot = d - 1;
            else op = op == CE_PUSH_L ? CE_PUSH_L0 : CE_PUSH_K0;
            d++;
            break;
        case CE_ADD_S: case CE_SUB_S: case CE_MUL_S:
            slot = d - 2;
            d--;
            break;
        case CE_DUP:
            slot = d - 1;
            d++;
            break;
        case CE_EMIT: {
            if (quotient) {
                const bool acc = last_emit[arg] == mi;
                if (d >= 2) {
                    slot = d - 2;
                    op = acc ? CE_ACC : CE_DROP;
                } else {
                    op = acc ? CE_ACC0 : CE_DROP0;
                }
                d--;
                break;
            }
            const uint32_t q = arg / PC, c = arg % PC;
            if (cur != (int64_t)q) {
                if (cur >= 0) flush();
                open(q);
            }
            cur_mask |= 1u << c;
            arg = L.out_off + c * shape.col_stride();
            if (d >= 2) slot = d - 2;
            else op = CE_EMIT0;
            d--;
            break;
        }
        }
        words.push_back(ce_word(op, slot, arg));
    }
    if (cur >= 0) flush();
    for (uint32_t q = 0; q < n_opanels && !quotient; q++)
        if (!flushed[q]) {  // never emitted into: all zeros
            open(q);
            flush();
        }
    words.push_back(ce_word(CE_END, 0, 0));
    words.push_back(ce_word(CE_END, 0, 0));  // read ahead by the last instruction
    return ZKG_OK;
}

// Translate at the preferred shape; a program whose stack does not fit the
// LDS there falls back to one row per lane.
int ce_translate_fit(const uint32_t *bc, uint32_t n, uint32_t width, uint32_t n_consts, uint32_t n_out,
                     CeShape preferred, bool stack_in_lds, uint32_t block_waves, uint32_t lds_budget, bool quotient,
                     CeLayout &L)
{
    CeShape shape;
    if (int rc = ce_pick_shape(preferred, &shape)) return rc;
    int rc = ce_translate(bc, n, width, n_consts, n_out, shape, stack_in_lds, block_waves, lds_budget, quotient, L);
    if (rc == ZKG_ENOMEM && shape.rows_per_lane > 1)
        rc = ce_translate(bc, n, width, n_consts, n_out, CeShape{1, 16}, stack_in_lds, block_waves, lds_budget,
                          quotient, L);
    if (rc == ZKG_ENOMEM) rc = set_err(ZKG_EINVAL, "program needs more than %u bytes of LDS per block", lds_budget);
    return rc;
}

}  // namespace

struct zkg_constraint_program {
    zkg_ctx *ctx = nullptr;
    uint32_t width = 0, n_constraints = 0, lds_budget = 0;
    std::vector<uint32_t> bytecode;  // the reference bytecode (host copy)
    std::vector<uint64_t> consts;    // constant pool (host copy)
    CeLayout interp, interp_q;       // what the interpreter runs (evaluation, quotient)
    void *d_prog = nullptr, *d_prog_q = nullptr, *d_consts = nullptr;
    CeLayout jit, jit_q;             // what the compiled kernels were generated from
    hipModule_t mod = nullptr, mod_q = nullptr;  // compiled straight-line kernels, if any
    hipFunction_t fn = nullptr, fn_q = nullptr;
};

// ------------------------------------------------------------------ compiled programs
// A program evaluated many times (one AIR, many traces) can be compiled to a
// straight-line kernel: the same staging, output panels and field arithmetic
// as the interpreter, with the stack resolved at compile time into registers,
// constants as immediates and no per-instruction dispatch.  The source is the
// device program above, rendered as HIP and compiled for gfx950 with hiprtc.
// One wave per block.

static const char kJitPrelude[] =
#include "_gen/zkg_jit_prelude.inc"
    ;

namespace {

constexpr uint32_t CE_JIT_LDS = 64u * 1024;  // per one-wave block

// The kernel for layout L: zkg_ce_jit (constraint values through output
// panels) or, for a quotient layout, zkg_ce_jit_q (one quotient value per row).
std::string ce_jit_source(const CeLayout &L, const std::vector<uint64_t> &consts, bool quotient)
{
    const uint32_t R = L.shape.rows_per_lane, PC = L.shape.panel, COLS = L.shape.col_stride();
    std::string b;
    char line[320];
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wformat-security"  // every format below is a literal
    auto emit = [&](const char *fmt, auto... a) {
        snprintf(line, sizeof line, fmt, a...);
        b += line;
        b += '\n';
    };
#pragma clang diagnostic pop
    b += kJitPrelude;
    if (quotient)
        emit("extern \"C\" __global__ __launch_bounds__(64) void zkg_ce_jit_q(const uint64_t *__restrict__ trace, "
             "uint64_t n_rows, uint32_t width, const uint64_t *__restrict__ coef, const uint64_t *__restrict__ invz, "
             "uint32_t log_block, uint64_t *__restrict__ qout)");
    else
        emit("extern \"C\" __global__ __launch_bounds__(64) void zkg_ce_jit(const uint64_t *__restrict__ trace, "
             "uint64_t n_rows, uint32_t width, uint32_t n_out, uint64_t *__restrict__ out)");
    emit("{");
    emit("    constexpr int R = %u, PC = %u;", R, PC);
    emit("    extern __shared__ __align__(16) unsigned char ce_lds[];");
    emit("    const uint32_t lane = threadIdx.x;");
    emit("    const uint64_t row0 = (uint64_t)blockIdx.x * (64 * R);");
    emit("    if (row0 >= n_rows) return;");
    emit("    unsigned char *const mine = ce_lds + lane * 8;");
    emit("#define LD(off, j) (*(const uint64_t *)(mine + (off) + 512 * (j)))");
    emit("#define EACH for (int j = 0; j < R; j++)");
    if (quotient) emit("    uint64_t acc[R]; EACH acc[j] = 0;");
    std::vector<uint32_t> st;  // the stack, as variable numbers
    uint32_t nv = 0;
    static const char *BIN[3] = {"add", "sub", "mul"};
    for (size_t i = 0; i < L.words.size(); i++) {
        const uint32_t w = L.words[i], op = w >> 27, slot = (w >> 22) & 31, arg = w & (CE_ARG_LIMIT - 1);
        if (op == CE_END) break;
        switch (op) {
        case CE_PUSH_L: case CE_PUSH_L0:
            emit("    uint64_t v%u[R]; EACH v%u[j] = LD(%uu, j);", nv, nv, arg);
            st.push_back(nv++);
            break;
        case CE_PUSH_K: case CE_PUSH_K0:
            emit("    uint64_t v%u[R]; EACH v%u[j] = 0x%llxull;", nv, nv, (unsigned long long)consts[arg]);
            st.push_back(nv++);
            break;
        case CE_ADD_S: case CE_SUB_S: case CE_MUL_S: {
            const uint32_t y = st.back();
            st.pop_back();
            const uint32_t x = st.back();
            st.pop_back();
            emit("    uint64_t v%u[R]; EACH v%u[j] = zkg::Gl::%s(v%u[j], v%u[j]);", nv, nv, BIN[op - CE_ADD_S], x, y);
            st.push_back(nv++);
            break;
        }
        case CE_DUP: st.push_back(st.back()); break;
        case CE_EMIT: case CE_EMIT0:
            emit("    EACH *(uint64_t *)(mine + %uu + 512 * j) = v%u[j];", arg, st.back());
            st.pop_back();
            break;
        case CE_ACC: case CE_ACC0:
            emit("    { const uint64_t k = coef[%uu]; EACH acc[j] = zkg::Gl::add(acc[j], zkg::Gl::mul(v%u[j], k)); }",
                 arg, st.back());
            st.pop_back();
            break;
        case CE_DROP: case CE_DROP0: st.pop_back(); break;
        case CE_ADD_L: case CE_SUB_L: case CE_MUL_L:
            emit("    uint64_t v%u[R]; EACH v%u[j] = zkg::Gl::%s(v%u[j], LD(%uu, j));", nv, nv, BIN[op - CE_ADD_L],
                 st.back(), arg);
            st.back() = nv++;
            break;
        case CE_ADD_K: case CE_SUB_K: case CE_MUL_K:
            emit("    uint64_t v%u[R]; EACH v%u[j] = zkg::Gl::%s(v%u[j], 0x%llxull);", nv, nv, BIN[op - CE_ADD_K],
                 st.back(), (unsigned long long)consts[arg]);
            st.back() = nv++;
            break;
        case CE_DBL: case CE_ZERO_T: case CE_SQR:
            emit("    uint64_t v%u[R]; EACH v%u[j] = zkg::Gl::%s(v%u[j], v%u[j]);", nv, nv, BIN[op - CE_DBL], st.back(),
                 st.back());
            st.back() = nv++;
            break;
        case CE_STAGE:
            emit("    zkg::stage_panel<R, PC>(trace, n_rows, width, row0, %uu, lane, ce_lds + %uu);", arg,
                 slot * L.shape.slot_bytes());
            break;
        case CE_OCLEAR:
            emit("    for (int c = 0; c < PC; c++) EACH *(uint64_t *)(mine + %uu + c * %uu + 512 * j) = 0;", arg, COLS);
            break;
        case CE_OFLUSH:
            emit("    zkg::flush_panel<R, PC>(out, n_rows, n_out, row0, %uu, 0x%xu, lane, ce_lds + %uu);", arg,
                 L.words[i + 1], L.out_off);
            i++;
            break;
        }
    }
    if (quotient)
        emit("    EACH { const uint64_t r = min(row0 + lane + 64 * j, n_rows - 1); "
             "qout[r] = zkg::Gl::mul(acc[j], invz[r >> log_block]); }");
    emit("}");
    return b;
}

int ce_jit_compile(const std::string &src, std::vector<char> &code, std::string &log)
{
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "zkg_ce_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return set_err(ZKG_EDEVICE, "hiprtcCreateProgram failed");
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops", "-Wno-pragma-once-outside-header"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof *opts), opts);
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    log.assign(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    int rc = ZKG_OK;
    if (r != HIPRTC_SUCCESS) {
        rc = set_err(ZKG_EDEVICE, "constraint program compile failed: %.400s", log.c_str());
    } else {
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
    }
    hiprtcDestroyProgram(&prog);
    return rc;
}

int ce_check_args(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions, uint32_t n_constraints)
{
    if (!bytecode && n_instructions) return set_err(ZKG_EINVAL, "null argument");
    if (!width || width > (1u << 24)) return set_err(ZKG_EINVAL, "width %u outside 1..2^24", width);
    if (!n_instructions) return set_err(ZKG_EINVAL, "empty bytecode");
    if (!n_constraints || n_constraints > CE_ARG_LIMIT)
        return set_err(ZKG_EINVAL, "n_constraints %u outside 1..2^22", n_constraints);
    return ZKG_OK;
}

void ce_info(const CeLayout &L, uint32_t info[8])
{
    info[0] = (uint32_t)L.words.size();
    info[1] = L.slots;
    info[2] = L.stages;
    info[3] = L.stack_slots;
    info[4] = L.wave_bytes;
    info[5] = L.shape.panel;
    info[6] = L.shape.rows_per_lane;
    info[7] = L.shape.col_stride();
}

// The compiled kernel's layout for a program (host only).
int ce_jit_layout(const uint32_t *bc, uint32_t n, uint32_t width, uint32_t n_consts, uint32_t n_out,
                  uint32_t lds_budget, bool quotient, CeLayout &L)
{
    return ce_translate_fit(bc, n, width, n_consts, n_out, CE_COMPILED_SHAPE, false, 1,
                            std::min(CE_JIT_LDS, lds_budget), quotient, L);
}

// Compile one kernel of the program (evaluation or quotient) and load it.
int ce_compile_kind(zkg_constraint_program *p, bool quotient)
{
    hipFunction_t &fn = quotient ? p->fn_q : p->fn;
    if (fn) return ZKG_OK;
    CeLayout L;
    if (int rc = ce_jit_layout(p->bytecode.data(), (uint32_t)p->bytecode.size(), p->width, (uint32_t)p->consts.size(),
                               p->n_constraints, p->lds_budget, quotient, L))
        return rc;
    std::vector<char> code;
    std::string log;
    if (int rc = ce_jit_compile(ce_jit_source(L, p->consts, quotient), code, log)) return rc;
    hipError_t e = hipSetDevice(p->ctx->device);
    if (e != hipSuccess) return hip_err(e, "hipSetDevice");
    hipModule_t mod = nullptr;
    hipFunction_t f = nullptr;
    e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) return hip_err(e, "load compiled constraint program");
    e = hipModuleGetFunction(&f, mod, quotient ? "zkg_ce_jit_q" : "zkg_ce_jit");
    if (e != hipSuccess) {
        (void)hipModuleUnload(mod);
        return hip_err(e, "compiled constraint program entry");
    }
    (quotient ? p->jit_q : p->jit) = std::move(L);
    (quotient ? p->mod_q : p->mod) = mod;
    fn = f;
    return ZKG_OK;
}

}  // namespace

extern "C" int zkg_constraint_translate(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                                        uint32_t n_constants, uint32_t n_constraints, int quotient, uint32_t *words,
                                        uint32_t cap, uint32_t info[8])
{
    if (!info || (!words && cap)) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = ce_check_args(width, bytecode, n_instructions, n_constraints)) return rc;
    CeLayout L;
    if (int rc = ce_translate_fit(bytecode, n_instructions, width, n_constants, n_constraints, CE_INTERP_SHAPE, true,
                                  CE_WAVES, CE_LDS_BYTES, quotient != 0, L))
        return rc;
    ce_info(L, info);
    std::memcpy(words, L.words.data(), std::min<size_t>(cap, L.words.size()) * 4);
    if (cap < L.words.size()) return set_err(ZKG_EINVAL, "%zu words needed, room for %u", L.words.size(), cap);
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
    p->bytecode.assign(bytecode, bytecode + n_instructions);
    p->consts.assign(constants, constants + n_constants);
    int lds_max = 0;
    int rc = ZKG_OK;
    hipError_t e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    if (e != hipSuccess) rc = hip_err(e, "device LDS size");
    p->lds_budget = std::min<uint32_t>(CE_LDS_BYTES, (uint32_t)std::max(lds_max, 0));
    if (!rc)
        rc = ce_translate_fit(bytecode, n_instructions, width, n_constants, n_constraints, CE_INTERP_SHAPE, true,
                              CE_WAVES, p->lds_budget, false, p->interp);
    if (!rc)
        rc = ce_translate_fit(bytecode, n_instructions, width, n_constants, n_constraints, CE_INTERP_SHAPE, true,
                              CE_WAVES, p->lds_budget, true, p->interp_q);
    const std::vector<uint32_t> &words = p->interp.words, &words_q = p->interp_q.words;
    if (!rc) rc = dev_alloc(ctx, words.size() * 4, &p->d_prog);
    if (!rc) rc = dev_alloc(ctx, words_q.size() * 4, &p->d_prog_q);
    if (!rc) rc = dev_alloc(ctx, std::max<uint32_t>(n_constants, 1) * 8, &p->d_consts);
    if (!rc) e = hipMemcpyAsync(p->d_prog, words.data(), words.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess)
        e = hipMemcpyAsync(p->d_prog_q, words_q.data(), words_q.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess && n_constants)
        e = hipMemcpyAsync(p->d_consts, constants, (uint64_t)n_constants * 8, hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // host vectors may change after return
    if (!rc && e != hipSuccess) rc = hip_err(e, "constraint program upload");
    if (rc) {
        zkg_constraint_program_free(p);
        return rc;
    }
    *out = p;
    return ZKG_OK;
}

extern "C" int zkg_constraint_program_info(const zkg_constraint_program *p, uint32_t info[8])
{
    if (!p || !info) return set_err(ZKG_EINVAL, "null argument");
    ce_info(p->fn ? p->jit : p->interp, info);
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
    const uint64_t *tr = (const uint64_t *)trace->ptr;
    uint64_t *o = (uint64_t *)out->ptr;
    if (p->fn) {  // compiled: one wave per block
        const uint64_t waves = (n_rows + p->jit.shape.wave_rows() - 1) / p->jit.shape.wave_rows();
        if (waves > 0x7FFFFFFFull) return set_err(ZKG_EINVAL, "too many rows");
        uint32_t width = p->width, n_out = p->n_constraints;
        void *args[] = {&tr, &n_rows, &width, &n_out, &o};
        hipError_t e = hipModuleLaunchKernel(p->fn, (uint32_t)waves, 1, 1, 64, 1, 1, p->jit.wave_bytes,
                                             p->ctx->stream, args, nullptr);
        if (e != hipSuccess) return hip_err(e, "compiled constraint eval launch");
        return ZKG_OK;
    }
    const CeLayout &L = p->interp;
    const uint64_t waves = (n_rows + L.shape.wave_rows() - 1) / L.shape.wave_rows();
    const uint64_t blocks = (waves + CE_WAVES - 1) / CE_WAVES;
    if (blocks > 0x7FFFFFFFull) return set_err(ZKG_EINVAL, "too many rows");
    hipError_t e = ce_launch(L.shape, dim3((uint32_t)blocks), CE_WAVES * L.wave_bytes, p->ctx->stream, tr, n_rows,
                             p->width, (const uint32_t *)p->d_prog, (const uint64_t *)p->d_consts, p->n_constraints,
                             L.wave_bytes, L.stack_off, L.out_off, o, p->lds_budget, CeQuot{});
    if (e != hipSuccess) return hip_err(e, "constraint eval launch");
    return ZKG_OK;
}

extern "C" int zkg_constraint_program_quotient(zkg_constraint_program *p, const zkg_buf *lde, uint64_t height,
                                               uint32_t log_blowup, uint64_t shift, const uint64_t *coef,
                                               zkg_buf *q_out)
{
    if (!p || !lde || !coef || !q_out) return set_err(ZKG_EINVAL, "null argument");
    if (!height || (height & (height - 1)) || height > (1ull << 32))
        return set_err(ZKG_EINVAL, "height %llu is not a power of two <= 2^32", (unsigned long long)height);
    int log_h = 0;
    while ((1ull << log_h) < height) log_h++;
    if ((int)log_blowup > log_h) return set_err(ZKG_EINVAL, "log_blowup %u exceeds log height %d", log_blowup, log_h);
    if (span_bytes(height, p->width, 8) > lde->bytes) return set_err(ZKG_EINVAL, "LDE buffer too small");
    if (height * 8 > q_out->bytes) return set_err(ZKG_EINVAL, "quotient buffer too small");
    if (lde->ctx != p->ctx || q_out->ctx != p->ctx) return set_err(ZKG_EINVAL, "buffers of another context");
    if (shift % Gl::P == 0) return set_err(ZKG_EINVAL, "zero coset shift");
    for (uint32_t k = 0; k < p->n_constraints; k++)
        if (coef[k] >= Gl::P) return set_err(ZKG_EINVAL, "coefficient %u is not canonical", k);
    // Row i sits at x_i = shift w_H^rev(i); with N = H / B, x_i^N = shift^N w_B^t,
    // t the bit reversal (log_blowup bits) of the top log_blowup bits of i.  So
    // 1 / Z_H(x_i) = 1 / (x_i^N - 1) takes B values, one per block of H / B rows.
    const uint32_t B = 1u << log_blowup;
    const uint64_t n = height >> log_blowup;
    auto gpow = [](uint64_t a, uint64_t e) {
        uint64_t r = 1;
        for (; e; e >>= 1, a = Gl::mul(a, a))
            if (e & 1) r = Gl::mul(r, a);
        return r;
    };
    uint64_t w_b = 0x185629dcda58878cULL;  // p3 Goldilocks two_adic_generator(32), squared down to order B
    for (uint32_t i = log_blowup; i < 32; i++) w_b = Gl::mul(w_b, w_b);
    const uint64_t sn = gpow(shift % Gl::P, n);
    std::vector<uint64_t> host(p->n_constraints + B);
    std::copy(coef, coef + p->n_constraints, host.begin());
    for (uint32_t blk = 0; blk < B; blk++) {
        uint32_t t = 0;
        for (uint32_t b = 0; b < log_blowup; b++) t |= ((blk >> b) & 1u) << (log_blowup - 1 - b);
        const uint64_t v = Gl::sub(Gl::mul(sn, gpow(w_b, t)), 1);
        if (!v) return set_err(ZKG_EINVAL, "the coset meets the trace domain (shift^N is a root of unity of order B)");
        host[p->n_constraints + blk] = gpow(v, Gl::P - 2);
    }
    void *d = nullptr;
    if (int rc = dev_alloc(p->ctx, host.size() * 8, &d)) return rc;
    hipError_t e = hipMemcpyAsync(d, host.data(), host.size() * 8, hipMemcpyHostToDevice, p->ctx->stream);
    CeQuot q;
    q.coef = (const uint64_t *)d;
    q.invz = q.coef + p->n_constraints;
    q.log_block = (uint32_t)(log_h - (int)log_blowup);
    q.qout = (uint64_t *)q_out->ptr;
    const uint64_t *tr = (const uint64_t *)lde->ptr;
    if (e == hipSuccess && p->fn_q) {
        const uint64_t waves = (height + p->jit_q.shape.wave_rows() - 1) / p->jit_q.shape.wave_rows();
        uint32_t width = p->width, lb = q.log_block;
        const uint64_t *cf = q.coef, *iz = q.invz;
        uint64_t *qo = q.qout;
        void *args[] = {&tr, &height, &width, &cf, &iz, &lb, &qo};
        e = hipModuleLaunchKernel(p->fn_q, (uint32_t)waves, 1, 1, 64, 1, 1, p->jit_q.wave_bytes, p->ctx->stream, args,
                                  nullptr);
    } else if (e == hipSuccess) {
        const CeLayout &L = p->interp_q;
        const uint64_t waves = (height + L.shape.wave_rows() - 1) / L.shape.wave_rows();
        const uint64_t blocks = (waves + CE_WAVES - 1) / CE_WAVES;
        e = ce_launch(L.shape, dim3((uint32_t)blocks), CE_WAVES * L.wave_bytes, p->ctx->stream, tr, height, p->width,
                      (const uint32_t *)p->d_prog_q, (const uint64_t *)p->d_consts, p->n_constraints, L.wave_bytes,
                      L.stack_off, L.out_off, nullptr, p->lds_budget, q);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(p->ctx->stream);  // host coefficients, temporary buffer
    dev_free(p->ctx, d);
    if (e != hipSuccess) return hip_err(e, "constraint quotient");
    return ZKG_OK;
}

extern "C" int zkg_constraint_program_compile(zkg_constraint_program *p)
{
    if (!p) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = ce_compile_kind(p, false)) return rc;
    return ce_compile_kind(p, true);
}

extern "C" void zkg_constraint_program_free(zkg_constraint_program *p)
{
    if (!p) return;
    if (p->d_prog || p->d_prog_q || p->d_consts || p->mod || p->mod_q) (void)hipStreamSynchronize(p->ctx->stream);
    if (p->d_prog) dev_free(p->ctx, p->d_prog);
    if (p->d_prog_q) dev_free(p->ctx, p->d_prog_q);
    if (p->d_consts) dev_free(p->ctx, p->d_consts);
    if (p->mod) (void)hipModuleUnload(p->mod);
    if (p->mod_q) (void)hipModuleUnload(p->mod_q);
    delete p;
}

// ------------------------------------------------------------------ per-context program cache
// zkg_constraint_eval_gl takes the program with every call, as the reference's
// accelerator does; a prover passes the same AIR program for every proof.  The
// context keeps the last `max_programs` programs (least recently used out) and
// compiles one once it has been seen `compile_after` times, if it has at most
// CE_AUTO_COMPILE_MAX instructions (compile time grows with the program; a
// longer one can still be compiled explicitly).  A program that fails to
// compile keeps running interpreted.
constexpr size_t CE_AUTO_COMPILE_MAX = 16384;

struct zkg::CeCache {
    struct Entry {
        uint32_t width = 0, n_out = 0;
        zkg_constraint_program *prog = nullptr;
        uint32_t uses = 0;
        uint64_t last = 0;
        bool no_compile = false;
    };
    uint32_t max_programs = 16, compile_after = 2;
    uint64_t tick = 0;
    std::vector<Entry> entries;
};

void zkg::ce_cache_release(zkg_ctx *ctx)
{
    if (!ctx->ce_cache) return;
    for (auto &e : ctx->ce_cache->entries) zkg_constraint_program_free(e.prog);
    delete ctx->ce_cache;
    ctx->ce_cache = nullptr;
}

extern "C" int zkg_constraint_cache_config(zkg_ctx *ctx, uint32_t max_programs, uint32_t compile_after_uses)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    (void)hipStreamSynchronize(ctx->stream);
    ce_cache_release(ctx);
    ctx->ce_cache_set = true;
    if (max_programs) {
        ctx->ce_cache = new CeCache;
        ctx->ce_cache->max_programs = max_programs;
        ctx->ce_cache->compile_after = compile_after_uses;
    }
    return ZKG_OK;
}

extern "C" int zkg_constraint_cache_info(const zkg_ctx *ctx, uint32_t info[3])
{
    if (!ctx || !info) return set_err(ZKG_EINVAL, "null argument");
    info[0] = info[1] = info[2] = 0;
    if (const CeCache *c = ctx->ce_cache) {
        info[0] = (uint32_t)c->entries.size();
        for (const auto &e : c->entries) {
            info[1] += e.prog->fn ? 1 : 0;
            info[2] += e.uses;
        }
    }
    return ZKG_OK;
}

static zkg_constraint_program *ce_cache_find(CeCache &c, uint32_t width, const uint32_t *bc, uint32_t n,
                                             const uint64_t *k, uint32_t nk, uint32_t n_out)
{
    for (auto &e : c.entries) {
        const zkg_constraint_program *p = e.prog;
        if (e.width == width && e.n_out == n_out && p->bytecode.size() == n && p->consts.size() == nk &&
            std::equal(bc, bc + n, p->bytecode.begin()) && std::equal(k, k + nk, p->consts.begin())) {
            e.uses++;
            e.last = ++c.tick;
            if (!e.prog->fn && !e.no_compile && c.compile_after && e.uses >= c.compile_after &&
                (n > CE_AUTO_COMPILE_MAX || ce_compile_kind(e.prog, false) != ZKG_OK))
                e.no_compile = true;
            return e.prog;
        }
    }
    return nullptr;
}

extern "C" int zkg_constraint_eval_gl(zkg_ctx *ctx, const zkg_buf *trace, uint64_t n_rows, uint32_t width,
                                      const uint32_t *bytecode, uint32_t n_instructions, const uint64_t *constants,
                                      uint32_t n_constants, uint32_t n_constraints, zkg_buf *out)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    if (!ctx->ce_cache_set) {
        ctx->ce_cache = new CeCache;  // 16 programs, compiled on the second use
        ctx->ce_cache_set = true;
    }
    zkg_constraint_program *p = nullptr;
    bool owned = true;
    int rc = ZKG_OK;
    if (ctx->ce_cache && bytecode && (constants || !n_constants)) {
        p = ce_cache_find(*ctx->ce_cache, width, bytecode, n_instructions, constants, n_constants, n_constraints);
        owned = p == nullptr;
    }
    if (!p) rc = zkg_constraint_program_create(ctx, width, bytecode, n_instructions, constants, n_constants,
                                               n_constraints, &p);
    if (!rc && owned && ctx->ce_cache) {
        CeCache &c = *ctx->ce_cache;
        if (c.entries.size() >= c.max_programs) {
            auto lru = std::min_element(c.entries.begin(), c.entries.end(),
                                        [](const CeCache::Entry &a, const CeCache::Entry &b) { return a.last < b.last; });
            zkg_constraint_program_free(lru->prog);
            c.entries.erase(lru);
        }
        CeCache::Entry e;
        e.width = width;
        e.n_out = n_constraints;
        e.prog = p;
        e.uses = 1;
        e.last = ++c.tick;
        c.entries.push_back(e);
        owned = false;
    }
    if (!rc) rc = zkg_constraint_program_eval(p, trace, n_rows, out);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_err(e, "constraint eval");
    }
    if (owned) zkg_constraint_program_free(p);
    return rc;
}

// Host-only views of the compiled path: its source, and a compile without a device.
static int ce_host_jit(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions, const uint64_t *constants,
                       uint32_t n_constants, uint32_t n_constraints, bool quotient, std::string &src)
{
    if (!constants && n_constants) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = ce_check_args(width, bytecode, n_instructions, n_constraints)) return rc;
    CeLayout L;
    if (int rc = ce_jit_layout(bytecode, n_instructions, width, n_constants, n_constraints, CE_LDS_BYTES, quotient, L))
        return rc;
    src = ce_jit_source(L, std::vector<uint64_t>(constants, constants + n_constants), quotient);
    return ZKG_OK;
}

extern "C" int zkg_constraint_jit_source(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                                         const uint64_t *constants, uint32_t n_constants, uint32_t n_constraints,
                                         char *src, uint64_t cap, uint64_t *size)
{
    if (!size || (!src && cap)) return set_err(ZKG_EINVAL, "null argument");
    std::string text;
    if (int rc = ce_host_jit(width, bytecode, n_instructions, constants, n_constants, n_constraints, false, text))
        return rc;
    *size = text.size() + 1;
    if (cap) snprintf(src, cap, "%s", text.c_str());
    if (cap && cap < text.size() + 1) return set_err(ZKG_EINVAL, "source needs %zu bytes", text.size() + 1);
    return ZKG_OK;
}

extern "C" int zkg_constraint_compile_check(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                                            const uint64_t *constants, uint32_t n_constants, uint32_t n_constraints,
                                            char *log, uint32_t log_cap)
{
    if (!log && log_cap) return set_err(ZKG_EINVAL, "null argument");
    std::string src, text;
    int rc = ZKG_OK;
    std::vector<char> code;
    for (int quotient = 0; quotient < 2 && !rc; quotient++) {
        rc = ce_host_jit(width, bytecode, n_instructions, constants, n_constants, n_constraints, quotient != 0, src);
        if (!rc) rc = ce_jit_compile(src, code, text);
    }
    if (log_cap) snprintf(log, log_cap, "%s", text.c_str());
    return rc;
}
