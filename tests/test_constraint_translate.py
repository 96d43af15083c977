"""CPU checks of libzkg's constraint-program translation (no device needed).

zkg_constraint_translate turns the reference's stack bytecode
(zkf-metal/src/shaders/constraint_eval.metal:1-74) into the device program the
HIP interpreter runs: fused "push x; op" instructions, LDS panel staging
planned with Belady's rule, zero-writes for never-emitted constraints.  An
emulator of the device program below runs the translated words wave by wave
(64 R rows with R rows per lane, LDS slots as column panels) and must agree
with the oracle on every row, so a wrong fusion, slot plan or panel offset
fails here.  Every work shape the library is built with is covered.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_constraint import random_program

import ziros_amd as Z

P = O.GL_P
SHAPES = ["1,16", "2,4", "2,8", "2,16", "4,4", "4,8"]
(PUSH_L, PUSH_L0, PUSH_K, PUSH_K0, ADD_S, SUB_S, MUL_S, DUP, EMIT, EMIT0, ADD_L, SUB_L, MUL_L, ADD_K, SUB_K, MUL_K,
 DBL, ZERO_T, SQR, STAGE, OCLEAR, OFLUSH, ACC, ACC0, DROP, DROP0, END) = range(27)


def _bin(k, a, b):
    return (a + b) % P if k == 0 else (a - b) % P if k == 1 else a * b % P


def emulate(words, info, trace, consts, n_c, coef=None):
    """Run the device program of the translation on the CPU: word =
    op << 27 | slot << 22 | arg, the stack below the top in LDS entries
    addressed by `slot`, trace panels in LDS slots, EMITs into an LDS output
    panel that OFLUSH (followed by a column-mask word) writes out.  With coef
    (a quotient program) returns the per-row accumulation sum coef[k] C_k."""
    assert words[-1] >> 27 == END
    PANEL, COL_STRIDE, WROWS = info["panel"], info["col_stride"], 64 * info["rows_per_lane"]
    assert COL_STRIDE == WROWS * 8 + 8
    SLOT_BYTES = PANEL * COL_STRIDE
    OUT_OFF = info["slots"] * SLOT_BYTES + info["stack_slots"] * WROWS * 8
    assert info["wave_bytes"] == (OUT_OFF + (0 if coef is not None else SLOT_BYTES) + 15) & ~15
    rows, width = trace.shape
    out = np.full((rows, n_c), -1, dtype=object)  # -1: never written
    accs = np.zeros(rows, dtype=object)
    for r0 in range(0, rows, WROWS):
        rr = range(r0, min(r0 + WROWS, rows))
        panels, stack, top = {}, {}, None
        acc = [0] * len(rr)
        opanel, fresh = {}, set()  # output panel columns; those written since the last flush

        def lds(arg):
            assert arg % COL_STRIDE == 0 and arg < info["slots"] * SLOT_BYTES
            return panels[arg // SLOT_BYTES][(arg % SLOT_BYTES) // COL_STRIDE]

        def spill(slot, v):
            assert slot < info["stack_slots"]
            stack[slot] = v

        it = iter(words)
        for w in it:
            op, slot, arg = w >> 27, (w >> 22) & 31, w & ((1 << 22) - 1)
            if op == END:
                break
            if op in (PUSH_L, PUSH_K):
                spill(slot, top)
            if op in (PUSH_L, PUSH_L0):
                top = lds(arg)
            elif op in (PUSH_K, PUSH_K0):
                top = [int(consts[arg])] * len(rr)
            elif op in (ADD_S, SUB_S, MUL_S):
                top = [_bin(op - ADD_S, x, y) for x, y in zip(stack.pop(slot), top)]
            elif op == DUP:
                spill(slot, top)
            elif op in (ADD_L, SUB_L, MUL_L):
                top = [_bin(op - ADD_L, x, y) for x, y in zip(top, lds(arg))]
            elif op in (ADD_K, SUB_K, MUL_K):
                top = [_bin(op - ADD_K, x, int(consts[arg])) for x in top]
            elif op in (DBL, ZERO_T, SQR):
                top = [_bin(op - DBL, x, x) for x in top]
            elif op in (EMIT, EMIT0):
                assert (arg - OUT_OFF) % COL_STRIDE == 0 and 0 <= arg - OUT_OFF < SLOT_BYTES
                c = (arg - OUT_OFF) // COL_STRIDE
                opanel[c] = top
                fresh.add(c)
                top = stack.pop(slot) if op == EMIT else None
            elif op in (ACC, ACC0, DROP, DROP0):
                assert coef is not None
                if op in (ACC, ACC0):
                    acc = [(a + int(coef[arg]) * t) % P for a, t in zip(acc, top)]
                top = stack.pop(slot) if op in (ACC, DROP) else None
            elif op == OCLEAR:
                assert arg == OUT_OFF
                for c in range(PANEL):
                    opanel[c] = [0] * len(rr)
                fresh = set(range(PANEL))
            elif op == OFLUSH:
                mask = next(it)
                for c in range(PANEL):
                    if mask >> c & 1:
                        assert c in fresh and arg * PANEL + c < n_c
                        out[r0:r0 + len(rr), arg * PANEL + c] = opanel[c]
                fresh = set()
            elif op == STAGE:
                assert slot < info["slots"]
                panels[slot] = [[int(trace[r, min(arg * PANEL + c, width - 1)]) for r in rr] for c in range(PANEL)]
            else:
                raise AssertionError("unknown device op %d" % op)
        assert not stack
        accs[r0:r0 + len(rr)] = acc
    return accs if coef is not None else out


def _trace(rng, rows, width):
    t = (rng.integers(0, 2**63, (rows, width), dtype=np.uint64) % np.uint64(P)).astype(np.uint64)
    t[0, :] = P - 1
    return t


def _check(trace, bc, consts, n_c):
    words, info = Z.translate_constraints(trace.shape[1], bc, len(consts), n_c)
    got = emulate([int(w) for w in words], info, trace, consts, n_c)
    exp = O.constraint_eval_gl(trace, bc, consts, n_c)
    assert (got == exp.astype(object)).all()
    return info


def test_reference_add_mul_program():
    n, w = 130, 2
    trace = np.array([(i * 7 + 1) % P for i in range(n * w)], dtype=np.uint64).reshape(n, w)
    cc = Z.ConstraintCompiler()
    cc.load_column(0); cc.load_column(1); cc.add(); cc.emit(0)
    cc.load_column(0); cc.load_column(1); cc.mul(); cc.emit(1)
    bc, consts, n_c = cc.finish()
    assert n_c == 2
    info = _check(trace, bc, consts, n_c)
    assert info["slots"] == 1 and info["stages"] == 1


@pytest.fixture(params=SHAPES)
def shape(request, monkeypatch):
    monkeypatch.setenv("ZKG_CE_SHAPE", request.param)
    return tuple(int(v) for v in request.param.split(","))


def test_shape_reported(shape):
    _, info = Z.translate_constraints(4, [O.instr(O.OP_LOAD, 3), O.instr(O.OP_EMIT, 0)], 0, 1)
    assert (info["rows_per_lane"], info["panel"]) == shape


def test_unknown_shape_rejected(monkeypatch):
    monkeypatch.setenv("ZKG_CE_SHAPE", "3,8")
    with pytest.raises(Z.ZkgError):
        Z.translate_constraints(4, [O.instr(O.OP_LOAD, 3), O.instr(O.OP_EMIT, 0)], 0, 1)


@pytest.mark.parametrize("seed", range(12))
def test_random_programs(seed, shape):
    rng = np.random.default_rng(100 + seed)
    w = int(rng.choice([1, 3, 16, 17, 40, 70]))
    n_const, n_c = int(rng.integers(1, 6)), int(rng.integers(1, 7))
    rows = int(rng.choice([1, 63, 64, 65, 150]))
    trace = _trace(rng, rows, w)
    consts = [int(v) % P for v in rng.integers(0, 2**63, n_const, dtype=np.uint64)]
    bc = random_program(rng, w, n_const, n_c, int(rng.integers(3, 120)))
    if seed % 3 == 0:  # unknown opcodes are skipped
        bc = [x for i in bc for x in ((i,) if rng.random() < 0.8 else (i, (int(rng.integers(7, 256)) << 24) | 5))]
    _check(trace, bc, consts, n_c)


def test_dup_fusions_and_deep_stack(shape):
    rng = np.random.default_rng(7)
    trace = _trace(rng, 70, 5)
    I = O.instr
    bc = [I(O.OP_LOAD, 2), I(O.OP_DUP), I(O.OP_MUL), I(O.OP_EMIT, 0),     # x^2
          I(O.OP_LOAD, 3), I(O.OP_DUP), I(O.OP_ADD), I(O.OP_EMIT, 1),     # 2x
          I(O.OP_LOAD, 4), I(O.OP_DUP), I(O.OP_SUB), I(O.OP_EMIT, 2)]     # 0
    bc += [I(O.OP_LOAD, i % 5) for i in range(32)] + [I(O.OP_MUL)] * 31 + [I(O.OP_EMIT, 3)]
    info = _check(trace, bc, [], 5)   # constraint 4 never emitted: zero
    assert info["stack_slots"] == 31
    if shape[0] == 4:  # a 32-deep stack at 4 rows per lane exceeds the LDS: one row per lane
        assert (info["rows_per_lane"], info["panel"]) == (1, 16)


def test_sequential_columns_stage_each_panel_once(shape):
    rng = np.random.default_rng(9)
    w = 80
    trace = _trace(rng, 64, w)
    I = O.instr
    bc = []
    for c in range(w - 1):
        bc += [I(O.OP_LOAD, c), I(O.OP_LOAD, c + 1), I(O.OP_MUL), I(O.OP_EMIT, c)]
    info = _check(trace, bc, [], w - 1)
    assert info["slots"] == 1 and info["stages"] == -(-w // info["panel"])


def test_interleaved_panels_take_more_slots(shape):
    rng = np.random.default_rng(11)
    w = 64
    trace = _trace(rng, 64, w)
    I = O.instr
    bc = []
    for c in range(16):   # every constraint reads columns 0..15 and 48..63 alternately
        bc += [I(O.OP_LOAD, c), I(O.OP_LOAD, 48 + c), I(O.OP_SUB), I(O.OP_EMIT, c)]
    info = _check(trace, bc, [], 16)
    slot_bytes = info["panel"] * info["col_stride"]
    stack_bytes = info["stack_slots"] * 64 * info["rows_per_lane"] * 8
    if 4 * (3 * slot_bytes + stack_bytes) <= 160 * 1024:   # two input slots and the output panel fit
        assert info["slots"] == 2 and info["stages"] == 2 * (16 // info["panel"])
    else:
        assert info["slots"] == 1 and info["stages"] == 32


@pytest.mark.parametrize("bc,width,n_const,n_c", [
    ([], 2, 0, 1),                                           # empty program
    ([O.instr(O.OP_ADD)], 2, 0, 1),                          # underflow
    ([O.instr(O.OP_LOAD, 0), O.instr(O.OP_EMIT, 1)], 2, 0, 1),  # constraint index
    ([O.instr(O.OP_LOAD, 2)], 2, 0, 1),                      # column index
    ([O.instr(O.OP_CONST, 0)], 2, 0, 1),                     # constant index
    ([O.instr(O.OP_LOAD, 0)] * 33, 2, 0, 1),                 # deeper than 32
    ([O.instr(O.OP_LOAD, 0)], 0, 0, 1),                      # width 0
    ([O.instr(O.OP_LOAD, 0)], 2, 0, 0),                      # no constraints
])
def test_invalid_programs_rejected(bc, width, n_const, n_c):
    with pytest.raises(Z.ZkgError):
        Z.translate_constraints(width, bc, n_const, n_c)


def test_compiler_mirror_encoding():
    cc = Z.ConstraintCompiler()
    cc.push_const(5); cc.load_column(0x1234567); cc.dup(); cc.sub(); cc.emit(3)
    bc, consts, n_c = cc.finish()
    assert consts == [5] and n_c == 4
    assert bc == [0 << 24 | 0, 1 << 24 | 0x234567, 5 << 24, 3 << 24, 6 << 24 | 3]


def test_words_carry_stack_slots():
    I = O.instr
    bc = [I(O.OP_LOAD, 0), I(O.OP_LOAD, 1), I(O.OP_LOAD, 2), I(O.OP_MUL), I(O.OP_ADD), I(O.OP_EMIT, 0)]
    words, _ = Z.translate_constraints(3, bc, 0, 1)
    ops = [(int(w) >> 27, (int(w) >> 22) & 31) for w in words]
    # stage, push onto empty, push (spill to 0), mul fused with its column, emit (stack empties),
    # flush of output panel 0 (mask word 1), end x2
    assert ops[:6] == [(STAGE, 0), (PUSH_L0, 0), (PUSH_L, 0), (MUL_L, 0), (ADD_S, 0), (EMIT0, 0)]
    assert ops[6] == (OFLUSH, 0) and int(words[7]) == 1 and ops[8:] == [(END, 0), (END, 0)]


def test_output_panels_flush_in_sessions(shape):
    """Emits that revisit an output panel, constraints never emitted, and a
    panel with no emits at all."""
    rng = np.random.default_rng(13)
    trace = _trace(rng, 100, 6)
    I = O.instr
    order = [0, 1, 20, 2, 21, 0, 40, 3, 1]   # 10..19, 22..39 and 41..49 never emitted
    bc = []
    for k, j in enumerate(order):
        bc += [I(O.OP_LOAD, k % 6), I(O.OP_CONST, 0), I(O.OP_MUL), I(O.OP_EMIT, j)]
    info = _check(trace, bc, [5], 50)
    assert info["words"] > 0



@pytest.mark.parametrize("seed", range(8))
def test_quotient_programs(seed, shape):
    """Quotient-mode translation: the last EMIT of each constraint accumulates
    coef[k] * value (earlier ones only pop), no output panels."""
    rng = np.random.default_rng(300 + seed)
    w = int(rng.choice([2, 17, 40]))
    n_const, n_c = int(rng.integers(1, 5)), int(rng.integers(1, 6))
    trace = _trace(rng, int(rng.choice([1, 64, 150])), w)
    consts = [int(v) % P for v in rng.integers(0, 2**63, n_const, dtype=np.uint64)]
    bc = random_program(rng, w, n_const, n_c, int(rng.integers(3, 100)))
    if seed % 2:  # emit some constraints twice: the last value counts
        bc = bc + [O.instr(O.OP_LOAD, 0), O.instr(O.OP_EMIT, 0)]
    coef = [int(v) % P for v in rng.integers(0, 2**63, n_c, dtype=np.uint64)]
    words, info = Z.translate_constraints(w, bc, n_const, n_c, quotient=True)
    assert not any((int(x) >> 27) in (EMIT, EMIT0, OCLEAR, OFLUSH) for x in words)
    got = emulate([int(x) for x in words], info, trace, consts, n_c, coef)
    c = O.constraint_eval_gl(trace, bc, consts, n_c).astype(object)
    exp = [sum(int(k) * int(v) for k, v in zip(coef, row)) % P for row in c]
    assert list(got) == exp
