"""CPU checks of libzkg's constraint-program translation (no device needed).

zkg_constraint_translate turns the reference's stack bytecode
(zkf-metal/src/shaders/constraint_eval.metal:1-74) into the device program the
HIP interpreter runs: fused "push x; op" instructions, LDS panel staging
planned with Belady's rule, zero-writes for never-emitted constraints.  An
emulator of the device program below runs the translated words wave by wave
(64 rows, LDS slots as 16-column panels) and must agree with the oracle on
every row, so a wrong fusion, slot plan or panel offset fails here.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_constraint import random_program

import ziros_amd as Z

P = O.GL_P
PANEL, ROWS, COL_STRIDE = 16, 64, 64 * 8 + 8
SLOT_BYTES = PANEL * COL_STRIDE
(PUSH_L, PUSH_K, ADD_S, SUB_S, MUL_S, DUP, EMIT, ADD_L, SUB_L, MUL_L, ADD_K, SUB_K, MUL_K, DBL, ZERO_T, SQR, STAGE,
 ZERO, END) = range(19)


def _bin(k, a, b):
    return (a + b) % P if k == 0 else (a - b) % P if k == 1 else a * b % P


def emulate(words, info, trace, consts, n_c):
    """Run the device program of the translation on the CPU."""
    assert len(words) % 4 == 0 and words[-1] >> 24 == END
    rows, width = trace.shape
    out = np.full((rows, n_c), -1, dtype=object)  # -1: never written
    for r0 in range(0, rows, ROWS):
        rr = range(r0, min(r0 + ROWS, rows))
        slots, stack, top, d = {}, [], None, 0

        def lds(arg):
            assert arg % COL_STRIDE == 0 and arg < info["slots"] * SLOT_BYTES
            return slots[arg // SLOT_BYTES][(arg % SLOT_BYTES) // COL_STRIDE]

        for w in words:
            op, arg = w >> 24, w & 0xFFFFFF
            if op == END:
                break
            if op in (PUSH_L, PUSH_K, DUP):
                x = lds(arg) if op == PUSH_L else [int(consts[arg])] * len(rr) if op == PUSH_K else top
                if d:
                    stack.append(top)
                    assert len(stack) <= info["stack_slots"]
                d += 1
                top = x
            elif op in (ADD_S, SUB_S, MUL_S):
                a = stack.pop()
                d -= 1
                top = [_bin(op - ADD_S, x, y) for x, y in zip(a, top)]
            elif op in (ADD_L, SUB_L, MUL_L):
                top = [_bin(op - ADD_L, x, y) for x, y in zip(top, lds(arg))]
            elif op in (ADD_K, SUB_K, MUL_K):
                top = [_bin(op - ADD_K, x, int(consts[arg])) for x in top]
            elif op in (DBL, ZERO_T, SQR):
                top = [_bin(op - DBL, x, x) for x in top]
            elif op == EMIT:
                out[r0:r0 + len(rr), arg] = top
                d -= 1
                top = stack.pop() if d else None
            elif op == ZERO:
                out[r0:r0 + len(rr), arg] = 0
            elif op == STAGE:
                slot, panel = arg >> 20, arg & 0xFFFFF
                assert slot < info["slots"]
                slots[slot] = [[int(trace[r, min(panel * PANEL + c, width - 1)]) for r in rr] for c in range(PANEL)]
            else:
                raise AssertionError("unknown device op %d" % op)
    return out


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


@pytest.mark.parametrize("seed", range(12))
def test_random_programs(seed):
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


def test_dup_fusions_and_deep_stack():
    rng = np.random.default_rng(7)
    trace = _trace(rng, 70, 5)
    I = O.instr
    bc = [I(O.OP_LOAD, 2), I(O.OP_DUP), I(O.OP_MUL), I(O.OP_EMIT, 0),     # x^2
          I(O.OP_LOAD, 3), I(O.OP_DUP), I(O.OP_ADD), I(O.OP_EMIT, 1),     # 2x
          I(O.OP_LOAD, 4), I(O.OP_DUP), I(O.OP_SUB), I(O.OP_EMIT, 2)]     # 0
    bc += [I(O.OP_LOAD, i % 5) for i in range(32)] + [I(O.OP_MUL)] * 31 + [I(O.OP_EMIT, 3)]
    info = _check(trace, bc, [], 5)   # constraint 4 never emitted: zero
    assert info["stack_slots"] == 31


def test_sequential_columns_stage_each_panel_once():
    rng = np.random.default_rng(9)
    w = 80
    trace = _trace(rng, 64, w)
    I = O.instr
    bc = []
    for c in range(w - 1):
        bc += [I(O.OP_LOAD, c), I(O.OP_LOAD, c + 1), I(O.OP_MUL), I(O.OP_EMIT, c)]
    info = _check(trace, bc, [], w - 1)
    assert info["slots"] == 1 and info["stages"] == 5


def test_interleaved_panels_take_more_slots():
    rng = np.random.default_rng(11)
    w = 64
    trace = _trace(rng, 64, w)
    I = O.instr
    bc = []
    for c in range(16):   # every constraint reads panels 0 and 3
        bc += [I(O.OP_LOAD, c), I(O.OP_LOAD, 48 + c), I(O.OP_SUB), I(O.OP_EMIT, c)]
    info = _check(trace, bc, [], 16)
    assert info["slots"] == 2 and info["stages"] == 2


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
