"""Pins for the oracle's constraint-evaluation stack machine (Goldilocks).

The semantics are those of zkf-metal/src/shaders/constraint_eval.metal:1-74.
Pinned by the reference's own test (zkf-metal/src/constraint_eval.rs:198-245:
2048 rows of (7 i + 1) mod p, c0 = col0 + col1, c1 = col0 * col1) and by
random programs checked against an independent Python big-integer evaluator.
"""
import numpy as np

import oracle as O

P = O.GL_P


def py_eval(trace, bc, consts, n_c):
    out = np.zeros((trace.shape[0], n_c), dtype=object)
    for r in range(trace.shape[0]):
        st = []
        for w in bc:
            op, arg = w >> 24, w & 0xFFFFFF
            if op == 0:
                st.append(int(consts[arg]))
            elif op == 1:
                st.append(int(trace[r, arg]))
            elif op in (2, 3, 4):
                b, a = st.pop(), st.pop()
                st.append((a + b) % P if op == 2 else (a - b) % P if op == 3 else a * b % P)
            elif op == 5:
                st.append(st[-1])
            elif op == 6:
                out[r, arg] = st.pop()
    return out


def random_program(rng, width, n_const, n_c, length):
    bc, sp = [], 0
    emitted = 0
    while len(bc) < length or sp:
        choices = []
        if sp < 30 and len(bc) < length:
            choices += ["load", "const", "dup"] if sp else ["load", "const"]
        if sp >= 2:
            choices += ["add", "sub", "mul"]
        if sp >= 1:
            choices += ["emit"]
        c = choices[rng.integers(len(choices))]
        if c == "load":
            bc.append(O.instr(O.OP_LOAD, int(rng.integers(width)))); sp += 1
        elif c == "const":
            bc.append(O.instr(O.OP_CONST, int(rng.integers(n_const)))); sp += 1
        elif c == "dup":
            bc.append(O.instr(O.OP_DUP)); sp += 1
        elif c in ("add", "sub", "mul"):
            bc.append(O.instr({"add": O.OP_ADD, "sub": O.OP_SUB, "mul": O.OP_MUL}[c])); sp -= 1
        else:
            bc.append(O.instr(O.OP_EMIT, emitted % n_c)); emitted += 1; sp -= 1
    return bc


def test_reference_add_mul_vector():
    n, w = 2048, 2
    trace = np.array([(i * 7 + 1) % P for i in range(n * w)], dtype=np.uint64).reshape(n, w)
    bc = [O.instr(O.OP_LOAD, 0), O.instr(O.OP_LOAD, 1), O.instr(O.OP_ADD), O.instr(O.OP_EMIT, 0),
          O.instr(O.OP_LOAD, 0), O.instr(O.OP_LOAD, 1), O.instr(O.OP_MUL), O.instr(O.OP_EMIT, 1)]
    out = O.constraint_eval_gl(trace, bc, [], 2)
    for r in range(n):
        a, b = int(trace[r, 0]), int(trace[r, 1])
        assert int(out[r, 0]) == (a + b) % P and int(out[r, 1]) == a * b % P


def test_random_programs_vs_python():
    rng = np.random.default_rng(5)
    for k in range(20):
        w, n_const, n_c = int(rng.integers(1, 9)), int(rng.integers(1, 6)), int(rng.integers(1, 5))
        trace = (rng.integers(0, 2**63, (17, w), dtype=np.uint64) % np.uint64(P)).astype(np.uint64)
        trace[0, :] = P - 1
        consts = [int(v) % P for v in rng.integers(0, 2**63, n_const, dtype=np.uint64)]
        bc = random_program(rng, w, n_const, n_c, int(rng.integers(3, 60)))
        out = O.constraint_eval_gl(trace, bc, consts, n_c)
        exp = py_eval(trace, bc, consts, n_c)
        assert (out.astype(object) == exp).all(), k


def test_unknown_opcodes_are_no_ops_and_bad_programs_rejected():
    trace = np.arange(8, dtype=np.uint64).reshape(4, 2)
    bc = [O.instr(O.OP_LOAD, 1), (9 << 24) | 3, O.instr(O.OP_EMIT, 0)]
    assert (O.constraint_eval_gl(trace, bc, [], 1)[:, 0] == trace[:, 1]).all()
    assert O.constraint_eval_gl(trace, [O.instr(O.OP_ADD)], [], 1) is None            # underflow
    assert O.constraint_eval_gl(trace, [O.instr(O.OP_LOAD, 2)], [], 1) is None        # column
    assert O.constraint_eval_gl(trace, [O.instr(O.OP_CONST, 0)], [], 1) is None       # constant
    assert O.constraint_eval_gl(trace, [O.instr(O.OP_LOAD, 0)] * 33, [], 1) is None   # depth 33
