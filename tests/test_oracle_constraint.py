"""Pins for the oracle's constraint-evaluation stack machine (Goldilocks).

The semantics are those of zkf-metal/src/shaders/constraint_eval.metal:1-74.
Pinned by the reference's own test (zkf-metal/src/constraint_eval.rs:198-245:
2048 rows of (7 i + 1) mod p, c0 = col0 + col1, c1 = col0 * col1) and by
random programs checked against an independent Python big-integer evaluator.

The quotient over a committed LDE (constraint values combined with caller
coefficients and divided by the trace domain's vanishing polynomial) is pinned
by what the mathematics fixes: for an AIR that holds on the trace domain the
quotient values interpolate a polynomial of degree N - 2, and they do not when
one trace row breaks the AIR, the coset shift is wrong or the rows are in
natural instead of bit-reversed order; and row by row against the formula
with points computed from the Goldilocks two-adic generator.
"""
import numpy as np
import pytest

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


# ---------------------------------------------------------------- quotient over the LDE
def _rev(i, bits):
    return int(format(i, "0%db" % bits)[::-1], 2) if bits else 0


def _square_air(rng, log_n, broken_row=None):
    """Columns a, b = a^2, c = a^3 on H (|H| = 2^log_n); constraints
    a*a - b and b*a - c vanish on H."""
    n = 1 << log_n
    a = (rng.integers(0, 2**63, n, dtype=np.uint64) % np.uint64(P)).astype(object)
    b = np.array([int(x) * int(x) % P for x in a], dtype=object)
    c = np.array([int(x) * int(y) % P for x, y in zip(a, b)], dtype=object)
    if broken_row is not None:
        b[broken_row] = (b[broken_row] + 1) % P
    trace = np.stack([a, b, c], axis=1).astype(np.uint64)
    bc = [O.instr(O.OP_LOAD, 0), O.instr(O.OP_DUP), O.instr(O.OP_MUL), O.instr(O.OP_LOAD, 1), O.instr(O.OP_SUB),
          O.instr(O.OP_EMIT, 0),
          O.instr(O.OP_LOAD, 1), O.instr(O.OP_LOAD, 0), O.instr(O.OP_MUL), O.instr(O.OP_LOAD, 2), O.instr(O.OP_SUB),
          O.instr(O.OP_EMIT, 1)]
    return trace, bc


def _degree_bound_holds(q, log_h, bound):
    """q: values in bit-reversed coset order; True if the interpolating
    polynomial has degree < bound (coefficients from index bound on vanish)."""
    h = 1 << log_h
    nat = np.array([q[_rev(j, log_h)] for j in range(h)], dtype=np.uint64).reshape(h, 1)
    coeffs = O.dft_batch(O.GL, nat, inverse=True)[:, 0]
    return not coeffs[bound:].any()


@pytest.mark.parametrize("log_blowup", [1, 2])
def test_quotient_is_low_degree_for_a_satisfied_air(log_blowup):
    rng = np.random.default_rng(3 + log_blowup)
    log_n = 6
    trace, bc = _square_air(rng, log_n)
    lde = O.coset_lde_batch(O.GL, trace, log_blowup, 7)
    coef = [int(v) % P for v in rng.integers(1, 2**63, 2, dtype=np.uint64)]
    q = O.constraint_quotient_gl(lde, log_blowup, 7, bc, [], 2, coef)
    # deg(a^2), deg(a b) = 2N - 2, minus deg Z_H = N: the quotient has degree N - 2
    assert _degree_bound_holds(q, log_n + log_blowup, (1 << log_n) - 1)
    assert not _degree_bound_holds(q, log_n + log_blowup, (1 << log_n) - 2)


def test_quotient_is_not_low_degree_when_the_air_fails_or_points_are_wrong():
    rng = np.random.default_rng(11)
    log_n, lb = 6, 1
    bound = (1 << log_n) - 1
    coef = [5, 9]
    trace, bc = _square_air(rng, log_n, broken_row=17)
    lde = O.coset_lde_batch(O.GL, trace, lb, 7)
    assert not _degree_bound_holds(O.constraint_quotient_gl(lde, lb, 7, bc, [], 2, coef), log_n + lb, bound)
    trace, bc = _square_air(rng, log_n)
    lde = O.coset_lde_batch(O.GL, trace, lb, 7)
    assert _degree_bound_holds(O.constraint_quotient_gl(lde, lb, 7, bc, [], 2, coef), log_n + lb, bound)
    # the wrong coset shift, or the rows taken in natural order, break it
    assert not _degree_bound_holds(O.constraint_quotient_gl(lde, lb, 49, bc, [], 2, coef), log_n + lb, bound)
    h = lde.shape[0]
    nat = lde[[_rev(i, log_n + lb) for i in range(h)]]
    assert not _degree_bound_holds(O.constraint_quotient_gl(nat, lb, 7, bc, [], 2, coef), log_n + lb, bound)


def test_quotient_matches_the_formula_row_by_row():
    rng = np.random.default_rng(12)
    log_h, lb, w = 5, 2, 4
    h, n = 1 << log_h, 1 << (log_h - lb)
    lde = (rng.integers(0, 2**63, (h, w), dtype=np.uint64) % np.uint64(P)).astype(np.uint64)
    consts = [3, P - 1]
    bc = random_program(rng, w, 2, 3, 40)
    coef = [int(v) % P for v in rng.integers(0, 2**63, 3, dtype=np.uint64)]
    q = O.constraint_quotient_gl(lde, lb, 7, bc, consts, 3, coef)
    c = py_eval(lde, bc, consts, 3)
    g = 1753635133440165772  # p3 Goldilocks TWO_ADIC_GENERATOR (2^32-th root of unity)
    w_h = pow(g, 1 << (32 - log_h), P)
    for i in range(h):
        x = 7 * pow(w_h, _rev(i, log_h), P) % P
        v = (pow(x, n, P) - 1) % P
        acc = sum(int(k) * int(cv) for k, cv in zip(coef, c[i])) % P
        assert int(q[i]) == acc * pow(v, P - 2, P) % P


def test_quotient_refuses_a_point_on_the_trace_domain():
    lde = np.ones((8, 1), dtype=np.uint64)
    bc = [O.instr(O.OP_LOAD, 0), O.instr(O.OP_EMIT, 0)]
    assert O.constraint_quotient_gl(lde, 1, 1, bc, [], 1, [1]) is None  # shift 1: x^N = 1 on half the rows
