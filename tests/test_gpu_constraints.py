"""GPU parity of constraint evaluation (HIP interpreter vs the oracle).

Semantics: ConstraintEvalAccelerator::eval_trace_goldilocks
(zkf-core/src/acceleration.rs:219-242) running the bytecode of
zkf-metal/src/shaders/constraint_eval.metal:1-74.  Bit-exact against the
oracle on every row.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_constraint import random_program

import ziros_amd as Z
from ziros_amd._lib import check

pytestmark = pytest.mark.gpu
P = O.GL_P
I = O.instr


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(0)
    yield c
    c.close()


def _trace(rng, rows, width):
    t = (rng.integers(0, 2**63, (rows, width), dtype=np.uint64) % np.uint64(P)).astype(np.uint64)
    t[0, :] = P - 1
    return t


def _run(ctx, trace, bc, consts, n_c):
    rows, w = trace.shape
    prog = Z.ConstraintProgram(ctx, w, bc, consts, n_c)
    tb = ctx.upload(trace)
    ob = ctx.alloc(rows * n_c * 8)
    ob.upload(np.full(rows * n_c, 0xDEADBEEF, dtype=np.uint64))  # every output must be written
    prog.eval(tb, rows, ob)
    got = ob.download(np.uint64, rows * n_c).reshape(rows, n_c)
    info = prog.info()
    prog.free(); tb.free(); ob.free()
    return got, info


def air_program(width, n_c, rng):
    """Degree-3 constraints of a round-function AIR: for constraint k,
    x^3 + a_k y + b_k - u v over four columns."""
    cc = Z.ConstraintCompiler()
    for k in range(n_c):
        x, y, u, v = (2 * k) % width, (2 * k + 1) % width, (2 * k + 2) % width, (2 * k + 3) % width
        cc.load_column(x); cc.dup(); cc.dup(); cc.mul(); cc.mul()
        cc.push_const(int(rng.integers(0, 2**63, dtype=np.uint64)) % P); cc.load_column(y); cc.mul(); cc.add()
        cc.push_const(int(rng.integers(0, 2**63, dtype=np.uint64)) % P); cc.add()
        cc.load_column(u); cc.load_column(v); cc.mul(); cc.sub()
        cc.emit(k)
    return cc.finish()


def test_reference_add_mul_vector(ctx):
    """zkf-metal/src/constraint_eval.rs:224-269 through the accelerator mirror."""
    n, w = 2048, 2
    trace = np.array([(i * 7 + 1) % P for i in range(n * w)], dtype=np.uint64)
    cc = Z.ConstraintCompiler()
    cc.load_column(0); cc.load_column(1); cc.add(); cc.emit(0)
    cc.load_column(0); cc.load_column(1); cc.mul(); cc.emit(1)
    bc, consts, n_c = cc.finish()
    ev = Z.HipConstraintEval(ctx)
    got = ev.eval_trace_goldilocks(trace, w, bc, consts, n_c)
    assert got.size == n * n_c
    t = trace.reshape(n, w)
    for r in range(n):
        a, b = int(t[r, 0]), int(t[r, 1])
        assert int(got[2 * r]) == (a + b) % P and int(got[2 * r + 1]) == a * b % P
    assert (got.reshape(n, n_c) == O.constraint_eval_gl(t, bc, consts, n_c)).all()


def test_accelerator_refuses_what_the_reference_refuses(ctx):
    ev = Z.HipConstraintEval(ctx)
    ok = [I(O.OP_LOAD, 0), I(O.OP_EMIT, 0)]
    t = np.arange(2048 * 2, dtype=np.uint64)
    assert ev.min_rows() == 1024
    for args in [(t[:1023 * 2], 2, ok, [], 1), (t, 0, ok, [], 1), (t[:5], 2, ok, [], 1), (t, 2, [], [], 1),
                 (t, 2, ok, [], 0), (t[:0], 2, ok, [], 1)]:
        with pytest.raises(Z.ZkgError):
            ev.eval_trace_goldilocks(*args)


SHAPES = ["1,16", "2,4", "2,8", "2,16", "4,4", "4,8"]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("seed", range(10))
def test_random_programs(ctx, seed, shape, monkeypatch):
    monkeypatch.setenv("ZKG_CE_SHAPE", shape)
    rng = np.random.default_rng(500 + seed)
    w = int(rng.choice([1, 2, 15, 16, 33, 64, 100]))
    rows = int(rng.choice([1, 63, 64, 65, 255, 256, 257, 3000]))
    n_const, n_c = int(rng.integers(1, 8)), int(rng.integers(1, 9))
    trace = _trace(rng, rows, w)
    consts = [int(v) % P for v in rng.integers(0, 2**63, n_const, dtype=np.uint64)]
    consts[0] = P - 1
    bc = random_program(rng, w, n_const, n_c, int(rng.integers(3, 300)))
    if seed % 2:
        bc = [x for i in bc for x in ((i,) if rng.random() < 0.8 else (i, (int(rng.integers(7, 256)) << 24) | 9))]
    got, info = _run(ctx, trace, bc, consts, n_c)
    assert "%d,%d" % (info["rows_per_lane"], info["panel"]) == shape or info["rows_per_lane"] == 1
    assert (got == O.constraint_eval_gl(trace, bc, consts, n_c)).all()


def test_air_program_large_trace(ctx):
    """2^18 rows x 64 columns, 32 degree-3 constraints: every row vs the oracle."""
    rng = np.random.default_rng(77)
    rows, w, n_c = 1 << 18, 64, 32
    trace = _trace(rng, rows, w)
    bc, consts, n_c = air_program(w, n_c, rng)
    got, info = _run(ctx, trace, bc, consts, n_c)
    assert info["slots"] == 1 and info["stages"] == 64 // info["panel"] + 1  # the last constraint wraps to panel 0
    assert (got == O.constraint_eval_gl(trace, bc, consts, n_c)).all()


def test_deep_stack_many_slots_large_lds(ctx, monkeypatch):
    """Depth-32 stack plus interleaved panels: more than 64 KiB of LDS per block."""
    monkeypatch.setenv("ZKG_CE_SHAPE", "1,16")
    rng = np.random.default_rng(3)
    w, rows = 64, 1000
    trace = _trace(rng, rows, w)
    bc = []
    for c in range(8):
        bc += [I(O.OP_LOAD, c), I(O.OP_LOAD, 16 + c), I(O.OP_LOAD, 32 + c), I(O.OP_LOAD, 48 + c)]
        bc += [I(O.OP_ADD), I(O.OP_MUL), I(O.OP_SUB), I(O.OP_EMIT, c)]
    bc += [I(O.OP_LOAD, (5 * i) % w) for i in range(32)] + [I(O.OP_MUL), I(O.OP_SUB)] * 15 + [I(O.OP_ADD)]
    bc += [I(O.OP_EMIT, 8)]
    got, info = _run(ctx, trace, bc, [], 10)
    assert info["stack_slots"] == 31 and info["slots"] >= 2
    assert info["wave_bytes"] * 4 > 64 * 1024
    assert (got == O.constraint_eval_gl(trace, bc, [], 10)).all()
    assert (got[:, 9] == 0).all()


def test_program_reuse_and_one_shot_entry(ctx):
    rng = np.random.default_rng(21)
    w, n_c = 20, 6
    bc, consts, n_c = air_program(w, n_c, rng)
    prog = Z.ConstraintProgram(ctx, w, bc, consts, n_c)
    for rows in (5000, 333):
        trace = _trace(rng, rows, w)
        tb, ob = ctx.upload(trace), ctx.alloc(rows * n_c * 8)
        prog.eval(tb, rows, ob)
        got = ob.download(np.uint64, rows * n_c).reshape(rows, n_c)
        assert (got == O.constraint_eval_gl(trace, bc, consts, n_c)).all()
        tb2, ob2 = ctx.upload(trace), ctx.alloc(rows * n_c * 8)
        bca = np.asarray(bc, dtype=np.uint32)
        csa = np.asarray(consts, dtype=np.uint64)
        check(ctx.lib.zkg_constraint_eval_gl(ctx.h, tb2.h, rows, w, bca.ctypes.data, bca.size, csa.ctypes.data,
                                               csa.size, n_c, ob2.h))
        assert (ob2.download(np.uint64, rows * n_c).reshape(rows, n_c) == got).all()
        for b in (tb, ob, tb2, ob2):
            b.free()
    prog.free()


def test_eval_checks_buffers(ctx):
    bc = [I(O.OP_LOAD, 3), I(O.OP_EMIT, 1)]
    prog = Z.ConstraintProgram(ctx, 4, bc, [], 2)
    tb, ob = ctx.alloc(100 * 4 * 8), ctx.alloc(100 * 2 * 8)
    with pytest.raises(Z.ZkgError):
        prog.eval(tb, 101, ob)      # trace too small
    small = ctx.alloc(100 * 2 * 8 - 8)
    with pytest.raises(Z.ZkgError):
        prog.eval(tb, 100, small)   # output too small
    with pytest.raises(Z.ZkgError):
        prog.eval(tb, 0, ob)
    with pytest.raises(Z.ZkgError):
        Z.ConstraintProgram(ctx, 4, bc, [P], 2)  # non-canonical constant
    for b in (tb, ob, small):
        b.free()
    prog.free()


@pytest.mark.parametrize("shape", SHAPES)
def test_output_panel_sessions(ctx, shape, monkeypatch):
    """Emits that revisit an output panel, constraints never emitted, a panel
    with no emits at all, and a constraint emitted twice (the last wins)."""
    monkeypatch.setenv("ZKG_CE_SHAPE", shape)
    rng = np.random.default_rng(31)
    rows = 777
    trace = _trace(rng, rows, 6)
    order = [0, 1, 20, 2, 21, 0, 40, 3, 1]
    bc = []
    for k, j in enumerate(order):
        bc += [I(O.OP_LOAD, k % 6), I(O.OP_CONST, 0), I(O.OP_MUL), I(O.OP_EMIT, j)]
    got, _ = _run(ctx, trace, bc, [5], 50)
    exp = O.constraint_eval_gl(trace, bc, [5], 50)
    assert (got == exp).all()
    assert (got[:, 10:20] == 0).all() and (got[:, 41:] == 0).all()


def _run_compiled(ctx, trace, bc, consts, n_c):
    rows, w = trace.shape
    prog = Z.ConstraintProgram(ctx, w, bc, consts, n_c, compiled=True)
    tb = ctx.upload(trace)
    ob = ctx.alloc(rows * n_c * 8)
    ob.upload(np.full(rows * n_c, 0xDEADBEEF, dtype=np.uint64))
    prog.eval(tb, rows, ob)
    got = ob.download(np.uint64, rows * n_c).reshape(rows, n_c)
    prog.free(); tb.free(); ob.free()
    return got


@pytest.mark.parametrize("seed", range(6))
def test_compiled_random_programs(ctx, seed, monkeypatch):
    """The straight-line kernel compiled for one program against the oracle."""
    monkeypatch.setenv("ZKG_CE_SHAPE", SHAPES[seed % len(SHAPES)])
    rng = np.random.default_rng(900 + seed)
    w = int(rng.choice([1, 15, 33, 64, 100]))
    rows = int(rng.choice([1, 65, 257, 3000]))
    n_const, n_c = int(rng.integers(1, 8)), int(rng.integers(1, 9))
    trace = _trace(rng, rows, w)
    consts = [int(v) % P for v in rng.integers(0, 2**63, n_const, dtype=np.uint64)]
    consts[0] = P - 1
    bc = random_program(rng, w, n_const, n_c, int(rng.integers(3, 200)))
    got = _run_compiled(ctx, trace, bc, consts, n_c)
    assert (got == O.constraint_eval_gl(trace, bc, consts, n_c)).all()


def test_compiled_air_program_and_sessions(ctx):
    rng = np.random.default_rng(78)
    rows, w = (1 << 16) + 37, 64
    trace = _trace(rng, rows, w)
    bc, consts, n_c = air_program(w, 32, rng)
    assert (_run_compiled(ctx, trace, bc, consts, n_c) == O.constraint_eval_gl(trace, bc, consts, n_c)).all()
    order = [0, 1, 20, 2, 21, 0, 40, 3, 1]
    bc = []
    for k, j in enumerate(order):
        bc += [I(O.OP_LOAD, k % 6), I(O.OP_CONST, 0), I(O.OP_MUL), I(O.OP_EMIT, j)]
    trace = _trace(rng, 777, 6)
    assert (_run_compiled(ctx, trace, bc, [5], 50) == O.constraint_eval_gl(trace, bc, [5], 50)).all()


def test_one_shot_entry_remembers_and_compiles(ctx):
    """The reference-shaped call with the same program again: the context
    compiles it on the second use; every result matches the oracle."""
    ev = Z.HipConstraintEval(ctx)
    ev.cache_config(4, 2)
    rng = np.random.default_rng(44)
    w = 12
    bc, consts, n_c = air_program(w, 5, rng)
    other, oconsts, on_c = air_program(w, 3, rng)
    for k in range(3):
        trace = _trace(rng, 1500 + 100 * k, w)
        got = ev.eval_trace_goldilocks(trace.reshape(-1), w, bc, consts, n_c).reshape(-1, n_c)
        assert (got == O.constraint_eval_gl(trace, bc, consts, n_c)).all()
        info = ev.cache_info()
        assert info["programs"] == 1 and info["uses"] == k + 1 and info["compiled"] == (1 if k >= 1 else 0)
    for _ in range(6):  # other programs push it out (4 kept, least recently used dropped)
        other = other + [I(O.OP_LOAD, 0), I(O.OP_EMIT, 0)]
        trace = _trace(rng, 1024, w)
        got = ev.eval_trace_goldilocks(trace.reshape(-1), w, other, oconsts, on_c).reshape(-1, on_c)
        assert (got == O.constraint_eval_gl(trace, other, oconsts, on_c)).all()
    assert ev.cache_info()["programs"] == 4
    ev.cache_config(0, 0)
    assert ev.cache_info()["programs"] == 0
    trace = _trace(rng, 2000, w)
    got = ev.eval_trace_goldilocks(trace.reshape(-1), w, bc, consts, n_c).reshape(-1, n_c)
    assert (got == O.constraint_eval_gl(trace, bc, consts, n_c)).all()
    assert ev.cache_info()["programs"] == 0
    ev.cache_config()


@pytest.mark.parametrize("compiled", [False, True])
def test_wide_trace_scattered_columns(ctx, compiled):
    """1000 columns read in a scattered order: panels cycle through the slots."""
    rng = np.random.default_rng(55)
    w, rows = 1000, 700
    trace = _trace(rng, rows, w)
    cols = rng.permutation(w)[:200]
    bc = []
    for k in range(40):
        a, b, c = (int(x) for x in cols[5 * k:5 * k + 3])
        bc += [I(O.OP_LOAD, a), I(O.OP_LOAD, b), I(O.OP_MUL), I(O.OP_LOAD, c), I(O.OP_SUB), I(O.OP_EMIT, k)]
    bc += [I(O.OP_LOAD, int(cols[(7 * i) % 200])) for i in range(20)] + [I(O.OP_ADD)] * 19 + [I(O.OP_EMIT, 40)]
    exp = O.constraint_eval_gl(trace, bc, [], 41)
    got = _run_compiled(ctx, trace, bc, [], 41) if compiled else _run(ctx, trace, bc, [], 41)[0]
    assert (got == exp).all()
    if not compiled:
        info = Z.translate_constraints(w, bc, 0, 41)[1]
        assert info["stages"] >= 100  # scattered columns: almost every read stages a panel


# ---------------------------------------------------------------- quotient over the LDE
def _quotient(ctx, lde, log_blowup, shift, bc, consts, n_c, coef, compiled=False):
    h, w = lde.shape
    prog = Z.ConstraintProgram(ctx, w, bc, consts, n_c, compiled=compiled)
    lb, qb = ctx.upload(lde), ctx.alloc(h * 8)
    qb.upload(np.full(h, 0xDEADBEEF, dtype=np.uint64))
    prog.quotient(lb, h, log_blowup, shift, coef, qb)
    q = qb.download(np.uint64, h)
    prog.free(); lb.free(); qb.free()
    return q


@pytest.mark.parametrize("compiled", [False, True])
@pytest.mark.parametrize("seed", range(5))
def test_quotient_random_programs(ctx, seed, compiled, monkeypatch):
    monkeypatch.setenv("ZKG_CE_SHAPE", SHAPES[seed % len(SHAPES)])
    rng = np.random.default_rng(700 + seed)
    log_h, lb = int(rng.integers(6, 15)), int(rng.integers(1, 4))
    w = int(rng.choice([3, 20, 64]))
    n_const, n_c = int(rng.integers(1, 5)), int(rng.integers(1, 7))
    lde = _trace(rng, 1 << log_h, w)
    consts = [int(v) % P for v in rng.integers(0, 2**63, n_const, dtype=np.uint64)]
    bc = random_program(rng, w, n_const, n_c, int(rng.integers(3, 120)))
    if seed % 2:
        bc = bc + [I(O.OP_LOAD, 0), I(O.OP_EMIT, n_c - 1)]  # emitted twice: the last value counts
    coef = [int(v) % P for v in rng.integers(0, 2**63, n_c, dtype=np.uint64)]
    got = _quotient(ctx, lde, lb, 7, bc, consts, n_c, coef, compiled)
    assert (got == O.constraint_quotient_gl(lde, lb, 7, bc, consts, n_c, coef)).all()


@pytest.mark.parametrize("compiled", [False, True])
def test_quotient_of_a_satisfied_air_is_low_degree(ctx, compiled):
    """a, b = a^2, c = a^3 on a 2^10-row domain, LDE on the GPU at blowup 2:
    the quotient of a*a - b and b*a - c has degree N - 2."""
    from test_oracle_constraint import _degree_bound_holds, _square_air
    rng = np.random.default_rng(91)
    log_n, lb = 10, 1
    trace, bc = _square_air(rng, log_n)
    n = 1 << log_n
    tb, lde_b = ctx.upload(trace), ctx.alloc((n << lb) * 3 * 8)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.GOLDILOCKS, tb.h, n, 3, lb, 7, lde_b.h))
    prog = Z.ConstraintProgram(ctx, 3, bc, [], 2, compiled=compiled)
    qb = ctx.alloc((n << lb) * 8)
    coef = [123456789, 987654321]
    prog.quotient(lde_b, n << lb, lb, 7, coef, qb)
    q = qb.download(np.uint64, n << lb)
    lde = lde_b.download(np.uint64, (n << lb) * 3).reshape(n << lb, 3)
    assert (q == O.constraint_quotient_gl(lde, lb, 7, bc, [], 2, coef)).all()
    assert _degree_bound_holds(q, log_n + lb, n - 1)
    prog.free(); tb.free(); lde_b.free(); qb.free()


def test_quotient_refusals(ctx):
    bc = [I(O.OP_LOAD, 0), I(O.OP_EMIT, 0)]
    prog = Z.ConstraintProgram(ctx, 1, bc, [], 1)
    lb, qb = ctx.alloc(64 * 8), ctx.alloc(64 * 8)
    with pytest.raises(Z.ZkgError):
        prog.quotient(lb, 64, 1, 1, [1], qb)       # shift 1: the coset meets the trace domain
    with pytest.raises(Z.ZkgError):
        prog.quotient(lb, 48, 1, 7, [1], qb)       # height not a power of two
    with pytest.raises(Z.ZkgError):
        prog.quotient(lb, 64, 7, 7, [1], qb)       # blowup above the height
    with pytest.raises(Z.ZkgError):
        prog.quotient(lb, 64, 1, 7, [P], qb)       # coefficient not canonical
    with pytest.raises(Z.ZkgError):
        prog.quotient(lb, 128, 1, 7, [1], qb)      # buffers too small
    prog.free(); lb.free(); qb.free()


@pytest.mark.parametrize("log_n,lb", [(14, 1), (13, 2)])
def test_quotient_low_degree_at_size(ctx, log_n, lb):
    """A satisfied AIR at 2^14 / 2^13 trace rows: GPU LDE, compiled quotient,
    and an inverse DFT on the host show degree exactly N - 2."""
    from test_oracle_constraint import _degree_bound_holds, _square_air
    rng = np.random.default_rng(log_n)
    trace, bc = _square_air(rng, log_n)
    n = 1 << log_n
    tb, lde_b, qb = ctx.upload(trace), ctx.alloc((n << lb) * 3 * 8), ctx.alloc((n << lb) * 8)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.GOLDILOCKS, tb.h, n, 3, lb, 7, lde_b.h))
    prog = Z.ConstraintProgram(ctx, 3, bc, [], 2, compiled=True)
    prog.quotient(lde_b, n << lb, lb, 7, [3, 5], qb)
    q = qb.download(np.uint64, n << lb)
    assert _degree_bound_holds(q, log_n + lb, n - 1)
    assert not _degree_bound_holds(q, log_n + lb, n - 2)
    prog.free(); tb.free(); lde_b.free(); qb.free()
