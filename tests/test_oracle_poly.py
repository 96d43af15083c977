"""The oracle's PolyOpsAccelerator restatement pinned by the reference's own
tests (zkf-metal/src/poly.rs:262-403), recomputed here with Python integers:
Horner evaluation at explicit points, evaluation on the coset shift * <g>,
and the quotient (f(x_i) - f(z)) / (x_i - z) over the subgroup, including a
z on the subgroup (whose entry is 0)."""
import numpy as np

import oracle as O

P = O.GL_P


def horner(coeffs, x):
    r = 0
    for c in reversed(coeffs):
        r = (r * x + c) % P
    return r


def subgroup_generator(log_n):
    # p3 two_adic_generator(log_n) = GENERATOR^((p - 1) / 2^log_n), GENERATOR = 7
    return pow(7, (P - 1) >> log_n, P)


def test_poly_eval_reference_vector():
    coeffs = [(i * 7 + 1) % P for i in range(64)]
    points = [(i * 13 + 3) % P for i in range(2048)]
    got = O.poly_eval_gl(coeffs, points)
    assert [int(v) for v in got] == [horner(coeffs, x) for x in points]


def test_poly_eval_edge_cases():
    pts = [0, 1, P - 1, 2**63, 12345]
    assert list(O.poly_eval_gl([], pts)) == [0] * 5
    assert list(O.poly_eval_gl([P - 5], pts)) == [P - 5] * 5
    c = [P - 1, P - 2, 3]
    assert [int(v) for v in O.poly_eval_gl(c, pts)] == [horner(c, x) for x in pts]


def test_poly_coset_eval_reference_vector():
    log_n = 11
    n = 1 << log_n
    coeffs = [(i * 5 + 9) % P for i in range(n)]
    got = O.poly_coset_eval_gl(coeffs, 7, log_n)
    g = subgroup_generator(log_n)
    for i in range(0, n, 61):
        assert int(got[i]) == horner(coeffs, 7 * pow(g, i, P) % P), i


def test_poly_coset_eval_shift_one_is_the_dft():
    n = 64
    coeffs = [int(v) for v in O.gen(O.GL, 3, n)]
    got = O.poly_coset_eval_gl(coeffs, 1, 6)
    g = subgroup_generator(6)
    assert [int(v) for v in got] == [horner(coeffs, pow(g, i, P)) for i in range(n)]


def test_poly_quotient_reference_vector():
    log_n = 11
    n = 1 << log_n
    coeffs = [(i * 17 + 4) % P for i in range(64)]
    g = subgroup_generator(log_n)
    xs = [pow(g, i, P) for i in range(n)]
    evals = [horner(coeffs, x) for x in xs]
    z = 9
    fz = horner(coeffs, z)
    got = O.poly_quotient_gl(evals, z, fz)
    exp = [(e - fz) * pow((x - z) % P, P - 2, P) % P for e, x in zip(evals, xs)]
    assert [int(v) for v in got] == exp
    # (f - f(z)) / (x - z) is a polynomial of degree 62, so its values on the
    # subgroup have no coefficient above x^62
    q_coeffs = O.dft_batch(O.GL, np.array(exp, dtype=np.uint64).reshape(n, 1), inverse=True).reshape(-1)
    assert all(int(c) == 0 for c in q_coeffs[63:])


def test_poly_quotient_z_on_the_subgroup():
    n = 256
    g = subgroup_generator(8)
    z = pow(g, 37, P)
    evals = [int(v) for v in O.gen(O.GL, 5, n)]
    got = O.poly_quotient_gl(evals, z, 77)
    assert int(got[37]) == 0
    for i in (0, 36, 38, 255):
        x = pow(g, i, P)
        assert int(got[i]) == (evals[i] - 77) * pow((x - z) % P, P - 2, P) % P
