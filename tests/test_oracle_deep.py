"""Pins for the oracle's DEEP quotient (PCS opening) over BabyBear / EF4.

p3-fri 0.4.2 TwoAdicFriPcs::open and p3-interpolation interpolate_coset are
not in the reference tree and no reference vector exercises them, so the
oracle is pinned by what the mathematics fixes:
  * the EF4 inverse agrees with an independent tower-formula inverse;
  * opened values equal Horner evaluation of the known coefficients;
  * the reduced codeword sum_k a^k (p_k(X) - p_k(z)) / (X - z) is the LDE of
    a polynomial of degree < N - 1, and stops being so when an opened value
    is off by one;
  * the alpha offset multiplies by alpha^off and the result accumulates.
"""
import numpy as np
import pytest

import oracle as O

BB, P = O.BB, O.BB_P
W11 = 11


def ef4_mul_py(a, b):
    c = [0] * 7
    for i in range(4):
        for j in range(4):
            c[i + j] += int(a[i]) * int(b[j])
    return [(c[k] + (W11 * c[k + 4] if k < 3 else 0)) % P for k in range(4)]


def ef4_inv_tower(a):
    """a = A(u) + X B(u), u = X^2, u^2 = 11: a^-1 = (A - X B) / (A^2 - u B^2)."""
    a0, a1, a2, a3 = (int(v) for v in a)
    A, B = (a0, a2), (a1, a3)

    def m2(x, y):  # EF2 product, u^2 = 11
        return ((x[0] * y[0] + W11 * x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)

    A2, B2 = m2(A, A), m2(B, B)
    uB2 = ((W11 * B2[1]) % P, B2[0])
    n = ((A2[0] - uB2[0]) % P, (A2[1] - uB2[1]) % P)
    m = (n[0] * n[0] - W11 * n[1] * n[1]) % P
    mi = pow(m, P - 2, P)
    ni = ((n[0] * mi) % P, (-n[1] * mi) % P)
    c_even = m2(A, ni)                       # A / n
    c_odd = m2(((-B[0]) % P, (-B[1]) % P), ni)  # -B / n
    return [c_even[0], c_odd[0], c_even[1], c_odd[1]]


def rand_ef4(rng):
    return [int(v) for v in rng.integers(0, P, 4)]


def horner(coeffs, z):
    acc = [0, 0, 0, 0]
    for c in reversed(coeffs):
        acc = ef4_mul_py(acc, z)
        acc[0] = (acc[0] + int(c)) % P
    return acc


def rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def make_lde(rng, n, w, lb):
    coeffs = rng.integers(0, P, (n, w)).astype(np.uint64)
    evals = O.dft_batch(BB, coeffs)
    return coeffs, O.coset_lde_batch(BB, evals, lb, 31)


def test_ef4_inverse_two_ways():
    rng = np.random.default_rng(1)
    for _ in range(50):
        a = rand_ef4(rng)
        inv = [int(v) for v in O.ef4_inv(a)]
        assert ef4_mul_py(a, inv) == [1, 0, 0, 0]
        assert inv == ef4_inv_tower(a)
    assert [int(v) for v in O.ef4_inv([1, 0, 0, 0])] == [1, 0, 0, 0]


@pytest.mark.parametrize("logn,w,lb", [(1, 2, 1), (5, 3, 1), (6, 2, 2), (4, 5, 3)])
def test_open_values_equal_horner(logn, w, lb):
    rng = np.random.default_rng(logn * 10 + lb)
    n = 1 << logn
    coeffs, lde = make_lde(rng, n, w, lb)
    z = rand_ef4(rng)
    ys = O.deep_open_values(lde, lb, 31, z)
    for k in range(w):
        assert [int(v) for v in ys[k]] == horner(coeffs[:, k], z), k


def natural_coeffs(ro, height):
    """ro rows are values at g w^rev(i); coefficients (times g^k) per EF4 component."""
    lg = height.bit_length() - 1
    nat = np.empty_like(ro)
    for i in range(height):
        nat[rev(i, lg)] = ro[i]
    return O.dft_batch(BB, nat, inverse=True)


@pytest.mark.parametrize("logn,w,lb", [(4, 3, 1), (5, 2, 2), (3, 6, 1)])
def test_reduced_codeword_is_low_degree(logn, w, lb):
    rng = np.random.default_rng(100 + logn)
    n = 1 << logn
    coeffs, lde = make_lde(rng, n, w, lb)
    z, alpha = rand_ef4(rng), rand_ef4(rng)
    ys = O.deep_open_values(lde, lb, 31, z)
    ro = O.deep_reduce(lde, 31, z, alpha, 0, ys)
    c = natural_coeffs(ro, n << lb)
    assert (c[n - 1:] == 0).all()
    assert (c[: n - 1] != 0).any()
    bad = ys.copy()
    bad[0][0] = (int(bad[0][0]) + 1) % P  # a planted wrong opened value
    c_bad = natural_coeffs(O.deep_reduce(lde, 31, z, alpha, 0, bad), n << lb)
    assert (c_bad[n - 1:] != 0).any()


def test_reduce_single_row_explicit():
    rng = np.random.default_rng(7)
    n, w, lb = 8, 4, 1
    coeffs, lde = make_lde(rng, n, w, lb)
    z, alpha = rand_ef4(rng), rand_ef4(rng)
    ys = O.deep_open_values(lde, lb, 31, z)
    ro = O.deep_reduce(lde, 31, z, alpha, 0, ys)
    h = n << lb
    lg = h.bit_length() - 1
    wh = int(O.two_adic_generator(BB, lg))
    for i in (0, 5, h - 1):
        x = (31 * pow(wh, rev(i, lg), P)) % P
        acc, ap = [0, 0, 0, 0], [1, 0, 0, 0]
        for k in range(w):
            term = [(int(lde[i, k]) * (1 if e == 0 else 0) - int(ys[k][e])) % P for e in range(4)]
            acc = [(acc[e] + v) % P for e, v in enumerate(ef4_mul_py(ap, term))]
            ap = ef4_mul_py(ap, alpha)
        d = [(x - z[0]) % P, (-z[1]) % P, (-z[2]) % P, (-z[3]) % P]
        assert [int(v) for v in ro[i]] == ef4_mul_py(acc, ef4_inv_tower(d)), i


def test_reduce_offset_and_accumulation():
    rng = np.random.default_rng(9)
    n, w, lb = 16, 3, 1
    _, lde = make_lde(rng, n, w, lb)
    z, alpha = rand_ef4(rng), rand_ef4(rng)
    ys = O.deep_open_values(lde, lb, 31, z)
    r0 = O.deep_reduce(lde, 31, z, alpha, 0, ys)
    r3 = O.deep_reduce(lde, 31, z, alpha, 3, ys)
    a3 = ef4_mul_py(ef4_mul_py(alpha, alpha), alpha)
    for i in range(n << lb):
        assert [int(v) for v in r3[i]] == ef4_mul_py(r0[i], a3)
    base = rng.integers(0, P, ((n << lb), 4)).astype(np.uint64)
    acc = O.deep_reduce(lde, 31, z, alpha, 0, ys, ro=base)
    assert (acc == (base + r0) % P).all()


# ---------------------------------------------------------------- Goldilocks, base-field challenges
GL, GP = O.GL, O.GL_P


def gl_lde(rng, n, w, lb):
    coeffs = (rng.integers(0, 2**63, (n, w), dtype=np.uint64) % np.uint64(GP)).astype(np.uint64)
    return coeffs, O.coset_lde_batch(GL, O.dft_batch(GL, coeffs), lb, 7)


@pytest.mark.parametrize("logn,w,lb", [(1, 2, 1), (5, 3, 1), (6, 2, 2), (4, 5, 3)])
def test_gl_open_values_equal_horner(logn, w, lb):
    rng = np.random.default_rng(500 + logn)
    coeffs, lde = gl_lde(rng, 1 << logn, w, lb)
    z = int(rng.integers(0, 2**62))
    ys = O.deep_gl_open_values(lde, lb, 7, z)
    for k in range(w):
        acc = 0
        for c in reversed(coeffs[:, k]):
            acc = (acc * z + int(c)) % GP
        assert int(ys[k]) == acc


@pytest.mark.parametrize("logn,w,lb", [(4, 3, 1), (5, 2, 2)])
def test_gl_reduced_codeword_is_low_degree(logn, w, lb):
    rng = np.random.default_rng(600 + logn)
    n = 1 << logn
    _, lde = gl_lde(rng, n, w, lb)
    z, alpha = int(rng.integers(0, 2**62)), int(rng.integers(0, 2**62))
    ys = O.deep_gl_open_values(lde, lb, 7, z)
    ro = O.deep_gl_reduce(lde, 7, z, alpha, 0, ys)
    h = n << lb
    lg = h.bit_length() - 1
    nat = np.empty_like(ro)
    for i in range(h):
        nat[rev(i, lg)] = ro[i]
    c = O.dft_batch(GL, nat.reshape(h, 1), inverse=True).reshape(h)
    assert (c[n - 1:] == 0).all() and (c[: n - 1] != 0).any()
    bad = ys.copy()
    bad[0] = (int(bad[0]) + 1) % GP
    nat_bad = np.empty_like(ro)
    ro_bad = O.deep_gl_reduce(lde, 7, z, alpha, 0, bad)
    for i in range(h):
        nat_bad[rev(i, lg)] = ro_bad[i]
    assert (O.dft_batch(GL, nat_bad.reshape(h, 1), inverse=True).reshape(h)[n - 1:] != 0).any()


def test_gl_reduce_row_explicit_and_offset():
    rng = np.random.default_rng(77)
    n, w, lb = 8, 4, 1
    _, lde = gl_lde(rng, n, w, lb)
    z, alpha = int(rng.integers(0, 2**62)), int(rng.integers(0, 2**62))
    ys = O.deep_gl_open_values(lde, lb, 7, z)
    ro = O.deep_gl_reduce(lde, 7, z, alpha, 2, ys)
    h = n << lb
    lg = h.bit_length() - 1
    wh = int(O.two_adic_generator(GL, lg))
    for i in (0, 3, h - 1):
        x = 7 * pow(wh, rev(i, lg), GP) % GP
        num = sum(pow(alpha, k, GP) * (int(lde[i, k]) - int(ys[k])) for k in range(w)) % GP
        assert int(ro[i]) == num * pow((x - z) % GP, GP - 2, GP) * alpha * alpha % GP
