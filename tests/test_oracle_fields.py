"""Oracle field arithmetic against Python big integers and closed forms."""
import random

import numpy as np
import pytest

import oracle as O

GL, BB = O.GL, O.BB


@pytest.mark.parametrize("f", [GL, BB])
def test_mul_add_sub_bruteforce(f):
    p = O.PRIME[f]
    rnd = random.Random(1234 + f)
    edge = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, (p + 1) // 2, 2**32 - 1 if f == GL else 2**27, 2**32 if f == GL else 2**30]
    vals = edge + [rnd.randrange(p) for _ in range(300)]
    for a in vals[:60]:
        for b in vals:
            assert O.mul(f, a, b) == a * b % p
            assert O.add(f, a, b) == (a + b) % p
            assert O.sub(f, a, b) == (a - b) % p


@pytest.mark.parametrize("f,adicity,gen", [(GL, 32, 7), (BB, 27, 31)])
def test_two_adic_generator_closed_form(f, adicity, gen):
    p = O.PRIME[f]
    # the top root equals GENERATOR^((p-1)/2^adicity) and has exact order 2^adicity
    top = pow(gen, (p - 1) >> adicity, p)
    assert O.two_adic_generator(f, adicity) == top
    assert pow(top, 1 << adicity, p) == 1 and pow(top, 1 << (adicity - 1), p) == p - 1
    for k in range(adicity + 1):
        w = O.two_adic_generator(f, k)
        assert w == pow(top, 1 << (adicity - k), p)
    # GENERATOR generates the multiplicative group: g^((p-1)/q) != 1 for prime q | p-1
    qs = [2, 3, 5, 17, 257, 65537] if f == GL else [2, 3, 5]
    for q in qs:
        assert (p - 1) % q == 0
        assert pow(gen, (p - 1) // q, p) != 1


def test_ef4_closed_form():
    p = BB = O.BB_P
    # 11 is a quadratic non-residue and p = 1 mod 4, so x^4 - 11 is irreducible
    assert pow(11, (p - 1) // 2, p) == p - 1 and p % 4 == 1
    rnd = np.random.default_rng(7)

    def polymul(a, b):
        c = [0] * 7
        for i in range(4):
            for j in range(4):
                c[i + j] += int(a[i]) * int(b[j])
        return [(c[k] + 11 * (c[k + 4] if k < 3 else 0)) % p for k in range(4)]

    for _ in range(200):
        a = rnd.integers(0, p, 4)
        b = rnd.integers(0, p, 4)
        assert list(O.ef4_mul(a, b)) == polymul(a, b)
    # x * x^3 = x^4 = 11
    assert list(O.ef4_mul([0, 1, 0, 0], [0, 0, 0, 1])) == [11, 0, 0, 0]
    # a^(p^4 - 1) = 1 for a != 0 (field of order p^4)
    a = [3, 5, 7, 9]
    e, r, base = p**4 - 1, [1, 0, 0, 0], a
    while e:
        if e & 1:
            r = list(O.ef4_mul(r, base))
        base = list(O.ef4_mul(base, base))
        e >>= 1
    assert r == [1, 0, 0, 0]


@pytest.mark.parametrize("f", [GL, BB])
def test_generator_canonical_and_slices(f):
    p = O.PRIME[f]
    a = O.gen(f, 5, 1000)
    assert (a < p).all()
    b = O.gen(f, 5, 300, start=700)
    assert (a[700:] == b).all()
    assert not (O.gen(f, 6, 1000) == a).all()
