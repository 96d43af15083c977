"""Oracle DFT / coset LDE against the O(n^2) definition and closed forms."""
import numpy as np
import pytest

import oracle as O


def naive_dft(f, col, inverse=False):
    p = O.PRIME[f]
    n = len(col)
    lg = n.bit_length() - 1
    w = O.two_adic_generator(f, lg)
    if inverse:
        w = pow(w, p - 2, p)
    out = []
    for j in range(n):
        acc = 0
        for k in range(n):
            acc += int(col[k]) * pow(w, j * k, p)
        out.append(acc % p)
    if inverse:
        ninv = pow(n, p - 2, p)
        out = [x * ninv % p for x in out]
    return out


@pytest.mark.parametrize("f", [O.GL, O.BB])
@pytest.mark.parametrize("lg", [0, 1, 2, 5, 8])
def test_dft_matches_definition(f, lg):
    n = 1 << lg
    m = O.gen(f, 100 + lg, n * 3).reshape(n, 3)
    fw = O.dft_batch(f, m)
    iv = O.dft_batch(f, m, inverse=True)
    for c in range(3):
        assert [int(x) for x in fw[:, c]] == naive_dft(f, m[:, c])
        assert [int(x) for x in iv[:, c]] == naive_dft(f, m[:, c], inverse=True)
    assert (O.dft_batch(f, fw, inverse=True) == m).all()


def test_dft_basic_goldilocks():
    # zkf-metal/src/ntt/p3_adapter.rs:604-619: dft of [1,2,3,4] has out[0] = 10
    out = O.dft_batch(O.GL, np.array([[1], [2], [3], [4]], dtype=np.uint64))
    assert out[0, 0] == 10


def patterned_goldilocks(h, w, seed):
    # zkf-metal/src/ntt/p3_adapter.rs:581-590 (from_u64 then canonical)
    i = np.arange(h * w, dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + i * np.uint64(0xBF58476D1CE4E5B9)
    v = np.where(v >= np.uint64(O.GL_P), v - np.uint64(O.GL_P), v)
    return v.reshape(h, w)


def test_patterned_roundtrip_2_13():
    # the reference's metal_ntt_matches_cpu inputs at 2^13 x 4, round trip
    m = patterned_goldilocks(1 << 13, 4, 42)
    assert (O.dft_batch(O.GL, O.dft_batch(O.GL, m), inverse=True) == m).all()


def _rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


@pytest.mark.parametrize("f", [O.GL, O.BB])
@pytest.mark.parametrize("lg,lb", [(0, 1), (1, 1), (3, 1), (4, 2), (6, 2), (5, 0)])
def test_coset_lde_closed_form(f, lg, lb):
    """Row j of the LDE = p(shift * w_{BN}^rev(j)) with p the interpolant."""
    p = O.PRIME[f]
    n = 1 << lg
    shift = 7 if f == O.GL else 31
    m = O.gen(f, 7 + lg, n * 2).reshape(n, 2)
    lde = O.coset_lde_batch(f, m, lb, shift)
    big = n << lb
    wb = O.two_adic_generator(f, lg + lb)
    for c in range(2):
        coeffs = naive_dft(f, m[:, c], inverse=True)
        for j in range(big):
            x = shift * pow(wb, _rev(j, lg + lb), p) % p
            val = sum(int(cf) * pow(x, k, p) for k, cf in enumerate(coeffs)) % p
            assert int(lde[j, c]) == val


@pytest.mark.parametrize("f", [O.GL, O.BB])
def test_coset_lde_invariant(f):
    """Undo the bit-reversal and the coset: coefficients beyond N vanish."""
    p = O.PRIME[f]
    n, lb = 1 << 7, 2
    shift = 7 if f == O.GL else 31
    m = O.gen(f, 3, n * 4).reshape(n, 4)
    lde = O.coset_lde_batch(f, m, lb, shift)
    big = n << lb
    nat = np.empty_like(lde)
    for j in range(big):
        nat[_rev(j, 9)] = lde[j]
    coeffs = O.dft_batch(f, nat, inverse=True)
    assert (coeffs[n:] == 0).all()
    sinv = pow(shift, p - 2, p)
    orig = O.dft_batch(f, m, inverse=True)
    for k in range(n):
        sk = pow(sinv, k, p)
        for c in range(4):
            assert int(coeffs[k, c]) * sk % p == int(orig[k, c])
