"""Oracle FRI fold: reference formula, p3 fold_matrix, degree halving."""
import numpy as np
import pytest

import oracle as O


def test_fold_reference_inputs_goldilocks():
    # zkf-metal/src/fri.rs:187-224: evals i*7+1, alpha 42, inv_twiddles i*13+3
    n = 2048
    ev = (np.arange(2 * n, dtype=np.uint64) * np.uint64(7) + np.uint64(1))
    tw = (np.arange(n, dtype=np.uint64) * np.uint64(13) + np.uint64(3))
    out = O.fri_fold_base(O.GL, ev, 42, tw)
    p = O.GL_P
    inv2 = (p + 1) // 2
    for i in (0, 1, 1000, n - 1):
        fe, fo = int(ev[2 * i]), int(ev[2 * i + 1])
        expect = ((fe + fo) * inv2 + 42 * ((fe - fo) % p) * int(tw[i])) % p
        assert int(out[i]) == expect


def _rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


@pytest.mark.parametrize("f", [O.GL, O.BB])
def test_fold_matrix_even_odd_identity(f):
    """g(x^2) = f_e(x^2) + beta f_o(x^2) on a bit-reversed codeword."""
    p = O.PRIME[f]
    lg = 6
    n2 = 1 << lg
    coeffs = [int(x) for x in O.gen(f, 4, n2 // 4)] + [0] * (3 * n2 // 4)
    w = O.two_adic_generator(f, lg)
    evals = np.array([sum(c * pow(w, _rev(j, lg) * k, p) for k, c in enumerate(coeffs)) % p
                      for j in range(n2)], dtype=np.uint64)
    beta = 123456789
    g = O.fri_fold_matrix_base(f, evals, beta)
    fe = coeffs[0::2]
    fo = coeffs[1::2]
    for i in range(n2 // 2):
        x = pow(w, _rev(2 * i, lg), p)
        y = x * x % p
        expect = (sum(c * pow(y, k, p) for k, c in enumerate(fe)) + beta * sum(c * pow(y, k, p) for k, c in enumerate(fo))) % p
        assert int(g[i]) == expect


def test_fold_ef4_matches_componentwise_base_when_beta_in_base():
    """With beta in the base field the EF4 fold acts per coefficient."""
    n = 64
    ev = O.gen(O.BB, 8, 2 * n * 4).reshape(2 * n, 4)
    beta = 987654
    g = O.fri_fold_matrix_ef4(ev, [beta, 0, 0, 0])
    for c in range(4):
        assert (g[:, c] == O.fri_fold_matrix_base(O.BB, ev[:, c], beta)).all()


def test_fold_ef4_low_degree():
    """Folding a low-degree EF4 codeword halves its degree bound."""
    p = O.BB_P
    lg, lb = 7, 2
    n = 1 << (lg - lb)
    coeffs = O.gen(O.BB, 12, n * 4).reshape(n, 4)
    cw = O.coset_lde_batch(O.BB, coeffs, lb, 31)  # 2^7 EF4 codeword, bit-reversed
    beta = O.fri_betas(99, 1)[0]
    g = O.fri_fold_matrix_ef4(cw, beta)
    nat = np.empty_like(g)
    for j in range(len(g)):
        nat[_rev(j, lg - 1)] = g[j]
    back = O.dft_batch(O.BB, nat, inverse=True)
    assert (back[n // 2:] == 0).all() and not (back[: n // 2] == 0).all()


def test_commit_phase_layers():
    p2 = O.Poseidon2(O.BB, 16)
    lg, lf = 9, 4
    cw = O.gen(O.BB, 13, (1 << lg) * 4).reshape(-1, 4)
    betas = O.fri_betas(5, lg - lf)
    roots, final = O.fri_commit_phase_ef4(p2, cw, lg, lf, betas)
    assert roots.shape == (lg - lf, 8) and final.shape == (1 << lf, 4)
    # layer 0 root = tree over width-2 EF4 rows (8 base elements)
    t0 = O.MerkleTree(p2, [cw.reshape(-1, 8)])
    assert (roots[0] == t0.root()).all()
    g = O.fri_fold_matrix_ef4(cw, betas[0])
    assert (roots[1] == O.MerkleTree(p2, [g.reshape(-1, 8)]).root()).all()


@pytest.mark.parametrize("f", [O.GL, O.BB])
def test_commit_phase_base_layers(f):
    """Base-field commit phase (Challenge = Val): layer r's root is the tree of
    layer r's codeword as a width-2 matrix, each layer is the fold of the one
    before, and the final codeword of a low-degree input is low degree."""
    p2 = O.Poseidon2(f, 16)
    lg, lf, lb = 9, 4, 2
    coeffs = O.gen(f, 21, 1 << (lg - lb)).reshape(-1, 1)
    cw = O.coset_lde_batch(f, coeffs, lb, O.lib().or_f_generator(f)).reshape(-1)  # bit-reversed, blowup 4
    betas = O.gen(f, 23, lg - lf)
    roots, final = O.fri_commit_phase_base(p2, cw, lg, lf, betas)
    assert roots.shape == (lg - lf, 8) and final.shape == (1 << lf,)
    layer = cw
    for r in range(lg - lf):
        assert (roots[r] == O.MerkleTree(p2, [layer.reshape(-1, 2)]).root()).all(), r
        layer = O.fri_fold_matrix_base(f, layer, int(betas[r]))
    assert (final == layer).all()
    # degree < 2^(lg - lb) folded lg - lf times: degree < 2^(lf - lb) on 2^lf points
    nat = np.empty_like(final)
    for j in range(len(final)):
        nat[_rev(j, lf)] = final[j]
    back = O.dft_batch(f, nat.reshape(-1, 1), inverse=True).reshape(-1)
    assert (back[1 << (lf - lb):] == 0).all() and not (back[: 1 << (lf - lb)] == 0).all()
    # a wrong beta in one round breaks the final codeword
    bad = betas.copy()
    bad[1] = (int(bad[1]) + 1) % O.PRIME[f]
    _, final_bad = O.fri_commit_phase_base(p2, cw, lg, lf, bad)
    assert not (final_bad == final).all()
