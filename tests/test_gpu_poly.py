"""PolyOpsAccelerator on the GPU (zkg_poly_*_goldilocks) against the oracle,
bit-exact: the reference's own test vectors (zkf-metal/src/poly.rs:262-403),
tile and run boundaries of the kernels, z on the subgroup, and refusals."""
import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu
P = O.GL_P


def test_reference_vectors(ctx):
    ops = Z.HipPolyOps(ctx)
    coeffs = np.array([(i * 7 + 1) % P for i in range(64)], dtype=np.uint64)
    points = np.array([(i * 13 + 3) % P for i in range(2048)], dtype=np.uint64)
    assert (ops.batch_eval_goldilocks(coeffs, points) == O.poly_eval_gl(coeffs, points)).all()
    c2 = np.array([(i * 5 + 9) % P for i in range(2048)], dtype=np.uint64)
    assert (ops.coset_eval_goldilocks(c2, 7, 11) == O.poly_coset_eval_gl(c2, 7, 11)).all()
    n = 2048
    c3 = [(i * 17 + 4) % P for i in range(64)]
    g = int(O.two_adic_generator(O.GL, 11))
    xs = np.array([pow(g, i, P) for i in range(n)], dtype=np.uint64)
    evals = O.poly_eval_gl(c3, xs)
    fz = int(O.poly_eval_gl(c3, [9])[0])
    assert (ops.quotient_goldilocks(evals, 9, fz) == O.poly_quotient_gl(evals, 9, fz)).all()


@pytest.mark.parametrize("n_coeffs,n_points", [(0, 5), (1, 1), (3, 300), (2047, 513), (2048, 256), (2049, 257),
                                               (5000, 1000), (100, 1 << 17)])
def test_batch_eval_vs_oracle(ctx, n_coeffs, n_points):
    c = O.gen(O.GL, 10 + n_coeffs, n_coeffs)
    x = O.gen(O.GL, 20 + n_points, n_points)
    x[:min(3, n_points)] = [0, 1, P - 1][:min(3, n_points)]
    assert (Z.HipPolyOps(ctx).batch_eval_goldilocks(c, x) == O.poly_eval_gl(c, x)).all()


@pytest.mark.parametrize("log_n", [0, 1, 4, 5, 11, 12, 16, 20])
def test_coset_eval_vs_oracle(ctx, log_n):
    c = O.gen(O.GL, 30 + log_n, 1 << log_n)
    shift = int(O.gen(O.GL, 99, 1)[0])
    assert (Z.HipPolyOps(ctx).coset_eval_goldilocks(c, shift, log_n) == O.poly_coset_eval_gl(c, shift, log_n)).all()


@pytest.mark.parametrize("log_n", [0, 1, 4, 5, 10, 15, 20])
def test_quotient_vs_oracle(ctx, log_n):
    n = 1 << log_n
    e = O.gen(O.GL, 40 + log_n, n)
    g = int(O.two_adic_generator(O.GL, log_n))
    for z in (int(O.gen(O.GL, 41, 1)[0]), pow(g, n // 3, P), 1):  # off and on the subgroup (index n/3, 0)
        got = Z.HipPolyOps(ctx).quotient_goldilocks(e, z, 12345)
        exp = O.poly_quotient_gl(e, z, 12345)
        assert (got == exp).all(), z


def test_refusals(ctx):
    ops = Z.HipPolyOps(ctx)
    with pytest.raises(Z.ZkgError):
        ops.batch_eval_goldilocks([P], [1])  # not canonical
    with pytest.raises(Z.ZkgError):
        ops.quotient_goldilocks(np.zeros(3, dtype=np.uint64), 1, 0)  # not a power of two
    with pytest.raises(Z.ZkgError):
        ops.coset_eval_goldilocks(np.zeros(1 << 23, dtype=np.uint64), 7, 23)  # above 2^22
