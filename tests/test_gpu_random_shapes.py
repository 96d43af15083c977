"""GPU parity on a seeded draw of transform and commit shapes.

The fixed-shape tests pin each tile size and plan; this draw adds the
combinations between them: heights 2^0 .. 2^17, widths 1 .. 130 (partial last
column chunks of every tile width), blowups 0 .. 2, random coset shifts, both
fields, each through the plan the tuning layers pick for its shape.  Every
value is compared with the oracle, bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}
P = {GL: O.GL_P, BB: O.BB_P}


def draw(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        field = GL if rng.random() < 0.5 else BB
        logn = int(rng.integers(0, 18))
        # keep each case to a few million elements so the oracle stays quick
        wmax = max(1, min(130, (3 << 20) >> (logn + 2)))
        width = int(rng.integers(1, wmax + 1))
        lb = int(rng.integers(0, 3))
        shift = int(rng.integers(2, 1 << 62)) % (P[field] - 2) + 2
        out.append((field, logn, width, lb, shift))
    return out


@pytest.mark.parametrize("case", range(24))
def test_transform_random_shape(ctx, case):
    field, logn, width, lb, shift = draw(24, 0x5EED00A1)[case]
    n = 1 << logn
    m = O.gen(field, 1000 + case, n * width).astype(DT[field]).reshape(n, width)
    dft = Z.HipDft(ctx, field)
    if lb == 0:
        assert (dft.dft_batch(m).astype(np.uint64) == O.dft_batch(field, m)).all(), (field, logn, width)
        assert (dft.idft_batch(m).astype(np.uint64) == O.dft_batch(field, m, inverse=True)).all(), (field, logn, width)
    else:
        got = dft.coset_lde_batch(m, lb, shift=shift)
        assert (got.astype(np.uint64) == O.coset_lde_batch(field, m, lb, shift)).all(), (field, logn, width, lb, shift)


@pytest.mark.parametrize("case", range(8))
def test_commit_random_shape(ctx, case):
    rng = np.random.default_rng(0x5EED00B0 + case)
    field, p2w = [(GL, 16), (GL, 8), (BB, 16)][case % 3]
    logn = int(rng.integers(1, 15))
    width = int(rng.integers(1, 100))
    lb = int(rng.integers(1, 3))
    n = 1 << logn
    p2 = O.Poseidon2(field, p2w)
    tr = O.gen(field, 2000 + case, n * width).astype(DT[field]).reshape(n, width)
    buf = ctx.upload(tr)
    tree, lde = Z.commit_trace(ctx, field, buf, n, width, lb, p2_width=p2w)
    root_o, lde_o = O.commit(p2, tr, lb)
    got = lde.download(DT[field], (n << lb) * width).reshape(n << lb, width)
    assert (got.astype(np.uint64) == lde_o).all(), (field, p2w, logn, width, lb)
    assert (tree.root().astype(np.uint64) == root_o).all(), (field, p2w, logn, width, lb)
    tree.free()
    lde.free()
    buf.free()
