"""Device-side Mmcs::verify_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:210-330).

Trees are committed and opened on the GPU. Every opening of a batch is
verified on the GPU, and the verdict is compared with the oracle's
verify_batch for untouched openings, tampered rows, tampered siblings, a
wrong index and a wrong root.
"""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import MMCS_SHAPES, rnd

import ziros_amd as Z

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
PRIME = {GL: O.GL_P, BB: O.BB_P}


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(0)
    for f, w in ((BB, 16), (GL, 16), (GL, 8)):
        c.load_constants(f, w, 42)
    yield c
    c.close()


def _valid_indices(shapes, rng, n):
    """Indices p3 can open: every matrix has the row index >> bits_reduced."""
    maxh = max(h for h, _ in shapes)
    logmax = (maxh - 1).bit_length()
    ok = [i for i in range(maxh)
          if all((i >> (logmax - (h - 1).bit_length())) < h for h, _ in shapes)]
    return np.array(rng.choice(ok, size=min(n, len(ok)), replace=False), dtype=np.uint64)


@pytest.mark.parametrize("shape_id", range(len(MMCS_SHAPES)))
@pytest.mark.parametrize("field,p2w", [(BB, 16), (GL, 16), (GL, 8)])
def test_verify_batch_matches_the_oracle(ctx, field, p2w, shape_id):
    shapes = MMCS_SHAPES[shape_id]
    rng = np.random.default_rng(40 + shape_id + 10 * p2w + field)
    mats = [rnd(field, h * w, 1300 + 5 * k + shape_id).reshape(h, w) for k, (h, w) in enumerate(shapes)]
    mm = Z.HipMerkleTreeMmcs(ctx, field, p2w)
    root, tree = mm.commit(mats)
    heights = [h for h, _ in shapes]
    widths = [w for _, w in shapes]
    tw = sum(widths)
    idx = _valid_indices(shapes, rng, 40)
    rows, paths = tree.open_mmcs(idx, tw)
    n = idx.size
    assert mm.verify_batch(root, heights, widths, idx, rows, paths).all()

    # tamper with a third of the openings, each in one way
    rows_t, paths_t, idx_t = rows.copy(), paths.copy(), idx.copy()
    bad = set()
    for q in range(0, n, 3):
        kind = (q // 3) % 3
        if kind == 0 and tw:
            c = int(rng.integers(0, tw))
            rows_t[q, c] = (int(rows_t[q, c]) + 1) % PRIME[field]
        elif kind == 1 and paths.shape[1]:
            l, e = int(rng.integers(0, paths.shape[1])), int(rng.integers(0, paths.shape[2]))
            paths_t[q, l, e] = (int(paths_t[q, l, e]) + 1) % PRIME[field]
        elif kind == 2 and n > 1:
            idx_t[q] = idx[(q + 1) % n]
        else:
            continue
        bad.add(q)
    got = mm.verify_batch(root, heights, widths, idx_t, rows_t, paths_t)
    p2 = O.Poseidon2(field, p2w)
    split = np.cumsum(widths)[:-1]
    for q in range(n):
        exp = O.verify(p2, root.astype(np.uint64), heights, widths, int(idx_t[q]),
                       np.split(rows_t[q].astype(np.uint64), split), paths_t[q].astype(np.uint64))
        assert got[q] == exp, q
        assert got[q] == (q not in bad), q
    wrong_root = root.copy()
    wrong_root[0] = (int(wrong_root[0]) + 1) % PRIME[field]
    assert not mm.verify_batch(wrong_root, heights, widths, idx, rows, paths).any()
    tree.free()


def test_verify_batch_refusals(ctx):
    mm = Z.HipMerkleTreeMmcs(ctx, BB, 16)
    m = rnd(BB, 16 * 3, 5).reshape(16, 3)
    root, tree = mm.commit([m])
    rows, paths = tree.open_mmcs([1, 2], 3)
    with pytest.raises(Z.ZkgError):   # index past the tallest height
        mm.verify_batch(root, [16], [3], [16, 2], rows, paths)
    with pytest.raises(Z.ZkgError):   # heights that round up to the same power of two
        mm.verify_batch(root, [16, 15], [3, 0], [1, 2], rows, np.zeros((2, 4, 8), np.uint32))
    bad = rows.copy()
    bad[0, 0] = O.BB_P                # not canonical
    with pytest.raises(Z.ZkgError):
        mm.verify_batch(root, [16], [3], [1, 2], bad, paths)
    tree.free()


def test_verify_batch_many_fri_queries(ctx):
    """80 queries against a 2^16-leaf tree of width-2 EF4 rows (the FRI layer
    shape): all verify, and flipping one sibling word fails exactly that one."""
    mm = Z.HipMerkleTreeMmcs(ctx, BB, 16)
    m = rnd(BB, (1 << 16) * 8, 99).reshape(1 << 16, 8)
    root, tree = mm.commit([m])
    rng = np.random.default_rng(5)
    idx = rng.integers(0, 1 << 16, 80).astype(np.uint64)
    rows, paths = tree.open_mmcs(idx, 8)
    assert mm.verify_batch(root, [1 << 16], [8], idx, rows, paths).all()
    paths[37, 9, 4] = (int(paths[37, 9, 4]) + 5) % O.BB_P
    ok = mm.verify_batch(root, [1 << 16], [8], idx, rows, paths)
    assert not ok[37] and ok.sum() == 79
    tree.free()
