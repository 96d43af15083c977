"""Batched openings gathered on the device (zkg_mmcs_open, zkg_fri_open): the
edge cases of the one-launch gather.  No queries, rows only, paths only and a
repeated index each give what the full call gives, and every full opening
verifies against the oracle's tree."""
import ctypes

import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


@pytest.mark.parametrize("field", [GL, BB])
def test_mmcs_open_partial_outputs(ctx, field):
    shapes = [(1024, 7), (256, 3)]
    p2 = O.Poseidon2(field, 16)
    mats = [O.gen(field, 70 + k, h * w).astype(DT[field]).reshape(h, w) for k, (h, w) in enumerate(shapes)]
    root, tree = Z.HipMerkleTreeMmcs(ctx, field, 16).commit(mats)
    exp = O.MerkleTree(p2, [m.astype(np.uint64) for m in mats])
    tw = sum(w for _, w in shapes)
    idx = np.array([5, 1023, 5, 0, 512], dtype=np.uint64)  # a repeated index
    rows, paths = tree.open_mmcs(idx, tw)
    for k, i in enumerate(idx):
        i = int(i)
        split = np.cumsum([w for _, w in shapes])[:-1]
        assert exp.verify(root.astype(np.uint64), i, np.split(rows[k].astype(np.uint64), split),
                          paths[k].astype(np.uint64))
    assert (rows[0] == rows[2]).all() and (paths[0] == paths[2]).all()
    lib, n = ctx.lib, idx.size
    r_only = np.empty_like(rows)
    assert lib.zkg_mmcs_open(tree.h, _p(idx), n, _p(r_only), None) == 0
    assert (r_only == rows).all()
    p_only = np.empty_like(paths)
    assert lib.zkg_mmcs_open(tree.h, _p(idx), n, None, _p(p_only)) == 0
    assert (p_only == paths).all()
    assert lib.zkg_mmcs_open(tree.h, _p(idx), 0, _p(r_only), _p(p_only)) == 0
    bad = np.array([1024], dtype=np.uint64)
    assert lib.zkg_mmcs_open(tree.h, _p(bad), 1, _p(r_only), _p(p_only)) != 0
    tree.free()


@pytest.mark.parametrize("field", [GL, BB])
def test_fri_open_partial_outputs(ctx, field):
    log_len, log_final = 12, 4
    nl = log_len - log_final
    cw = O.gen(field, 91, 1 << log_len).astype(DT[field])
    betas = O.gen(field, 92, nl)
    buf = ctx.upload(cw)
    fc = Z.HipFri(ctx).commit_phase_base(field, buf, log_len, log_final, betas)
    idx = np.array([7, (1 << log_len) - 1, 7, 2048], dtype=np.uint64)
    rows, paths = fc.open(idx)
    assert (rows[0] == rows[2]).all() and (paths[0] == paths[2]).all()
    lib, n = ctx.lib, idx.size
    r_only = np.empty_like(rows)
    assert lib.zkg_fri_open(fc.h, _p(idx), n, _p(r_only), None) == 0
    assert (r_only == rows).all()
    p_only = np.empty_like(paths)
    assert lib.zkg_fri_open(fc.h, _p(idx), n, None, _p(p_only)) == 0
    assert (p_only == paths).all()
    assert lib.zkg_fri_open(fc.h, _p(idx), 0, None, None) == 0
    bad = np.array([1 << log_len], dtype=np.uint64)
    assert lib.zkg_fri_open(fc.h, _p(bad), 1, _p(r_only), _p(p_only)) != 0
    # every opening verifies against the oracle's layer trees
    p2 = O.Poseidon2(field, 16)
    roots_o, _ = O.fri_commit_phase_base(p2, cw, log_len, log_final, betas)
    layers = [cw.astype(np.uint64)]
    for r in range(nl - 1):
        layers.append(O.fri_fold_matrix_base(field, layers[-1], int(betas[r])))
    for k, qi in enumerate(idx):
        off = 0
        for r in range(nl):
            leaf = int(qi) >> (r + 1)
            plen = log_len - 1 - r
            assert O.verify(p2, roots_o[r], [1 << (log_len - 1 - r)], [2], leaf, [rows[k, r]],
                            paths[k, off:off + plen])
            off += plen
    fc.free()
    buf.free()
