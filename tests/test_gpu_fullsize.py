"""Full-size parity at the benchmark shapes.

Config 2 (BabyBear 2^20 x 64, blowup 2, Poseidon2-16): the GPU root equals the
oracle's root bit for bit, 64 random authentication paths verify, and the
commit is deterministic.  Config 4 (FRI over EF4, 2^22 -> 2^8): every layer
root and the final codeword equal the oracle's, 80 query openings verify.  The
same phase over Goldilocks with base-field challenges (the reference STARK's
configuration), in one call and in transcript order.
Inputs come from the seeded generator on the GPU; the oracle regenerates the
same stream on the CPU.
"""
import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu


def test_config2_root_and_paths(ctx):
    n, w, lb = 1 << 20, 64, 1
    seed = 0x5EED0002
    tr = ctx.random(Z.BABYBEAR, n * w, seed)
    tree, lde = Z.commit_trace(ctx, Z.BABYBEAR, tr, n, w, lb)
    root = tree.root().astype(np.uint64)
    tree2, lde2 = Z.commit_trace(ctx, Z.BABYBEAR, tr, n, w, lb)
    assert (tree2.root().astype(np.uint64) == root).all()
    p2 = O.Poseidon2(O.BB, 16)
    root_o, _ = O.commit(p2, O.gen(O.BB, seed, n * w).reshape(n, w), lb)
    assert (root == root_o).all()
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, n << lb, 64).astype(np.uint64)
    rows, paths = tree.open_batch(idx, w)
    for k, i in enumerate(idx):
        assert O.verify(p2, root, [n << lb], [w], int(i), [rows[k]], paths[k])
    bad = rows[0].astype(np.uint64).copy()
    bad[3] = (bad[3] + 1) % O.BB_P
    assert not O.verify(p2, root, [n << lb], [w], int(idx[0]), [bad], paths[0])


def test_config4_fri_commit_phase(ctx):
    log_len, log_final = 22, 8
    seed = 0x5EED0004
    # 2^20 EF4 coefficients at blowup 4 (coset LDE of 4 base columns)
    coeff = ctx.random(Z.BABYBEAR, (1 << 20) * 4, seed)
    cw = ctx.alloc((1 << log_len) * 16)
    assert ctx.lib.zkg_coset_lde_batch(ctx.h, Z.BABYBEAR, coeff.h, 1 << 20, 4, 2, 31, cw.h) == 0
    betas = O.fri_betas(seed, log_len - log_final)
    fc = Z.HipFri(ctx).commit_phase(cw, log_len, log_final, betas.astype(np.uint32))
    p2 = O.Poseidon2(O.BB, 16)
    cw_o = O.coset_lde_batch(O.BB, O.gen(O.BB, seed, (1 << 20) * 4).reshape(-1, 4), 2, 31)
    cw_gpu = cw.download(np.uint32, (1 << log_len) * 4).reshape(-1, 4)
    assert (cw_gpu.astype(np.uint64) == cw_o).all()
    roots_o, final_o = O.fri_commit_phase_ef4(p2, cw_o, log_len, log_final, betas)
    assert (fc.layer_roots().astype(np.uint64) == roots_o).all()
    assert (fc.final().astype(np.uint64) == final_o).all()
    q = np.random.default_rng(seed).integers(0, 1 << log_len, 80).astype(np.uint64)
    rows, paths = fc.open(q)
    # every opening verifies against its layer root
    nl = log_len - log_final
    for k in range(80):
        off = 0
        for r in range(nl):
            plen = log_len - 1 - r
            leaf = int(q[k]) >> (r + 1)
            assert O.verify(p2, roots_o[r], [1 << (log_len - 1 - r)], [8], leaf, [rows[k, r]],
                            paths[k, off:off + plen])
            off += plen


@pytest.mark.parametrize("stepwise", [False, True])
def test_goldilocks_fri_commit_phase_full_size(ctx, stepwise):
    log_len, log_final = 22, 8
    seed = 0x5EED0044
    nl = log_len - log_final
    # 2^20 Goldilocks coefficients at blowup 4 (coset LDE of one column)
    coeff = ctx.random(Z.GOLDILOCKS, 1 << 20, seed)
    cw = ctx.alloc((1 << log_len) * 8)
    assert ctx.lib.zkg_coset_lde_batch(ctx.h, Z.GOLDILOCKS, coeff.h, 1 << 20, 1, 2, 7, cw.h) == 0
    betas = O.gen(Z.GOLDILOCKS, seed + 1, nl)
    fri = Z.HipFri(ctx)
    if stepwise:
        fc = fri.begin(cw, log_len, log_final, field=Z.GOLDILOCKS)
        for r in range(nl):
            fc.fold_next(int(betas[r]))
    else:
        fc = fri.commit_phase_base(Z.GOLDILOCKS, cw, log_len, log_final, betas)
    p2 = O.Poseidon2(O.GL, 16)
    cw_o = O.coset_lde_batch(O.GL, O.gen(O.GL, seed, 1 << 20).reshape(-1, 1), 2, 7).reshape(-1)
    assert (cw.download(np.uint64, 1 << log_len) == cw_o).all()
    roots_o, final_o = O.fri_commit_phase_base(p2, cw_o, log_len, log_final, betas)
    assert (fc.layer_roots() == roots_o).all()
    assert (fc.final() == final_o).all()
    q = np.random.default_rng(seed).integers(0, 1 << log_len, 80).astype(np.uint64)
    rows, paths = fc.open(q)
    for k in range(80):
        off = 0
        for r in range(nl):
            plen = log_len - 1 - r
            leaf = int(q[k]) >> (r + 1)
            assert O.verify(p2, roots_o[r], [1 << (log_len - 1 - r)], [2], leaf, [rows[k, r]],
                            paths[k, off:off + plen])
            off += plen
    fc.free()
    cw.free()
    coeff.free()
