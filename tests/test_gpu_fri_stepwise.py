"""GPU parity of the layer-by-layer FRI commit phase (zkg_fri_begin /
zkg_fri_fold_next) in transcript order: each beta_r is drawn from layer r's
root as the GPU returns it, before the fold that needs it.  The oracle, run
with the betas so drawn, reproduces every root (so each beta also follows from
the oracle's own root), the final codeword and the openings, bit for bit; the
one-call commit phase with the same betas gives the same commitment.  EF4 over
BabyBear, Goldilocks with base-field challenges (the reference STARK's
configuration) and BabyBear with base-field challenges."""
import ctypes

import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
EF4 = "ef4"
DT = {GL: np.uint64, BB: np.uint32}
P = {GL: O.GL_P, BB: O.BB_P}


def draw_beta(kind, root, r):
    """A stand-in for the challenger: FNV-style mixing of the root words and
    the layer index, reduced to canonical challenge words."""
    h = 0xCBF29CE484222325 ^ r
    for w in root:
        h = ((h ^ int(w)) * 0x100000001B3) % (1 << 64)
    if kind == EF4:
        out = []
        for i in range(4):
            h = ((h ^ (i + 1)) * 0x100000001B3) % (1 << 64)
            out.append(h % O.BB_P)
        return np.array(out, dtype=np.uint64)
    return h % P[kind]


def codeword(kind, log_len, seed):
    if kind == EF4:
        return O.gen(BB, seed, (1 << log_len) * 4).reshape(-1, 4)
    return O.gen(kind, seed, 1 << log_len)


def run_stepwise(ctx, kind, cw, log_len, log_final):
    buf = ctx.upload(cw.astype(np.uint32 if kind == EF4 else DT[kind]))
    fc = Z.HipFri(ctx).begin(buf, log_len, log_final, field=None if kind == EF4 else kind)
    betas, roots = [], []
    for r in range(log_len - log_final):
        roots.append(fc.layer_root(r).astype(np.uint64))
        b = draw_beta(kind, roots[-1], r)
        betas.append(b)
        fc.fold_next(b)
    return fc, buf, np.array(betas, dtype=np.uint64), np.array(roots)


@pytest.mark.parametrize("kind", [EF4, GL, BB])
@pytest.mark.parametrize("log_len,log_final", [(2, 1), (10, 3), (16, 8), (20, 8)])
def test_fri_stepwise_transcript_vs_oracle(ctx, kind, log_len, log_final):
    nl = log_len - log_final
    cw = codeword(kind, log_len, 0x5EED0050 + log_len)
    fc, buf, betas, roots = run_stepwise(ctx, kind, cw, log_len, log_final)
    assert fc.num_layers == nl
    if kind == EF4:
        p2 = O.Poseidon2(BB, 16)
        roots_o, final_o = O.fri_commit_phase_ef4(p2, cw, log_len, log_final, betas.reshape(nl, 4))
    else:
        p2 = O.Poseidon2(kind, 16)
        roots_o, final_o = O.fri_commit_phase_base(p2, cw, log_len, log_final, betas)
    assert (roots == roots_o).all(), kind
    assert (fc.layer_roots().astype(np.uint64) == roots_o).all()
    assert (fc.final().astype(np.uint64) == final_o).all()
    # the one-call phase with the same betas: same roots, final and openings
    fri = Z.HipFri(ctx)
    if kind == EF4:
        one = fri.commit_phase(buf, log_len, log_final, betas.reshape(nl, 4).astype(np.uint32))
    else:
        one = fri.commit_phase_base(kind, buf, log_len, log_final, betas)
    assert (one.layer_roots() == fc.layer_roots()).all()
    assert (one.final() == fc.final()).all()
    q = np.unique(np.array([0, 1, (1 << log_len) - 1, 4321 % (1 << log_len)], dtype=np.uint64))
    rows_s, paths_s = fc.open(q)
    rows_1, paths_1 = one.open(q)
    assert (rows_s == rows_1).all() and (paths_s == paths_1).all()
    # openings verify against the oracle's trees
    if kind == EF4:
        layers = [cw.astype(np.uint64)]
        for r in range(nl - 1):
            layers.append(O.fri_fold_matrix_ef4(layers[-1], betas.reshape(nl, 4)[r]))
        mats = [layers[r].reshape(-1, 8) for r in range(nl)]
    else:
        layers = [cw.astype(np.uint64)]
        for r in range(nl - 1):
            layers.append(O.fri_fold_matrix_base(kind, layers[-1], int(betas[r])))
        mats = [layers[r].reshape(-1, 2) for r in range(nl)]
    trees = [O.MerkleTree(p2, [mats[r]]) for r in range(nl)]
    for k, qi in enumerate(q):
        off = 0
        for r in range(nl):
            leaf = int(qi) >> (r + 1)
            plen = log_len - 1 - r
            path = paths_s[k, off:off + plen].astype(np.uint64)
            off += plen
            assert (rows_s[k, r].astype(np.uint64) == mats[r][leaf]).all(), (k, r)
            assert trees[r].verify(roots_o[r], leaf, [mats[r][leaf]], path), (k, r)
    one.free()
    fc.free()
    buf.free()


def test_fri_stepwise_order_and_refusals(ctx):
    log_len, log_final = 8, 5
    cw = codeword(GL, log_len, 7)
    buf = ctx.upload(cw.astype(np.uint64))
    fc = Z.HipFri(ctx).begin(buf, log_len, log_final, field=GL)
    with pytest.raises(Z.ZkgError):
        fc.layer_root(1)  # not committed before the first fold
    with pytest.raises(Z.ZkgError):
        fc.final()  # phase not complete
    with pytest.raises(Z.ZkgError):
        fc.open(np.array([0], dtype=np.uint64))
    with pytest.raises(Z.ZkgError):
        fc.fold_next(O.GL_P)  # beta not canonical
    fc.fold_next(5)
    fc.layer_root(1)
    # constants reloaded mid-phase (the same seed: the session's tables stay as they were)
    ctx.load_constants(GL, 16, 42)
    with pytest.raises(Z.ZkgError):
        fc.fold_next(6)
    fc.free()
    fc = Z.HipFri(ctx).begin(buf, log_len, log_final, field=GL)
    for b in (1, 2, 3):
        fc.fold_next(b)
    with pytest.raises(Z.ZkgError):
        fc.fold_next(4)  # complete
    roots_o, final_o = O.fri_commit_phase_base(O.Poseidon2(GL, 16), cw, log_len, log_final,
                                               np.array([1, 2, 3], dtype=np.uint64))
    assert (fc.layer_roots().astype(np.uint64) == roots_o).all()
    assert (fc.final().astype(np.uint64) == final_o).all()
    fc.free()
    # a one-call phase cannot be stepped; EF4 needs BabyBear
    one = Z.HipFri(ctx).commit_phase_base(GL, buf, log_len, log_final, np.array([1, 2, 3], dtype=np.uint64))
    with pytest.raises(Z.ZkgError):
        one.fold_next(1)
    one.free()
    h = ctypes.c_void_p()
    assert ctx.lib.zkg_fri_begin(ctx.h, 1, GL, buf.h, log_len, log_final, ctypes.byref(h)) != 0
    buf.free()
