"""GPU parity of the base-field FRI commit phase (zkg_fri_commit_phase_base):
Goldilocks, the reference STARK's configuration (Challenge = Val), and
BabyBear with base-field challenges.  Every layer root, the final codeword and
query openings (rows and sibling paths, verified against the oracle's trees)
are compared with the oracle bit for bit."""
import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}


@pytest.mark.parametrize("field", [GL, BB])
@pytest.mark.parametrize("log_len,log_final", [(2, 1), (10, 3), (16, 8), (20, 8)])
def test_fri_commit_phase_base_vs_oracle(ctx, field, log_len, log_final):
    p2 = O.Poseidon2(field, 16)
    cw = O.gen(field, 0x5EED0040 + log_len, 1 << log_len)
    nl = log_len - log_final
    betas = O.gen(field, 0x5EED0041, nl)
    buf = ctx.upload(cw.astype(DT[field]))
    fc = Z.HipFri(ctx).commit_phase_base(field, buf, log_len, log_final, betas)
    roots_o, final_o = O.fri_commit_phase_base(p2, cw, log_len, log_final, betas)
    assert fc.num_layers == nl
    assert (fc.layer_roots().astype(np.uint64) == roots_o).all()
    assert (fc.final().astype(np.uint64) == final_o).all()
    layers = [cw.astype(np.uint64)]
    for r in range(nl - 1):
        layers.append(O.fri_fold_matrix_base(field, layers[-1], int(betas[r])))
    trees = [O.MerkleTree(p2, [layers[r].reshape(-1, 2)]) for r in range(nl)]
    q = np.unique(np.array([0, 1, (1 << log_len) - 1, 12345 % (1 << log_len), 77 % (1 << log_len)],
                           dtype=np.uint64))
    rows, paths = fc.open(q)
    for k, qi in enumerate(q):
        off = 0
        for r in range(nl):
            leaf = int(qi) >> (r + 1)
            mat = layers[r].reshape(-1, 2)
            assert (rows[k, r].astype(np.uint64) == mat[leaf]).all(), (k, r)
            plen = log_len - 1 - r
            path = paths[k, off:off + plen].astype(np.uint64)
            off += plen
            assert (path == trees[r].path(leaf)).all(), (k, r)
            assert trees[r].verify(roots_o[r], leaf, [mat[leaf]], path)
    fc.free()
    buf.free()


def test_fri_commit_phase_base_rejects(ctx):
    buf = ctx.upload(O.gen(GL, 1, 1 << 8).astype(np.uint64))
    fri = Z.HipFri(ctx)
    with pytest.raises(Z.ZkgError):
        fri.commit_phase_base(GL, buf, 8, 8, np.zeros(1, dtype=np.uint64))  # no fold
    with pytest.raises(Z.ZkgError):
        fri.commit_phase_base(GL, buf, 9, 3, np.zeros(6, dtype=np.uint64))  # buffer too small
    with pytest.raises(Z.ZkgError):
        fri.commit_phase_base(GL, buf, 8, 3, np.full(5, O.GL_P, dtype=np.uint64))  # beta not canonical
    buf.free()
