"""Oracle Merkle tree: p3 layer layout, padding rules and openings."""
import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def bb():
    return O.Poseidon2(O.BB, 16)


@pytest.fixture(scope="module")
def gl():
    return O.Poseidon2(O.GL, 16)


def test_power_of_two_layers(bb):
    m = O.gen(O.BB, 1, 64 * 10).reshape(64, 10)
    t = O.MerkleTree(bb, [m])
    assert [len(l) for l in t.layers] == [64, 32, 16, 8, 4, 2, 1]
    assert (t.layers[0][5] == bb.hash(m[5])).all()
    assert (t.layers[3][2] == bb.compress(t.layers[2][4], t.layers[2][5])).all()


def test_padding_rules_from_reference(bb):
    """vendor/p3-merkle-tree-gpu/src/merkle_tree.rs:427-506: layers are padded
    with zero digests to an even length unless the next one is the root."""
    # height 6: digest layer lengths 6 -> 4 (3 real + zero) -> 2 -> 1
    m = O.gen(O.BB, 2, 6 * 3).reshape(6, 3)
    t = O.MerkleTree(bb, [m])
    assert [len(l) for l in t.layers] == [6, 4, 2, 1]
    assert (t.layers[1][3] == 0).all() and not (t.layers[1][2] == 0).all()
    # height 3: first layer padded to 4 with a zero digest
    t3 = O.MerkleTree(bb, [m[:3]])
    assert [len(l) for l in t3.layers] == [4, 2, 1]
    assert (t3.layers[0][3] == 0).all()
    # height 1: the single row digest is the root
    t1 = O.MerkleTree(bb, [m[:1]])
    assert len(t1.layers) == 1 and (t1.root() == bb.hash(m[0])).all()
    # heights 2: root formation case, one compression
    t2 = O.MerkleTree(bb, [m[:2]])
    assert [len(l) for l in t2.layers] == [2, 1]


def test_mixed_heights_inject_and_verify(bb):
    a = O.gen(O.BB, 3, 16 * 4).reshape(16, 4)
    b = O.gen(O.BB, 4, 4 * 9).reshape(4, 9)
    c = O.gen(O.BB, 5, 16 * 2).reshape(16, 2)
    t = O.MerkleTree(bb, [b, a, c])
    # tallest matrices are hashed together, row by row (a then c in input order)
    assert (t.layers[0][3] == bb.hash(np.concatenate([a[3], c[3]]))).all()
    # at the layer of length 4 the small matrix rows are injected
    lvl = bb.compress(t.layers[1][2], t.layers[1][3])
    assert (t.layers[2][1] == bb.compress(lvl, bb.hash(b[1]))).all()
    for idx in range(16):
        rows = [b[idx >> 2], a[idx], c[idx]]
        assert t.verify(t.root(), idx, rows, t.path(idx))
        bad = [r.copy() for r in rows]
        bad[1][0] = (int(bad[1][0]) + 1) % O.BB_P
        assert not t.verify(t.root(), idx, bad, t.path(idx))


def test_incompatible_heights_rejected(bb):
    with pytest.raises(ValueError):
        O.MerkleTree(bb, [np.zeros((8, 1), np.uint64), np.zeros((7, 1), np.uint64)])


@pytest.mark.parametrize("f,w", [(O.GL, 16), (O.GL, 8), (O.BB, 16)])
def test_path_verify_roundtrip(f, w):
    p2 = O.Poseidon2(f, w)
    m = O.gen(f, 9, 32 * 13).reshape(32, 13)
    t = O.MerkleTree(p2, [m])
    for idx in (0, 1, 17, 31):
        assert t.verify(t.root(), idx, [m[idx]], t.path(idx))
        assert not t.verify(t.root(), idx ^ 1, [m[idx]], t.path(idx))


def test_commit_matches_lde_then_tree(gl):
    tr = O.gen(O.GL, 21, 64 * 8).reshape(64, 8)
    root, lde = O.commit(gl, tr, 1)
    assert (lde == O.coset_lde_batch(O.GL, tr, 1, 7)).all()
    assert (root == O.MerkleTree(gl, [lde]).root()).all()
