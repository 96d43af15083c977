"""Oracle Poseidon2 pinned by the reference golden vector and closed forms."""
import numpy as np
import pytest

import oracle as O

# zkf-backends/src/wrapping/poseidon2_goldilocks.rs:546-564: Poseidon2Goldilocks<16>
# from SmallRng::seed_from_u64(42), input [12345, 67890, 0 x 14], p3-goldilocks 0.4.2.
GOLDEN_GL16 = [
    0xdcfb2f7d8d417b5b, 0x6876895bf9575193, 0x90bd412a43f20077, 0x074f4e474e56c486,
    0x1f83bf002930935b, 0x75c3419055d75624, 0x2f63a8055d62f0b3, 0xf3dd5e4ba55f7e63,
    0x9c72c243250cf50e, 0x206df890ef41a718, 0xbd7f02f242e585b7, 0x3e1249b7c2a0a12e,
    0xc59030e61bfe32bb, 0x47f2231a2b674637, 0x112a64e1d2c5aea4, 0xa02dd5cfce1ceaf5,
]
# same file :566-587, the LE byte serialisation of the golden state
GOLDEN_GL16_LE = bytes.fromhex(
    "5b7b418d7d2ffbdc935157f95b89766877 00f2432a41bd9086c4564e474e4f075b93302900bf831f"
    "2456d7559041c375b3f0625d05a8632f637e5fa54b5eddf30ef50c2543c2729c18a741ef90f86d20"
    "b785e542f2027fbd2ea1a0c2b749123ebb32fe1be63090c53746672b1a23f247a4aec5d2e1642a11"
    "f5ea1ccecfd52da0".replace(" ", ""))


def test_goldilocks16_golden_vector():
    p2 = O.Poseidon2(O.GL, 16, seed=42)
    out = p2.permute([12345, 67890] + [0] * 14)
    assert [int(x) for x in out] == GOLDEN_GL16
    assert out.astype("<u8").tobytes() == GOLDEN_GL16_LE


def test_golden_vector_fails_on_planted_mistakes():
    """The vector pins the RNG seed, the constant order and the field."""
    good = O.Poseidon2(O.GL, 16, seed=42).permute([12345, 67890] + [0] * 14)
    assert [int(x) for x in O.Poseidon2(O.GL, 16, seed=43).permute([12345, 67890] + [0] * 14)] != GOLDEN_GL16
    # swapping the initial/terminal constant blocks changes the output
    p2 = O.Poseidon2(O.GL, 16, seed=42)
    rc = p2.flat_constants()
    half = 4 * 16
    swapped = np.concatenate([rc[-half:], rc[half:-half], rc[:half]])
    for i, v in enumerate(swapped):
        p2.p.rc[i] = int(v)
    assert [int(x) for x in p2.permute([12345, 67890] + [0] * 14)] != [int(x) for x in good]


def test_constant_counts():
    # zkf-metal/src/poseidon2/goldilocks.rs:89-98 and babybear.rs:121-130:
    # rc.len() == rounds_f * WIDTH + rounds_p
    for f, w, rf, rp in [(O.GL, 16, 8, 22), (O.BB, 16, 8, 13), (O.GL, 8, 8, 22)]:
        p2 = O.Poseidon2(f, w)
        assert (p2.p.rounds_f, p2.p.rounds_p) == (rf, rp)
        assert p2.flat_constants().size == rf * w + rp
        assert (p2.flat_constants() < O.PRIME[f]).all()


def test_babybear_diag_reference_values():
    # zkf-metal/src/poseidon2/babybear.rs:139-157
    d = O.Poseidon2(O.BB, 16).diag
    p = O.BB_P
    assert d[1] == 1 and d[2] == 2 and d[4] == 3 and d[5] == 4
    assert d[0] == p - 2 and d[7] == p - 3 and d[8] == p - 4
    # closed forms of the remaining entries
    inv = lambda x: pow(x, p - 2, p)
    assert d[3] == inv(2) and d[6] == p - inv(2)
    assert d[9] == inv(256) and d[10] == inv(4) and d[11] == inv(8) and d[12] == inv(2**27)
    assert d[13] == p - inv(256) and d[14] == p - inv(16) and d[15] == p - inv(2**27)


def test_babybear_constant_convention_switch():
    """BabyBear constants are parity unpinned: the Montgomery interpretation of
    the sampled u31 (default) and the canonical one differ by 2^-32."""
    m = O.Poseidon2(O.BB, 16, bb_rc_mode=0).flat_constants()
    c = O.Poseidon2(O.BB, 16, bb_rc_mode=1).flat_constants()
    p = O.BB_P
    rinv = pow(2, -32, p)
    assert all(int(a) == int(b) * rinv % p for a, b in zip(m, c))


@pytest.mark.parametrize("f,w", [(O.GL, 16), (O.BB, 16), (O.GL, 8)])
def test_sponge_semantics(f, w):
    """PaddingFreeSponge (poseidon2_goldilocks.rs:372-398): overwrite mode,
    partial chunk keeps the previous rate values, empty input -> zero."""
    p2 = O.Poseidon2(f, w)
    r = w // 2
    assert (p2.hash([]) == 0).all()
    data = O.gen(f, 11, 3 * r + 2)
    # manual sponge with explicit permutations
    s = np.zeros(w, dtype=np.uint64)
    for k in range(0, len(data), r):
        chunk = data[k:k + r]
        s[: len(chunk)] = chunk
        s = p2.permute(s)
    assert (p2.hash(data) == s[:r]).all()
    # an exact multiple of the rate does not permute again
    s = np.zeros(w, dtype=np.uint64)
    s[:r] = data[:r]
    s = p2.permute(s)
    assert (p2.hash(data[:r]) == s[:r]).all()
    # compression = permute(l || r)[:r]
    lft, rgt = data[:r], data[r:2 * r]
    assert (p2.compress(lft, rgt) == p2.permute(np.concatenate([lft, rgt]))[:r]).all()


def test_permutation_is_bijective_sample():
    p2 = O.Poseidon2(O.BB, 16)
    seen = set()
    for i in range(200):
        s = np.zeros(16, dtype=np.uint64)
        s[0] = i
        seen.add(tuple(p2.permute(s)))
    assert len(seen) == 200
