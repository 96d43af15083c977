"""Tuning layers (CPU): the transform plan and the hash / Merkle split levels
resolve override > learned > static table > built-in, as the reference's
device tuning does (zkf-metal/src/tuning.rs:349-365, its test
runtime_override_beats_learned_and_static at :458-480)."""
import pytest

from ziros_amd import host as H
from ziros_amd._lib import ZkgError

GL, BB = H.GOLDILOCKS, H.BABYBEAR


@pytest.fixture(autouse=True)
def clean_layers():
    for layer in (H.TUNE_OVERRIDE, H.TUNE_LEARNED):
        H.set_lde_plans(layer, [])
        H.set_splits(layer, [])
    yield
    for layer in (H.TUNE_OVERRIDE, H.TUNE_LEARNED):
        H.set_lde_plans(layer, [])
        H.set_splits(layer, [])


def test_plans_are_listed():
    names = H.lde_plans()
    assert len(names) == 4 and len(set(names)) == 4


def test_override_beats_learned_beats_static():
    base, src = H.lde_plan_for(BB, 1 << 20, 64, 1)
    assert src in ("static", "default")
    H.set_lde_plans(H.TUNE_LEARNED, [(BB, 20, 6, 1, 1)])
    assert H.lde_plan_for(BB, 1 << 20, 64, 1) == (1, "learned")
    assert H.lde_plan_for(BB, 1 << 20, 65, 1)[1] != "learned"  # width class 7, not 6
    H.set_lde_plans(H.TUNE_OVERRIDE, [(BB, 20, 6, 1, 2)])
    assert H.lde_plan_for(BB, 1 << 20, 64, 1) == (2, "override")
    H.set_lde_plans(H.TUNE_OVERRIDE, [])
    assert H.lde_plan_for(BB, 1 << 20, 64, 1) == (1, "learned")
    H.set_lde_plans(H.TUNE_LEARNED, [])
    assert H.lde_plan_for(BB, 1 << 20, 64, 1) == (base, src)


def test_wildcards_match_any_shape():
    H.set_lde_plans(H.TUNE_OVERRIDE, [(-1, -1, -1, -1, 3)])
    for f, h, w, lb in ((GL, 1 << 10, 3, 0), (BB, 1 << 22, 256, 2)):
        assert H.lde_plan_for(f, h, w, lb) == (3, "override")
    H.set_lde_plans(H.TUNE_OVERRIDE, [(GL, -1, -1, 2, 1)])
    assert H.lde_plan_for(GL, 1 << 16, 32, 2) == (1, "override")
    assert H.lde_plan_for(GL, 1 << 16, 32, 1)[1] != "override"
    assert H.lde_plan_for(BB, 1 << 16, 32, 2)[1] != "override"


def test_bad_entries_are_refused():
    with pytest.raises(ZkgError):
        H.set_lde_plans(H.TUNE_OVERRIDE, [(BB, 20, 6, 1, len(H.lde_plans()))])
    with pytest.raises(ZkgError):
        H.set_lde_plans(7, [])
    with pytest.raises(ZkgError):
        H.set_splits(H.TUNE_OVERRIDE, [(BB, 16, 31, -1, -1, -1)])
    with pytest.raises(ZkgError):
        H.set_splits(H.TUNE_OVERRIDE, [(BB, 16, -1, -1, -1, 10)])


def test_split_values_fall_through_layer_by_layer():
    static = H.splits(GL, 16)  # the static table's values (or the built-in ones)
    assert all(v > 0 for v in static.values())
    H.set_splits(H.TUNE_LEARNED, [(GL, 16, 14, 18, -1, -1)])
    H.set_splits(H.TUNE_OVERRIDE, [(-1, -1, 12, -1, -1, 4)])
    got = H.splits(GL, 16)
    # lanes and fused from the override, chunks from the learned layer, hash chunks from below
    assert got == dict(merkle_lanes_log2=12, merkle_chunks_log2=18, hash_chunks_log2=static["hash_chunks_log2"],
                       fused_levels=4)
    assert H.splits(BB, 16)["merkle_lanes_log2"] == 12
    gl8 = H.splits(GL, 8)["merkle_chunks_log2"]
    H.set_splits(H.TUNE_LEARNED, [])
    assert H.splits(GL, 8)["merkle_chunks_log2"] == gl8  # the learned entry was for width 16 only
