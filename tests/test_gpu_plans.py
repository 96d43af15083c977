"""Every transform tile plan gives the oracle's values (GPU).  The plans
differ in radix split, tile-store unrolling and pass-B occupancy; a transform's
plan is chosen per shape at run time, so each plan must be exact on every tile
size it can meet.  Each test forces one plan through the override layer."""
import numpy as np
import pytest

import oracle as O
import ziros_amd as Z
from ziros_amd import host as H

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}

# in-LDS tiles of 2^1 .. 2^11 points: single-pass sizes and both four-step passes
LDE_SHAPES = [(1, 1, 3), (4, 2, 7), (8, 1, 5), (9, 1, 33), (11, 1, 16), (12, 1, 8), (13, 2, 20), (15, 1, 9),
              (16, 1, 4), (18, 1, 3), (20, 1, 2), (22, 1, 1)]
NTT_SHAPES = [(3, 5), (10, 7), (13, 33), (17, 4), (21, 1)]


@pytest.fixture
def forced():
    def force(plan):
        H.set_lde_plans(H.TUNE_OVERRIDE, [(-1, -1, -1, -1, plan)])
    yield force
    H.set_lde_plans(H.TUNE_OVERRIDE, [])
    H.set_lde_plans(H.TUNE_LEARNED, [])


@pytest.mark.parametrize("plan", range(4))
@pytest.mark.parametrize("field", [GL, BB])
def test_coset_lde_every_plan(ctx, forced, plan, field):
    forced(plan)
    for logn, lb, width in LDE_SHAPES:
        n = 1 << logn
        assert H.lde_plan_for(field, n, width, lb) == (plan, "override")
        m = O.gen(field, 500 + logn, n * width).astype(DT[field]).reshape(n, width)
        got = Z.HipDft(ctx, field).coset_lde_batch(m, lb)
        assert (got.astype(np.uint64) == O.coset_lde_batch(field, m, lb, Z.GENERATOR[field])).all(), (logn, lb, width)


@pytest.mark.parametrize("plan", range(4))
@pytest.mark.parametrize("field", [GL, BB])
def test_ntt_every_plan(ctx, forced, plan, field):
    forced(plan)
    dft = Z.HipDft(ctx, field)
    for logn, width in NTT_SHAPES:
        n = 1 << logn
        m = O.gen(field, 600 + logn, n * width).astype(DT[field]).reshape(n, width)
        assert (dft.dft_batch(m).astype(np.uint64) == O.dft_batch(field, m)).all(), (logn, width)
        assert (dft.idft_batch(m).astype(np.uint64) == O.dft_batch(field, m, inverse=True)).all(), (logn, width)


@pytest.mark.parametrize("plan", range(4))
def test_commit_every_plan(ctx, forced, plan):
    forced(plan)
    for field, logn, width, lb in ((BB, 13, 24, 1), (GL, 12, 16, 2)):
        n = 1 << logn
        tr = O.gen(field, 700 + logn, n * width).reshape(n, width)
        tree, _ = Z.commit_trace(ctx, field, ctx.upload(tr.astype(DT[field])), n, width, lb)
        assert (tree.root().astype(np.uint64) == O.commit(O.Poseidon2(field, 16, 42), tr, lb)[0]).all()


def test_learned_plan_is_recorded_and_used(ctx, forced):
    H.set_lde_plans(H.TUNE_OVERRIDE, [])
    p = H.learn_lde_plan(ctx, BB, 1 << 16, 64, 1)
    assert 0 <= p < len(H.lde_plans())
    assert H.lde_plan_for(BB, 1 << 16, 64, 1) == (p, "learned")
    m = O.gen(BB, 77, (1 << 16) * 64).astype(np.uint32).reshape(1 << 16, 64)
    got = Z.HipDft(ctx, BB).coset_lde_batch(m, 1)
    assert (got.astype(np.uint64) == O.coset_lde_batch(BB, m, 1, 31)).all()


def test_split_overrides_keep_results(ctx):
    """The split levels only choose the work form: any setting gives the same tree."""
    try:
        for field, p2w, lanes, fused in ((GL, 16, 11, 2), (GL, 8, 16, 4), (BB, 16, 12, 5)):
            H.set_splits(H.TUNE_OVERRIDE, [(field, p2w, lanes, -1, -1, fused)])
            n, w = 1 << 13, 12
            tr = O.gen(field, 800 + p2w, n * w).reshape(n, w)
            tree, _ = Z.commit_trace(ctx, field, ctx.upload(tr.astype(DT[field])), n, w, 1, p2_width=p2w)
            assert (tree.root().astype(np.uint64) == O.commit(O.Poseidon2(field, p2w, 42), tr, 1)[0]).all()
    finally:
        H.set_splits(H.TUNE_OVERRIDE, [])
