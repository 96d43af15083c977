"""The batched and served commit workloads at their own shapes.

* Batched: the bench's draw of 48 BabyBear traces (heights 2^U{12..20}, widths
  U{8..256}, blowup 2).  With heights capped at 2^16 every root is checked
  against the oracle, on the captured graph and on its replay.  At full size
  (1.59e9 cells) every root equals the lone fused commit's root, and the widest
  full-height trace (2^20 x 256) is checked against the oracle on its own.
* Served: the bench's request mix (60% BabyBear / 40% Goldilocks, widths
  16..128, blowup 2 or 4) with heights capped at 2^16, plus one Goldilocks and
  one BabyBear request of 2^22 rows at blowup 4: 2^24-row LDEs, the largest
  the resident twiddle tables cover.  Every result is checked against the
  oracle.

Traces are generated on the device (zkg_fill_random) and on the host by the
oracle's copy of the same element stream.
"""
import time

import numpy as np
import pytest

import oracle as O
import ziros_amd as Z
from ziros_amd import workloads as W

pytestmark = pytest.mark.gpu

GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
ES = {GL: 8, BB: 4}


def _device_trace(ctx, field, n, w, seed):
    return ctx.random(field, n * w, seed)


def _oracle_root(field, n, w, lb, seed):
    tr = O.gen(field, seed, n * w).reshape(n, w)
    return O.commit(O.Poseidon2(field, 16, 42), tr, lb)[0]


def test_batch_draw_capped_at_2p16_vs_oracle(ctx):
    shapes = [(min(lg, 16), w) for lg, w in W.batch_shapes()]
    jobs, keep = [], []
    for k, (lg, w) in enumerate(shapes):
        n = 1 << lg
        tr = _device_trace(ctx, BB, n, w, W.SEED_BATCH + k)
        lde = ctx.alloc((n << 1) * w * 4)
        jobs.append(Z.make_job(BB, tr, n, w, 1, lde))
        keep += [tr, lde]
    expect = [_oracle_root(BB, 1 << lg, w, 1, W.SEED_BATCH + k) for k, (lg, w) in enumerate(shapes)]
    for rep in range(2):  # capture, then a replay of the cached graph
        trees = Z.commit_batch(ctx, jobs)
        for k, (t, e) in enumerate(zip(trees, expect)):
            assert (t.root().astype(np.uint64) == e).all(), (rep, k, shapes[k])
            t.free()


def test_batch_draw_full_size_equals_lone_commits(ctx):
    shapes = W.batch_shapes()
    jobs, bufs = [], []
    for k, (lg, w) in enumerate(shapes):
        n = 1 << lg
        tr = _device_trace(ctx, BB, n, w, W.SEED_BATCH + k)
        lde = ctx.alloc((n << 1) * w * 4)
        jobs.append(Z.make_job(BB, tr, n, w, 1, lde))
        bufs.append((tr, lde))
    roots = []
    for rep in range(2):
        trees = Z.commit_batch(ctx, jobs)
        r = [t.root().astype(np.uint64) for t in trees]
        for t in trees:
            t.free()
        if rep:
            assert all((a == b).all() for a, b in zip(roots, r))
        roots = r
    for k, ((lg, w), (tr, lde)) in enumerate(zip(shapes, bufs)):
        tree, _ = Z.commit_trace(ctx, BB, tr, 1 << lg, w, 1)
        assert (tree.root().astype(np.uint64) == roots[k]).all(), (k, lg, w)
        tree.free()
    for tr, lde in bufs:
        tr.free()
        lde.free()


def test_widest_full_height_batch_trace_vs_oracle(ctx):
    shapes = W.batch_shapes()
    k = max(range(len(shapes)), key=lambda i: (shapes[i][0], shapes[i][1]))
    lg, w = shapes[k]
    assert (lg, w) == (20, 256)
    n = 1 << lg
    tr = _device_trace(ctx, BB, n, w, W.SEED_BATCH + k)
    tree, lde = Z.commit_trace(ctx, BB, tr, n, w, 1)
    got = tree.root().astype(np.uint64)
    tree.free()
    lde.free()
    tr.free()
    assert (got == _oracle_root(BB, n, w, 1, W.SEED_BATCH + k)).all()


def test_served_mix_with_2p24_row_ldes_vs_oracle(ctx):
    reqs, _ = W.serve_requests(48)
    mix = [(f, min(lg, 16), w, lb) for f, lg, w, lb in reqs[:14]]
    mix += [(GL, 22, 16, 2), (BB, 22, 34, 2)]
    assert {f for f, *_ in mix} == {GL, BB} and {lb for *_, lb in mix} == {1, 2}
    jobs, keep = [], []
    for k, (f, lg, w, lb) in enumerate(mix):
        n = 1 << lg
        tr = _device_trace(ctx, f, n, w, W.SEED_SERVE + k)
        lde = ctx.alloc((n << lb) * w * ES[f])
        jobs.append(Z.make_job(f, tr, n, w, lb, lde))
        keep += [tr, lde]
    srv = Z.Server(ctx, slots=2, max_batch=8)
    try:
        tickets = [srv.submit(j) for j in jobs]
        done = {}
        t0 = time.time()
        while len(done) < len(tickets) and time.time() - t0 < 60:
            for i, t in enumerate(tickets):
                if i not in done:
                    r = srv.poll(t)
                    if r is not None:
                        done[i] = r
            time.sleep(0.001)
        assert len(done) == len(tickets)
        roots = {}
        for i, (h, _) in done.items():
            t = Z.MerkleTree(ctx, h, mix[i][0])
            roots[i] = t.root().astype(np.uint64)
            t.free()
    finally:
        srv.stop()
    for b in keep:
        b.free()
    for i, (f, lg, w, lb) in enumerate(mix):
        assert (roots[i] == _oracle_root(f, 1 << lg, w, lb, W.SEED_SERVE + i)).all(), (i, mix[i])
