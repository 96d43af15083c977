"""Poseidon2 constants on the device: tables handed over as p3 builds them,
and what a reload may and may not touch.

* Injected constants.  The reference hands its accelerators p3's own round
  constants, flattened (zkf-metal/src/poseidon2/babybear.rs:76-111,
  goldilocks.rs:42-83).  A caller of libzkg does the same through
  zkg_constants_set.  Every commit path (single commit, mixed-height MMCS, FRI
  commit phase, batch) is checked bit for bit against the oracle built from
  the same table, for both readings of the BabyBear seed convention the oracle
  knows (bb_rc_mode 0 and 1), so the result does not depend on which reading
  p3-monty-31 uses.
* Reloads.  A cached batch graph holds the addresses of the constant tables it
  was captured with; a reload drops it.  The trait's explicit-constant
  permutation (and the dispatch driver's Poseidon batch and Merkle layer arms
  built on it) leaves the resident tables alone.  A reload while the serving
  loop runs is refused.
"""
import ctypes
import time

import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu

GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}
EINVAL, ESTATE = -1, -4


def rnd(field, n, seed):
    return O.gen(field, seed, n).astype(DT[field])


@pytest.fixture(scope="module")
def cctx():
    """A context of its own: these tests replace its constants."""
    c = Z.Context(0)
    yield c
    c.close()


def _check_commit(c, p2, field, logn, width, lb, seed):
    n = 1 << logn
    tr = rnd(field, n * width, seed).reshape(n, width)
    buf = c.upload(tr)
    tree, lde = Z.commit_trace(c, field, buf, n, width, lb, p2_width=p2.width)
    root_o, lde_o = O.commit(p2, tr, lb)
    exp = O.MerkleTree(p2, [lde_o])
    for i in range(len(exp.layers)):
        assert (tree.layer(i).astype(np.uint64) == exp.layers[i]).all(), i
    assert (tree.root().astype(np.uint64) == root_o).all()
    return root_o


def _check_mmcs(c, p2, field):
    shapes = [(4096, 5), (1000, 3), (256, 9)]
    mats = [rnd(field, h * w, 500 + k).reshape(h, w) for k, (h, w) in enumerate(shapes)]
    root, tree = Z.HipMerkleTreeMmcs(c, field, p2.width).commit(mats)
    exp = O.MerkleTree(p2, mats)
    for i in range(len(exp.layers)):
        assert (tree.layer(i).astype(np.uint64) == exp.layers[i]).all(), i
    assert (root.astype(np.uint64) == exp.root()).all()
    tree.free()


def _check_fri(c, p2, log_len=13, log_final=3):
    cw = rnd(BB, (1 << log_len) * 4, 77).reshape(-1, 4)
    betas = O.fri_betas(0xBE7A, log_len - log_final)
    fc = Z.HipFri(c).commit_phase(c.upload(cw), log_len, log_final, betas.astype(np.uint32))
    roots_o, final_o = O.fri_commit_phase_ef4(p2, cw, log_len, log_final, betas)
    assert (fc.layer_roots().astype(np.uint64) == roots_o).all()
    assert (fc.final().astype(np.uint64) == final_o).all()
    fc.free()


def _batch_jobs(c, specs, seed0):
    jobs, keep, mats = [], [], []
    for k, (field, p2w, logn, width, lb) in enumerate(specs):
        n = 1 << logn
        tr = rnd(field, n * width, seed0 + k).reshape(n, width)
        b = c.upload(tr)
        lde = c.alloc((n << lb) * width * np.dtype(DT[field]).itemsize)
        jobs.append(Z.make_job(field, b, n, width, lb, lde, p2w))
        keep += [b, lde]
        mats.append(tr)
    return jobs, keep, mats


@pytest.mark.parametrize("bb_rc_mode", [0, 1])
def test_injected_babybear_constants_on_every_commit_path(cctx, bb_rc_mode):
    p2 = O.Poseidon2(BB, 16, seed=0xC0FFEE, bb_rc_mode=bb_rc_mode)
    cctx.set_constants(BB, 16, p2.flat_constants())
    _check_commit(cctx, p2, BB, 12, 24, 1, 3)
    _check_commit(cctx, p2, BB, 10, 7, 2, 4)
    _check_mmcs(cctx, p2, BB)
    _check_fri(cctx, p2)
    specs = [(BB, 16, 11, 16, 1), (BB, 16, 9, 40, 2), (BB, 16, 13, 8, 1)]
    jobs, keep, mats = _batch_jobs(cctx, specs, 600)
    for rep in range(2):
        for t, m, (_, _, _, _, lb) in zip(Z.commit_batch(cctx, jobs), mats, specs):
            assert (t.root().astype(np.uint64) == O.commit(p2, m, lb)[0]).all(), rep


def test_the_two_babybear_readings_give_different_roots(cctx):
    """The check above is not vacuous: the two readings of the seed differ."""
    roots = []
    for mode in (0, 1):
        p2 = O.Poseidon2(BB, 16, seed=0xC0FFEE, bb_rc_mode=mode)
        cctx.set_constants(BB, 16, p2.flat_constants())
        roots.append(_check_commit(cctx, p2, BB, 8, 8, 1, 9))
    assert not (roots[0] == roots[1]).all()


@pytest.mark.parametrize("width", [16, 8])
def test_injected_goldilocks_constants(cctx, width):
    p2 = O.Poseidon2(GL, width, seed=0xD1CE)
    cctx.set_constants(GL, width, p2.flat_constants())
    _check_commit(cctx, p2, GL, 12, 8, 1, 5)
    _check_mmcs(cctx, p2, GL)


def test_set_constants_rejects_bad_tables(cctx):
    p2 = O.Poseidon2(BB, 16, seed=1)
    flat = p2.flat_constants().astype(np.uint32)
    with pytest.raises(Z.ZkgError) as e:
        cctx.set_constants(BB, 16, flat, n_external_rounds=6)
    assert e.value.code == EINVAL
    bad = flat.copy()
    bad[5] = O.BB_P
    with pytest.raises(Z.ZkgError) as e:
        cctx.set_constants(BB, 16, bad)
    assert e.value.code == EINVAL


def test_batch_graph_is_dropped_when_constants_change(cctx):
    """The cached graph's kernels read the tables of their capture.  Batching the
    same jobs again after a reload must hash with the new constants."""
    specs = [(BB, 16, 12, 8, 1), (GL, 16, 10, 16, 1), (BB, 16, 10, 33, 2), (GL, 8, 11, 8, 1)]
    jobs, keep, mats = _batch_jobs(cctx, specs, 700)
    for seed in (42, 7, 42, 1234):
        for f, w in ((BB, 16), (GL, 16), (GL, 8)):
            cctx.load_constants(f, w, seed)
        for rep in range(2):
            trees = Z.commit_batch(cctx, jobs)
            for t, m, (f, w, _, _, lb) in zip(trees, mats, specs):
                assert (t.root().astype(np.uint64) == O.commit(O.Poseidon2(f, w, seed), m, lb)[0]).all(), (seed, rep)


def test_explicit_constant_calls_leave_resident_constants(ctx):
    """On the session context (seed 42): the trait's batch permutation with
    another table, and the dispatch driver's Merkle layer with another seed,
    must not change what a later commit hashes with."""
    acc = Z.HipPoseidon2(ctx)
    for field, seed in ((GL, 19), (BB, 23)):
        other = O.Poseidon2(field, 16, seed=seed)
        st = rnd(field, 8 * 16, 2).reshape(8, 16)
        fn = acc.batch_permute_goldilocks if field == GL else acc.batch_permute_babybear
        got = fn(st.copy(), other.flat_constants().astype(DT[field]), 8, other.p.rounds_p)
        assert (got.reshape(8, 16)[3].astype(np.uint64) == other.permute(st[3])).all()
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    for field in (GL, BB):
        br.write(7, rnd(field, 16 * 8, 3))
        drv.execute(br, op=Z.host.OP_MERKLE_LAYER, field=field, size=16, seed=1234, in_slot=7, out_slot=8,
                    out_bytes=8 * 8 * np.dtype(DT[field]).itemsize)
    br.destroy()
    for field, w in ((GL, 16), (BB, 16)):
        _check_commit(ctx, O.Poseidon2(field, w, 42), field, 10, 12, 1, 11)


def test_driver_refuses_a_short_constant_slot(ctx):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    try:
        br.write(5, rnd(GL, 4 * 16, 1))
        flat = O.Poseidon2(GL, 16, 7).flat_constants().astype(np.uint64)
        br.write(9, flat[:-3])
        node = dict(op=Z.host.OP_POSEIDON_BATCH, field=GL, size=4, width=16, in_slot=5, out_slot=6, aux_slot=9,
                    n_ext=8, n_int=22)
        with pytest.raises(Z.ZkgError) as e:
            drv.execute(br, **node)
        assert e.value.code == EINVAL
        br.write(9, flat)  # the full table goes through
        drv.execute(br, **node)
    finally:
        br.destroy()


@pytest.mark.parametrize("op", ["poseidon", "merkle"])
def test_driver_refuses_counts_beyond_the_budget(ctx, op):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    try:
        br.write(5, rnd(BB, 64, 1))
        for size in (1 << 62, (1 << 64) - 1):
            if op == "poseidon":
                node = dict(op=Z.host.OP_POSEIDON_BATCH, field=BB, size=size, width=4, in_slot=5, out_slot=6)
            else:
                node = dict(op=Z.host.OP_MERKLE_LAYER, field=BB, size=size, seed=1, in_slot=5, out_slot=6,
                            out_bytes=32)
            with pytest.raises(Z.ZkgError) as e:
                drv.execute(br, **node)
            assert e.value.code == EINVAL
    finally:
        br.destroy()


def test_reload_refused_while_serving(cctx):
    for f, w in ((BB, 16), (GL, 16)):
        cctx.load_constants(f, w, 42)
    specs = [(BB, 16, 13, 16, 1), (GL, 16, 12, 24, 1), (BB, 16, 12, 64, 2)] * 2
    jobs, keep, mats = _batch_jobs(cctx, specs, 800)
    srv = Z.Server(cctx, slots=2, max_batch=2)
    try:
        tickets = [srv.submit(j) for j in jobs]
        for f in (BB, GL):
            with pytest.raises(Z.ZkgError) as e:
                cctx.load_constants(f, 16, 99)
            assert e.value.code == ESTATE
            with pytest.raises(Z.ZkgError) as e:
                cctx.set_constants(f, 16, O.Poseidon2(f, 16, 99).flat_constants())
            assert e.value.code == ESTATE
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
        for i, (h, _) in done.items():
            f, w, _, _, lb = specs[i]
            got = Z.MerkleTree(cctx, h, f).root().astype(np.uint64)
            assert (got == O.commit(O.Poseidon2(f, w, 42), mats[i], lb)[0]).all(), i
    finally:
        srv.stop()
    cctx.load_constants(BB, 16, 5)  # allowed again once the loop has stopped
