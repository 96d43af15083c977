"""GPU parity: libzkg (HIP, gfx950) against the CPU oracle, bit-exact.

Every test here runs the product path through the C ABI (ziros_amd binds
libzkg.so with ctypes) and compares with oracle/liboracle.so on the same
seeded inputs.  Integer/field work: the bar is bit-exact equality.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
import ziros_amd as Z
from test_oracle_poseidon2 import GOLDEN_GL16

pytestmark = pytest.mark.gpu

GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}


def rnd(field, n, seed):
    return O.gen(field, seed, n).astype(DT[field])


# ----------------------------------------------------------------- Poseidon2
def test_permute_golden_vector(ctx):
    s = np.zeros(16, dtype=np.uint64)
    s[0], s[1] = 12345, 67890
    out = Z.HipPoseidon2(ctx).permute(GL, 16, s.reshape(1, 16))
    assert [int(x) for x in out[0]] == GOLDEN_GL16


@pytest.mark.parametrize("field,width", [(GL, 16), (GL, 8), (BB, 16)])
def test_permute_vs_oracle(ctx, field, width):
    p2 = O.Poseidon2(field, width, seed=42)
    n = 777
    st = rnd(field, n * width, 3).reshape(n, width)
    p = O.PRIME[field]
    st[0] = 0
    st[1] = p - 1
    st[2] = np.arange(width)
    got = Z.HipPoseidon2(ctx).permute(field, width, st)
    for i in list(range(8)) + [n // 2, n - 1]:
        assert (got[i].astype(np.uint64) == p2.permute(st[i])).all(), i
    # the rest in bulk
    exp = np.stack([p2.permute(st[i]) for i in range(0, n, 7)])
    assert (got[::7].astype(np.uint64) == exp).all()


def test_batch_permute_explicit_constants(ctx):
    """Poseidon2Accelerator form: flat [initial][internal][terminal] table."""
    acc = Z.HipPoseidon2(ctx)
    for field, seed in ((GL, 7), (BB, 19)):
        p2 = O.Poseidon2(field, 16, seed=seed)
        rc = p2.flat_constants().astype(DT[field])
        st = rnd(field, 64 * 16, 5).reshape(64, 16)
        fn = acc.batch_permute_goldilocks if field == GL else acc.batch_permute_babybear
        got = fn(st.copy(), rc, 8, p2.p.rounds_p)
        exp = np.stack([p2.permute(st[i]) for i in range(64)])
        assert (got.reshape(64, 16).astype(np.uint64) == exp).all()


def test_generator_matches_oracle(ctx):
    for field in (GL, BB):
        b = ctx.random(field, 5000, seed=77, start=123)
        got = b.download(DT[field], 5000)
        assert (got.astype(np.uint64) == O.gen(field, 77, 5000, start=123)).all()


# ----------------------------------------------------------------- NTT / LDE
@pytest.mark.parametrize("field", [GL, BB])
@pytest.mark.parametrize("logn,width", [(0, 3), (1, 5), (3, 1), (6, 16), (10, 7), (11, 64), (12, 3), (13, 33),
                                        (16, 8)])
def test_ntt_vs_oracle(ctx, field, logn, width):
    n = 1 << logn
    m = rnd(field, n * width, 11 + logn).reshape(n, width)
    dft = Z.HipDft(ctx, field)
    assert (dft.dft_batch(m).astype(np.uint64) == O.dft_batch(field, m)).all()
    assert (dft.idft_batch(m).astype(np.uint64) == O.dft_batch(field, m, inverse=True)).all()


@pytest.mark.parametrize("field", [GL, BB])
@pytest.mark.parametrize("logn,lb,width", [(0, 1, 4), (1, 1, 3), (4, 2, 7), (8, 0, 5), (11, 1, 16), (11, 2, 9),
                                           (12, 1, 8), (13, 2, 20), (14, 1, 64), (15, 1, 3)])
def test_coset_lde_vs_oracle(ctx, field, logn, lb, width):
    n = 1 << logn
    m = rnd(field, n * width, 21 + logn).reshape(n, width)
    shift = O.lib().or_f_generator(field)
    got = Z.HipDft(ctx, field).coset_lde_batch(m, lb)
    assert (got.astype(np.uint64) == O.coset_lde_batch(field, m, lb, shift)).all()


@pytest.mark.parametrize("field", [GL, BB])
@pytest.mark.parametrize("logn,lb,width", [(18, 1, 3), (19, 2, 2), (20, 1, 9), (21, 1, 2), (22, 1, 1)])
def test_coset_lde_large_vs_oracle(ctx, field, logn, lb, width):
    """Four-step tiles of 2^9 .. 2^11 points (both passes)."""
    n = 1 << logn
    m = rnd(field, n * width, 70 + logn).reshape(n, width)
    shift = O.lib().or_f_generator(field)
    got = Z.HipDft(ctx, field).coset_lde_batch(m, lb)
    assert (got.astype(np.uint64) == O.coset_lde_batch(field, m, lb, shift)).all()


@pytest.mark.parametrize("field", [GL, BB])
@pytest.mark.parametrize("logn,width", [(18, 2), (21, 1), (22, 1)])
def test_ntt_large_vs_oracle(ctx, field, logn, width):
    n = 1 << logn
    m = rnd(field, n * width, 90 + logn).reshape(n, width)
    dft = Z.HipDft(ctx, field)
    assert (dft.dft_batch(m).astype(np.uint64) == O.dft_batch(field, m)).all()
    assert (dft.idft_batch(m).astype(np.uint64) == O.dft_batch(field, m, inverse=True)).all()


def test_coset_lde_other_shift(ctx):
    m = rnd(BB, 1024 * 4, 5).reshape(1024, 4)
    got = Z.HipDft(ctx, BB).coset_lde_batch(m, 1, shift=5)
    assert (got.astype(np.uint64) == O.coset_lde_batch(BB, m, 1, 5)).all()


# ----------------------------------------------------------------- hashing / Merkle
@pytest.mark.parametrize("field,p2w", [(GL, 16), (GL, 8), (BB, 16)])
@pytest.mark.parametrize("width", [1, 4, 7, 8, 9, 16, 33, 64])
def test_hash_rows_vs_oracle(ctx, field, p2w, width):
    p2 = O.Poseidon2(field, p2w)
    h = 512
    m = rnd(field, h * width, width).reshape(h, width)
    root, tree = Z.HipMerkleTreeMmcs(ctx, field, p2w).commit_matrix(m)
    exp = O.MerkleTree(p2, [m])
    assert (tree.layer(0, h).astype(np.uint64) == exp.layers[0]).all()
    assert (root.astype(np.uint64) == exp.root()).all()


@pytest.mark.parametrize("p2w", [8, 16])
@pytest.mark.parametrize("h", [1 << 15, (1 << 15) + 1])
def test_hash_rows_at_the_chunk_split(ctx, p2w, h):
    # Goldilocks row hashes of up to 2^15 rows run with W/4 lanes per row, longer ones with a row per lane
    # (a one-matrix MMCS commit takes the same row hash, at any height)
    p2 = O.Poseidon2(GL, p2w)
    m = rnd(GL, h * 6, 77 + p2w).reshape(h, 6)
    root, tree = Z.HipMerkleTreeMmcs(ctx, GL, p2w).commit([m])
    exp = O.MerkleTree(p2, [m])
    assert (tree.layer(0).astype(np.uint64) == exp.layers[0]).all()
    assert (root.astype(np.uint64) == exp.root()).all()
    tree.free()


@pytest.mark.parametrize("field,p2w,logh,width", [(BB, 16, 0, 5), (BB, 16, 1, 5), (BB, 16, 10, 24), (GL, 16, 12, 8),
                                                  (GL, 8, 11, 8), (BB, 16, 13, 8), (BB, 16, 9, 1),
                                                  (BB, 16, 19, 8), (GL, 16, 18, 4), (GL, 8, 16, 4),
                                                  (GL, 8, 1, 3), (GL, 8, 7, 5), (GL, 8, 18, 2),
                                                  (GL, 16, 16, 3), (GL, 16, 14, 2)])
def test_merkle_layers_and_openings(ctx, field, p2w, logh, width):
    p2 = O.Poseidon2(field, p2w)
    h = 1 << logh
    m = rnd(field, h * width, 31 + logh).reshape(h, width)
    root, tree = Z.HipMerkleTreeMmcs(ctx, field, p2w).commit_matrix(m)
    exp = O.MerkleTree(p2, [m])
    assert tree.num_layers == len(exp.layers)
    for i in range(tree.num_layers):
        assert (tree.layer(i, h).astype(np.uint64) == exp.layers[i]).all(), i
    idx = np.array(sorted({0, h - 1, h // 2, (h * 3) // 7}), dtype=np.uint64)
    rows, paths = tree.open_batch(idx, width)
    for k, i in enumerate(idx):
        i = int(i)
        assert (rows[k].astype(np.uint64) == m[i]).all()
        assert (paths[k].astype(np.uint64) == exp.path(i).reshape(paths[k].shape)).all()
        assert exp.verify(root.astype(np.uint64), i, [rows[k]], paths[k].astype(np.uint64))


# Mixed-height MMCS (compress_and_inject): the oracle's layer lengths and
# padding follow vendor/p3-merkle-tree-gpu/src/merkle_tree.rs:84-350.
MMCS_SHAPES = [
    [(100, 3), (13, 7), (100, 5), (50, 2), (1, 4)],  # injections at 64, 16 and the root; odd padding
    [(4, 3), (1, 2)],  # injection at the root layer
    [(7, 9)],  # one non-power-of-two matrix
    [(1, 5)],  # a single row: the root is the row digest
    [(33, 1), (17, 12), (9, 8), (5, 3), (3, 2), (2, 6)],
    [(64, 17), (64, 0), (32, 1)],  # a zero-width matrix contributes nothing
    # larger trees: plain runs between injections take the single-matrix tree kernels
    [(4096, 5), (4096, 3), (512, 7), (64, 2)],
    [(16384, 4)],  # one matrix: the single-matrix row hash and tree
    [(8192, 3), (1000, 2)],  # a non-power-of-two height injected at 1024
]


@pytest.mark.parametrize("shape_id", range(len(MMCS_SHAPES)))
@pytest.mark.parametrize("field,p2w", [(BB, 16), (GL, 16), (GL, 8)])
def test_mmcs_mixed_heights(ctx, field, p2w, shape_id):
    shapes = MMCS_SHAPES[shape_id]
    p2 = O.Poseidon2(field, p2w)
    mats = [rnd(field, h * w, 900 + 7 * k + shape_id).reshape(h, w) for k, (h, w) in enumerate(shapes)]
    root, tree = Z.HipMerkleTreeMmcs(ctx, field, p2w).commit(mats)
    exp = O.MerkleTree(p2, mats)
    assert tree.num_layers == len(exp.layers)
    for i in range(tree.num_layers):
        assert tree.layer_len(i) == exp.layers[i].shape[0], i
        assert (tree.layer(i).astype(np.uint64) == exp.layers[i]).all(), i
    assert (root.astype(np.uint64) == exp.root()).all()
    maxh = max(h for h, _ in shapes)
    logmax = (maxh - 1).bit_length()
    tw = sum(w for _, w in shapes)
    shift = [logmax - (h - 1).bit_length() for h, _ in shapes]
    cand = {0, maxh - 1, maxh // 2, (maxh * 3) // 7} | {(h - 1) << sh for (h, _), sh in zip(shapes, shift)}
    # an index whose row lies in the padding of a shorter matrix has no opening (p3 panics on it)
    idx = np.array(sorted(i for i in cand if all((i >> sh) < h for (h, _), sh in zip(shapes, shift))),
                   dtype=np.uint64)
    rows, paths = tree.open_mmcs(idx, tw)
    for k, i in enumerate(idx):
        i = int(i)
        want = [m[i >> (logmax - (m.shape[0] - 1).bit_length())] for m in mats]
        assert (rows[k].astype(np.uint64) == np.concatenate(want).astype(np.uint64)).all()
        assert (paths[k].astype(np.uint64) == exp.path(i).reshape(paths[k].shape)).all()
        split = np.cumsum([m.shape[1] for m in mats])[:-1]
        assert exp.verify(root.astype(np.uint64), i, np.split(rows[k].astype(np.uint64), split),
                          paths[k].astype(np.uint64))
    tree.free()


def test_mmcs_rejects_heights_sharing_a_power_of_two(ctx):
    a, b = rnd(BB, 100 * 2, 1).reshape(100, 2), rnd(BB, 120 * 2, 2).reshape(120, 2)
    with pytest.raises(Z.ZkgError):
        Z.HipMerkleTreeMmcs(ctx, BB, 16).commit([a, b])


def test_mmcs_single_matrix_matches_merkle_commit(ctx):
    m = rnd(GL, 256 * 5, 77).reshape(256, 5)
    mm = Z.HipMerkleTreeMmcs(ctx, GL, 16)
    r1, t1 = mm.commit([m])
    r2, t2 = mm.commit_matrix(m)
    assert (r1 == r2).all()
    rows, paths = t2.open_mmcs([3, 200], 5)
    r_a, p_a = t2.open_batch([3, 200], 5)
    assert (rows == r_a).all() and (paths == p_a).all()
    t1.free()
    t2.free()


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("field,p2w,logn,width,lb", [(GL, 8, 14, 8, 1), (GL, 16, 14, 8, 1), (BB, 16, 12, 16, 1),
                                                     (BB, 16, 10, 40, 2), (GL, 16, 9, 3, 2)])
def test_commit_vs_oracle(ctx, plain, field, p2w, logn, width, lb):
    """Config 1 shape (Goldilocks 2^14 x 8, blowup 2) and smaller mixes."""
    n = 1 << logn
    p2 = O.Poseidon2(field, p2w)
    tr = rnd(field, n * width, 0x5EED0001).reshape(n, width)
    buf = ctx.upload(tr)
    tree, lde = Z.commit_trace(ctx, field, buf, n, width, lb, p2_width=p2w, plain=plain)
    root_o, lde_o = O.commit(p2, tr, lb)
    got_lde = lde.download(DT[field], (n << lb) * width).reshape(n << lb, width)
    assert (got_lde.astype(np.uint64) == lde_o).all()
    assert (tree.root().astype(np.uint64) == root_o).all()


# ----------------------------------------------------------------- FRI
def test_fri_fold_reference_inputs(ctx):
    """zkf-metal/src/fri.rs:187-224 inputs, FriAccelerator semantics."""
    n = 2048
    fri = Z.HipFri(ctx)
    for field, fn in ((GL, fri.fold_goldilocks), (BB, fri.fold_babybear)):
        ev = (np.arange(2 * n, dtype=np.uint64) * 7 + 1) % O.PRIME[field]
        tw = (np.arange(n, dtype=np.uint64) * 13 + 3) % O.PRIME[field]
        got = fn(ev.astype(DT[field]), 42, tw.astype(DT[field]))
        assert (got.astype(np.uint64) == O.fri_fold_base(field, ev, 42, tw)).all()


@pytest.mark.parametrize("logn", [0, 1, 5, 12, 16])
def test_fri_fold_matrix_vs_oracle(ctx, logn):
    n = 1 << logn
    fri = Z.HipFri(ctx)
    ev = rnd(BB, 2 * n * 4, 40 + logn).reshape(2 * n, 4)
    beta = O.fri_betas(9, 1)[0]
    got = fri.fold_matrix_ef4(ev, beta.astype(np.uint32))
    assert (got.astype(np.uint64) == O.fri_fold_matrix_ef4(ev, beta)).all()
    for field in (GL, BB):
        e = rnd(field, 2 * n, 50 + logn)
        b = int(O.gen(field, 3, 1)[0])
        assert (fri.fold_matrix_base(field, e, b).astype(np.uint64) == O.fri_fold_matrix_base(field, e, b)).all()


# (18, 1): 17 layer trees of at most 2^17 leaves, built together in two groups of at most 16
@pytest.mark.parametrize("log_len,log_final", [(12, 4), (9, 8), (14, 8), (19, 12), (18, 1)])
def test_fri_commit_phase_vs_oracle(ctx, log_len, log_final):
    p2 = O.Poseidon2(BB, 16)
    cw = rnd(BB, (1 << log_len) * 4, 60 + log_len).reshape(-1, 4)
    nl = log_len - log_final
    betas = O.fri_betas(0x5EED0004, nl)
    buf = ctx.upload(cw)
    fc = Z.HipFri(ctx).commit_phase(buf, log_len, log_final, betas.astype(np.uint32))
    roots_o, final_o = O.fri_commit_phase_ef4(p2, cw, log_len, log_final, betas)
    assert (fc.layer_roots().astype(np.uint64) == roots_o).all()
    assert (fc.final().astype(np.uint64) == final_o).all()
    # query openings verify against every layer root
    layers = [cw.astype(np.uint64)]
    for r in range(nl - 1):
        layers.append(O.fri_fold_matrix_ef4(layers[-1], betas[r]))
    trees = [O.MerkleTree(p2, [layers[r].reshape(-1, 8)]) for r in range(nl)]
    q = np.array([0, 1, (1 << log_len) - 1, 12345 % (1 << log_len)], dtype=np.uint64)
    rows, paths = fc.open(q)
    for k, qi in enumerate(q):
        off = 0
        for r in range(nl):
            leaf = int(qi) >> (r + 1)
            mat = layers[r].reshape(-1, 8)
            assert (rows[k, r].astype(np.uint64) == mat[leaf]).all()
            plen = log_len - 1 - r
            path = paths[k, off:off + plen].astype(np.uint64)
            off += plen
            assert (path == trees[r].path(leaf)).all()
            assert trees[r].verify(roots_o[r], leaf, [mat[leaf]], path)


# ----------------------------------------------------------------- batch and serve
def _jobs(ctx, specs):
    jobs, keep, expect = [], [], []
    for k, (field, p2w, logn, width, lb) in enumerate(specs):
        n = 1 << logn
        tr = rnd(field, n * width, 100 + k).reshape(n, width)
        b = ctx.upload(tr)
        lde = ctx.alloc((n << lb) * width * np.dtype(DT[field]).itemsize)
        jobs.append(Z.make_job(field, b, n, width, lb, lde, p2w))
        keep += [b, lde]
        expect.append(O.commit(O.Poseidon2(field, p2w), tr, lb)[0])
    return jobs, keep, expect


SPECS = [(BB, 16, 12, 8, 1), (GL, 16, 10, 16, 1), (BB, 16, 14, 24, 1), (BB, 16, 8, 3, 2), (GL, 8, 11, 8, 1),
         (BB, 16, 13, 64, 1)]


def test_commit_batch_matches_lone_commits(ctx):
    jobs, keep, expect = _jobs(ctx, SPECS)
    for rep in range(3):  # capture, then replays of the cached graph
        trees = Z.commit_batch(ctx, jobs)
        for t, e in zip(trees, expect):
            assert (t.root().astype(np.uint64) == e).all(), rep


def test_serve_results_identical_in_mix(ctx):
    import time

    jobs, keep, expect = _jobs(ctx, SPECS * 2)
    srv = Z.Server(ctx, slots=2, max_batch=4)
    try:
        tickets = [srv.submit(j) for j in jobs]
        done = {}
        t0 = time.time()
        while len(done) < len(tickets) and time.time() - t0 < 60:
            for i, t in enumerate(tickets):
                if i in done:
                    continue
                r = srv.poll(t)
                if r is not None:
                    done[i] = r
            time.sleep(0.001)
        assert len(done) == len(tickets)
        for i, (h, lat) in done.items():
            tree = Z.MerkleTree(ctx, h, jobs[i].field)
            assert (tree.root().astype(np.uint64) == expect[i]).all(), i
            assert lat > 0
    finally:
        srv.stop()


def test_commit_batch_trees_open(ctx):
    """Trees handed out by a (replayed) batch graph are device copies: their
    openings verify against the oracle's paths."""
    specs = [(BB, 16, 10, 8, 1), (GL, 16, 9, 5, 1)]
    jobs, keep, expect = _jobs(ctx, specs)
    for rep in range(2):
        trees = Z.commit_batch(ctx, jobs)
    for k, ((field, p2w, logn, width, lb), t) in enumerate(zip(specs, trees)):
        n = 1 << logn
        tr = rnd(field, n * width, 100 + k).reshape(n, width)
        lde = O.coset_lde_batch(field, tr, lb, Z.GENERATOR[field])
        exp = O.MerkleTree(O.Poseidon2(field, p2w), [lde])
        idx = [0, 5, (n << lb) - 1]
        rows, paths = t.open_batch(idx, width)
        for q, i in enumerate(idx):
            assert (rows[q].astype(np.uint64) == lde[i]).all()
            assert (paths[q].astype(np.uint64) == exp.path(i).reshape(paths[q].shape)).all()


def test_serve_under_a_tight_hbm_budget():
    """Admission against the context's HBM budget: small requests complete and
    are bit-exact; a request whose tree and scratch can never fit fails with an
    error instead of waiting forever."""
    import time

    c2 = Z.Context(0, hbm_budget=96 << 20)
    try:
        c2.load_constants(BB, 16, 42)
        small = [(BB, 16, 12, 8, 1)] * 4
        jobs, keep, expect = _jobs(c2, small)
        n = 1 << 18
        big_tr = c2.upload(rnd(BB, n * 16, 7).reshape(n, 16))
        big_lde = c2.alloc((n << 1) * 16 * 4)
        big = Z.make_job(BB, big_tr, n, 16, 1, big_lde, 16)
        srv = Z.Server(c2, slots=2, max_batch=2)
        try:
            tickets = [srv.submit(j) for j in jobs] + [srv.submit(big)]
            done, failed = {}, []
            t0 = time.time()
            while len(done) + len(failed) < len(tickets) and time.time() - t0 < 60:
                for i, t in enumerate(tickets):
                    if i in done or i in failed:
                        continue
                    try:
                        r = srv.poll(t)
                    except Z.ZkgError:
                        failed.append(i)
                        continue
                    if r is not None:
                        done[i] = r
                time.sleep(0.001)
            assert failed == [len(tickets) - 1] and sorted(done) == list(range(len(jobs)))
            for i, (h, _) in done.items():
                assert (Z.MerkleTree(c2, h, BB).root().astype(np.uint64) == expect[i]).all()
        finally:
            srv.stop()
    finally:
        c2.close()


def test_errors_are_reported(ctx):
    with pytest.raises(Z.ZkgError):
        Z.HipDft(ctx, BB).coset_lde_batch(np.zeros((3, 2), dtype=np.uint32), 1)  # not a power of two
    with pytest.raises(Z.ZkgError):
        Z.HipDft(ctx, BB).coset_lde_batch(np.zeros((4, 2), dtype=np.uint32), 3)  # blowup too large


def test_sizes_that_wrap_are_rejected(ctx):
    """Counts whose byte size exceeds 64 bits must fail the buffer check, not
    wrap below the buffer size.  Every call is refused on the host, before any
    copy or launch."""
    EINVAL = -1
    lib, h = ctx.lib, ctx.h
    buf = ctx.alloc(4096)
    out = ctypes.c_void_p()
    src = (ctypes.c_uint8 * 16)()
    try:
        assert lib.zkg_buf_upload(buf.h, (1 << 64) - 8, src, 16) == EINVAL
        assert lib.zkg_buf_download(buf.h, 4090, src, 16) == EINVAL
        assert lib.zkg_poseidon2_permute(h, GL, 16, buf.h, 1 << 61) == EINVAL
        assert lib.zkg_hash_rows(h, BB, 16, buf.h, 1 << 62, 8, buf.h) == EINVAL
        assert lib.zkg_merkle_commit(h, GL, 16, buf.h, 1 << 31, (1 << 32) - 1, ctypes.byref(out)) == EINVAL
        heights = (ctypes.c_uint64 * 1)(1 << 31)
        widths = (ctypes.c_uint32 * 1)((1 << 32) - 1)
        mats = (ctypes.c_void_p * 1)(buf.h.value)
        assert lib.zkg_mmcs_commit(h, GL, 16, 1, mats, heights, widths, ctypes.byref(out)) == EINVAL
        assert not out.value
    finally:
        buf.free()


# ----------------------------------------------------------------- DEEP quotient
def _deep_case(ctx, logn, w, lb, seed):
    n = 1 << logn
    rng = np.random.default_rng(seed)
    tr = rnd(BB, n * w, seed).reshape(n, w)
    lde = O.coset_lde_batch(BB, tr, lb, 31)
    z = [int(v) for v in rng.integers(0, O.BB_P, 4)]
    alpha = [int(v) for v in rng.integers(0, O.BB_P, 4)]
    buf = ctx.upload(lde.astype(np.uint32))
    return lde, z, alpha, buf, Z.HipDeep(ctx, BB, buf, n << lb, w, lb, 31)


@pytest.mark.parametrize("logn,w,lb", [(0, 3, 1), (3, 5, 1), (7, 16, 1), (10, 64, 1), (9, 33, 2), (12, 8, 3),
                                       (11, 300, 1)])
def test_deep_open_values_vs_oracle(ctx, logn, w, lb):
    lde, z, _, buf, deep = _deep_case(ctx, logn, w, lb, 40 + logn)
    ys = deep.open_values(z)
    assert (ys.astype(np.uint64) == O.deep_open_values(lde, lb, 31, z)).all()


@pytest.mark.parametrize("logn,w,lb,off", [(0, 2, 1, 0), (4, 7, 1, 3), (8, 64, 1, 0), (10, 17, 2, 5), (11, 64, 1, 1)])
def test_deep_reduce_vs_oracle(ctx, logn, w, lb, off):
    lde, z, alpha, buf, deep = _deep_case(ctx, logn, w, lb, 70 + logn)
    h = (1 << logn) << lb
    base = rnd(BB, h * 4, 5).reshape(h, 4)
    ro = ctx.upload(base.astype(np.uint32))
    ys = deep.open(z, alpha, off, ro)
    exp_ys = O.deep_open_values(lde, lb, 31, z)
    assert (ys.astype(np.uint64) == exp_ys).all()
    got = ro.download(np.uint32, h * 4).reshape(h, 4)
    assert (got.astype(np.uint64) == O.deep_reduce(lde, 31, z, alpha, off, exp_ys, ro=base)).all()
    # the separate entry point accumulates the same quotient again
    deep.reduce(z, alpha, off, ys, ro)
    twice = O.deep_reduce(lde, 31, z, alpha, off, exp_ys, ro=O.deep_reduce(lde, 31, z, alpha, off, exp_ys, ro=base))
    assert (ro.download(np.uint32, h * 4).reshape(h, 4).astype(np.uint64) == twice).all()


def test_deep_rejects_point_on_domain(ctx):
    lde, _, alpha, buf, deep = _deep_case(ctx, 4, 3, 1, 3)
    x = 31 * pow(int(O.two_adic_generator(BB, 5)), 7, O.BB_P) % O.BB_P  # a coset point
    with pytest.raises(Z.ZkgError):
        deep.open_values([x, 0, 0, 0])


def test_deep_base_field_challenges_babybear(ctx):
    """The reference's Challenge<F> = F: z, alpha in BabyBear are EF4 values with
    zero upper coefficients, and everything stays in the base field."""
    lde, _, _, buf, deep = _deep_case(ctx, 9, 12, 1, 123)
    z, alpha = [987654321, 0, 0, 0], [13579, 0, 0, 0]
    h = (1 << 9) << 1
    ro = ctx.upload(np.zeros(h * 4, dtype=np.uint32))
    ys = deep.open(z, alpha, 2, ro)
    exp_ys = O.deep_open_values(lde, 1, 31, z)
    assert (ys.astype(np.uint64) == exp_ys).all() and (exp_ys[:, 1:] == 0).all()
    got = ro.download(np.uint32, h * 4).reshape(h, 4).astype(np.uint64)
    assert (got == O.deep_reduce(lde, 31, z, alpha, 2, exp_ys)).all() and (got[:, 1:] == 0).all()


@pytest.mark.parametrize("logn,w,lb,off", [(0, 2, 1, 0), (4, 7, 1, 3), (8, 64, 1, 0), (10, 17, 2, 5), (11, 64, 1, 1),
                                           (9, 100, 3, 0)])
def test_deep_goldilocks_vs_oracle(ctx, logn, w, lb, off):
    n = 1 << logn
    h = n << lb
    rng = np.random.default_rng(300 + logn)
    tr = rnd(GL, n * w, 300 + logn).reshape(n, w)
    lde = O.coset_lde_batch(GL, tr, lb, 7)
    z, alpha = int(rng.integers(0, 2**62)), int(rng.integers(0, 2**62))
    buf = ctx.upload(lde)
    deep = Z.HipDeep(ctx, GL, buf, h, w, lb, 7)
    base = rnd(GL, h, 9)
    ro = ctx.upload(base)
    ys = deep.open(z, alpha, off, ro)
    exp_ys = O.deep_gl_open_values(lde, lb, 7, z)
    assert (ys == exp_ys).all()
    assert (ro.download(np.uint64, h) == O.deep_gl_reduce(lde, 7, z, alpha, off, exp_ys, ro=base)).all()
    assert (deep.open_values(z) == exp_ys).all()


def test_deep_goldilocks_rejects_point_on_domain(ctx):
    lde = O.coset_lde_batch(GL, rnd(GL, 16 * 2, 1).reshape(16, 2), 1, 7)
    buf = ctx.upload(lde)
    x = 7 * pow(int(O.two_adic_generator(GL, 5)), 3, O.GL_P) % O.GL_P
    with pytest.raises(Z.ZkgError):
        Z.HipDeep(ctx, GL, buf, 32, 2, 1, 7).open_values(x)


def test_batch_replays_interleaved_with_eager_work(ctx):
    """The batch graph's persistent row-hash launch resets its own tile counter;
    replays separated by other eager launches (single commits, a FRI commit
    phase, permutations) and synchronisations must each give the oracle's roots."""
    specs = [(BB, 16, 12, 40, 1), (GL, 16, 11, 9, 2), (BB, 16, 10, 7, 1), (GL, 8, 12, 16, 1)]
    jobs, keep, expect = _jobs(ctx, specs)
    other = rnd(BB, (1 << 11) * 24, 3).reshape(1 << 11, 24)
    obuf = ctx.upload(other)
    cw = ctx.upload(rnd(BB, (1 << 12) * 4, 4).reshape(-1, 4))
    betas = O.fri_betas(1, 6).astype(np.uint32)
    root_other = O.commit(O.Poseidon2(BB, 16), other, 1)[0]
    for rep in range(5):
        trees = Z.commit_batch(ctx, jobs)
        for k, (t, e) in enumerate(zip(trees, expect)):
            assert (t.root().astype(np.uint64) == e).all(), (rep, k)
            t.free()
        tree, _ = Z.commit_trace(ctx, BB, obuf, 1 << 11, 24, 1)
        assert (tree.root().astype(np.uint64) == root_other).all()
        tree.free()
        Z.HipFri(ctx).commit_phase(cw, 12, 6, betas).free()
        ctx.synchronize()
