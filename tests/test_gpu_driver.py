"""The runtime dispatch driver over libzkg (GpuDispatchDriver of
zkf-runtime/src/metal_driver.rs:55-93, node payloads of execution.rs:51-106).

Each arm is checked against the oracle composition of what the Metal
driver's arm does (metal_dispatch_macos.rs:305-779):
- NTT, LDE and FRI fold against the oracle transforms;
- the Poseidon batch and Merkle layer against oracle permutations of the same
  zero-padded 16-lane states;
- nodes the lane does not take raise the fallback signal.
"""
import numpy as np
import pytest

import oracle as O

import ziros_amd as Z

pytestmark = pytest.mark.gpu
GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(0)
    yield c
    c.close()


def _rand(field, n, seed):
    return O.gen(field, seed, n).astype(DT[field])


@pytest.mark.parametrize("field,n,w", [(GL, 1 << 12, 1), (BB, 1 << 10, 4), (GL, 4, 1)])
@pytest.mark.parametrize("inverse", [0, 1])
def test_ntt(ctx, field, n, w, inverse):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    v = _rand(field, n * w, 11 + n + inverse) if n > 4 else np.array([1, 2, 3, 4], dtype=DT[field])
    br.write(1, v)
    tel = drv.execute(br, op=Z.host.OP_NTT, field=field, size=n, width=w, inverse=inverse, in_slot=1, out_slot=2)
    exp = O.dft_batch(field, v.reshape(n, w).astype(np.uint64), inverse=bool(inverse))
    assert (br.read(2, DT[field]).reshape(n, w).astype(np.uint64) == exp).all()
    assert tel["kernel"].startswith("hip-ntt") and tel["output_bytes"] == v.nbytes
    if n == 4 and not inverse:  # the reference's test: forward NTT of [1, 2, 3, 4] starts with 10
        assert int(br.read(2, DT[field])[0]) == 10
    # in place
    drv.execute(br, op=Z.host.OP_NTT, field=field, size=n, width=w, inverse=inverse, in_slot=1, out_slot=1)
    assert (br.read(1, DT[field]).reshape(n, w).astype(np.uint64) == exp).all()
    br.destroy()


@pytest.mark.parametrize("field,blowup", [(GL, 2), (BB, 4)])
def test_lde(ctx, field, blowup):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    n, w = 1 << 10, 3
    m = _rand(field, n * w, 21).reshape(n, w)
    br.write(3, m)
    drv.execute(br, op=Z.host.OP_LDE, field=field, size=n, width=w, blowup=blowup, in_slot=3, out_slot=4)
    shift = 7 if field == GL else 31
    exp = O.coset_lde_batch(field, m.astype(np.uint64), blowup.bit_length() - 1, shift)
    assert (br.read(4, DT[field]).reshape(n * blowup, w).astype(np.uint64) == exp).all()
    br.destroy()


@pytest.mark.parametrize("field", [GL, BB])
def test_lde_in_place(ctx, field):
    # the output slot is the input slot: the input stays valid until the LDE is done
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    n, w = 1 << 11, 2
    m = _rand(field, n * w, 23).reshape(n, w)
    br.write(3, m)
    drv.execute(br, op=Z.host.OP_LDE, field=field, size=n, width=w, blowup=4, in_slot=3, out_slot=3)
    exp = O.coset_lde_batch(field, m.astype(np.uint64), 2, 7 if field == GL else 31)
    assert (br.read(3, DT[field]).reshape(n * 4, w).astype(np.uint64) == exp).all()
    for bad in (3, 0, 1 << 20):  # not a power of two, zero, beyond any LDE
        with pytest.raises(Z.ZkgError):
            drv.execute(br, op=Z.host.OP_LDE, field=field, size=n, width=w, blowup=bad, in_slot=3, out_slot=4)
    br.destroy()


@pytest.mark.parametrize("field,count,width,explicit", [(GL, 37, 12, True), (GL, 20, 16, False), (BB, 25, 5, False),
                                                        (BB, 9, 16, True)])
def test_poseidon_batch(ctx, field, count, width, explicit):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    x = _rand(field, count * width, 31 + count)
    br.write(5, x)
    seed = 7 if explicit else 42
    p2 = O.Poseidon2(field, 16, seed)
    node = dict(op=Z.host.OP_POSEIDON_BATCH, field=field, size=count, width=width, in_slot=5, out_slot=6)
    if explicit:
        br.write(9, p2.flat_constants().astype(DT[field]))
        node.update(aux_slot=9, n_ext=8, n_int=22 if field == GL else 13)
    drv.execute(br, **node)
    got = br.read(6, DT[field]).reshape(count, width).astype(np.uint64)
    for p in range(count):
        s = np.zeros(16, dtype=np.uint64)
        s[:width] = x[p * width:(p + 1) * width]
        assert (got[p] == p2.permute(s)[:width]).all(), p
    br.destroy()


@pytest.mark.parametrize("field,leaves,lw,ow", [(GL, 9, 8, 8), (GL, 16, 3, 4), (BB, 7, 8, 8), (BB, 6, 12, 2)])
def test_merkle_layer(ctx, field, leaves, lw, ow):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    x = _rand(field, leaves * lw, 41 + leaves)
    br.write(7, x)
    parents = (leaves + 1) // 2
    es = np.dtype(DT[field]).itemsize
    drv.execute(br, op=Z.host.OP_MERKLE_LAYER, field=field, size=leaves, seed=1234, in_slot=7, out_slot=8,
                out_bytes=parents * ow * es)
    got = br.read(8, DT[field]).reshape(parents, ow).astype(np.uint64)
    p2 = O.Poseidon2(field, 16, 1234)
    for p in range(parents):
        s = np.zeros(16, dtype=np.uint64)
        left, right = 2 * p, min(2 * p + 1, leaves - 1)
        k = min(lw, 8)
        s[:k] = x[left * lw:left * lw + k]
        s[8:8 + k] = x[right * lw:right * lw + k]
        assert (got[p] == p2.permute(s)[:ow]).all(), p
    if field == GL and lw == 8:  # full-width leaves: the p3 TruncatedPermutation compression
        assert (got[0] == p2.compress(x[:8], x[8:16])).all()
    br.destroy()


@pytest.mark.parametrize("field", [GL, BB])
def test_fri_fold(ctx, field):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    n = 1 << 12
    ev, tw = _rand(field, 2 * n, 51), _rand(field, n, 52)
    alpha = int(_rand(field, 1, 53)[0])
    br.write(10, ev)
    br.write(11, np.array([alpha, 99], dtype=DT[field]))  # alpha is the slot's first element
    br.write(12, tw)
    drv.execute(br, op=Z.host.OP_FRI_FOLD, field=field, size=2 * n, in_slot=10, aux_slot=11, aux2_slot=12,
                out_slot=13)
    exp = O.fri_fold_base(field, ev.astype(np.uint64), alpha, tw.astype(np.uint64))
    assert (br.read(13, DT[field]).astype(np.uint64) == exp).all()
    br.destroy()


@pytest.mark.parametrize("alias", ["evals", "alpha", "twiddles"])
def test_fri_fold_output_aliases_an_input(ctx, alias):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    n = 1 << 11
    ev, tw = _rand(BB, 2 * n, 54), _rand(BB, n, 55)
    alpha = int(_rand(BB, 1, 56)[0])
    br.write(10, ev)
    br.write(11, np.array([alpha], dtype=np.uint32))
    br.write(12, tw)
    out = {"evals": 10, "alpha": 11, "twiddles": 12}[alias]
    drv.execute(br, op=Z.host.OP_FRI_FOLD, field=BB, size=2 * n, in_slot=10, aux_slot=11, aux2_slot=12, out_slot=out)
    exp = O.fri_fold_base(BB, ev.astype(np.uint64), alpha, tw.astype(np.uint64))
    assert (br.read(out, np.uint32).astype(np.uint64) == exp).all()
    br.destroy()


def test_fri_fold_needs_a_resident_alpha_slot(ctx):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    n = 1 << 10
    br.write(10, _rand(GL, 2 * n, 57))
    br.write(12, _rand(GL, n, 58))
    with pytest.raises(Z.ZkgError):
        drv.execute(br, op=Z.host.OP_FRI_FOLD, field=GL, size=2 * n, in_slot=10, aux_slot=11, aux2_slot=12,
                    out_slot=13)
    br.write(11, np.zeros(0, dtype=np.uint64))  # resident but empty: alpha = 0 (the reference's unwrap_or(0))
    drv.execute(br, op=Z.host.OP_FRI_FOLD, field=GL, size=2 * n, in_slot=10, aux_slot=11, aux2_slot=12, out_slot=13)
    ev, tw = br.read(10, np.uint64), br.read(12, np.uint64)
    assert (br.read(13, np.uint64) == O.fri_fold_base(GL, ev, 0, tw)).all()
    br.destroy()


def test_fallback_signals_and_thresholds(ctx):
    drv, br = Z.HipDispatchDriver(ctx), Z.BufferBridge(ctx)
    br.write(1, np.arange(32, dtype=np.uint64))
    with pytest.raises(Z.FallbackSignal):
        drv.execute(br, op=99, field=GL, size=32, in_slot=1, out_slot=2)            # not on this lane
    with pytest.raises(Z.FallbackSignal):
        drv.execute(br, op=Z.host.OP_NTT, field=7, size=32, in_slot=1, out_slot=2)  # no payload variant
    with pytest.raises(Z.FallbackSignal):
        drv.execute(br, op=Z.host.OP_POSEIDON_BATCH, field=GL, size=1, width=17, in_slot=1, out_slot=2)
    with pytest.raises(Z.ZkgError):  # a hard error: NTT of a size that is not a power of two
        br.write(3, np.arange(24, dtype=np.uint64))
        drv.execute(br, op=Z.host.OP_NTT, field=GL, size=24, in_slot=3, out_slot=4)
    assert drv.should_try_on_either(op=Z.host.OP_NTT, field=GL, size=1 << 16)
    assert not drv.should_try_on_either(op=Z.host.OP_NTT, field=GL, size=64)
    br.free_shared(1)
    with pytest.raises(Z.ZkgError):
        br.read(1, np.uint64)
    br.destroy()
