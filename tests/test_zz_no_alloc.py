"""zkg_ntt_batch_noalloc against zkg_ntt_batch, through the C ABI.

The entry point keeps the four-step scratch in a fixed per-GPU device array
and allocates nothing.  Each case runs both entry points on copies of the same
seeded matrix and requires bit-identical results; two calls in a row (the
second smaller) show that the shared array's leftover contents do not leak
into a later result.  Sizes beyond the array are refused.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
import ziros_amd as Z

pytestmark = pytest.mark.gpu

GL, BB = Z.GOLDILOCKS, Z.BABYBEAR
DT = {GL: np.uint64, BB: np.uint32}
EINVAL = -1
SCRATCH = 64 << 20


@pytest.fixture(scope="module")
def nctx():
    c = Z.Context(0)
    yield c
    c.close()


def _both(c, field, logn, width, inverse, seed):
    n = 1 << logn
    m = O.gen(field, seed, n * width).astype(DT[field])
    a, b = c.upload(m), c.upload(m)
    Z._lib.check(c.lib.zkg_ntt_batch(c.h, field, a.h, n, width, inverse))
    used = c.hbm_used()
    Z._lib.check(c.lib.zkg_ntt_batch_noalloc(c.h, field, b.h, n, width, inverse))
    c.synchronize()
    assert c.hbm_used() == used  # nothing allocated through the context
    ref = a.download(DT[field], n * width)
    got = b.download(DT[field], n * width)
    a.free()
    b.free()
    return ref, got, m


@pytest.mark.parametrize("field", [GL, BB])
@pytest.mark.parametrize("inverse", [0, 1])
def test_noalloc_matches_ntt_batch_two_calls(nctx, field, inverse):
    # first call: a large transform that fills most of the array; second: smaller
    big_log = 20 if field == GL else 21
    for logn, width, seed in ((big_log, 8 if field == GL else 7, 11), (13, 5, 12)):
        assert (1 << logn) * width * np.dtype(DT[field]).itemsize <= SCRATCH
        ref, got, m = _both(nctx, field, logn, width, inverse, seed + inverse)
        assert (got == ref).all(), (logn, width)
    # and the oracle on a small case
    ref, got, m = _both(nctx, field, 12, 3, inverse, 5)
    exp = O.dft_batch(field, m.reshape(-1, 3), inverse=bool(inverse)).reshape(-1)
    assert (got.astype(np.uint64) == exp).all()


def test_noalloc_small_transforms_need_no_scratch(nctx):
    for logn in (0, 1, 11):
        ref, got, _ = _both(nctx, BB, logn, 4, 0, 20 + logn)
        assert (got == ref).all()


def test_noalloc_refuses_sizes_beyond_the_array(nctx):
    n = 1 << 22
    buf = nctx.alloc(n * 16 * 4)
    try:
        rc = nctx.lib.zkg_ntt_batch_noalloc(nctx.h, BB, buf.h, n, 16, 0)  # 256 MiB of scratch
        assert rc == EINVAL
        assert b"scratch" in nctx.lib.zkg_last_error()
    finally:
        buf.free()
