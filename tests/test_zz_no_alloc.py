"""GPU parity of bi_mul_u32_dev_noalloc, the multiply whose scratch is a
static device array (no allocation), against the existing entry point
bi_mul_u32_dev on the same inputs, bit-exact, through the C ABI: two calls of
the new entry point in a row, the second with a smaller input (so the second
runs over the first one's leftovers in the shared array)."""
import numpy as np
import pytest

import bigint_ffi as F
import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu


def _mul_dev(ctx, name, da, db, dr, m, n):
    ctx.check(getattr(F.lib(), name)(ctx.h, m, n, da.ptr, db.ptr, dr.ptr), name)
    return dr.download(np.uint32, n * m).reshape(n, m)


def _pair(ctx, n, m, seed, top_zero):
    a, b = W.plain_pairs(n, m, seed=seed, top_zero=top_zero)
    return a, b, ctx.upload(a), ctx.upload(b), ctx.alloc(a.nbytes)


@pytest.mark.parametrize("big,small", [
    ((131072, 512, True), (300, 200, False)),     # full measured batch, then a truncated-schedule batch
    ((64, 4096, False), (1000, 64, True)),        # teams spanning a workgroup, then one-wave teams
])
def test_noalloc_matches_mul_dev_twice(ctx, big, small):
    (n1, m1, z1), (n2, m2, z2) = big, small
    a1, b1, da1, db1, dr1 = _pair(ctx, n1, m1, 71 + m1, z1)
    a2, b2, da2, db2, dr2 = _pair(ctx, n2, m2, 72 + m2, z2)
    want1 = _mul_dev(ctx, "bi_mul_u32_dev", da1, db1, dr1, m1, n1)
    want2 = _mul_dev(ctx, "bi_mul_u32_dev", da2, db2, dr2, m2, n2)
    # two calls in a row of the allocation-free entry point, larger first
    ctx.check(F.lib().bi_mul_u32_dev_noalloc(ctx.h, m1, n1, da1.ptr, db1.ptr, dr1.ptr), "noalloc 1")
    ctx.check(F.lib().bi_mul_u32_dev_noalloc(ctx.h, m2, n2, da2.ptr, db2.ptr, dr2.ptr), "noalloc 2")
    got1 = dr1.download(np.uint32, n1 * m1).reshape(n1, m1)
    got2 = dr2.download(np.uint32, n2 * m2).reshape(n2, m2)
    for x in (da1, db1, dr1, da2, db2, dr2):
        ctx._release(x)
    assert (got1 == want1).all()
    assert (got2 == want2).all()
    idx = np.random.default_rng(5).choice(n1, min(n1, 256), replace=False)
    assert (got1[idx] == O.mul(a1[idx], b1[idx])).all()
    assert (got2 == O.mul(a2, b2)).all()


def test_noalloc_limits(ctx):
    x = np.zeros((1, 8), np.uint32)
    p = x.ctypes.data_as(__import__("ctypes").c_void_p)
    L = F.lib()
    assert L.bi_mul_u32_dev_noalloc(ctx.h, 4097, 1, p, p, p) == -2   # beyond the multiply's size limit
    assert L.bi_mul_u32_dev_noalloc(ctx.h, 0, 1, p, p, p) == -1
    assert L.bi_mul_u32_dev_noalloc(ctx.h, 8, 0, p, p, p) == 0       # empty batch: nothing launched
