"""GPU parity of the compare / precision / subtract / shift / multiply-by-
digit / divide-by-digit helpers against the oracle (bit-exact)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


def batch(n, m, seed):
    rng = np.random.default_rng(seed)
    a, b = rnd(rng, (n, m)), rnd(rng, (n, m))
    b[::4] = a[::4]                       # equal pairs
    b[1::4, m // 2:] = a[1::4, m // 2:]   # equal top halves: decided lower down
    if m > 1:
        b[2::8, -1] = a[2::8, -1]
        b[2::8, : m - 1] = a[2::8, : m - 1]
        b[2::8, 0] ^= 1                   # differ only in the lowest digit
    a[5::9, m // 3:] = 0                  # short operands
    a[6::9] = 0
    return a, b


@pytest.mark.parametrize("m,n", [(1, 50), (3, 70), (8, 200), (16, 300), (17, 300), (64, 300), (255, 65), (512, 200), (1000, 40),
                                 (4096, 9)])
def test_cmp_sub_prec(ctx, m, n):
    a, b = batch(n, m, m + n)
    assert (ctx.cmp(a, b) == O.cmp(a, b)).all()
    r, s = ctx.sub(a, b)
    r2, s2 = O.sub(a, b)
    assert (s == s2).all() and (r == r2).all()
    assert (ctx.prec(a) == O.prec(a)).all()


@pytest.mark.parametrize("m,n", [(8, 300), (512, 200), (1000, 40)])
@pytest.mark.parametrize("into", ["a", "b"])
def test_sub_in_place(ctx, m, n, into):
    """|a - b| written over a or over b: each instance's compare reads its top
    digits before the same wave stores over them (m = 8 takes the separate
    compare pass)."""
    import bigint_ffi as F
    a, b = batch(n, m, 7 * m + n)
    da, db, ds = ctx.upload(a), ctx.upload(b), ctx.alloc(max(n, 4))
    dr = da if into == "a" else db
    ctx.check(F.lib().bi_sub_u32_dev(ctx.h, m, n, da.ptr, db.ptr, dr.ptr, ds.ptr), "bi_sub_u32_dev")
    r2, s2 = O.sub(a, b)
    assert (ds.download(np.uint8, n) == s2).all()
    assert (dr.download(np.uint32, n * m).reshape(n, m) == r2).all()


@pytest.mark.parametrize("m,n", [(16, 3000), (64, 2000), (512, 400), (1000, 70)])
@pytest.mark.parametrize("span,tiles,nt", [(256, 1, 1), (1024, 2, 0), (2048, 4, 1), (8192, 1, 1), (8192, 4, 0)])
def test_sub_tuned_shapes(ctx, m, n, span, tiles, nt):
    """|a - b| at the shapes the |a-b| tuner sweeps (spans of 1..63 instances
    compare inside the span waves; longer spans take the compare pass)."""
    a, b = batch(n, m, 3 * m + span + tiles)
    keep = ctx.tune_get_sub(m)
    ctx.tune_set_sub(m, span, tiles, nt)
    try:
        r, s = ctx.sub(a, b)
    finally:
        ctx.tune_set_sub(m, *keep.values())
    r2, s2 = O.sub(a, b)
    assert (s == s2).all() and (r == r2).all()


def test_sub_tuner_keeps_results(ctx):
    """bi_tune_run(what = 3) and (what = 0) change shapes, never digits."""
    m, n = 512, 300
    a, b = batch(n, m, 99)
    keep_sub, keep_add = ctx.tune_get_sub(m), ctx.tune_get(m)
    try:
        ctx.tune(m, 8 << 20, 3)
        ctx.tune(m, 8 << 20, 0)
        r, s = ctx.sub(a, b)
        ra = ctx.add(a, b)
    finally:
        ctx.tune_set_sub(m, *keep_sub.values())
        ctx.tune_set_add(m, keep_add["add_span_digits"], keep_add["add_tiles_in_flight"], keep_add["add_nontemporal"])
    r2, s2 = O.sub(a, b)
    assert (s == s2).all() and (r == r2).all()
    assert (ra == O.add(a, b)).all()


@pytest.mark.parametrize("m", [1, 5, 64, 300])
@pytest.mark.parametrize("s", [-400, -64, -7, -1, 0, 1, 3, 64, 299, 400])
def test_shift(ctx, m, s):
    a, _ = batch(33, m, 3)
    assert (ctx.shift(s, a) == O.shift(s, a)).all()


def test_shift_in_place(ctx):
    a, _ = batch(40, 128, 4)
    assert (ctx.shift(5, a, in_place=True) == O.shift(5, a)).all()
    assert (ctx.shift(-9, a, in_place=True) == O.shift(-9, a)).all()


@pytest.mark.parametrize("m,n", [(1, 40), (4, 100), (63, 50), (64, 37), (65, 37), (100, 41), (128, 33), (129, 35),
                                 (256, 64), (257, 33), (260, 30), (512, 30), (513, 17), (516, 17), (1024, 20),
                                 (2048, 9), (3001, 5), (4100, 4)])
def test_muld_quod(ctx, m, n):
    rng = np.random.default_rng(m)
    a, _ = batch(n, m, m * 3)
    d = rnd(rng, n)
    d[0], d[1], d[2] = 1, 0xFFFFFFFF, 2
    a[3] = 0xFFFFFFFF
    assert (ctx.muld(a, d) == O.muld(a, d)).all()
    q, rem = ctx.quod(a, d)
    q2, rem2 = O.quod(a, d)
    assert (q == q2).all() and (rem == rem2).all()


@pytest.mark.parametrize("m,n", [(1, 70), (2, 70), (255, 9), (256, 9), (511, 9), (4096, 12), (5000, 5)])
def test_quod_divisors(ctx, m, n):
    """The wave-parallel one-digit division on every divisor shape: 1, 2,
    powers of two, small odd, 2^32-1, 2^31, zero (quotient 0, remainder 0),
    random; all-ones and zero dividends; checked against the oracle and the
    closed form floor(a/d), a mod d."""
    rng = np.random.default_rng(m + n)
    a, _ = batch(n, m, m * 7 + 1)
    d = rnd(rng, n)
    fixed = [1, 2, 4, 3, 7, 0xFFFFFFFF, 0x80000000, 0, 65537]
    for j, v in enumerate(fixed[:n]):
        d[j] = v
    a[n - 1] = 0xFFFFFFFF
    if n > 2:
        a[n - 2] = 0
    q, rem = ctx.quod(a, d)
    nz = d != 0   # the oracle restates div.c's quod, which has no zero divisor
    q2, rem2 = O.quod(a[nz], d[nz])
    assert (q[nz] == q2).all() and (rem[nz] == rem2).all()
    for i in range(n):
        if d[i]:
            x = O.to_int(a[i])
            assert O.to_int(q[i]) == x // int(d[i]) and int(rem[i]) == x % int(d[i])
        else:
            assert not q[i].any() and rem[i] == 0
