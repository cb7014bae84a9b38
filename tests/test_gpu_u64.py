"""GPU parity of the u64-limb device entry points (bi_add_u64_dev,
bi_mul_u64_dev, bi_divmod_u64_dev) through the C ABI.

The reference's u64 runs use the same bytes as its u32 runs with half the
limb count (cuda/main.cu:255-258 validates them as u32 words), and its
division prototype is specified over u64 limbs for the large division
workload (prototypes/div.c:427).  Each case here hands u64 device buffers to
the u64 entry point and checks the bytes against the oracle (u32 view) or
GMP."""
import numpy as np
import pytest

import bigint_ffi as F
import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu


def _u64_call(ctx, name, m64, a64, b64):
    n = a64.shape[0]
    da, db, dr = ctx.upload(a64), ctx.upload(b64), ctx.alloc(a64.nbytes)
    ctx.check(getattr(F.lib(), name)(ctx.h, m64, n, da.ptr, db.ptr, dr.ptr), name)
    out = dr.download(np.uint64, n * m64).reshape(n, m64)
    for x in (da, db, dr):
        ctx._release(x)
    return out


def test_add_u64_dev_full_batch(ctx):
    """131072 pairs of 256 u64 limbs (the large uniform batch in u64 limbs),
    5% all-ones + 1 chains, every instance against the oracle."""
    n, m = 131072, 512
    a, b = W.plain_pairs(n, m, seed=61, chain_frac=0.05)
    r = _u64_call(ctx, "bi_add_u64_dev", m // 2, a.view(np.uint64), b.view(np.uint64))
    assert (r.view(np.uint32) == O.add(a, b)).all()


def test_mul_u64_dev_full_batch(ctx):
    """131072 pairs of 256 u64 limbs with the top 128 limbs zero: low 256
    limbs of every product against the oracle."""
    n, m = 131072, 512
    a, b = W.plain_pairs(n, m, seed=62, top_zero=True)
    r = _u64_call(ctx, "bi_mul_u64_dev", m // 2, a.view(np.uint64), b.view(np.uint64))
    assert (r.view(np.uint32) == O.mul(a, b)).all()


@pytest.mark.parametrize("m64,n", [(1, 7), (3, 100), (33, 64), (2048, 6)])
def test_add_mul_u64_dev_sizes(ctx, m64, n):
    rng = np.random.default_rng(m64 + n)
    a = rng.integers(0, 1 << 64, (n, m64), dtype=np.uint64)
    b = rng.integers(0, 1 << 64, (n, m64), dtype=np.uint64)
    a[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    b[0] = 0
    b[0, 0] = 1
    a32, b32 = a.view(np.uint32), b.view(np.uint32)
    assert (_u64_call(ctx, "bi_add_u64_dev", m64, a, b).view(np.uint32) == O.add(a32, b32)).all()
    assert (_u64_call(ctx, "bi_mul_u64_dev", m64, a, b).view(np.uint32) == O.mul(a32, b32)).all()


def _divmod_u64(ctx, u64, v64):
    n, mu = u64.shape
    mv = v64.shape[1]
    du, dv = ctx.upload(u64), ctx.upload(v64)
    dq, dr = ctx.alloc(u64.nbytes), ctx.alloc(v64.nbytes)
    ctx.check(F.lib().bi_divmod_u64_dev(ctx.h, n, mu, mv, du.ptr, dv.ptr, dq.ptr, dr.ptr), "bi_divmod_u64_dev")
    q = dq.download(np.uint64, n * mu).reshape(n, mu)
    r = dr.download(np.uint64, n * mv).reshape(n, mv)
    for x in (du, dv, dq, dr):
        ctx._release(x)
    return q, r


def test_divmod_u64_dev_division_workload_sample(ctx):
    """32 divisions at the large division workload's u64 shape: numerators of
    4096 u64 limbs by divisors of 2048 u64 limbs (top limb nonzero), through
    bi_divmod_u64_dev; quotient and remainder bit-exact against GMP
    (mpz_tdiv_qr), else Python divmod."""
    n, mu, mv = 32, 8192, 4096
    u, v = W.div_pairs(n, mu, mv, seed=63)
    q, r = _divmod_u64(ctx, u.view(np.uint64), v.view(np.uint64))
    if O.Gmp.available():
        gq, gr = O.Gmp().batch("div", u, v)
    else:
        gq, gr = O.py_div(u, v)
    assert (q.view(np.uint32) == gq).all() and (r.view(np.uint32) == gr).all()


@pytest.mark.parametrize("mu64,mv64,n", [(1, 1, 40), (4, 2, 64), (100, 37, 16)])
def test_divmod_u64_dev_small(ctx, mu64, mv64, n):
    """Short u64 divisions, divisors with zero top limbs included, against
    Python divmod over the same bytes."""
    rng = np.random.default_rng(mu64 * 3 + mv64)
    u = rng.integers(0, 1 << 64, (n, mu64), dtype=np.uint64)
    v = np.zeros((n, mv64), np.uint64)
    for i in range(n):
        k = int(rng.integers(1, mv64 + 1))
        v[i, :k] = rng.integers(1, 1 << 64, k, dtype=np.uint64)
    q, r = _divmod_u64(ctx, u, v)
    gq, gr = O.py_div(u.view(np.uint32), v.view(np.uint32))
    assert (q.view(np.uint32) == gq).all() and (r.view(np.uint32) == gr).all()


@pytest.mark.parametrize("m64,n", [(1, 50), (8, 300), (256, 200), (2048, 9)])
def test_u64_helpers(ctx, m64, n):
    """|a-b|, compare, precision and shift through the u64-limb entry points:
    the same bytes as the u32 helpers; precision counted in u64 limbs
    (index of the top nonzero limb + 1, 1 for zero) and shifts by whole u64
    limbs, both checked against numpy on the u64 arrays."""
    rng = np.random.default_rng(m64 * 5 + n)
    a = rng.integers(0, 1 << 64, (n, m64), dtype=np.uint64)
    b = rng.integers(0, 1 << 64, (n, m64), dtype=np.uint64)
    b[::4] = a[::4]
    a[1::5, m64 // 2:] = 0
    a[2::7] = 0
    if m64 > 1:
        b[3::6, 1:] = a[3::6, 1:]        # decided by the lowest limb
    L = F.lib()
    da, db = ctx.upload(a), ctx.upload(b)
    dr, ds, dc, dp = ctx.alloc(a.nbytes), ctx.alloc(max(n, 4)), ctx.alloc(max(n, 4)), ctx.alloc(4 * n)
    ctx.check(L.bi_sub_u64_dev(ctx.h, m64, n, da.ptr, db.ptr, dr.ptr, ds.ptr), "bi_sub_u64_dev")
    r2, s2 = O.sub(a.view(np.uint32), b.view(np.uint32))
    assert (dr.download(np.uint64, n * m64).reshape(n, m64).view(np.uint32) == r2).all()
    assert (ds.download(np.uint8, n) == s2).all()
    ctx.check(L.bi_cmp_u64_dev(ctx.h, m64, n, da.ptr, db.ptr, dc.ptr), "bi_cmp_u64_dev")
    assert (dc.download(np.int8, n) == O.cmp(a.view(np.uint32), b.view(np.uint32))).all()
    ctx.check(L.bi_prec_u64_dev(ctx.h, m64, n, da.ptr, dp.ptr), "bi_prec_u64_dev")
    nz = a != 0
    want = np.where(nz.any(axis=1), m64 - np.argmax(nz[:, ::-1], axis=1), 1).astype(np.uint32)
    assert (dp.download(np.uint32, n) == want).all()
    for s in (-m64 - 1, -3, -1, 0, 1, 2, m64 // 2, m64):
        ctx.check(L.bi_shift_u64_dev(ctx.h, m64, n, s, da.ptr, dr.ptr), "bi_shift_u64_dev")
        got = dr.download(np.uint64, n * m64).reshape(n, m64)
        want = np.zeros_like(a)
        if 0 <= s < m64:
            want[:, s:] = a[:, : m64 - s]
        elif -m64 < s < 0:
            want[:, : m64 + s] = a[:, -s:]
        assert (got == want).all(), s
    for x in (da, db, dr, ds, dc, dp):
        ctx._release(x)
