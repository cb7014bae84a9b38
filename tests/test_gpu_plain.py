"""GPU parity of the plain batched add / multiply (bi_add_*, bi_mul_*) against
the oracle, bit-exact, through the C ABI.  Cases follow the reference's size
sweep (cuda/main.cu:476-513) plus what it never generates: odd digits,
all-ones carry chains, zero operands, non-multiple-of-4 sizes, unaligned
device pointers, tails that are not a multiple of any block shape."""
import os

import numpy as np
import pytest

import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


def edge_batch(n, m, seed):
    rng = np.random.default_rng(seed)
    a, b = rnd(rng, (n, m)), rnd(rng, (n, m))
    if n > 6:
        a[0], b[0] = 0xFFFFFFFF, 0
        b[0, 0] = 1
        a[1], b[1] = 0xFFFFFFFF, 0xFFFFFFFF
        a[2] = 0
        b[3, m // 2:] = 0
        a[3, m // 2:] = 0
        a[4, : m - 1] = 0xFFFFFFFF
        b[4] = 0
        b[4, 0] = 1
        a[5] = 0
        a[5, 0] = 1
    return a, b


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7, 8, 16, 31, 64, 67, 512])
def test_golden_vectors(ctx, m):
    v = np.load(os.path.join(GOLD, "vectors.npz"))
    a, b = v[f"a_{m}"], v[f"b_{m}"]
    assert (ctx.add(a, b) == v[f"add_{m}"]).all()
    assert (ctx.mul(a, b) == v[f"mul_{m}"]).all()


@pytest.mark.parametrize("m,n", [(1, 1), (2, 33), (3, 100), (4, 257), (5, 64), (16, 1000), (63, 77), (64, 1024),
                                 (65, 300), (100, 129), (255, 40), (256, 513), (257, 31), (511, 9), (512, 700),
                                 (1000, 37), (1024, 64), (2048, 20), (4096, 6), (8192, 3), (12345, 2)])
def test_add_sizes(ctx, m, n):
    a, b = edge_batch(n, m, seed=m * 7 + n)
    assert (ctx.add(a, b) == O.add(a, b)).all()


@pytest.mark.parametrize("m,n", [(1, 1), (2, 33), (3, 100), (4, 257), (5, 64), (8, 500), (16, 1000), (63, 77),
                                 (64, 1024), (65, 300), (100, 129), (255, 40), (256, 513), (257, 31), (511, 9),
                                 (512, 700), (700, 33), (863, 10), (864, 50), (865, 20), (1000, 37), (1024, 64),
                                 (2048, 20), (4096, 6)])
def test_mul_sizes(ctx, m, n):
    a, b = edge_batch(n, m, seed=m * 11 + n)
    assert (ctx.mul(a, b) == O.mul(a, b)).all()


@pytest.mark.parametrize("m", [64, 512, 864, 2048])
def test_mul_top_half_zero_and_mixed(ctx, m):
    """Top-half-zero operands take the full-product schedule; a batch mixing
    them with full-width operands exercises both schedules in one launch."""
    a, b = W.plain_pairs(301, m, seed=m, top_zero=True)
    a[::3] = rnd(np.random.default_rng(1), (len(a[::3]), m))
    b[1::7, : m // 3] = 0
    assert (ctx.mul(a, b) == O.mul(a, b)).all()


@pytest.mark.parametrize("m,n", [(1, 5), (7, 100), (64, 1024), (257, 31), (430, 64), (512, 700), (863, 9),
                                 (864, 20)])
def test_mul_limb_convolution(ctx, m, n):
    """The 27-bit limb convolution (the tuner's other choice below 865
    digits) gives the same digits as the 32-bit digit convolution, on both
    of its schedules (top halves zero: full product; else truncated)."""
    a, b = edge_batch(n, m, seed=m * 13 + n)
    a[n // 2:, m // 2:] = 0
    b[n // 2:, m // 2:] = 0
    a[n // 2, : m // 2] = 0xFFFFFFFF    # largest column sums of the full product
    b[n // 2, : m // 2] = 0xFFFFFFFF
    digit = ctx.mul(a, b)
    bits = ctx.tune_get_mul_limb_bits()
    ctx.tune_set_mul_limb_bits(27)
    try:
        limb = ctx.mul(a, b)
    finally:
        ctx.tune_set_mul_limb_bits(bits)
    assert (limb == digit).all()
    assert (digit == O.mul(a, b)).all()


def test_u64_view_is_u32_bytes(ctx):
    a, b = W.plain_pairs(200, 512, seed=4, chain_frac=0.05)
    a64, b64 = a.view(np.uint64), b.view(np.uint64)
    assert (ctx.add_u64(a64, b64).view(np.uint32) == O.add(a, b)).all()
    a, b = W.plain_pairs(200, 512, seed=5, top_zero=True)
    assert (ctx.mul_u64(a.view(np.uint64), b.view(np.uint64)).view(np.uint32) == O.mul(a, b)).all()


def test_unaligned_device_pointers(ctx):
    """Buffers offset by one digit take the scalar-load path."""
    import bigint_ffi as F
    m, n = 64, 97
    a, b = edge_batch(n, m, seed=3)
    da = ctx.upload(np.concatenate([[7], a.ravel()]).astype(np.uint32))
    db = ctx.upload(np.concatenate([[9], b.ravel()]).astype(np.uint32))
    dr = ctx.alloc((n * m + 1) * 4)
    import ctypes
    off = lambda buf: ctypes.c_void_p(buf.ptr.value + 4)
    ctx.check(F.lib().bi_add_u32_dev(ctx.h, m, n, off(da), off(db), off(dr)), "add")
    r = dr.download(np.uint32, n * m + 1)[1:].reshape(n, m)
    assert (r == O.add(a, b)).all()
    ctx.check(F.lib().bi_mul_u32_dev(ctx.h, m, n, off(da), off(db), off(dr)), "mul")
    r = dr.download(np.uint32, n * m + 1)[1:].reshape(n, m)
    assert (r == O.mul(a, b)).all()


def test_in_place_aliasing(ctx):
    m, n = 512, 300
    a, b = edge_batch(n, m, seed=8)
    da, db = ctx.upload(a), ctx.upload(b)
    ctx.add_dev(m, n, da, db, da)        # rs aliases as
    assert (da.download(np.uint32, n * m).reshape(n, m) == O.add(a, b)).all()
    da.upload(a)
    ctx.mul_dev(m, n, da, db, db)        # rs aliases bs
    assert (db.download(np.uint32, n * m).reshape(n, m) == O.mul(a, b)).all()


def test_config_small_full(ctx):
    """1024 pairs of 64-digit integers: add, and mul with the top 32 digits zero."""
    a, b = W.plain_pairs(1024, 64, seed=1)
    assert (ctx.add(a, b) == O.add(a, b)).all()
    a, b = W.plain_pairs(1024, 64, seed=2, top_zero=True)
    assert (ctx.mul(a, b) == O.mul(a, b)).all()
    g = O.Gmp() if O.Gmp.available() else None
    if g:
        assert (ctx.mul(a[:64], b[:64]) == g.batch("mul", a[:64], b[:64])).all()


def test_config_large_full(ctx):
    """131072 pairs of 512 digits, 5% all-ones + 1 chains: whole-batch add
    parity; multiply (top halves zero) checked on every instance."""
    n, m = 131072, 512
    a, b = W.plain_pairs(n, m, seed=11, chain_frac=0.05)
    r = ctx.add(a, b)
    assert (r == O.add(a, b)).all()
    chains = (a == 0xFFFFFFFF).all(axis=1)
    assert chains.sum() == 6554 and (r[chains] == 0).all()
    a, b = W.plain_pairs(n, m, seed=12, top_zero=True)
    assert (ctx.mul(a, b) == O.mul(a, b)).all()


def test_error_codes(ctx):
    import bigint_ffi as F
    L = F.lib()
    x = np.zeros((1, 8), np.uint32)
    p = x.ctypes.data_as(__import__("ctypes").c_void_p)
    assert L.bi_add_u32(ctx.h, 0, 1, p, p, p) == -1          # m == 0
    assert L.bi_mul_u32(ctx.h, 5000, 1, p, p, p) == -2       # beyond the multiply's size limit
    assert L.bi_add_u32(ctx.h, 8, 0, p, p, p) == 0           # empty batch is a no-op


def _long_edge_batch(n, m, seed):
    """Instances whose carries cross chunk boundaries: all-ones + 1 (one
    generating chunk, then all-ones chunks), b = ~a with a carry planted in
    the lowest digit (a chain through every chunk), sums all-ones except one
    digit at a chunk boundary, and random digits."""
    rng = np.random.default_rng(seed)
    a, b = rnd(rng, (n, m)), rnd(rng, (n, m))
    a[0], b[0] = 0xFFFFFFFF, 0
    b[0, 0] = 1
    if n > 1:
        b[1] = ~a[1]
        a[1, 0], b[1, 0] = 0xFFFFFFFF, 1
    if n > 2:
        b[2] = ~a[2]
        b[2, 0] = 0xFFFFFFFF - a[2, 0] + 1 if a[2, 0] else 1
        k = min(m - 1, 1024)
        b[2, k] = 0 if a[2, k] != 0xFFFFFFFF else 1       # the chain stops at a chunk start
    if n > 3:
        a[3] = 0xFFFFFFFF
        b[3] = 0                                          # all-ones chunks, no carry at all
    return a, b


@pytest.mark.parametrize("m,n", [(1025, 9), (2048, 40), (3000, 7), (4096, 33), (8192, 12), (20000, 3),
                                 (1 << 20, 2)])
@pytest.mark.parametrize("span,tiles", [(256, 1), (1024, 1), (1024, 4), (2048, 2), (4096, 1)])
def test_add_long_instances_in_chunks(ctx, m, n, span, tiles):
    """Instances longer than the span are added in span-sized chunks by
    separate waves and the carries between chunks resolved by a second
    pass: every digit against the oracle, for chunk sizes that do and do not
    divide m."""
    a, b = _long_edge_batch(n, m, seed=m + span + tiles)
    keep = ctx.tune_get(m)
    ctx.tune_set_add(m, span, tiles, 1)
    try:
        r = ctx.add(a, b)
    finally:
        ctx.tune_set_add(m, keep["add_span_digits"], keep["add_tiles_in_flight"], keep["add_nontemporal"])
    assert (r == O.add(a, b)).all()
    assert (r[0] == 0).all()


def test_add_long_chunks_unaligned_and_in_place(ctx):
    """The chunked add with buffers one digit past a 16-byte boundary (scalar
    loads) and with the result written over an operand."""
    import ctypes
    import bigint_ffi as F
    m, n = 5000, 6
    a, b = _long_edge_batch(n, m, seed=5)
    want = O.add(a, b)
    da = ctx.upload(np.concatenate([[7], a.ravel()]).astype(np.uint32))
    db = ctx.upload(np.concatenate([[9], b.ravel()]).astype(np.uint32))
    dr = ctx.alloc((n * m + 1) * 4)
    off = lambda buf: ctypes.c_void_p(buf.ptr.value + 4)
    ctx.check(F.lib().bi_add_u32_dev(ctx.h, m, n, off(da), off(db), off(dr)), "add")
    assert (dr.download(np.uint32, n * m + 1)[1:].reshape(n, m) == want).all()
    da2, db2 = ctx.upload(a), ctx.upload(b)
    ctx.add_dev(m, n, da2, db2, db2)        # rs aliases bs
    assert (db2.download(np.uint32, n * m).reshape(n, m) == want).all()


@pytest.mark.parametrize("m,n", [(3000, 2500), (64, 50000), (4096, 1100)])
def test_host_pointer_pipeline_chunks(ctx, m, n):
    """The host-pointer entry points move batches larger than one 8 MiB chunk
    through pinned staging in chunks on two streams (a last partial chunk
    included): add and multiply, u32 and u64 views, against the oracle."""
    a, b = _long_edge_batch(n, m, seed=m + n)
    assert (ctx.add(a, b) == O.add(a, b)).all()
    assert (ctx.add_u64(a.view(np.uint64), b.view(np.uint64)).view(np.uint32) == O.add(a, b)).all()
    want = O.mul(a, b)
    assert (ctx.mul(a, b) == want).all()
    assert (ctx.mul_u64(a.view(np.uint64), b.view(np.uint64)).view(np.uint32) == want).all()


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 8, 11, 12, 16, 17, 24, 31, 32, 33, 36, 41, 48, 57, 63, 64])
def test_mul_short_one_lane_per_instance(ctx, m):
    """Batches of at least 16384 instances of at most 32 digits (131072 of up
    to 64) multiply one lane per instance with the operands in registers:
    every m (the 4-digit
    rounding included), all-ones operands (the largest column sums), zero
    operands, against the oracle."""
    n = 20000 if m <= 32 else 131072 + 77     # above 32 digits from 131072 instances
    a, b = edge_batch(n, m, seed=m * 3)
    a[7], b[7] = 0xFFFFFFFF, 0xFFFFFFFF
    assert (ctx.mul(a, b) == O.mul(a, b)).all()


@pytest.mark.parametrize("m", [8, 16, 31, 32, 40, 63, 64])
def test_mul_short_top_half_zero_waves(ctx, m):
    """Waves whose instances all have zero top halves (both operands below
    B^(M/2), M = m rounded up to 4) take the whole product of the low halves;
    the other waves the low product.  A batch of such instances with every
    64th instance full width (so some waves take each path), all-ones low
    halves (the largest whole products), against the oracle."""
    n = (16384 if m <= 32 else 131072) + 300
    M = (m + 3) // 4 * 4
    a, b = W.plain_pairs(n, m, seed=m + 11)
    a[:, M // 2:] = 0
    b[:, M // 2:] = 0
    a[5, : M // 2] = 0xFFFFFFFF
    b[5, : M // 2] = 0xFFFFFFFF
    full = np.arange(4000, n, 64 * 5)
    a[full] = rnd(np.random.default_rng(m), (len(full), m))
    assert (ctx.mul(a, b) == O.mul(a, b)).all()


def test_mul_short_unaligned_and_in_place(ctx):
    """The one-lane-per-instance multiply with buffers one digit past a
    16-byte boundary (scalar loads) and with the result over an operand."""
    import ctypes
    import bigint_ffi as F
    m, n = 16, 17000
    a, b = edge_batch(n, m, seed=99)
    want = O.mul(a, b)
    da = ctx.upload(np.concatenate([[7], a.ravel()]).astype(np.uint32))
    db = ctx.upload(np.concatenate([[9], b.ravel()]).astype(np.uint32))
    dr = ctx.alloc((n * m + 1) * 4)
    off = lambda buf: ctypes.c_void_p(buf.ptr.value + 4)
    ctx.check(F.lib().bi_mul_u32_dev(ctx.h, m, n, off(da), off(db), off(dr)), "mul")
    assert (dr.download(np.uint32, n * m + 1)[1:].reshape(n, m) == want).all()
    da2, db2 = ctx.upload(a), ctx.upload(b)
    ctx.mul_dev(m, n, da2, db2, da2)        # rs aliases as
    assert (da2.download(np.uint32, n * m).reshape(n, m) == want).all()
