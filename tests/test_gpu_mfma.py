"""GPU parity of the matrix-core multiply (byte convolution, limb_bits 8:
bi_tune_set_mul_limb_bits(ctx, 8)) against the oracle, bit-exact, through
the C ABI.  The products are those of the reference's multiply (low m digits,
cuda/ker-mul.cu.h:147-666); these cases cover what the byte convolution adds:
the balanced signed-byte rewrite of each operand (all-ones and 0x7F../0x80..
byte patterns, whose rewrite carries through every digit), every tile count
and partial tile (m from 33 to 1024), operands whose significant lengths cut
the tile and block ranges (top halves zero, single nonzero digits), zero
operands, unaligned and aliased buffers."""
import ctypes

import numpy as np
import pytest

import bigint_ffi as F
import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


def byte_edge_batch(n, m, seed):
    rng = np.random.default_rng(seed)
    a, b = rnd(rng, (n, m)), rnd(rng, (n, m))
    rows = [
        (0xFFFFFFFF, 0xFFFFFFFF),   # largest byte sums, carries through every digit
        (0x7F7F7F7F, 0x7F7F7F7F),   # rewrite propagates: every byte lands on 0xFF
        (0x80808080, 0x80808080),   # every signed byte -128
        (0x7F7F7F80, 0x80000000),
        (0, 0xFFFFFFFF),
    ]
    for i, (x, y) in enumerate(rows):
        if i < n:
            a[i], b[i] = x, y
    if n > 8:
        a[5, m // 2:] = 0            # top halves zero
        b[5, m // 2:] = 0
        a[6] = 0                     # single nonzero digits at the ends
        a[6, m - 1] = 0xFFFFFFFF
        b[6] = 0
        b[6, 0] = 0x80000001
        a[7, 1:] = 0                 # one-digit operand against a full one
        b[8, : m - 1] = 0            # only the top digit
    return a, b


def mul_bits(ctx, bits, a, b):
    old = ctx.tune_get_mul_limb_bits()
    ctx.tune_set_mul_limb_bits(bits)
    try:
        return ctx.mul(a, b)
    finally:
        ctx.tune_set_mul_limb_bits(old)


@pytest.mark.parametrize("m,n", [(33, 257), (40, 100), (63, 64), (64, 1000), (65, 300), (100, 129), (128, 500),
                                 (255, 40), (256, 513), (257, 31), (300, 64), (384, 99), (511, 9), (512, 700),
                                 (513, 20), (700, 33), (768, 64), (1000, 37), (1023, 12), (1024, 64)])
def test_mfma_mul_sizes(ctx, m, n):
    a, b = byte_edge_batch(n, m, seed=m * 17 + n)
    got = mul_bits(ctx, 8, a, b)
    assert (got == O.mul(a, b)).all()


@pytest.mark.parametrize("m", [64, 256, 512, 1024])
def test_mfma_mul_top_half_zero(ctx, m):
    """The reference benchmark's operands (top halves zero: the whole
    product), mixed in one launch with full-width and zero operands."""
    a, b = W.plain_pairs(777, m, seed=m + 5, top_zero=True)
    a[::5] = rnd(np.random.default_rng(m), (len(a[::5]), m))
    b[3::11] = 0
    assert (mul_bits(ctx, 8, a, b) == O.mul(a, b)).all()


def test_mfma_mul_full_batch(ctx):
    """131072 pairs of 512 digits with the top halves zero (the multiply
    workload at full size), every instance against the oracle."""
    n, m = 131072, 512
    a, b = W.plain_pairs(n, m, seed=72, top_zero=True)
    a[::97] = 0xFFFFFFFF
    a[::97, m // 2:] = 0
    b[::97] = 0xFFFFFFFF
    b[::97, m // 2:] = 0
    da, db, dr = ctx.upload(a), ctx.upload(b), ctx.alloc(a.nbytes)
    old = ctx.tune_get_mul_limb_bits()
    ctx.tune_set_mul_limb_bits(8)
    try:
        ctx.mul_dev(m, n, da, db, dr)
    finally:
        ctx.tune_set_mul_limb_bits(old)
    got = dr.download(np.uint32, n * m).reshape(n, m)
    assert (got == O.mul(a, b)).all()


def test_mfma_mul_unaligned_and_in_place(ctx):
    """Buffers one digit past a 16-byte boundary (scalar loads and stores)
    and the result written over an operand."""
    m, n = 512, 300
    a, b = byte_edge_batch(n, m, seed=4242)
    want = O.mul(a, b)
    old = ctx.tune_get_mul_limb_bits()
    ctx.tune_set_mul_limb_bits(8)
    try:
        da = ctx.upload(np.concatenate([[7], a.ravel()]).astype(np.uint32))
        db = ctx.upload(np.concatenate([[9], b.ravel()]).astype(np.uint32))
        dr = ctx.alloc((n * m + 1) * 4)
        off = lambda buf: ctypes.c_void_p(buf.ptr.value + 4)
        ctx.check(F.lib().bi_mul_u32_dev(ctx.h, m, n, off(da), off(db), off(dr)), "mul")
        assert (dr.download(np.uint32, n * m + 1)[1:].reshape(n, m) == want).all()
        da2, db2 = ctx.upload(a), ctx.upload(b)
        ctx.mul_dev(m, n, da2, db2, da2)    # rs aliases as
        assert (da2.download(np.uint32, n * m).reshape(n, m) == want).all()
    finally:
        ctx.tune_set_mul_limb_bits(old)


def test_mfma_outside_its_range_uses_the_digit_convolution(ctx):
    """limb_bits 8 outside 33..1024 digits falls back to the other kernels
    with the same results."""
    for m, n in [(16, 50), (2048, 5), (1025, 7)]:
        a, b = byte_edge_batch(n, m, seed=m)
        assert (mul_bits(ctx, 8, a, b) == O.mul(a, b)).all()
