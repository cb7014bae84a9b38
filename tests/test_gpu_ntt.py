"""GPU parity of the full products of long operands (bi_mul_full_u32_dev):
the transform (algo 2) and the quadratic long product (algo 1) against the
oracle's schoolbook and GMP, at every transform size, with all-ones operands
(the largest coefficients and the longest carry chains), zeros, unequal
lengths and more pairs than the persistent grid."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


SHAPES = [(1, 1), (3, 5), (64, 64), (100, 27), (128, 128), (300, 212), (1000, 24), (513, 511),
          (1024, 1024), (2048, 2048), (1500, 2596), (4096, 4096), (4095, 4097), (8191, 1), (1, 8191)]


@pytest.mark.parametrize("ma,mb", SHAPES)
@pytest.mark.parametrize("algo", [1, 2])
def test_full_product(ctx, ma, mb, algo):
    rng = np.random.default_rng(ma * 31 + mb)
    n = 6
    a, b = rnd(rng, (n, ma)), rnd(rng, (n, mb))
    a[1], b[1] = 0xFFFFFFFF, 0xFFFFFFFF          # largest coefficients, full carry chains
    a[2] = 0                                     # zero operand
    b[3] = 0
    b[3, 0] = 1                                  # times one
    a[4, :] = 0xFFFFFFFF
    b[4, :] = 0
    b[4, -1] = 1                                 # (B^ma - 1) B^(mb-1)
    r = ctx.mul_full(a, b, algo=algo)
    assert (r == O.mul_full(a, b, threads=8)).all()


@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("ma,mb", [(1024, 1024), (3000, 1000), (4096, 4096)])
def test_transform_vs_gmp(ctx, ma, mb):
    rng = np.random.default_rng(ma + 7 * mb)
    a, b = rnd(rng, (16, ma)), rnd(rng, (16, mb))
    r = ctx.mul_full(a, b, algo=2)
    assert (r == O.Gmp().batch("mul", a, b, m_out=ma + mb)).all()


def test_transform_many_pairs(ctx):
    """More pairs than the persistent grid holds (each workgroup loops)."""
    rng = np.random.default_rng(5)
    a, b = rnd(rng, (3000, 1024)), rnd(rng, (3000, 1024))
    a[::7] = 0xFFFFFFFF
    b[::7] = 0xFFFFFFFF
    r = ctx.mul_full(a, b, algo=2)
    sample = np.r_[0:40, 2960:3000, 7:3000:97]
    assert (r[sample] == O.mul_full(a[sample], b[sample], threads=8)).all()
    # the two algorithms agree on every pair
    assert (r == ctx.mul_full(a, b, algo=1)).all()


def test_transform_many_pairs_parked(ctx):
    """The 2^13-point transform (first operand parked in global memory) with
    more pairs than its persistent grid."""
    rng = np.random.default_rng(6)
    a, b = rnd(rng, (700, 2048)), rnd(rng, (700, 2000))
    a[::5] = 0xFFFFFFFF
    b[::5] = 0xFFFFFFFF
    r = ctx.mul_full(a, b, algo=2)
    sample = np.r_[0:10, 690:700, 3:700:61]
    assert (r[sample] == O.mul_full(a[sample], b[sample], threads=8)).all()
    assert (r == ctx.mul_full(a, b, algo=1)).all()


@pytest.mark.parametrize("ma,mb", [(1536, 1536), (3000, 1100), (3072, 3072), (8191, 4101), (8192, 4096),
                                   (6000, 6000), (12796, 3), (4097, 4096)])
def test_transform_25bit_chunks(ctx, ma, mb):
    """Sizes where 25-bit chunks need fewer points than 16-bit ones, up to
    the division's largest product (8191 x 4101 digits) and the 12800-digit
    limit of a 2^14-point transform."""
    rng = np.random.default_rng(ma ^ (mb << 3))
    n = 4
    a, b = rnd(rng, (n, ma)), rnd(rng, (n, mb))
    a[1], b[1] = 0xFFFFFFFF, 0xFFFFFFFF
    a[2, : ma // 2] = 0
    b[3] = 0
    b[3, -1] = 1
    r = ctx.mul_full(a, b, algo=2)
    assert (r == O.mul_full(a, b, threads=8)).all()


def test_auto_and_shape_errors(ctx):
    import bigint_ffi as F
    rng = np.random.default_rng(9)
    a, b = rnd(rng, (4, 2000)), rnd(rng, (4, 1500))
    assert (ctx.mul_full(a, b, algo=0) == O.mul_full(a, b)).all()
    big_a, big_b = rnd(rng, (1, 9000)), rnd(rng, (1, 4000))
    with pytest.raises(F.BigIntError) as ei:
        ctx.mul_full(big_a, big_b, algo=2)          # 13000 digits: beyond the largest transform
    assert ei.value.code == -2
    assert (ctx.mul_full(big_a, big_b, algo=0) == O.mul_full(big_a, big_b)).all()
