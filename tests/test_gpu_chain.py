"""GPU parity of the fused chains (bi_add_chain_u32_dev, bi_mul_chain_u32_dev)
against the oracle's restatement of the reference's bench chains
(cuda/ker-add.cu.h:88-112, cuda/ker-mul.cu.h:223-250), bit-exact, through the
C ABI: register/LDS-resident paths (add m <= 2048, mul m <= 512) and the
per-link fallbacks above them, carry chains, zero operands, ragged tails."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


def batch(n, m, seed):
    rng = np.random.default_rng(seed)
    a, b = rnd(rng, (n, m)), rnd(rng, (n, m))
    if n > 5:
        a[0], b[0] = 0xFFFFFFFF, 0
        b[0, 0] = 1
        a[1], b[1] = 0xFFFFFFFF, 0xFFFFFFFF
        a[2] = 0
        a[3, m // 2:] = 0
        b[3, m // 2:] = 0
        a[4] = 0
        a[4, 0] = 1
    return a, b


@pytest.mark.parametrize("m", [1, 3, 4, 8, 64, 100, 256, 300, 512, 1024, 2048, 2100])
@pytest.mark.parametrize("p", [1, 2, 10])
def test_add_chain(ctx, m, p):
    n = 37 if m > 1024 else 301
    a, b = batch(n, m, seed=m * 7 + p)
    r = ctx.add_chain(a, b, p)
    assert (r == O.add_chain(a, b, p)).all()


@pytest.mark.parametrize("m", [1, 3, 8, 16, 64, 67, 128, 256, 500, 512, 513, 1024])
@pytest.mark.parametrize("p", [1, 2, 6])
def test_mul_chain(ctx, m, p):
    n = 23 if m > 512 else 131
    a, b = batch(n, m, seed=m * 11 + p)
    r = ctx.mul_chain(a, b, p)
    assert (r == O.mul_chain(a, b, p)).all()


def test_chain_closed_forms_small(ctx):
    a, b = batch(40, 24, seed=5)
    assert (ctx.add_chain(a, b, 10) == O.py_add_chain(a, b, 10)).all()
    assert (ctx.mul_chain(a, b, 6) == O.py_mul_chain(a, b, 6)).all()


def test_chain_bench_shape_sample(ctx):
    """131072 x 512 (the measured shape): every instance against the oracle
    for the ten-add chain; a 2048-instance slice for the six-product chain."""
    import bigint_workloads as W
    a, b = W.plain_pairs(131072, 512, seed=77, chain_frac=0.05)
    r = ctx.add_chain(a, b, 10)
    assert (r == O.add_chain(a, b, 10, threads=8)).all()
    a2, b2 = a[:2048], b[:2048]
    assert (ctx.mul_chain(a2, b2, 6) == O.mul_chain(a2, b2, 6, threads=8)).all()
