"""GPU parity of the host-pointer forms of the division and of the ragged
add and multiply (bi_divmod_u32 / bi_divmod_u64, bi_add_csr_u32,
bi_mul_csr_u32): operands and results in host memory, one round trip a call,
against GMP (or Python divmod) and the oracle."""
import ctypes

import numpy as np
import pytest

import bigint_ffi as F
import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu

MARK = 0x5A5A5A5A


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def expect_div(u, v):
    return O.gmp_div(u, v, 8) if O.gmp_available() else O.py_div(u, v)


@pytest.mark.parametrize("n,mu,mv", [(1, 5, 3), (7, 64, 32), (300, 4096, 2048), (20, 8192, 4096)])
def test_divmod_u32_host(ctx, n, mu, mv):
    u, v = W.div_pairs(n, mu, mv, seed=n + mu)
    q = np.zeros((n, mu), np.uint32)
    r = np.zeros((n, mv), np.uint32)
    ctx.check(F.lib().bi_divmod_u32(ctx.h, n, mu, mv, ptr(u), ptr(v), ptr(q), ptr(r)), "bi_divmod_u32")
    gq, gr = expect_div(u, v)
    assert (q == gq).all() and (r == gr).all()


def test_divmod_u64_host_c4_sample(ctx):
    """The C4 shape in u64 limbs (u 4096, v 2048 limbs), a sample of 16, with
    a zero divisor in it: BI_ERR_DIVZERO, that instance q = r = 0, every
    other one exact."""
    n, mu, mv = 16, 4096, 2048
    u32, v32 = W.div_pairs(n, 2 * mu, 2 * mv, seed=44)
    v32[3] = 0
    u, v = u32.view(np.uint64), v32.view(np.uint64)
    q = np.full((n, mu), 7, np.uint64)
    r = np.full((n, mv), 7, np.uint64)
    rc = F.lib().bi_divmod_u64(ctx.h, n, mu, mv, ptr(u), ptr(v), ptr(q), ptr(r))
    assert rc == -5
    assert not q[3].any() and not r[3].any()
    keep = [i for i in range(n) if i != 3]
    gq, gr = expect_div(u32[keep], v32[keep])
    assert (q.view(np.uint32)[keep] == gq).all() and (r.view(np.uint32)[keep] == gr).all()


@pytest.mark.parametrize("lead", [0, 37, 3000])
def test_add_csr_host(ctx, lead):
    o, a0, b0 = W.ragged(700, 1, 2048, seed=lead + 5)
    off = o + lead
    rng = np.random.default_rng(lead)
    a = np.concatenate([rng.integers(0, 1 << 32, lead, dtype=np.uint64).astype(np.uint32), a0])
    b = np.concatenate([np.full(lead, 0xFFFFFFFF, np.uint32), b0])
    r = np.full(len(a), MARK, np.uint32)
    ctx.check(F.lib().bi_add_csr_u32(ctx.h, len(off) - 1, ptr(off), ptr(a), ptr(b), ptr(r)), "bi_add_csr_u32")
    assert (r[:lead] == MARK).all()
    assert (r[lead:] == O.add_csr(o, a0, b0)).all()


@pytest.mark.parametrize("mode", [0, 1])
def test_mul_csr_host(ctx, mode):
    off, a, b = W.ragged(400, 1, 1500, seed=17 + mode)
    sizes = np.diff(off)
    want = O.mul_csr(off, a, b, mode=mode)
    if mode == 0:
        r = np.full(len(a), MARK, np.uint32)
        ctx.check(F.lib().bi_mul_csr_u32(ctx.h, len(off) - 1, ptr(off), None, ptr(a), ptr(b), ptr(r), 0),
                  "bi_mul_csr_u32")
        assert (r == want).all()
        return
    # full products into windows 3 digits longer than they need: the gaps keep the caller's digits
    off_out = np.zeros_like(off)
    np.cumsum(2 * sizes + 3, out=off_out[1:])
    r = np.full(int(off_out[-1]), MARK, np.uint32)
    ctx.check(F.lib().bi_mul_csr_u32(ctx.h, len(off) - 1, ptr(off), ptr(off_out), ptr(a), ptr(b), ptr(r), 1),
              "bi_mul_csr_u32")
    want, dense = want
    for i in range(len(sizes)):
        s, k = int(off_out[i]), int(2 * sizes[i])
        assert (r[s:s + k] == want[dense[i]:dense[i] + k]).all(), i
        assert (r[s + k:off_out[i + 1]] == MARK).all(), i


def test_csr_host_rejects_bad_offsets(ctx):
    """Offsets are checked on the host before anything runs: decreasing,
    negative, or (mode 1) output windows too short for 2 n_i digits."""
    a = np.ones(64, np.uint32)
    r = np.zeros(64, np.uint32)
    bad = np.array([0, 10, 5, 20], np.int64)
    neg = np.array([-1, 10, 20, 30], np.int64)
    assert F.lib().bi_add_csr_u32(ctx.h, 3, ptr(bad), ptr(a), ptr(a), ptr(r)) == -1
    assert F.lib().bi_add_csr_u32(ctx.h, 3, ptr(neg), ptr(a), ptr(a), ptr(r)) == -1
    off = np.array([0, 10, 20, 30], np.int64)
    short = np.array([0, 20, 39, 60], np.int64)
    rr = np.zeros(64, np.uint32)
    assert F.lib().bi_mul_csr_u32(ctx.h, 3, ptr(off), ptr(short), ptr(a), ptr(a), ptr(rr), 1) == -1
    assert F.lib().bi_mul_csr_u32(ctx.h, 3, ptr(off), None, ptr(a), ptr(a), ptr(rr), 2) == -1
