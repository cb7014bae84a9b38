"""GPU parity of the ragged (CSR) add and multiply against the oracle,
including the single persistent launch that does both."""
import numpy as np
import pytest

import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,lo,hi,seed", [(1, 1, 1, 1), (7, 1, 9, 2), (500, 1, 40, 3), (2000, 8, 2048, 4),
                                          (300, 500, 4096, 5), (4000, 1, 300, 6)])
def test_add_csr(ctx, N, lo, hi, seed):
    off, a, b = W.ragged(N, lo, hi, seed=seed)
    assert (ctx.add_csr(off, a, b) == O.add_csr(off, a, b)).all()


def test_add_csr_empty_instances(ctx):
    sizes = np.array([3, 0, 0, 5, 0, 1, 0, 300, 0], np.int64)
    off, a, b = W.ragged(len(sizes), sizes=sizes, seed=9)
    assert (ctx.add_csr(off, a, b) == O.add_csr(off, a, b)).all()


@pytest.mark.parametrize("N,lo,hi,seed", [(1, 5, 5, 1), (64, 1, 9, 2), (500, 1, 100, 3), (1500, 8, 2048, 4),
                                          (40, 1000, 4096, 5)])
@pytest.mark.parametrize("lanes", [0, 64, 128, 256, 512, 1024])
@pytest.mark.parametrize("rows", [4, 8])
def test_mul_csr_low(ctx, N, lo, hi, seed, lanes, rows):
    """Every lane budget and row count the ragged tuner can pick."""
    off, a, b = W.ragged(N, lo, hi, seed=seed)
    # a third of the instances with zero top halves take the full schedule
    for i in range(0, N, 3):
        s, e = off[i], off[i + 1]
        a[s + (e - s) // 2:e] = 0
        b[s + (e - s) // 2:e] = 0
    keep = ctx.tune_get_csr()
    ctx.tune_set_csr(keep["csr_add_span_digits"], keep["csr_add_tiles_in_flight"], keep["csr_add_nontemporal"],
                     keep["csr_mul_lanes"], rows)
    try:
        r = ctx.mul_csr(off, a, b, mode=0, lane_budget=lanes)
    finally:
        ctx.tune_set_csr(*keep.values())
    assert (r == O.mul_csr(off, a, b, mode=0)).all()


@pytest.mark.parametrize("span,tiles", [(512, 1), (768, 2), (1024, 1), (1536, 4), (2048, 2), (4096, 4)])
def test_add_csr_tuned_shapes(ctx, span, tiles):
    """The ragged add at every span / tiles-in-flight class the tuner sweeps,
    alone and inside the one-launch add+multiply."""
    off, a, b = W.ragged(3000, 1, 2048, seed=span + tiles)
    keep = ctx.tune_get_csr()
    ctx.tune_set_csr(span, tiles, 1, keep["csr_mul_lanes"], keep["csr_mul_rows"])
    try:
        r = ctx.add_csr(off, a, b)
        rm, ra = ctx.mul_csr(off, a, b, mode=0, with_add=True)
    finally:
        ctx.tune_set_csr(*keep.values())
    want = O.add_csr(off, a, b)
    assert (r == want).all() and (ra == want).all()
    assert (rm == O.mul_csr(off, a, b, mode=0)).all()


def test_ragged_tuner_keeps_results(ctx):
    """bi_tune_run(what = 2) picks shapes from a sample; results after it are
    unchanged, and the chosen shape is one the tests above cover."""
    off, a, b = W.ragged(2000, 8, 2048, seed=77)
    before = ctx.mul_csr(off, a, b, mode=0, with_add=True)
    keep = ctx.tune_get_csr()
    try:
        ctx.tune(2048, 4 << 20, 2)
        got = ctx.tune_get_csr()
        assert got["csr_mul_lanes"] in (64, 128, 256, 512, 1024) and got["csr_mul_rows"] in (4, 8)
        assert got["csr_add_tiles_in_flight"] in (1, 2, 4) and 512 <= got["csr_add_span_digits"] <= 4096
        after = ctx.mul_csr(off, a, b, mode=0, with_add=True)
    finally:
        ctx.tune_set_csr(*keep.values())
    assert (after[0] == before[0]).all() and (after[1] == before[1]).all()
    assert (after[1] == O.add_csr(off, a, b)).all()


@pytest.mark.parametrize("N,lo,hi,seed", [(1, 3, 3, 1), (300, 1, 64, 2), (800, 8, 2048, 3)])
def test_mul_csr_full(ctx, N, lo, hi, seed):
    off, a, b = W.ragged(N, lo, hi, seed=seed)
    r, oo = ctx.mul_csr(off, a, b, mode=1)
    r2, oo2 = O.mul_csr(off, a, b, mode=1)
    assert (oo == oo2).all() and (r == r2).all()


def test_addmul_one_launch(ctx):
    off, a, b = W.ragged(3000, 8, 2048, seed=12)
    rm, ra = ctx.mul_csr(off, a, b, mode=0, with_add=True)
    assert (ra == O.add_csr(off, a, b)).all()
    assert (rm == O.mul_csr(off, a, b, mode=0)).all()


def test_csr_unaligned_device_pointers(ctx):
    """Operand and result buffers one digit past a 16-byte boundary take the
    scalar-load instantiation of the span walk, in the add-only launch and in
    the launch that adds and multiplies together."""
    import ctypes
    import bigint_ffi as F
    off, a, b = W.ragged(700, 1, 600, seed=21)
    total = int(off[-1])
    doff = ctx.upload(off)
    da = ctx.upload(np.concatenate([[7], a]).astype(np.uint32))
    db = ctx.upload(np.concatenate([[9], b]).astype(np.uint32))
    dr, dm = ctx.alloc((total + 1) * 4), ctx.alloc((total + 1) * 4)
    plus1 = lambda buf: ctypes.c_void_p(buf.ptr.value + 4)
    ctx.check(F.lib().bi_add_csr_u32_dev(ctx.h, len(off) - 1, doff.ptr, total, plus1(da), plus1(db), plus1(dr)),
              "add_csr")
    want = O.add_csr(off, a, b)
    assert (dr.download(np.uint32, total + 1)[1:] == want).all()
    plan = ctx.csr_plan(off, 0)
    ctx.check(F.lib().bi_addmul_csr_u32_dev(ctx.h, plan.h, doff.ptr, None, plus1(da), plus1(db), plus1(dr),
                                            plus1(dm)), "addmul_csr")
    assert (dr.download(np.uint32, total + 1)[1:] == want).all()
    assert (dm.download(np.uint32, total + 1)[1:] == O.mul_csr(off, a, b, mode=0)).all()


def test_mul_csr_empty_instances(ctx):
    sizes = np.array([0, 4, 0, 0, 17, 0, 1, 0], np.int64)
    off, a, b = W.ragged(len(sizes), sizes=sizes, seed=13)
    assert (ctx.mul_csr(off, a, b) == O.mul_csr(off, a, b)).all()


def test_ragged_config_every_instance(ctx):
    """The measured ragged batch (2^17 instances, 8..2048 digits, carry chains,
    zero operands): add and low-product of every instance against the oracle,
    from the one-launch add+multiply entry point."""
    N = 1 << 17
    off, a, b = W.ragged(N, 8, 2048, seed=31)
    p, s = ctx.mul_csr(off, a, b, mode=0, with_add=True)
    assert (s == O.add_csr(off, a, b, threads=16)).all()
    assert (p == O.mul_csr(off, a, b, threads=16)).all()
