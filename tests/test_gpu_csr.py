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


def _chain_batch(span, seed):
    """Ragged operands whose instance boundaries and carry chains sit where
    the fixed-chunk add decides things: chains that cross several chunk
    boundaries (all-ones + 1) and all-ones sums with no carry to run (the
    look-back below a chunk walks down to the instance start), instances
    that start exactly on a chunk boundary or end one digit before it, runs
    of empty instances, and hundreds of one-digit instances in one tile."""
    rng = np.random.default_rng(seed)
    sizes, kinds = [], []

    def add(sz, kind):
        sizes.append(sz)
        kinds.append(kind)

    add(3 * span + 5, "chain")
    add(span - 5 - 1, "rand")          # the next instance ends one digit before a chunk boundary
    add(1, "ones")
    add(2 * span + 17, "ones")         # all-ones sums, no carry anywhere
    add(span - 17, "chain")            # starts mid-chunk, ends on the boundary
    for _ in range(100):
        add(0, "rand")
    add(4 * span, "chain")             # starts on a chunk boundary
    for _ in range(300):
        add(1, "rand")
    for _ in range(70):
        add(int(rng.integers(0, 4)), "chain")
    add(5000, "ones_gen_top")          # all-ones sums with a carry generated only at the top digit
    for _ in range(200):
        add(int(rng.integers(1, 3 * span)), "rand" if rng.random() < 0.7 else "chain")
    off = np.zeros(len(sizes) + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    total = int(off[-1])
    a = rng.integers(0, 1 << 32, total, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 1 << 32, total, dtype=np.uint64).astype(np.uint32)
    for i, k in enumerate(kinds):
        s, e = off[i], off[i + 1]
        if e == s or k == "rand":
            continue
        a[s:e] = 0xFFFFFFFF
        b[s:e] = 0
        if k == "chain":
            b[s] = 1
        elif k == "ones_gen_top":
            b[e - 1] = 1
    return off, a, b


@pytest.mark.parametrize("span,tiles", [(256, 1), (768, 4), (1024, 2), (4096, 1)])
def test_add_csr_chunk_boundaries_and_chains(ctx, span, tiles):
    off, a, b = _chain_batch(span, seed=span + tiles)
    keep = ctx.tune_get_csr()
    ctx.tune_set_csr(span, tiles, 1, keep["csr_mul_lanes"], keep["csr_mul_rows"])
    try:
        r = ctx.add_csr(off, a, b)
    finally:
        ctx.tune_set_csr(*keep.values())
    assert (r == O.add_csr(off, a, b)).all()


@pytest.mark.parametrize("lead", [37, 1024, 5000])
def test_add_csr_leading_gap(ctx, lead):
    """offs[0] > 0: the digits before the first instance belong to none;
    they are not written, and the instances after them are exact."""
    o, a0, b0 = W.ragged(400, 1, 700, seed=lead)
    off = o + lead
    rng = np.random.default_rng(lead)
    pad = rng.integers(0, 1 << 32, lead, dtype=np.uint64).astype(np.uint32)
    a = np.concatenate([pad, a0])
    b = np.concatenate([np.full(lead, 0xFFFFFFFF, np.uint32), b0])
    total = int(off[-1])
    doff, da, db = ctx.upload(off), ctx.upload(a), ctx.upload(b)
    mark = np.full(total, 0x5A5A5A5A, np.uint32)
    dr = ctx.upload(mark)
    ctx.add_csr_dev(len(off) - 1, total, doff, da, db, dr)
    r = dr.download(np.uint32, total)
    for x in (doff, da, db, dr):
        ctx._release(x)
    want = O.add_csr(o, a0, b0)
    assert (r[:lead] == 0x5A5A5A5A).all()
    assert (r[lead:] == want).all()
