"""Serving: continuous batching on one GPU.  Every result must be identical
alone and in a mixed stream (bit-exact; integer results have no tolerance),
and equal to the oracle / GMP."""
import time

import numpy as np
import pytest

import bigint_ffi as F
import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu


def mk_requests(n, seed, lo=16, hi=600, div_frac=0.1):
    rng = np.random.default_rng(seed)
    sizes = W.ragged_sizes(n, lo, hi, seed=seed)
    ops, As, Bs = [], [], []
    for i in range(n):
        u = rng.random()
        op = F.OP_ADD if u < 0.5 else (F.OP_MUL if u < 1 - div_frac else F.OP_DIV)
        m = int(sizes[i])
        a = rng.integers(0, 1 << 32, m, dtype=np.uint64).astype(np.uint32)
        if op == F.OP_DIV:
            mb = max(1, m // 2)
            b = rng.integers(0, 1 << 32, mb, dtype=np.uint64).astype(np.uint32)
            if b[-1] == 0:
                b[-1] = 1
        else:
            b = rng.integers(0, 1 << 32, m, dtype=np.uint64).astype(np.uint32)
        if i % 17 == 0 and op == F.OP_ADD:
            a[:] = 0xFFFFFFFF
            b[:] = 0
            b[0] = 1
        ops.append(op)
        As.append(a)
        Bs.append(b)
    return ops, As, Bs


def expected(op, a, b):
    if op == F.OP_ADD:
        return O.add(a[None], b[None])[0]
    if op == F.OP_MUL:
        return O.mul(a[None], b[None])[0]
    q, r = O.py_div(a[None], b[None])
    return q[0], r[0]


def same(x, y):
    if isinstance(x, tuple):
        return all((p == q).all() for p, q in zip(x, y))
    return (x == y).all()


def test_each_request_alone_equals_mix(ctx):
    ops, As, Bs = mk_requests(120, seed=1)
    srv = F.Server(ctx, arena_bytes=64 << 20, slots=2)
    try:
        alone = []
        for op, a, b in zip(ops, As, Bs):
            r = srv.wait(srv.submit(op, a, b))
            alone.append((r[0], r[1]) if op == F.OP_DIV else r[0])
        # all at once: requests arrive while others are on the GPU
        tks = [srv.submit(op, a, b) for op, a, b in zip(ops, As, Bs)]
        mixed = []
        for tk, op in zip(tks, ops):
            r = srv.wait(tk)
            mixed.append((r[0], r[1]) if op == F.OP_DIV else r[0])
        st = srv.stats()
    finally:
        assert srv.close() == 0
    for i, (op, a, b) in enumerate(zip(ops, As, Bs)):
        assert same(alone[i], mixed[i]), i
        assert same(mixed[i], expected(op, a, b)), i
    assert st["batches"] >= 2 and st["max_batch"] > 1   # requests were batched together


def test_trace_open_loop_small_arena(ctx):
    """Poisson arrivals through a deliberately small arena: admission must
    wait for memory to free up, and every result must still be exact."""
    ops, As, Bs = mk_requests(300, seed=2, lo=16, hi=1024)
    rng = np.random.default_rng(3)
    arrival = np.cumsum(rng.exponential(1 / 5000.0, len(ops)))
    srv = F.Server(ctx, arena_bytes=2 << 20, slots=3)
    try:
        outs, lat, status, wall = srv.run_trace(ops, As, Bs, arrival)
        st = srv.stats()
    finally:
        assert srv.close() == 0
    assert (status == 0).all()
    assert st["peak_arena_bytes"] <= 2 << 20
    for i, (op, a, b) in enumerate(zip(ops, As, Bs)):
        assert same(outs[i], expected(op, a, b)), i
    assert (lat > 0).all() and np.percentile(lat, 99) < 30


def test_zero_divisor_request(ctx):
    srv = F.Server(ctx, arena_bytes=8 << 20)
    try:
        a = np.arange(1, 9, dtype=np.uint32)
        tk = srv.submit(F.OP_DIV, a, np.zeros(3, np.uint32))
        q, r, rc, _ = srv.wait(tk)
        assert rc == -5 and (q == 0).all() and (r == 0).all()
        tk = srv.submit(F.OP_ADD, a, a)
        r, rc, _ = srv.wait(tk)
        assert rc == 0 and (r == 2 * a).all()
    finally:
        srv.close()


def test_concurrent_submitters(ctx):
    """Four host threads submit and wait at the same time (the API is
    thread-safe); every result is exact."""
    import threading
    ops, As, Bs = mk_requests(200, seed=4, lo=8, hi=2048)
    srv = F.Server(ctx, arena_bytes=32 << 20, slots=3)
    results = [None] * len(ops)
    errors = []

    def worker(k):
        try:
            mine = list(range(k, len(ops), 4))
            tks = [(i, srv.submit(ops[i], As[i], Bs[i])) for i in mine]
            for i, tk in tks:
                r = srv.wait(tk)
                results[i] = (r[0], r[1]) if ops[i] == F.OP_DIV else r[0]
        except Exception as e:   # reported below, in the test's own thread
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    try:
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
    finally:
        assert srv.close() == 0
    assert not errors, errors
    for i, (op, a, b) in enumerate(zip(ops, As, Bs)):
        assert same(results[i], expected(op, a, b)), i


def test_edge_sizes(ctx):
    """One-digit and 4096-digit operands (the server's largest), divisors of
    one digit, divisors longer than the dividend, all in one stream."""
    rng = np.random.default_rng(8)
    big = lambda m: rng.integers(0, 1 << 32, m, dtype=np.uint64).astype(np.uint32)
    reqs = [(F.OP_ADD, big(1), big(1)), (F.OP_MUL, big(1), big(1)),
            (F.OP_ADD, np.full(4096, 0xFFFFFFFF, np.uint32), np.r_[np.uint32(1), np.zeros(4095, np.uint32)]),
            (F.OP_MUL, big(4096), big(4096)), (F.OP_DIV, big(4096), big(4096)),
            (F.OP_DIV, big(4096), np.array([7], np.uint32)), (F.OP_DIV, big(5), big(300)),
            (F.OP_DIV, big(4096), big(2048)), (F.OP_MUL, big(3), big(3))]
    srv = F.Server(ctx, arena_bytes=64 << 20, slots=2)
    try:
        tks = [srv.submit(op, a, b) for op, a, b in reqs]
        outs = [srv.wait(tk) for tk in tks]
    finally:
        assert srv.close() == 0
    for (op, a, b), o in zip(reqs, outs):
        got = (o[0], o[1]) if op == F.OP_DIV else o[0]
        assert same(got, expected(op, a, b))


def test_rejected_requests(ctx):
    srv = F.Server(ctx, arena_bytes=8 << 20)
    try:
        a = np.arange(1, 9, dtype=np.uint32)
        for op, x, y, code in [(F.OP_ADD, a, a[:4], -1),                      # add of unequal lengths
                               (7, a, a, -1),                                  # unknown operation
                               (F.OP_MUL, np.zeros(4097, np.uint32), np.zeros(4097, np.uint32), -2)]:
            with pytest.raises(F.BigIntError) as ei:
                srv.submit(op, x, y)
            assert ei.value.code == code
        r, rc, _ = srv.wait(srv.submit(F.OP_ADD, a, a))   # still serving
        assert rc == 0 and (r == 2 * a).all()
    finally:
        assert srv.close() == 0
