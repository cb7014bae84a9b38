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
