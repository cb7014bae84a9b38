"""GPU parity of the GMP-format export / import (bi_export_sizes_u32_dev,
bi_export_u32_dev, bi_import_u32_dev) against the packed-format closed form
(oracle py_export / py_import, itself pinned to mpz_export / mpz_import) and
GMP directly, bit-exact: zero integers, leading zeros, odd digit counts in
8-byte words, many instances (multi-block offset scan), truncating import."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def batch(n, m, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 1 << 32, (n, m), dtype=np.uint64).astype(np.uint32)
    lens = rng.integers(0, m + 1, n)
    for i in range(n):
        x[i, lens[i]:] = 0
    if n > 3:
        x[0] = 0
        x[1] = 0xFFFFFFFF
    return x


@pytest.mark.parametrize("wsize", [4, 8])
@pytest.mark.parametrize("n,m", [(1, 1), (37, 3), (100, 8), (257, 64), (40, 511), (9, 4096), (5000, 5)])
def test_export_import(ctx, wsize, n, m):
    x = batch(n, m, seed=n * 13 + m + wsize)
    offs, words = ctx.export(x, wsize)
    want_offs, want_words = O.py_export(x, wsize)
    assert (offs == want_offs).all()
    assert (words == want_words).all()
    assert (ctx.import_(offs, words, m, wsize) == x).all()
    if m > 2:
        assert (ctx.import_(offs, words, m // 2, wsize) == x[:, : m // 2]).all()


def test_export_all_zero(ctx):
    x = np.zeros((50, 7), np.uint32)
    offs, words = ctx.export(x, 8)
    assert (offs == 0).all() and words.size == 0
    assert (ctx.import_(offs, words, 7, 8) == 0).all()


def test_export_many_instances(ctx):
    x = batch(300000, 4, seed=3)
    offs, words = ctx.export(x, 4)
    want_offs, want_words = O.py_export(x, 4)
    assert (offs == want_offs).all() and (words == want_words).all()


@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
def test_export_against_gmp(ctx):
    x = batch(64, 33, seed=5)
    g = O.Gmp()
    for wsize in (4, 8):
        offs, words = ctx.export(x, wsize)
        q = wsize // 4
        for i in range(64):
            cnt, w = g.export_words(x[i], wsize)
            assert cnt == int(offs[i + 1] - offs[i])
            assert (w == words[int(offs[i]) * q:int(offs[i + 1]) * q]).all()
