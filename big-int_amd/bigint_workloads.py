"""Seeded synthetic inputs for the batched big-integer workloads.

Shapes follow the workloads the engine is measured on:

* plain pairs -- n instances of m u32 digits, uniform random digits, with an
  optional fraction of "carry chain" pairs (a = all 0xFFFFFFFF, b = 1, whose sum
  carries through every digit) and an optional zero top half (so the full
  product fits in m digits);
* ragged CSR batches -- N instances with digit counts drawn log-uniformly in
  [lo, hi], stored back to back with an int64 offset table (N+1 entries);
* division pairs -- numerators of mu digits and divisors of mv digits with a
  nonzero top digit.

Digits are uniform over the full u32 range (odd digits and 0xFFFFFFFF
included), unlike the reference's generator whose digits are always even
(cuda/helpers.h:96-111).
"""
import numpy as np


def _digits(rng, shape):
    return rng.integers(0, 1 << 32, size=shape, dtype=np.uint64).astype(np.uint32)


def plain_pairs(n, m, seed=1, chain_frac=0.0, top_zero=False):
    rng = np.random.default_rng(seed)
    a = _digits(rng, (n, m))
    b = _digits(rng, (n, m))
    if top_zero:
        a[:, m // 2:] = 0
        b[:, m // 2:] = 0
    if chain_frac > 0:
        k = int(round(chain_frac * n))
        idx = rng.choice(n, size=k, replace=False)
        a[idx] = 0xFFFFFFFF
        b[idx] = 0
        b[idx, 0] = 1
    return a, b


def ragged_sizes(N, lo, hi, seed=1):
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=N))).astype(np.int64)
    return np.clip(sizes, lo, hi)


def ragged(N, lo=8, hi=2048, seed=1, chain_frac=0.10, zero_frac=0.05, sizes=None):
    """Returns (off[int64 N+1], a[flat u32], b[flat u32])."""
    rng = np.random.default_rng(seed + 7919)
    if sizes is None:
        sizes = ragged_sizes(N, lo, hi, seed)
    off = np.zeros(N + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    total = int(off[-1])
    a = _digits(rng, total)
    b = _digits(rng, total)
    pick = rng.random(N)
    chain = np.nonzero(pick < chain_frac)[0]
    zero = np.nonzero((pick >= chain_frac) & (pick < chain_frac + zero_frac))[0]
    for i in chain:
        s, e = off[i], off[i + 1]
        if e == s:
            continue
        a[s:e] = 0xFFFFFFFF
        b[s:e] = 0
        b[s] = 1
    for i in zero:
        s, e = off[i], off[i + 1]
        (a if (i & 1) else b)[s:e] = 0
    return off, a, b


def div_pairs(n, mu, mv, seed=1):
    """u: [n][mu] full random; v: [n][mv] with a nonzero top digit."""
    rng = np.random.default_rng(seed)
    u = _digits(rng, (n, mu))
    v = _digits(rng, (n, mv))
    top = v[:, -1]
    while (top == 0).any():
        z = top == 0
        top[z] = _digits(rng, int(z.sum()))
    v[:, -1] = top
    return u, v
