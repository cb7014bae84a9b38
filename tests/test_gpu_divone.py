"""GPU parity of the one-launch division (bi_divone.hip: one workgroup per
instance, the whole Newton pipeline in LDS, each instance on its own
precision schedule), taken by bi_divmod_u32_dev for batches of up to
bi_tune_get_div_one() instances, against GMP (mpz_tdiv_qr) or Python divmod,
and against the batched pipeline on the same inputs."""
import contextlib

import numpy as np
import pytest

import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def div_one(ctx, n_max):
    old = ctx.tune_get_div_one()
    ctx.tune_set_div_one(n_max)
    try:
        yield
    finally:
        ctx.tune_set_div_one(old)


def expect(u, v):
    if O.gmp_available():
        return O.gmp_div(u, v, 8)
    return O.py_div(u, v)


def edge_rows(u, v, rng):
    """Rows at the algorithm's edges, planted in place (rows past n skipped)."""
    n, mu = u.shape
    mv = v.shape[1]
    rows = []

    def put(i, uu=None, vv=None):
        if i < n:
            if uu is not None:
                u[i] = uu
            if vv is not None:
                v[i] = vv
            rows.append(i)

    put(0, vv=O.from_int(1, mv))                                  # v = 1
    put(1, vv=np.full(mv, 0xFFFFFFFF, np.uint32))                 # v all ones
    put(2, vv=O.from_int(1 << (32 * (mv - 1)), mv))               # v = B^(mv-1)
    put(3, uu=np.full(mu, 0xFFFFFFFF, np.uint32))                 # u all ones
    put(4, uu=np.zeros(mu, np.uint32))                            # u = 0
    vi = O.to_int(v[5]) if n > 5 else 1
    if vi and n > 5:
        k = (1 << (32 * mu)) // (vi * 3) if vi < (1 << (32 * mu)) else 0
        put(5, uu=O.from_int(k * vi, mu))                         # exact multiple
    if n > 6:
        vi = O.to_int(v[6])
        k = ((1 << (32 * mu)) - vi) // (vi * 2)
        put(6, uu=O.from_int(k * vi + vi - 1, mu))                # remainder v - 1
    if n > 7 and mv > 1:
        vv = np.zeros(mv, np.uint32)
        vv[0] = 7                                                 # one significant digit in a long row
        put(7, vv=vv)
    if n > 8:
        vv = v[8].copy()
        vv[mv // 2:] = 0                                          # half-length divisor
        if not vv.any():
            vv[0] = 3
        put(8, vv=vv)
    if n > 9:
        uu = u[9].copy()
        uu[max(1, mv // 2):] = 0                                  # numerator shorter than the divisor
        put(9, uu=uu)
    if n > 10:
        put(10, uu=u[10].copy(), vv=O.from_int(0x80000000 << (32 * (mv - 1)), mv))   # 2^31 B^(mv-1)
    return rows


@pytest.mark.parametrize("mu,mv,n", [(1, 1, 12), (2, 1, 12), (3, 3, 12), (5, 2, 16), (8, 3, 16), (40, 17, 16),
                                     (64, 64, 12), (100, 3, 12), (300, 1, 12), (511, 256, 12), (1000, 999, 12),
                                     (2048, 1024, 12), (4096, 2048, 12), (4096, 16, 12), (4096, 4096, 12),
                                     (5400, 2700, 12), (6700, 3350, 8), (6700, 2, 8)])
def test_divone_shapes_and_edges(ctx, mu, mv, n):
    rng = np.random.default_rng(mu * 131 + mv)
    u, v = W.div_pairs(n, mu, mv, seed=mu * 7 + mv)
    edge_rows(u, v, rng)
    with div_one(ctx, 256):
        q, r = ctx.divmod(u, v)
    gq, gr = expect(u, v)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("mu,mv", [(12, 10), (40, 35), (64, 40), (200, 64), (1000, 600), (3000, 2990)])
def test_divone_start_edges(ctx, mu, mv):
    """The Newton steps start from floor(B^(2 p0) / (T + 1)), T = the divisor's
    top p0 <= 31 digits, by a one-wave long division: divisors whose top k
    digits are all ones (T + 1 = B^p0 when k >= p0, a carry chain through T
    otherwise), or end in 0xFFFFFFFE, or are 1 followed by zeros (the smallest
    T), for every k up to 40, with quotients short enough that the start is
    the whole reciprocal (p0 = P + 2) and long enough that it is not."""
    rng = np.random.default_rng(mu * 7 + mv)
    rows_u, rows_v = [], []
    for k in range(1, min(mv, 40) + 1):
        for kind in range(3):
            v = rng.integers(0, 1 << 32, mv, dtype=np.uint64).astype(np.uint32)
            if kind == 0:
                v[mv - k:] = 0xFFFFFFFF
            elif kind == 1:
                v[mv - k:] = 0xFFFFFFFF
                v[mv - k] = 0xFFFFFFFE
            else:
                v[mv - k:] = 0
                v[mv - 1] = 1
            lu = int(rng.integers(mv, mu + 1))   # quotient lengths 1 .. mu - mv + 1 digits
            u = np.zeros(mu, np.uint32)
            u[:lu] = rng.integers(0, 1 << 32, lu, dtype=np.uint64).astype(np.uint32)
            if kind == 2 and k % 2:
                u[:lu] = 0xFFFFFFFF
            rows_u.append(u)
            rows_v.append(v)
    u, v = np.stack(rows_u), np.stack(rows_v)
    with div_one(ctx, 4096):
        q, r = ctx.divmod(u, v)
    gq, gr = expect(u, v)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("n,mu,mv", [(16, 128, 67), (16, 128, 66), (16, 128, 65), (300, 700, 300),
                                     (16, 1024, 259), (16, 1024, 258), (16, 1024, 257), (16, 1500, 700)])
def test_divone_two_half_quotient_edges(ctx, n, mu, mv):
    """Quotients of P - 1 digits, P = prec(u) - prec(v) + 2, come in two
    halves from P = 64 on (256-lane workgroups) or P = 768 on (1024 lanes):
    full-length rows at P = threshold - 1, threshold, threshold + 1, the
    edge rows (a numerator of all ones, an exact multiple, remainder v - 1,
    a divisor of all ones, v = 1: the longest quotient), and numerators
    whose top or bottom half is all ones or zero (the second half's
    numerator r1 B^k + (u mod B^k) then has a zero or all-ones tail)."""
    u, v = W.div_pairs(n, mu, mv, seed=n * 7 + mu + mv)
    u[:, mu - 1] |= 1
    v[:, mv - 1] |= 1
    edge_rows(u, v, np.random.default_rng(mu))
    h = (mu - mv) // 2
    u[11, :h] = 0xFFFFFFFF
    u[12, :h] = 0
    u[13, h:] = 0xFFFFFFFF
    u[14, :mv] = v[14]
    with div_one(ctx, 4096):
        q, r = ctx.divmod(u, v)
    gq, gr = expect(u, v)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("mu,mv", [(4096, 2048), (512, 256), (64, 32)])
def test_divone_ragged_lengths_in_one_batch(ctx, mu, mv):
    """Every instance a different significant length (log-uniform, as a
    serving round holds): each runs its own precision schedule."""
    n = 64
    rng = np.random.default_rng(mu + 3 * mv)
    u = np.zeros((n, mu), np.uint32)
    v = np.zeros((n, mv), np.uint32)
    for i in range(n):
        lu = int(np.exp(rng.uniform(0, np.log(mu)))) or 1
        lv = min(mv, int(np.exp(rng.uniform(0, np.log(mv)))) or 1)
        u[i, :lu] = rng.integers(0, 1 << 32, lu, dtype=np.uint64).astype(np.uint32)
        v[i, :lv] = rng.integers(0, 1 << 32, lv, dtype=np.uint64).astype(np.uint32)
        v[i, lv - 1] |= 1
    with div_one(ctx, 256):
        q, r = ctx.divmod(u, v)
    gq, gr = expect(u, v)
    assert (q == gq).all() and (r == gr).all()


def test_divone_zero_divisor(ctx):
    u, v = W.div_pairs(8, 300, 100, seed=9)
    v[3] = 0
    with div_one(ctx, 256):
        q, r, rc = ctx.divmod(u, v, allow_zero=True)
    assert rc == -5
    assert not q[3].any() and not r[3].any()
    keep = [i for i in range(8) if i != 3]
    gq, gr = expect(u[keep], v[keep])
    assert (q[keep] == gq).all() and (r[keep] == gr).all()


@pytest.mark.parametrize("n", [1, 255, 256, 257])
def test_divone_switch_and_engines_agree(ctx, n):
    """Up to bi_tune_get_div_one() instances the one-launch path, above it the
    batched pipeline: the same inputs give the same exact results either way
    (and the batched pipeline is forced for the same batch as well)."""
    u, v = W.div_pairs(n, 2048, 1024, seed=n + 100)
    if n > 2:
        v[1] = 0xFFFFFFFF
        u[2] = 0
    gq, gr = expect(u, v)
    with div_one(ctx, 256):
        q, r = ctx.divmod(u, v)
    assert (q == gq).all() and (r == gr).all()
    with div_one(ctx, 0):
        q0, r0 = ctx.divmod(u, v)
    assert (q0 == gq).all() and (r0 == gr).all()


def test_divone_repeat_calls_and_reshape(ctx):
    """Two shapes in turn, twice: nothing carries over between launches."""
    a = W.div_pairs(16, 1024, 512, seed=1)
    b = W.div_pairs(5, 77, 9, seed=2)
    ea, eb = expect(*a), expect(*b)
    with div_one(ctx, 256):
        for _ in range(2):
            for (u, v), (gq, gr) in ((a, ea), (b, eb)):
                q, r = ctx.divmod(u, v)
                assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("mu,mv", [(6800, 100), (6800, 3400)])
def test_divone_past_lds_limit_takes_batched_pipeline(ctx, mu, mv):
    """Numerators past what one workgroup's LDS holds (about 6760 digits)
    go to the batched pipeline even in small batches: still exact."""
    u, v = W.div_pairs(4, mu, mv, seed=mu + mv)
    u[1] = 0xFFFFFFFF
    with div_one(ctx, 256):
        q, r = ctx.divmod(u, v)
    gq, gr = expect(u, v)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("n,mu,mv", [(1024, 2048, 1024), (600, 300, 150), (3000, 64, 32), (300, 4096, 2048)])
def test_divone_larger_batches(ctx, n, mu, mv):
    """Past 256 instances: 256-lane workgroups for operands up to 2048 digits
    (four a CU), 1024-lane ones beyond; the default switch (4096, at most 2048
    past 2048 digits) keeps these on the one-launch path.  Against GMP."""
    u, v = W.div_pairs(n, mu, mv, seed=n + mu)
    edge_rows(u, v, np.random.default_rng(n))
    q, r = ctx.divmod(u, v)
    gq, gr = expect(u, v)
    assert (q == gq).all() and (r == gr).all()
