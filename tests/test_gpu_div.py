"""GPU parity of the batched Newton division (bi_divmod_*) against the
oracle's restatement of prototypes/div.c, GMP and Python divmod."""
import contextlib
import json
import os

import numpy as np
import pytest

import bigint_workloads as W
import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


@contextlib.contextmanager
def div_one(ctx, n_max):
    """Batches of at most n_max instances take the one-launch division
    (bi_divone.hip); 0 keeps every batch on the batched pipeline."""
    old = ctx.tune_get_div_one()
    ctx.tune_set_div_one(n_max)
    try:
        yield
    finally:
        ctx.tune_set_div_one(old)


def test_kat_fixed(ctx):
    with open(os.path.join(GOLD, "kat.json")) as f:
        k = json.load(f)["div_fixed"]
    q, r = ctx.divmod(np.array([k["u"]], np.uint32), np.array([k["v"]], np.uint32))
    assert q[0].tolist() == k["q"] and r[0].tolist() == k["r"]


def test_kat_thesis_shinv_example(ctx):
    with open(os.path.join(GOLD, "kat.json")) as f:
        k = json.load(f)["shinv_thesis"]
    q, r = ctx.divmod(np.array([k["u"]], np.uint32), np.array([k["v"]], np.uint32))
    assert q[0].tolist() == k["q"]


@pytest.mark.parametrize("mu,mv", [(4, 2), (8, 3), (16, 16), (64, 33), (256, 128)])
def test_golden_vectors(ctx, mu, mv):
    v = np.load(os.path.join(GOLD, "vectors.npz"))
    q, r = ctx.divmod(v[f"du_{mu}_{mv}"], v[f"dv_{mu}_{mv}"])
    assert (q == v[f"dq_{mu}_{mv}"]).all() and (r == v[f"dr_{mu}_{mv}"]).all()


@pytest.mark.parametrize("mu,mv,n", [(1, 1, 50), (2, 1, 50), (3, 2, 60), (8, 8, 100), (17, 5, 100), (64, 64, 80),
                                     (100, 37, 64), (512, 256, 40), (1024, 1000, 12), (2048, 700, 8)])
@pytest.mark.parametrize("one", [0, 256])
def test_divmod_mixed_lengths(ctx, mu, mv, n, one):
    """Divisors of every significant length in one batch (so the batch-wide
    precision is set by its longest quotient), all-ones, powers of the base,
    divisor > dividend, equal operands; on the batched pipeline (one = 0) and
    the one-launch division (one = 256)."""
    rng = np.random.default_rng(mu * 7 + mv)
    u = rnd(rng, (n, mu))
    v = np.zeros((n, mv), np.uint32)
    for i in range(n):
        nz = int(rng.integers(1, mv + 1))
        v[i, :nz] = rnd(rng, nz)
        if v[i, nz - 1] == 0:
            v[i, nz - 1] = 1
    v[0] = 0
    v[0, 0] = 1
    if n > 4:
        v[1] = 0xFFFFFFFF
        v[2] = 0
        v[2, mv - 1] = 1                     # B^(mv-1)
        u[3, :] = 0
        u[3, : min(mu, mv)] = v[3, : min(mu, mv)]   # u == v (when they fit)
        u[4] = 0                             # zero dividend
    with div_one(ctx, one):
        q, r = ctx.divmod(u, v)
    q2, r2 = O.py_div(u, v)
    assert (q == q2).all() and (r == r2).all()


@pytest.mark.parametrize("mu,mv,n", [(8, 8, 200), (64, 32, 200), (300, 150, 30)])
def test_divmod_vs_c_prototype_restatement(ctx, mu, mv, n):
    u, v = W.div_pairs(n, mu, mv, seed=mu + mv)
    q, r = ctx.divmod(u, v)
    q2, r2 = O.div(u, v)
    assert (q == q2).all() and (r == r2).all()


def test_zero_divisor_reported(ctx):
    u, v = W.div_pairs(6, 16, 8, seed=3)
    v[2] = 0
    q, r, rc = ctx.divmod(u, v, allow_zero=True)
    assert rc == -5
    assert (q[2] == 0).all() and (r[2] == 0).all()
    keep = [0, 1, 3, 4, 5]
    q2, r2 = O.py_div(u[keep], v[keep])
    assert (q[keep] == q2).all() and (r[keep] == r2).all()


def test_graph_replay_and_reshape(ctx):
    """Same shape twice (cached graph), then a different shape (the batched
    pipeline, forced: these batches would take the one-launch path)."""
    u, v = W.div_pairs(64, 128, 64, seed=5)
    q2, r2 = O.py_div(u, v)
    with div_one(ctx, 0):
        for _ in range(3):
            q, r = ctx.divmod(u, v)
            assert (q == q2).all() and (r == r2).all()
        u, v = W.div_pairs(16, 40, 9, seed=6)
        q, r = ctx.divmod(u, v)
    q2, r2 = O.py_div(u, v)
    assert (q == q2).all() and (r == r2).all()


def test_division_config_sample(ctx):
    """The measured shape (numerator 8192 u32 digits, divisor 4096 with a
    nonzero top digit) on a sample of 64 instances, checked against GMP."""
    u, v = W.div_pairs(64, 8192, 4096, seed=77)
    q, r = ctx.divmod(u, v)
    if O.Gmp.available():
        gq, gr = O.Gmp().batch("div", u, v)
    else:
        gq, gr = O.py_div(u, v)
    assert (q == gq).all() and (r == gr).all()
    q8, r8 = O.div(u[:2], v[:2])
    assert (q[:2] == q8).all() and (r[:2] == r8).all()


@pytest.mark.parametrize("n,mu,mv", [(1, 4096, 2048), (3, 4096, 2048), (16, 4096, 2048), (16, 2048, 1024),
                                     (5, 4096, 4000)])
def test_division_small_batch_row_split(ctx, n, mu, mv):
    """Small batches of long operands (the serving shapes): fewer instances
    than CUs, so each long product's rows are split over several workgroups
    whose column sums are added before the carry pass (the batched pipeline,
    forced).  Checked against GMP."""
    u, v = W.div_pairs(n, mu, mv, seed=n * 31 + mu)
    u[0, : mu // 2] = 0xFFFFFFFF
    with div_one(ctx, 0):
        q, r = ctx.divmod(u, v)
    if O.Gmp.available():
        gq, gr = O.Gmp().batch("div", u, v)
    else:
        gq, gr = O.py_div(u, v)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("n", [255, 256, 257])
def test_division_wide_pass_boundary(ctx, n):
    """Around the count (256) where the carry passes and windowed adds switch
    between one workgroup per instance and one wave per instance, and the long
    products stop splitting their rows (the batched pipeline, forced): same
    shapes on both sides, vs GMP."""
    u, v = W.div_pairs(n, 2048, 1024, seed=n)
    v[1, :] = 0xFFFFFFFF
    u[2, :] = 0
    with div_one(ctx, 0):
        q, r = ctx.divmod(u, v)
    if O.Gmp.available():
        gq, gr = O.Gmp().batch("div", u, v)
    else:
        gq, gr = O.py_div(u, v)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.parametrize("n", [511, 512])
def test_division_wrap_remainder(ctx, n):
    """From 512 instances with a divisor of 2048+ digits the remainder comes
    from the cyclic transform of Q V mod (B^6400 - 1) folded back by
    k_wrap_rem; 511 takes the low triangle.  Exact quotients and remainders,
    with remainders 0 and v - 1 planted (u = q v and u = q v + v - 1)."""
    u, v = W.div_pairs(n, 4096, 2048, seed=n + 7)
    for i, extra in ((0, 0), (1, -1)):
        qi = O.to_int(u[i]) // O.to_int(v[i])
        val = qi * O.to_int(v[i]) + (O.to_int(v[i]) - 1 if extra else 0)
        u[i] = [(val >> (32 * k)) & 0xFFFFFFFF for k in range(4096)]
    with div_one(ctx, 0):   # the batched pipeline's remainder paths
        q, r = ctx.divmod(u, v)
    if O.Gmp.available():
        gq, gr = O.Gmp().batch("div", u, v)
    else:
        gq, gr = O.py_div(u, v)
    assert (q == gq).all() and (r == gr).all()
    assert not r[0].any()


@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
def test_division_full_batch_identity(ctx):
    """The full measured batch (16384 x 8192 / 4096 digits): for EVERY
    instance u = q v + r and r < v, checked with GMP (a size-independent
    property; exact quotients of a sample are compared in the test above)."""
    u, v = W.div_pairs(16384, 8192, 4096, seed=2024)
    q, r = ctx.divmod(u, v)
    assert O.Gmp().divmod_violations(u, v, q, r) == []


@pytest.mark.skipif(not O.gmp_available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("n,mu,mv", [(512, 8192, 4096), (600, 6000, 2500)])
def test_division_kept_divisor_transform(ctx, n, mu, mv):
    """From 512 instances V's 2^13-point transform is computed by the first
    Newton step that takes its V X by transform, stored, and loaded by the
    later ones and by the cyclic remainder instead of being recomputed:
    every quotient and remainder exact against GMP (mpz_tdiv_qr)."""
    u, v = W.div_pairs(n, mu, mv, seed=n + mu)
    v[3] = 0xFFFFFFFF
    u[4, :] = 0xFFFFFFFF
    with div_one(ctx, 0):   # the batched pipeline's transform products
        q, r = ctx.divmod(u, v)
    gq, gr = O.gmp_div(u, v, 8)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.skipif(not O.gmp_available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("n,mu,mv", [(512, 8192, 4096), (64, 8192, 4096), (200, 2000, 1000), (300, 700, 600)])
def test_division_newton_window_edges(ctx, n, mu, mv):
    """The Newton steps multiply X by V's digits from c - p - 1 up only
    (E / B^c at most one unit high, X' - 1 after the step): divisors whose
    dropped low digits are all ones (the dropped product largest), a top
    digit 1 over all ones, B^(mv-1) + 1 and B^(mv-1) (E smallest), all ones,
    and numerators of all ones against them; exact against GMP on the
    batched pipeline."""
    u, v = W.div_pairs(n, mu, mv, seed=3 * n + mu)
    v[0, : mv // 2] = 0xFFFFFFFF
    v[1, :] = 0xFFFFFFFF
    v[1, mv - 1] = 1
    v[2, :] = 0
    v[2, mv - 1] = 1
    v[2, 0] = 1
    v[3, :] = 0
    v[3, mv - 1] = 1
    v[4, :] = 0xFFFFFFFF
    v[5, : mv - 3] = 0xFFFFFFFF
    for row in range(6, 12):
        u[row, :] = 0xFFFFFFFF
        v[row] = v[row - 6]
    with div_one(ctx, 0):
        q, r = ctx.divmod(u, v)
    gq, gr = O.gmp_div(u, v, 8)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.skipif(not O.gmp_available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("n,mu,mv", [(512, 8192, 4096), (600, 6000, 2500), (300, 900, 700), (200, 600, 100)])
def test_division_two_half_quotient_edges(ctx, n, mu, mv):
    """Long quotients come in two halves, q = q1 B^k + q0 (k = (P - 1) / 2,
    P = the batch's quotient digits + 1): numerators built as q v + r with
    q all ones, q = B^k - 1 (q1 = 0), q = B^k (q0 = 0), q1 all ones over q0
    = 0, q0 all ones under q1 = 1, random q with r = v - 1 and r = 0, and
    all-ones numerators; exact against GMP on the batched pipeline."""
    u, v = W.div_pairs(n, mu, mv, seed=11 * n + mv)
    P = mu - mv + 2          # the batch's quotient digits + 1 (top digits nonzero)
    k = (P - 1) // 2
    qdig = P - 1
    rng = np.random.default_rng(n + mu)
    qs = [(1 << (32 * qdig)) - 1, (1 << (32 * k)) - 1, 1 << (32 * k),
          ((1 << (32 * (qdig - k))) - 1) << (32 * k), (1 << (32 * k)) + (1 << (32 * k)) - 1]
    for row, q in enumerate(qs):
        vi = O.to_int(v[row])
        if q * vi < (1 << (32 * mu)):
            u[row] = O.from_int(q * vi, mu)
    for row in (5, 6):
        vi = O.to_int(v[row])
        q = int(rng.integers(1, 1 << 62)) << (32 * max(0, k - 1))
        r = vi - 1 if row == 5 else 0
        if q * vi + r < (1 << (32 * mu)):
            u[row] = O.from_int(q * vi + r, mu)
    u[7] = 0xFFFFFFFF
    with div_one(ctx, 0):
        q, r = ctx.divmod(u, v)
    gq, gr = O.gmp_div(u, v, 8)
    assert (q == gq).all() and (r == gr).all()


@pytest.mark.skipif(not O.gmp_available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("n,mu,mv", [(512, 8192, 4096), (600, 6000, 2500)])
def test_division_wraparound_quotient_edges(ctx, n, mu, mv):
    """From 512 instances of these shapes the quotient's digits come from a
    cyclic transform product plus a low triangle (the wrap-around identity)
    instead of the top triangle.  Rows at its edges, in one batch: quotient
    zero (numerator below the divisor), one-digit quotients, numerator all
    ones, exact multiples (remainder zero), a power-of-two divisor, a divisor
    of all ones; exact against GMP (mpz_tdiv_qr)."""
    u, v = W.div_pairs(n, mu, mv, seed=7 * n + mv)
    u[0, mv - 1:] = 0                      # u < v: q = 0
    u[1, mv + 1:] = 0                      # short quotient
    u[2, :] = 0xFFFFFFFF
    v[3, :] = 0xFFFFFFFF
    v[4, :] = 0
    v[4, mv - 1] = 0x80000000              # B^(mv-1) * 2^31
    u[5, :] = 0                            # u = 0
    # exact multiples: u = v * k with k of mu - mv digits (fits in mu digits)
    k = O.to_int(W.div_pairs(1, mu - mv, 1, seed=3)[0][0])
    for row in (6, 7):
        u[row] = O.from_int(O.to_int(v[row]) * k, mu)
    with div_one(ctx, 0):   # the batched pipeline's wrap-around quotient
        q, r = ctx.divmod(u, v)
    gq, gr = O.gmp_div(u, v, 8)
    assert (q == gq).all() and (r == gr).all()
    assert not r[6].any() and not r[7].any() and not q[0].any() and not q[5].any()
