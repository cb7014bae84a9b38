"""Pins the CPU oracle (oracle/bigint_oracle.c) before anything is checked
against it: reference known-answer vectors, committed golden vectors computed
with Python int, GMP as a second oracle, exhaustive checks of the carry
operator, and mathematical invariants.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def kat():
    with open(os.path.join(GOLD, "kat.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def vec():
    return np.load(os.path.join(GOLD, "vectors.npz"))


def rnd(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


# ---------------------------------------------------------------- carry operator
def test_carry_operator_identity(kat):
    e = kat["carry_identity"]["identity"]
    for x in kat["carry_identity"]["values"]:
        assert O.carry_op(e, x) == x
        assert O.carry_op(x, e) == x


def test_carry_operator_associative_and_meaning():
    vals = [0, 1, 2]   # a digit sum either carries, is all-ones, or neither
    for a in vals:
        for b in vals:
            for c in vals:
                assert O.carry_op(O.carry_op(a, b), c) == O.carry_op(a, O.carry_op(b, c))
    # meaning: (lower, upper) composed carries out iff upper generates, or
    # upper propagates and lower carries
    for lo in vals:
        for up in vals:
            want = (up & 1) | ((up >> 1) & lo & 1)
            assert O.carry_op(lo, up) & 1 == want


# ---------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7, 8, 16, 31, 64, 67, 512])
def test_golden_add_sub_mul(vec, m):
    a, b = vec[f"a_{m}"], vec[f"b_{m}"]
    assert (O.add(a, b) == vec[f"add_{m}"]).all()
    assert (O.mul(a, b) == vec[f"mul_{m}"]).all()
    r, s = O.sub(a, b)
    assert (r == vec[f"sub_{m}"]).all() and (s == vec[f"sgn_{m}"]).all()


@pytest.mark.parametrize("mu,mv", [(4, 2), (8, 3), (16, 16), (64, 33), (256, 128)])
def test_golden_div(vec, mu, mv):
    u, v = vec[f"du_{mu}_{mv}"], vec[f"dv_{mu}_{mv}"]
    q, r = O.div(u, v)
    assert (q == vec[f"dq_{mu}_{mv}"]).all()
    assert (r == vec[f"dr_{mu}_{mv}"]).all()


# ---------------------------------------------------------------- reference KATs
def test_kat_div_fixed(kat):
    k = kat["div_fixed"]
    u = np.array([k["u"]], np.uint32)
    v = np.array([k["v"]], np.uint32)
    q, r = O.div(u, v)
    assert q[0].tolist() == k["q"] and r[0].tolist() == k["r"]


def test_kat_div_futhark_base16(kat):
    k = kat["div_futhark_u16"]
    # regroup base-2^16 digits into base-2^32 digits
    def to32(d16):
        x = sum(int(d) << (16 * i) for i, d in enumerate(d16))
        return O.from_int(x, 4)
    u = to32(k["u16"])[None]
    v = to32(k["v16"])[None]
    q, r = O.div(u, v)
    assert O.to_int(q[0]) == sum(int(d) << (16 * i) for i, d in enumerate(k["q16"]))
    assert O.to_int(u[0]) == O.to_int(q[0]) * O.to_int(v[0]) + O.to_int(r[0])


def test_kat_shinv_thesis(kat):
    """thesis/sections/division.tex:110-131: v = B + 1 (k = 1), h = 3.
    div.c shifts u and v by one digit when k == 1 (prototypes/div.c:443-448),
    so the quotient is the pinned output; shinv_3(v) is the closed form
    floor(B^3 / v) and u*shinv/B^h is the raw estimate before the delta fix."""
    k = kat["shinv_thesis"]
    v = O.to_int(np.array(k["v"], np.uint32))
    assert (1 << 96) // v == O.to_int(np.array(k["shinv"], np.uint32))
    u = np.array([k["u"]], np.uint32)
    q, r = O.div(u, np.array([k["v"]], np.uint32))
    assert q[0].tolist() == k["q"]
    raw = (O.to_int(u[0]) * O.to_int(np.array(k["shinv"], np.uint32))) >> 96
    assert O.from_int(raw, 3).tolist() == k["raw_q"]


@pytest.mark.parametrize("h,vdig", [(6, [5, 7, 9]), (10, [1, 2, 3, 4]), (9, [0, 0, 0, 1]), (8, [3, 0, 0, 0, 7]),
                                    (7, [0xFFFFFFFF] * 4), (12, [123, 0xFFFFFFFF, 0x80000000])])
def test_shinv_closed_form(h, vdig):
    v = np.array(vdig, np.uint32)
    k = int(O.lib().or_findk(len(v), v.ctypes.data_as(__import__("ctypes").c_void_p)))
    if k < 2:
        pytest.skip("shinv requires k >= 2 (k <= 1 is handled by div)")
    w = O.shinv(v, h, k, 4 * 16)
    exact = (1 << (32 * h)) // O.to_int(v)
    assert exact - O.to_int(w) in (0, 1)


# ---------------------------------------------------------------- GMP second oracle
@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("m", [1, 3, 17, 64, 255, 512])
def test_against_gmp(m):
    rng = np.random.default_rng(m)
    a, b = rnd(rng, (40, m)), rnd(rng, (40, m))
    a[0] = 0xFFFFFFFF
    b[0] = 0
    b[0, 0] = 1
    b[1] = 0
    g = O.Gmp()
    assert (O.add(a, b) == g.batch("add", a, b)).all()
    assert (O.mul(a, b) == g.batch("mul", a, b)).all()
    assert (O.sub(a, b)[0] == g.batch("sub", a, b)).all()


@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("mu,mv", [(8, 8), (33, 17), (128, 64), (300, 7)])
def test_div_against_gmp(mu, mv):
    rng = np.random.default_rng(mu * 1000 + mv)
    u = rnd(rng, (30, mu))
    v = np.zeros((30, mv), np.uint32)
    for i in range(30):
        nz = int(rng.integers(1, mv + 1))
        v[i, :nz] = rnd(rng, nz)
        v[i, 0] |= 1
    q, r = O.div(u, v)
    gq, gr = O.Gmp().batch("div", u, v)
    assert (q == gq).all() and (r == gr).all()


# ---------------------------------------------------------------- helpers vs closed forms
def test_helpers_closed_forms():
    rng = np.random.default_rng(5)
    m = 23
    a = rnd(rng, (50, m))
    a[3] = 0
    a[4, 5:] = 0
    d = rnd(rng, 50)
    d[d == 0] = 7
    for s in (-30, -23, -5, -1, 0, 1, 4, 22, 23, 40):
        r = O.shift(s, a)
        for i in (0, 4, 17):
            x = O.to_int(a[i])
            want = (x << (32 * s)) % (1 << (32 * m)) if s >= 0 else x >> (32 * -s)
            assert O.to_int(r[i]) == want
    p = O.prec(a)
    for i in range(50):
        x = O.to_int(a[i])
        assert p[i] == max(1, (x.bit_length() + 31) // 32)
    rm = O.muld(a, d)
    q, rem = O.quod(a, d)
    for i in range(50):
        x = O.to_int(a[i])
        assert O.to_int(rm[i]) == (x * int(d[i])) % (1 << (32 * m))
        assert O.to_int(q[i]) == x // int(d[i]) and rem[i] == x % int(d[i])
    b = a.copy()
    b[::3, 0] ^= 1
    c = O.cmp(a, b)
    for i in range(50):
        x, y = O.to_int(a[i]), O.to_int(b[i])
        assert c[i] == (x > y) - (x < y)


def test_csr_matches_uniform():
    import bigint_workloads as W
    off, a, b = W.ragged(300, 1, 80, seed=3)
    r = O.add_csr(off, a, b)
    lo = O.mul_csr(off, a, b, mode=0)
    full, off_out = O.mul_csr(off, a, b, mode=1)
    for i in range(300):
        s, e = off[i], off[i + 1]
        n = int(e - s)
        x, y = O.to_int(a[s:e]), O.to_int(b[s:e])
        assert O.to_int(r[s:e]) == (x + y) % (1 << (32 * n))
        assert O.to_int(lo[s:e]) == (x * y) % (1 << (32 * n))
        assert O.to_int(full[off_out[i]:off_out[i + 1]]) == x * y


def test_planted_mistake_is_detected():
    """A wrong sum (one flipped digit) must not pass the comparison the GPU
    tests use."""
    rng = np.random.default_rng(9)
    a, b = rnd(rng, (8, 64)), rnd(rng, (8, 64))
    r = O.add(a, b)
    bad = r.copy()
    bad[5, 33] ^= 0x10
    assert not (bad == O.py_add(a, b)).all()
    assert (r == O.py_add(a, b)).all()


@pytest.mark.parametrize("m,p", [(1, 10), (5, 3), (64, 10), (67, 6), (512, 6)])
def test_chains_closed_forms(m, p):
    """The bench chains of the reference: r1 = a+b, r = r1 + r (p additions)
    is p(a+b); r1 = ab, r = r1 r (p products) is (ab)^p, both mod B^m.  All-ones
    operands exercise full-length carries."""
    rng = np.random.default_rng(m * 31 + p)
    a, b = rnd(rng, (6, m)), rnd(rng, (6, m))
    a[0] = 0xFFFFFFFF
    b[0] = 0
    b[0, 0] = 1
    a[1] = 0xFFFFFFFF
    b[1] = 0xFFFFFFFF
    assert (O.add_chain(a, b, p) == O.py_add_chain(a, b, p)).all()
    assert (O.mul_chain(a, b, p) == O.py_mul_chain(a, b, p)).all()
    # p = 1 is the plain operation
    assert (O.add_chain(a, b, 1) == O.py_add(a, b)).all()
    assert (O.mul_chain(a, b, 1) == O.py_mul(a, b)).all()


@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("wsize", [4, 8])
def test_export_format_against_gmp(wsize):
    """The packed export closed form (py_export) is exactly what mpz_export
    writes (order -1, native endian, no nails), incl. zero and odd digit
    counts in 8-byte words; py_import inverts it."""
    rng = np.random.default_rng(wsize)
    m = 9
    x = rnd(rng, (12, m))
    x[0] = 0
    x[1, 1:] = 0
    x[2, 3:] = 0
    x[3, 8] = 0
    x[4, :] = 0xFFFFFFFF
    offs, words = O.py_export(x, wsize)
    g = O.Gmp()
    for i in range(len(x)):
        cnt, w = g.export_words(x[i], wsize)
        assert cnt == int(offs[i + 1] - offs[i])
        q = wsize // 4
        assert (w == words[int(offs[i]) * q:int(offs[i + 1]) * q]).all()
        assert (g.import_words(w, cnt, wsize, m) == x[i]).all()
    assert (O.py_import(offs, words, m, wsize) == x).all()
    # truncation on import into fewer digits
    assert (O.py_import(offs, words, 3, wsize) == x[:, :3]).all()


@pytest.mark.skipif(not O.Gmp.available(), reason="libgmp.so.10 not present")
def test_divmod_identity_checker_catches_mistakes():
    """The identity check used at full division size flags a wrong quotient
    digit and a remainder that is not reduced."""
    rng = np.random.default_rng(4)
    u, v = rnd(rng, (6, 12)), rnd(rng, (6, 5))
    v[:, -1] |= 1
    q, r = O.py_div(u, v)
    g = O.Gmp()
    assert g.divmod_violations(u, v, q, r) == []
    q2 = q.copy()
    q2[2, 0] ^= 1
    assert g.divmod_violations(u, v, q2, r) == [2]
    # q - 1 with r + v: the identity holds but the remainder is not below v
    i = 3
    assert O.to_int(q[i]) > 0
    qm = O.from_int(O.to_int(q[i]) - 1, q.shape[1])[None]
    rp = O.from_int(O.to_int(r[i]) + O.to_int(v[i]), 6)[None]
    v6 = O.from_int(O.to_int(v[i]), 6)[None]
    assert g.divmod_violations(u[i:i + 1], v6, qm, rp) == [0]


@pytest.mark.skipif(not O.gmp_available(), reason="libgmp.so.10 not present")
def test_gmp_batches_match_oracle():
    """The all-core GMP leg of the CPU baseline (gmp_batch.c) gives the same
    digits as the oracle and the Python-int closed forms: add, low product,
    division, ragged add and ragged low product."""
    import bigint_workloads as W
    a, b = W.plain_pairs(300, 64, seed=3, chain_frac=0.1)
    assert (O.gmp_add(a, b, 4) == O.add(a, b)).all()
    assert (O.gmp_mul(a, b, 4) == O.mul(a, b)).all()
    assert (O.gmp_mul(a[:5], b[:5], 1) == np.stack([O.from_int(O.to_int(x) * O.to_int(y) % (1 << 2048), 64)
                                                    for x, y in zip(a[:5], b[:5])])).all()
    u, v = W.div_pairs(40, 96, 33, seed=4)
    q, r = O.gmp_div(u, v, 4)
    q2, r2 = O.py_div(u, v)
    assert (q == q2).all() and (r == r2).all()
    f = O.gmp_mul_full(a[:6, :40], b[:6], 2)
    assert (f == O.mul_full(a[:6, :40], b[:6])).all()
    off, x, y = W.ragged(500, 1, 300, seed=5)
    assert (O.gmp_csr(0, off, x, y, 4) == O.add_csr(off, x, y)).all()
    assert (O.gmp_csr(1, off, x, y, 4) == O.mul_csr(off, x, y)).all()
