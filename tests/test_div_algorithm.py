"""The division algorithm of the GPU kernels (bi_div.hip, bi_divone.hip), run
on exact Python integers and checked against divmod: the Newton iteration for
the whole-shifted inverse with the per-instance or batch-wide precision
schedule, the start computed directly at up to 31 digits or at 3 digits, the
divisor window (V's digits from c - p - 1 up, then X' - 1), the top-columns
quotient and the delta fix, and the long quotient in two halves from an
inverse of half the precision.  Also a lane-by-lane emulation of the one-wave
long division that computes the direct start (64 lanes, the kernel's borrow
look-aheads and Knuth's two-digit estimate).  These pin the arithmetic the
kernels implement; the GPU tests pin the kernels against GMP.  CPU only."""
import random

import pytest

B = 1 << 32
M = B - 1


def digits(x, n):
    return [(x >> (32 * i)) & M for i in range(n)]


def sig(x):
    return (x.bit_length() + 31) // 32


def top_columns(a, b, lo, ln):
    """Digits [lo, lo + ln) of a*b from the columns lo - 2 (rounded down to
    4) and up only: the exact window or one less, as the kernels form it."""
    klo = (lo - 2 if lo >= 2 else 0) & ~3
    t = 0
    for k in range(klo, len(a) + len(b) - 1):
        s = sum(a[i] * b[k - i] for i in range(max(0, k - len(b) + 1), min(len(a) - 1, k) + 1))
        t += s << (32 * k)
    return (t >> (32 * lo)) % B**ln


def divide(u, v, start_max=31, window=True, extra_p=0):
    """q, r and the number of delta corrections; extra_p raises the working
    precision above what the instance needs (a batch-wide schedule)."""
    su, sv = sig(u), sig(v)
    NV = max(sv, 3)
    s = NV - sv
    V, U, NU = v << (32 * s), u << (32 * s), su + s
    P = max(su - sv + 1, 1) + 1 + extra_p
    PP = P + 2
    p = PP
    while p > start_max:
        p = (p + 3) // 2
    T = V >> (32 * (NV - p)) if NV >= p else V << (32 * (p - NV))
    X = B ** (2 * p) // (T + 1)
    while p < PP:
        p2 = PP
        while (p2 + 3) // 2 > p:
            p2 = (p2 + 3) // 2
        E = B ** (NV + p) - V * X
        assert 0 <= E < B ** (NV + 2)          # the invariant every step keeps
        c = NV - p - 2 if NV >= p + 2 else 0
        cw = c - p - 1 if (window and c > p + 1) else 0
        Ehi = (-(V >> (32 * cw)) * X * B**cw) % B ** (NV + 2)
        lenE = NV + 2 - c
        d = NV + 2 * p - p2 - c
        assert d >= 5
        Y = top_columns(digits(X, p + 1), digits(Ehi >> (32 * c), lenE), d, p2 + 2)
        X = ((X << (32 * (p2 - p))) + Y - (1 if cw else 0)) % B ** (p2 + 2)
        p = p2
    assert 0 <= B ** (NV + PP) - V * X < B ** (NV + 2)
    LU = max(1, min(NU, P + NV - 2))
    Q = top_columns(digits(U, LU), digits(X, PP + 1), NV + PP, P)
    R = (U - (Q * V) % B ** (NV + 1)) % B ** (NV + 1)
    fixes = 0
    for _ in range(2):
        if R >= V:
            Q, R, fixes = Q + 1, R - V, fixes + 1
    return Q, R >> (32 * s), fixes


def newton_inverse(V, NV, PP):
    """The batched pipeline's inverse: 3-digit start, divisor window."""
    X = B**6 // ((V >> (32 * (NV - 3))) + 1)
    p = 3
    while p < PP:
        p2 = PP
        while (p2 + 3) // 2 > p:
            p2 = (p2 + 3) // 2
        c = NV - p - 2 if NV >= p + 2 else 0
        cw = c - p - 1 if c > p + 1 else 0
        Ehi = (-(V >> (32 * cw)) * X * B**cw) % B ** (NV + 2)
        d = NV + 2 * p - p2 - c
        Y = top_columns(digits(X, p + 1), digits(Ehi >> (32 * c), NV + 2 - c), d, p2 + 2)
        X = ((X << (32 * (p2 - p))) + Y - (1 if cw else 0)) % B ** (p2 + 2)
        p = p2
    return X


def quot_rem(U, NU, V, NV, W, Ph, Pq):
    LU = max(1, min(NU, Pq + NV - 2))
    Q = top_columns(digits(U, LU), digits(W, Ph + 1), NV + Ph, Pq)
    R = (U - (Q * V) % B ** (NV + 1)) % B ** (NV + 1)
    fixes = 0
    for _ in range(2):
        if R >= V:
            Q, R, fixes = Q + 1, R - V, fixes + 1
    return Q, R, fixes


def divide_two_halves(u, v, extra_p=0):
    """The batched pipeline with a long quotient: q = q1 B^k + q0, both halves
    from the inverse at max(P1, P0) + 2 digits."""
    su, sv = sig(u), sig(v)
    NV = max(sv, 3)
    s = NV - sv
    V, U, NU = v << (32 * s), u << (32 * s), su + s
    P = max(su - sv + 1, 1) + 1 + extra_p
    k = (P - 1) // 2
    P1, P0 = P - k, k + 2
    h = max(P1, P0) + 2
    W = newton_inverse(V, NV, h)
    q1, R1, f1 = quot_rem(U >> (32 * k), max(NU - k, 1), V, NV, W, h, P1)
    assert R1 < V
    U3 = (R1 << (32 * k)) | (U % B**k)
    q0, R, f0 = quot_rem(U3, NV + 1 + k, V, NV, W, h, P0)
    assert q0 < B**k or k == 0
    return (q1 << (32 * k)) + q0, R >> (32 * s), max(f1, f0)


def cases(rng, count, max_digits):
    for _ in range(count):
        mv = rng.randint(1, max_digits)
        mu = rng.randint(1, max_digits + 20)
        v = rng.getrandbits(32 * mv)
        u = rng.getrandbits(32 * mu)
        k = rng.random()
        if k < 0.1:
            v = B ** (mv - 1)                                  # the smallest top digits
        elif k < 0.2:
            v = B**mv - 1                                      # all ones
        elif k < 0.3:
            v = B ** (mv - 1) + rng.getrandbits(32 * rng.randint(0, 2))
        elif k < 0.4:
            v = B**mv - 1 - rng.getrandbits(32 * rng.randint(0, mv))
        elif k < 0.5:
            v = (B ** (mv - 1)) * (1 + rng.getrandbits(31)) + (B ** max(0, mv - 3) - 1)   # low digits all ones
        if rng.random() < 0.15:
            u = B**mu - 1
        elif rng.random() < 0.15 and v:
            u = v * rng.getrandbits(32 * max(1, mu - mv)) + v - 1
        if v:
            yield u, v


@pytest.mark.parametrize("start_max,window,extra", [(31, True, 0), (3, True, 0), (3, True, 7), (31, False, 0),
                                                    (3, False, 0)])
def test_newton_division_exact(start_max, window, extra):
    """The per-instance schedule (one-launch kernel: start at up to 31
    digits, window), the batch schedule (3-digit start, window, a precision
    above the instance's), and both without the window: exact quotient and
    remainder, at most one delta correction, every invariant on the way."""
    rng = random.Random(start_max * 100 + window * 10 + extra)
    worst = 0
    for u, v in cases(rng, 400, 48):
        q, r, fixes = divide(u, v, start_max, window, extra)
        assert (q, r) == divmod(u, v), (hex(u), hex(v))
        worst = max(worst, fixes)
    assert worst <= 1


@pytest.mark.parametrize("extra", [0, 1, 9])
def test_two_half_quotient_exact(extra):
    """The long-quotient tail of the batched pipeline: the quotient in two
    halves from an inverse of half the precision, with the batch's working
    precision at or above an instance's; quotients with all-ones and empty
    halves among the cases."""
    rng = random.Random(40 + extra)
    worst = 0
    for u, v in cases(rng, 300, 40):
        q, r, fixes = divide_two_halves(u, v, extra)
        assert (q, r) == divmod(u, v), (hex(u), hex(v))
        worst = max(worst, fixes)
    for mv in (3, 7, 20):
        v = rng.getrandbits(32 * mv) | (1 << (32 * mv - 1))
        for qd in (9, 30):
            k = qd // 2
            for q in (B**qd - 1, B**k - 1, B**k, (B ** (qd - k) - 1) * B**k, 2 * B**k - 1):
                for r in (0, v - 1):
                    assert divide_two_halves(q * v + r, v, extra)[:2] == (q, r)
    assert worst <= 1


def test_newton_division_longer_operands():
    """Operands of 100-260 digits, where the window drops hundreds of digits."""
    rng = random.Random(5)
    for _ in range(12):
        mv = rng.randint(100, 200)
        mu = rng.randint(mv, 2 * mv + 60)
        v = rng.getrandbits(32 * mv) | (1 << (32 * mv - 1))
        if rng.random() < 0.3:
            v = B ** (mv - 1) + rng.getrandbits(32)
        u = rng.getrandbits(32 * mu)
        for start_max in (31, 3):
            q, r, _ = divide(u, v, start_max)
            assert (q, r) == divmod(u, v)


def wave_carry(g, p, cin):
    e, c = [0] * 64, cin
    for i in range(64):
        e[i] = c
        c = 1 if g[i] else (c if p[i] else 0)
    return e, c


def recip_top(Vd, NV, p0):
    """Lane-level emulation of wave_recip_top (bi_divone.hip)."""
    lanes = range(64)
    t = [(Vd[NV - p0 + l] if NV - p0 + l >= 0 else 0) if l < p0 else 0 for l in lanes]
    inc, _ = wave_carry([False] * 64, [l < p0 and t[l] == M for l in lanes], 1)
    if inc[p0]:
        return B**p0
    dd = [((t[l] + inc[l]) & M) if l < p0 else 0 for l in lanes]
    s = 32 - dd[p0 - 1].bit_length()
    below = [0] + dd[:63]
    dnv = [((dd[l] << s) & M) | (below[l] >> (32 - s) if s else 0) for l in lanes]
    dn = [0] + dnv[:p0]
    vt, v2 = dnv[p0 - 1], dnv[p0 - 2]
    r = [(1 << s) if l == 2 * p0 else 0 for l in lanes]
    X = [0] * (p0 + 1)
    for j in range(p0, -1, -1):
        da = [dn[l - j + 1] if 0 <= l - j < p0 else 0 for l in lanes]
        db = [dn[l - j] if 1 <= l - j <= p0 else 0 for l in lanes]
        top, nx, nx2 = r[j + p0], r[j + p0 - 1], r[j + p0 - 2]
        if top >= vt:
            qh, rh = M, nx + vt
        else:
            qh, rh = divmod((top << 32) | nx, vt)
        for _ in range(2):
            if rh < B and qh * v2 > ((rh << 32) | nx2):
                qh, rh = qh - 1, rh + vt
        lo = [(qh * da[l]) & M for l in lanes]
        hi = [(qh * db[l]) >> 32 for l in lanes]
        e1, c1 = wave_carry([r[l] < lo[l] for l in lanes], [r[l] == lo[l] for l in lanes], 0)
        y = [(r[l] - lo[l] - e1[l]) & M for l in lanes]
        e2, c2 = wave_carry([y[l] < hi[l] for l in lanes], [y[l] == hi[l] for l in lanes], 0)
        r = [(y[l] - hi[l] - e2[l]) & M for l in lanes]
        if c1 | c2:
            qh -= 1
            sm = [r[l] + da[l] for l in lanes]
            e3, _ = wave_carry([sm[l] > M for l in lanes], [(sm[l] & M) == M for l in lanes], 0)
            r = [(sm[l] + e3[l]) & M for l in lanes]
        X[j] = qh
    return sum(x << (32 * i) for i, x in enumerate(X))


def test_direct_start_long_division_emulated():
    """floor(B^(2 p0) / (T + 1)) for p0 = 3..31 and divisors with the
    smallest, largest and carry-chain tops, against Python's division."""
    rng = random.Random(9)
    n = 0
    while n < 600:
        p0 = rng.randint(3, 31)
        NV = rng.randint(3, 40)
        k = rng.random()
        V = rng.getrandbits(32 * NV) | (1 << (32 * NV - 1 - rng.randint(0, 31)))
        if k < 0.15:
            V = B ** (NV - 1)
        elif k < 0.3:
            V = B**NV - 1
        elif k < 0.4:
            V = B**NV - 1 - B ** max(0, NV - p0)
        Vd = digits(V, NV)
        T = V >> (32 * (NV - p0)) if NV >= p0 else V << (32 * (p0 - NV))
        assert recip_top(Vd, NV, p0) == B ** (2 * p0) // (T + 1)
        n += 1
