"""ORACLE / TEST INFRASTRUCTURE ONLY.

ctypes wrapper around ``oracle/liboracle.so`` (the sequential C restatement in
``bigint_oracle.c``) plus two independent exact references used to pin it:

* ``py_*`` -- closed forms on Python ``int`` (arbitrary precision);
* ``Gmp``  -- GNU MP through ``libgmp.so.10`` (dlopen, documented ``mpz_*``
  entry points), the reference's own validation library
  (cuda/helpers.h:161-187, prototypes/div.c:470-486).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u32p = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        U32, U64, I32, INT = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int
        sig = {
            "or_add_batch": (None, [U32, U64, P, P, P, INT]),
            "or_sub_batch": (None, [U32, U64, P, P, P, P, INT]),
            "or_mul_batch": (None, [U32, U64, P, P, P, INT]),
            "or_cmp_batch": (None, [U32, U64, P, P, P, INT]),
            "or_prec_batch": (None, [U32, U64, P, P, INT]),
            "or_shift_batch": (None, [I32, U32, U64, P, P, INT]),
            "or_muld_batch": (None, [U32, U64, P, P, P, INT]),
            "or_quod_batch": (None, [U32, U64, P, P, P, P, INT]),
            "or_div_batch": (INT, [U32, U32, U64, P, P, P, P, INT]),
            "or_add_csr": (None, [U64, P, P, P, P, INT]),
            "or_mul_csr": (None, [U64, P, P, P, P, P, INT, INT]),
            "or_carry_op": (U32, [U32, U32]),
            "or_mul_low": (None, [U32, P, P, P]),
            "or_mul_full": (None, [U32, P, U32, P, P]),
            "or_mul_full_batch": (None, [U32, U32, U64, P, P, P, INT]),
            "or_shinv": (INT, [P, INT, INT, P, U32]),
            "or_div": (INT, [U32, P, P, P, P]),
            "or_findk": (U32, [U32, P]),
            "og_available": (INT, []),
            "og_batch": (INT, [INT, U32, U32, U32, U64, P, P, P, P, INT]),
            "og_batch_csr": (INT, [INT, U64, P, P, P, P, INT]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype=np.uint32):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- GMP over batches (all cores)
def gmp_available():
    return bool(lib().og_available())


def gmp_add(a, b, threads=0, out=None):
    """(a + b) mod B^m per instance with GMP (mpz_add), OpenMP over instances."""
    return _gmp_op(0, a, b, threads, out)


def gmp_mul(a, b, threads=0, out=None):
    """low m digits of a*b per instance with GMP (mpz_mul)."""
    return _gmp_op(1, a, b, threads, out)


def _gmp_op(op, a, b, threads, out=None):
    a, b = _c(a), _c(b)
    n, m = a.shape
    r = np.empty_like(a) if out is None else out
    rc = lib().og_batch(op, m, m, m, n, _p(a), _p(b), _p(r), None, threads)
    if rc:
        raise RuntimeError(f"og_batch: {rc}")
    return r


def gmp_mul_full(a, b, threads=0):
    """whole products a*b (ma + mb digits) per instance with GMP (mpz_mul)."""
    a, b = _c(a), _c(b)
    n, ma = a.shape
    mb = b.shape[1]
    r = np.empty((n, ma + mb), np.uint32)
    rc = lib().og_batch(1, ma, mb, ma + mb, n, _p(a), _p(b), _p(r), None, threads)
    if rc:
        raise RuntimeError(f"og_batch: {rc}")
    return r


def gmp_div(u, v, threads=0):
    """(floor(u/v), u mod v) per instance with GMP (mpz_tdiv_qr)."""
    u, v = _c(u), _c(v)
    n, mu = u.shape
    mv = v.shape[1]
    q, r = np.empty_like(u), np.empty_like(v)
    rc = lib().og_batch(2, mu, mv, mu, n, _p(u), _p(v), _p(q), _p(r), threads)
    if rc:
        raise RuntimeError(f"og_batch: {rc}")
    return q, r


def gmp_csr(op, off, a, b, threads=0):
    """ragged add (op 0) or low-n_i multiply (op 1) with GMP."""
    off, a, b = _c(off, np.int64), _c(a), _c(b)
    r = np.zeros_like(a)
    rc = lib().og_batch_csr(op, len(off) - 1, _p(off), _p(a), _p(b), _p(r), threads)
    if rc:
        raise RuntimeError(f"og_batch_csr: {rc}")
    return r


# ---------------------------------------------------------------- batched C oracle
def add(a, b, threads=0, out=None):
    """out: optional result array of a's shape (reused, e.g. by timed passes)"""
    a, b = _c(a), _c(b)
    n, m = a.shape
    r = np.empty_like(a) if out is None else out
    lib().or_add_batch(m, n, _p(a), _p(b), _p(r), threads)
    return r


def sub(a, b, threads=0):
    a, b = _c(a), _c(b)
    n, m = a.shape
    r = np.empty_like(a)
    s = np.empty(n, np.uint8)
    lib().or_sub_batch(m, n, _p(a), _p(b), _p(r), _p(s), threads)
    return r, s


def mul(a, b, threads=0, out=None):
    a, b = _c(a), _c(b)
    n, m = a.shape
    r = np.empty_like(a) if out is None else out
    lib().or_mul_batch(m, n, _p(a), _p(b), _p(r), threads)
    return r


def mul_full(a, b, threads=0):
    """Whole products a*b (ma + mb digits) of n pairs, schoolbook."""
    a, b = _c(a), _c(b)
    n, ma = a.shape
    mb = b.shape[1]
    r = np.empty((n, ma + mb), np.uint32)
    lib().or_mul_full_batch(ma, mb, n, _p(a), _p(b), _p(r), threads)
    return r


def add_chain(a, b, p, threads=0):
    """The reference's add bench chain (cuda/ker-add.cu.h:88-112): r1 = a + b,
    then r = r1 + r until p additions are done."""
    r1 = add(a, b, threads)
    r = r1
    for _ in range(1, p):
        r = add(r1, r, threads)
    return r


def mul_chain(a, b, p, threads=0):
    """The reference's multiply bench chain (cuda/ker-mul.cu.h:223-250):
    r1 = a b, then r = r1 r until p products are done."""
    r1 = mul(a, b, threads)
    r = r1
    for _ in range(1, p):
        r = mul(r1, r, threads)
    return r


def cmp(a, b, threads=0):
    a, b = _c(a), _c(b)
    n, m = a.shape
    o = np.empty(n, np.int8)
    lib().or_cmp_batch(m, n, _p(a), _p(b), _p(o), threads)
    return o


def prec(a, threads=0):
    a = _c(a)
    n, m = a.shape
    o = np.empty(n, np.uint32)
    lib().or_prec_batch(m, n, _p(a), _p(o), threads)
    return o


def shift(s, a, threads=0):
    a = _c(a)
    n, m = a.shape
    r = np.empty_like(a)
    lib().or_shift_batch(int(s), m, n, _p(a), _p(r), threads)
    return r


def muld(a, d, threads=0):
    a, d = _c(a), _c(d)
    n, m = a.shape
    r = np.empty_like(a)
    lib().or_muld_batch(m, n, _p(a), _p(d), _p(r), threads)
    return r


def quod(a, d, threads=0):
    a, d = _c(a), _c(d)
    n, m = a.shape
    r = np.empty_like(a)
    rem = np.empty(n, np.uint32)
    lib().or_quod_batch(m, n, _p(a), _p(d), _p(r), _p(rem), threads)
    return r, rem


def div(u, v, threads=0):
    """u: [n][mu], v: [n][mv] -> (q [n][mu], r [n][mv]); restates prototypes/div.c."""
    u, v = _c(u), _c(v)
    n, mu = u.shape
    mv = v.shape[1]
    q = np.empty_like(u)
    r = np.empty((n, mv), np.uint32)
    rc = lib().or_div_batch(mu, mv, n, _p(u), _p(v), _p(q), _p(r), threads)
    if rc != 0:
        raise ZeroDivisionError("oracle div: a divisor is zero")
    return q, r


def shinv(v, h, k, p):
    v = _c(np.concatenate([_c(v), np.zeros(max(0, p - len(v)), np.uint32)])[:p])
    w = np.zeros(p, np.uint32)
    rc = lib().or_shinv(_p(v), int(h), int(k), _p(w), p)
    if rc != 0:
        raise ZeroDivisionError
    return w


def add_csr(off, a, b, threads=0):
    off, a, b = _c(off, np.int64), _c(a), _c(b)
    r = np.zeros_like(a)
    lib().or_add_csr(len(off) - 1, _p(off), _p(a), _p(b), _p(r), threads)
    return r


def mul_csr(off, a, b, mode=0, threads=0):
    """mode 0: low n_i digits in place of the operands; mode 1: full 2 n_i product."""
    off, a, b = _c(off, np.int64), _c(a), _c(b)
    n = len(off) - 1
    sizes = np.diff(off)
    off_out = np.zeros(n + 1, np.int64)
    np.cumsum(2 * sizes, out=off_out[1:])
    r = np.zeros(int(off_out[-1]) if mode else len(a), np.uint32)
    lib().or_mul_csr(n, _p(off), _p(off_out), _p(a), _p(b), _p(r), int(mode), threads)
    return (r, off_out) if mode else r


def carry_op(c1, c2):
    return int(lib().or_carry_op(int(c1), int(c2)))


# ---------------------------------------------------------------- Python int closed forms
def to_int(row):
    return int.from_bytes(np.ascontiguousarray(row, np.uint32).tobytes(), "little")


def from_int(x, m):
    return np.frombuffer(int(x).to_bytes(4 * m, "little"), np.uint32).copy()


def py_add(a, b):
    n, m = a.shape
    mod = 1 << (32 * m)
    return np.stack([from_int((to_int(a[i]) + to_int(b[i])) % mod, m) for i in range(n)])


def py_mul(a, b):
    n, m = a.shape
    mod = 1 << (32 * m)
    return np.stack([from_int((to_int(a[i]) * to_int(b[i])) % mod, m) for i in range(n)])


def py_add_chain(a, b, p):
    n, m = a.shape
    mod = 1 << (32 * m)
    return np.stack([from_int(p * (to_int(a[i]) + to_int(b[i])) % mod, m) for i in range(n)])


def py_mul_chain(a, b, p):
    n, m = a.shape
    mod = 1 << (32 * m)
    return np.stack([from_int(pow(to_int(a[i]) * to_int(b[i]), p, mod), m) for i in range(n)])


def py_export(x, wsize):
    """Closed form of the packed export: per integer its significant words
    (little-endian, wsize bytes each), back to back; offsets in words."""
    offs, parts = [0], []
    for row in x:
        v = to_int(row)
        nbytes = (v.bit_length() + 7) // 8
        words = (nbytes + wsize - 1) // wsize
        parts.append(v.to_bytes(words * wsize, "little"))
        offs.append(offs[-1] + words)
    packed = b"".join(parts)
    return np.array(offs, np.uint64), np.frombuffer(packed, np.uint32).copy() if packed else np.zeros(0, np.uint32)


def py_import(offs, words, m, wsize):
    raw = np.ascontiguousarray(words, np.uint32).tobytes()
    out = []
    for i in range(len(offs) - 1):
        v = int.from_bytes(raw[int(offs[i]) * wsize:int(offs[i + 1]) * wsize], "little")
        out.append(from_int(v % (1 << (32 * m)), m))
    return np.stack(out)


def py_sub(a, b):
    n, m = a.shape
    out, sign = [], []
    for i in range(n):
        d = to_int(a[i]) - to_int(b[i])
        out.append(from_int(abs(d), m))
        sign.append(1 if d < 0 else 0)
    return np.stack(out), np.array(sign, np.uint8)


def py_div(u, v):
    n, mu = u.shape
    mv = v.shape[1]
    qs, rs = [], []
    for i in range(n):
        q, r = divmod(to_int(u[i]), to_int(v[i]))
        qs.append(from_int(q, mu))
        rs.append(from_int(r, mv))
    return np.stack(qs), np.stack(rs)


# ---------------------------------------------------------------- GMP (second oracle)
class Gmp:
    """GNU MP via dlopen of libgmp.so.10.  mpz_t is {int alloc; int size; limb* d}."""

    _lib = None

    @classmethod
    def available(cls):
        try:
            cls._load()
            return True
        except OSError:
            return False

    @classmethod
    def _load(cls):
        if cls._lib is None:
            L = ctypes.CDLL("libgmp.so.10")
            P, S = ctypes.c_void_p, ctypes.c_size_t
            for nm, args in {
                "__gmpz_init": [P], "__gmpz_clear": [P],
                "__gmpz_import": [P, S, ctypes.c_int, S, ctypes.c_int, S, P],
                "__gmpz_export": [P, P, ctypes.c_int, S, ctypes.c_int, S, P],
                "__gmpz_add": [P, P, P], "__gmpz_sub": [P, P, P], "__gmpz_mul": [P, P, P],
                "__gmpz_abs": [P, P], "__gmpz_tdiv_qr": [P, P, P, P], "__gmpz_cmp": [P, P],
            }.items():
                f = getattr(L, nm)
                f.argtypes = args
                f.restype = P if nm == "__gmpz_export" else (ctypes.c_int if nm == "__gmpz_cmp" else None)
            cls._lib = {nm[2:]: getattr(L, nm) for nm in (
                "__gmpz_init", "__gmpz_clear", "__gmpz_import", "__gmpz_export", "__gmpz_add",
                "__gmpz_sub", "__gmpz_mul", "__gmpz_abs", "__gmpz_tdiv_qr", "__gmpz_cmp")}
        return cls._lib

    def __init__(self):
        self.L = self._load()
        self._z = []

    def z(self):
        buf = ctypes.create_string_buffer(16)
        self.L["gmpz_init"](buf)
        self._z.append(buf)
        return buf

    def close(self):
        for b in self._z:
            self.L["gmpz_clear"](b)
        self._z = []

    def imp(self, zbuf, row):
        row = np.ascontiguousarray(row, np.uint32)
        self.L["gmpz_import"](zbuf, len(row), -1, 4, 0, 0, row.ctypes.data_as(ctypes.c_void_p))

    def exp(self, zbuf, m):
        out = np.zeros(4 * m + 8, np.uint32)   # GMP writes the full value; truncate afterwards
        cnt = ctypes.c_size_t(0)
        self.L["gmpz_export"](out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cnt), -1, 4, 0, 0, zbuf)
        return out[:m].copy()

    def export_words(self, row, wsize):
        """mpz_import(order -1, 4-byte words) then mpz_export(order -1,
        wsize-byte words, native endian): (count, words as uint32)."""
        z = self.z()
        self.imp(z, row)
        out = np.zeros(len(row) + 2, np.uint32)
        cnt = ctypes.c_size_t(0)
        self.L["gmpz_export"](out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cnt), -1, wsize, 0, 0, z)
        self.close()
        return int(cnt.value), out[: int(cnt.value) * (wsize // 4)].copy()

    def import_words(self, words, count, wsize, m):
        """mpz_import(order -1, wsize-byte words) then export truncated to m digits."""
        z = self.z()
        w = np.ascontiguousarray(words, np.uint32)
        self.L["gmpz_import"](z, count, -1, wsize, 0, 0, w.ctypes.data_as(ctypes.c_void_p))
        r = self.exp(z, m)
        self.close()
        return r

    def divmod_violations(self, u, v, q, r):
        """Indices i where u_i != q_i v_i + r_i or r_i >= v_i (v_i != 0)."""
        zu, zv, zq, zr, zt = self.z(), self.z(), self.z(), self.z(), self.z()
        bad = []
        for i in range(u.shape[0]):
            self.imp(zu, u[i]); self.imp(zv, v[i]); self.imp(zq, q[i]); self.imp(zr, r[i])
            self.L["gmpz_mul"](zt, zq, zv)
            self.L["gmpz_add"](zt, zt, zr)
            if self.L["gmpz_cmp"](zt, zu) != 0 or self.L["gmpz_cmp"](zr, zv) >= 0:
                bad.append(i)
        self.close()
        return bad

    def batch(self, op, a, b, m_out=None):
        """op in add/mul/sub/div; results truncated to m digits like
        cuda/helpers.h:161-187 (sub returns |a-b|, div returns (q, r))."""
        n, m = a.shape
        m_out = m_out or m
        za, zb, zr, zs = self.z(), self.z(), self.z(), self.z()
        res, res2 = [], []
        for i in range(n):
            self.imp(za, a[i]); self.imp(zb, b[i])
            if op == "add":
                self.L["gmpz_add"](zr, za, zb)
            elif op == "mul":
                self.L["gmpz_mul"](zr, za, zb)
            elif op == "sub":
                self.L["gmpz_sub"](zs, za, zb); self.L["gmpz_abs"](zr, zs)
            elif op == "div":
                self.L["gmpz_tdiv_qr"](zr, zs, za, zb)
                res2.append(self.exp(zs, b.shape[1]))
            res.append(self.exp(zr, m_out))
        self.close()
        return (np.stack(res), np.stack(res2)) if op == "div" else np.stack(res)
