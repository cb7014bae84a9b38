"""Writes tests/golden/vectors.npz: seeded inputs with expected outputs
computed by Python int (exact closed forms), independently of the C oracle.

Cases: add / sub / low-m mul on uniform batches (m in 1..67 plus 512),
carry-chain and zero edge cases, and divmod at a few sizes.
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def to_int(row):
    return int.from_bytes(np.ascontiguousarray(row, np.uint32).tobytes(), "little")


def from_int(x, m):
    return np.frombuffer(int(x).to_bytes(4 * m, "little"), np.uint32).copy()


def main():
    rng = np.random.default_rng(20240501)
    out = {}
    for m in (1, 2, 3, 4, 5, 7, 8, 16, 31, 64, 67, 512):
        n = 12
        a = rng.integers(0, 1 << 32, (n, m), dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, 1 << 32, (n, m), dtype=np.uint64).astype(np.uint32)
        a[0], b[0] = 0xFFFFFFFF, 0
        b[0, 0] = 1                      # all-ones + 1 -> 0 (full carry chain)
        a[1], b[1] = 0xFFFFFFFF, 0xFFFFFFFF
        a[2] = 0                         # zero operand
        b[3] = a[3]                      # equal operands (sub -> 0)
        mod = 1 << (32 * m)
        add = np.stack([from_int((to_int(a[i]) + to_int(b[i])) % mod, m) for i in range(n)])
        mul = np.stack([from_int((to_int(a[i]) * to_int(b[i])) % mod, m) for i in range(n)])
        sub = np.stack([from_int(abs(to_int(a[i]) - to_int(b[i])), m) for i in range(n)])
        sgn = np.array([1 if to_int(a[i]) < to_int(b[i]) else 0 for i in range(n)], np.uint8)
        out.update({f"a_{m}": a, f"b_{m}": b, f"add_{m}": add, f"mul_{m}": mul, f"sub_{m}": sub, f"sgn_{m}": sgn})
    for (mu, mv) in ((4, 2), (8, 3), (16, 16), (64, 33), (256, 128)):
        n = 10
        u = rng.integers(0, 1 << 32, (n, mu), dtype=np.uint64).astype(np.uint32)
        v = np.zeros((n, mv), np.uint32)
        for i in range(n):
            nz = int(rng.integers(1, mv + 1))
            v[i, :nz] = rng.integers(0, 1 << 32, nz, dtype=np.uint64).astype(np.uint32)
            if v[i].sum() == 0:
                v[i, 0] = 1
        v[0] = 0
        v[0, 0] = 1                      # divide by one
        v[1] = 0xFFFFFFFF                # all-ones divisor
        q = np.stack([from_int(to_int(u[i]) // to_int(v[i]), mu) for i in range(n)])
        r = np.stack([from_int(to_int(u[i]) % to_int(v[i]), mv) for i in range(n)])
        out.update({f"du_{mu}_{mv}": u, f"dv_{mu}_{mv}": v, f"dq_{mu}_{mv}": q, f"dr_{mu}_{mv}": r})
    np.savez_compressed(os.path.join(HERE, "vectors.npz"), **out)


if __name__ == "__main__":
    main()
