"""ctypes binding of libbigint's C ABI (include/bigint.h).

This is the host-side mirror a Python caller (the tests, bench.py, the smoke
check) uses; it performs no arithmetic itself.  Device buffers are allocated
through the library (bi_dev_alloc) so no torch tensor is involved.  When the
shared library is missing the import of a call fails loudly: there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbigint.so")

BI_OK = 0
ERRORS = {
    -1: "BI_ERR_ARG", -2: "BI_ERR_SHAPE", -3: "BI_ERR_HIP", -4: "BI_ERR_NOMEM",
    -5: "BI_ERR_DIVZERO", -6: "BI_ERR_BUSY", -7: "BI_ERR_NODEV", -8: "BI_ERR_STATE",
}

P = ctypes.c_void_p
U32, U64, I32, INT, SZ = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int, ctypes.c_size_t
FLT_P = ctypes.POINTER(ctypes.c_float)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "bi_create": [ctypes.POINTER(P), INT],
    "bi_destroy": [P],
    "bi_strerror": [INT],
    "bi_last_hip_error": [P],
    "bi_sync": [P],
    "bi_stream": [P],
    "bi_dev_alloc": [P, ctypes.POINTER(P), SZ],
    "bi_dev_free": [P, P],
    "bi_copy_h2d": [P, P, P, SZ],
    "bi_copy_d2h": [P, P, P, SZ],
    "bi_copy_d2d": [P, P, P, SZ],
    "bi_timer_start": [P],
    "bi_timer_stop": [P, FLT_P],
    "bi_device_info": [P, ctypes.POINTER(INT), ctypes.POINTER(INT), ctypes.c_char_p, SZ],
    "bi_probe_mac_rate": [P, ctypes.POINTER(ctypes.c_double)],
    "bi_copy_kernel": [P, P, P, SZ],
    "bi_tune_run": [P, U32, U64, INT, FLT_P],
    "bi_tune_set_add": [P, U32, U32, U32, U32],
    "bi_tune_set_sub": [P, U32, U32, U32, U32],
    "bi_tune_set_mul": [P, U32, U32],
    "bi_tune_set_csr_lanes": [P, U32],
    "bi_tune_set_csr": [P, U32, U32, U32, U32, U32],
    "bi_tune_get_sub": [P, U32, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32)],
    "bi_tune_get_csr": [P, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32),
                        ctypes.POINTER(U32)],
    "bi_tune_set_mul_limb_bits": [P, U32],
    "bi_tune_get_mul_limb_bits": [P, ctypes.POINTER(U32)],
    "bi_tune_get": [P, U32, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32),
                    ctypes.POINTER(U32), ctypes.POINTER(U32)],
    "bi_tune_save": [P, ctypes.c_char_p],
    "bi_tune_load": [P, ctypes.c_char_p],
    "bi_add_u32": [P, U32, U64, P, P, P],
    "bi_add_u64": [P, U32, U64, P, P, P],
    "bi_add_u32_dev": [P, U32, U64, P, P, P],
    "bi_add_u64_dev": [P, U32, U64, P, P, P],
    "bi_mul_u32": [P, U32, U64, P, P, P],
    "bi_mul_u64": [P, U32, U64, P, P, P],
    "bi_mul_u32_dev": [P, U32, U64, P, P, P],
    "bi_mul_u64_dev": [P, U32, U64, P, P, P],
    "bi_mul_u32_dev_noalloc": [P, U32, U64, P, P, P],
    "bi_add_chain_u32_dev": [P, U32, U64, U32, P, P, P],
    "bi_export_sizes_u32_dev": [P, U32, U64, P, U32, P],
    "bi_export_u32_dev": [P, U32, U64, P, U32, P, P],
    "bi_import_u32_dev": [P, U32, U64, P, U32, P, P],
    "bi_mul_chain_u32_dev": [P, U32, U64, U32, P, P, P],
    "bi_mul_full_u32_dev": [P, U64, U32, U32, P, P, P, INT],
    "bi_add_csr_u32_dev": [P, U64, P, U64, P, P, P],
    "bi_csr_plan_create": [P, U64, P, INT, U32, ctypes.POINTER(P)],
    "bi_csr_plan_destroy": [P],
    "bi_csr_plan_info": [P, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32),
                         ctypes.POINTER(ctypes.c_double)],
    "bi_mul_csr_u32_dev": [P, P, P, P, P, P, P],
    "bi_addmul_csr_u32_dev": [P, P, P, P, P, P, P, P],
    "bi_divmod_u32_dev": [P, U64, U32, U32, P, P, P, P],
    "bi_divmod_u32": [P, U64, U32, U32, P, P, P, P],
    "bi_divmod_u64": [P, U64, U32, U32, P, P, P, P],
    "bi_add_csr_u32": [P, U64, P, P, P, P],
    "bi_mul_csr_u32": [P, U64, P, P, P, P, P, INT],
    "bi_divmod_u64_dev": [P, U64, U32, U32, P, P, P, P],
    "bi_div_last_info": [P, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(ctypes.c_double)],
    "bi_div_last_work": [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "bi_tune_set_div_one": [P, U32],
    "bi_tune_get_div_one": [P, ctypes.POINTER(U32)],
    "bi_srv_start": [P, SZ, U32, U32, ctypes.POINTER(P)],
    "bi_srv_submit": [P, INT, U32, P, U32, P, ctypes.POINTER(U64)],
    "bi_srv_poll": [P, U64, P, P, ctypes.POINTER(INT), ctypes.POINTER(ctypes.c_double)],
    "bi_srv_stats": [P, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(SZ)],
    "bi_srv_stop": [P],
    "bi_srv_run_trace": [P, U64, P, P, P, P, P, P, P, P, P, P, P, P, P, P, ctypes.POINTER(ctypes.c_double)],
    "bi_sub_u32_dev": [P, U32, U64, P, P, P, P],
    "bi_cmp_u32_dev": [P, U32, U64, P, P, P],
    "bi_prec_u32_dev": [P, U32, U64, P, P],
    "bi_shift_u32_dev": [P, U32, U64, I32, P, P],
    "bi_muld_u32_dev": [P, U32, U64, P, P, P],
    "bi_quod_u32_dev": [P, U32, U64, P, P, P, P],
    "bi_sub_u64_dev": [P, U32, U64, P, P, P, P],
    "bi_cmp_u64_dev": [P, U32, U64, P, P, P],
    "bi_prec_u64_dev": [P, U32, U64, P, P],
    "bi_shift_u64_dev": [P, U32, U64, I32, P, P],
}
_RESTYPE = {"bi_strerror": ctypes.c_char_p, "bi_last_hip_error": ctypes.c_char_p, "bi_stream": P}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built; run __graft_entry__.build() (make -C big-int_amd)")
        L = ctypes.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, INT)
        _lib = L
    return _lib


class BigIntError(RuntimeError):
    def __init__(self, code, where, detail=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({lib().bi_strerror(code).decode()}) {detail}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DevBuf:
    """A device allocation owned by a Context (freed with it or explicitly)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx.check(lib().bi_dev_alloc(ctx.h, ctypes.byref(p), self.nbytes), "bi_dev_alloc")
        self.ptr = p

    def free(self):
        if self.ptr:
            lib().bi_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(lib().bi_copy_h2d(self.ctx.h, self.ptr, _ptr(arr), arr.nbytes), "bi_copy_h2d")

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(lib().bi_copy_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes), "bi_copy_d2h")
        return out


class CsrPlan:
    def __init__(self, h, n, mode):
        self.h, self.n, self.mode = h, n, mode

    def info(self):
        nb, blk, lds = U32(), U32(), U32()
        um = ctypes.c_double()
        lib().bi_csr_plan_info(self.h, ctypes.byref(nb), ctypes.byref(blk), ctypes.byref(lds), ctypes.byref(um))
        return {"bundles": nb.value, "block": blk.value, "lds_bytes": lds.value, "useful_macs": um.value}

    def close(self):
        if self.h:
            lib().bi_csr_plan_destroy(self.h)
            self.h = None


OP_ADD, OP_MUL, OP_DIV = 0, 1, 2


class Server:
    """Continuous-batching server (bi_srv_*) on a Context."""

    def __init__(self, ctx, arena_bytes=1 << 30, max_batch=4096, slots=2):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(lib().bi_srv_start(ctx.h, int(arena_bytes), int(max_batch), int(slots), ctypes.byref(h)),
                  "bi_srv_start")
        self.h = h

    def submit(self, op, a, b):
        a = np.ascontiguousarray(a, np.uint32)
        b = np.ascontiguousarray(b, np.uint32)
        t = U64()
        self.ctx.check(lib().bi_srv_submit(self.h, int(op), len(a), _ptr(a), len(b), _ptr(b), ctypes.byref(t)),
                       "bi_srv_submit")
        return t.value, len(a), len(b), op

    def wait(self, tk, timeout=60.0):
        import time
        ticket, ma, mb, op = tk
        r0 = np.zeros(ma, np.uint32)
        r1 = np.zeros(mb, np.uint32)
        done = INT()
        lat = ctypes.c_double()
        t0 = time.time()
        while True:
            rc = lib().bi_srv_poll(self.h, ticket, _ptr(r0), _ptr(r1), ctypes.byref(done), ctypes.byref(lat))
            if done.value:
                return (r0, r1, rc, lat.value) if op == OP_DIV else (r0, rc, lat.value)
            if rc:
                raise BigIntError(rc, "bi_srv_poll")
            if time.time() - t0 > timeout:
                raise TimeoutError("request not completed")
            time.sleep(0.0005)

    def stats(self):
        b, mb = U64(), U64()
        peak = SZ()
        rc = lib().bi_srv_stats(self.h, ctypes.byref(b), ctypes.byref(mb), ctypes.byref(peak))
        return {"batches": b.value, "max_batch": mb.value, "peak_arena_bytes": peak.value, "status": rc}

    def run_trace(self, ops, a_list, b_list, arrival_s):
        n = len(ops)
        ops = np.ascontiguousarray(ops, np.int32)
        ma = np.array([len(x) for x in a_list], np.uint32)
        mb = np.array([len(x) for x in b_list], np.uint32)
        a_off = np.zeros(n, np.uint64)
        b_off = np.zeros(n, np.uint64)
        if n > 1:
            a_off[1:] = np.cumsum(ma[:-1])
            b_off[1:] = np.cumsum(mb[:-1])
        a_pool = np.concatenate(a_list).astype(np.uint32) if n else np.zeros(1, np.uint32)
        b_pool = np.concatenate(b_list).astype(np.uint32) if n else np.zeros(1, np.uint32)
        r0 = np.zeros(max(1, int(ma.sum())), np.uint32)
        r1 = np.zeros(max(1, int(mb.sum())), np.uint32)
        lat = np.zeros(n, np.float64)
        status = np.zeros(n, np.int32)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        wall = ctypes.c_double()
        self.ctx.check(lib().bi_srv_run_trace(self.h, n, _ptr(ops), _ptr(ma), _ptr(mb), _ptr(a_off), _ptr(a_pool),
                                              _ptr(b_off), _ptr(b_pool), _ptr(arr), _ptr(r0), _ptr(a_off), _ptr(r1),
                                              _ptr(b_off), _ptr(lat), _ptr(status), ctypes.byref(wall)),
                       "bi_srv_run_trace")
        outs = []
        for i in range(n):
            q = r0[a_off[i]:a_off[i] + ma[i]]
            outs.append((q, r1[b_off[i]:b_off[i] + mb[i]]) if ops[i] == OP_DIV else q)
        return outs, lat, status, wall.value

    def close(self):
        if self.h:
            rc = lib().bi_srv_stop(self.h)
            self.h = None
            return rc
        return 0


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().bi_create(ctypes.byref(h), int(device))
        if rc != BI_OK:
            raise BigIntError(rc, "bi_create")
        self.h = h
        self._bufs = []

    def check(self, rc, where):
        if rc != BI_OK:
            detail = lib().bi_last_hip_error(self.h).decode() if rc == -3 else ""
            raise BigIntError(rc, where, detail)
        return rc

    def close(self):
        for b in self._bufs:
            b.free()
        self._bufs = []
        if self.h:
            lib().bi_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes):
        b = DevBuf(self, nbytes)
        self._bufs.append(b)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = self.alloc(max(arr.nbytes, 4))
        b.upload(arr)
        return b

    def sync(self):
        self.check(lib().bi_sync(self.h), "bi_sync")

    def timer_start(self):
        self.check(lib().bi_timer_start(self.h), "bi_timer_start")

    def timer_stop(self):
        ms = ctypes.c_float()
        self.check(lib().bi_timer_stop(self.h, ctypes.byref(ms)), "bi_timer_stop")
        return ms.value

    def device_info(self):
        cus, clk = ctypes.c_int(), ctypes.c_int()
        name = ctypes.create_string_buffer(256)
        self.check(lib().bi_device_info(self.h, ctypes.byref(cus), ctypes.byref(clk), name, 256), "bi_device_info")
        return {"cus": cus.value, "clock_khz": clk.value, "name": name.value.decode()}

    def mac_rate(self):
        r = ctypes.c_double()
        self.check(lib().bi_probe_mac_rate(self.h, ctypes.byref(r)), "bi_probe_mac_rate")
        return r.value

    def copy_kernel(self, dst, src, nbytes):
        self.check(lib().bi_copy_kernel(self.h, dst.ptr, src.ptr, nbytes), "bi_copy_kernel")

    # ---- autotuner
    def tune(self, m, bytes_per_operand, what):
        ms = ctypes.c_float()
        self.check(lib().bi_tune_run(self.h, m, int(bytes_per_operand), int(what), ctypes.byref(ms)), "bi_tune_run")
        return ms.value

    def tune_get(self, m):
        v = [U32() for _ in range(6)]
        self.check(lib().bi_tune_get(self.h, m, *[ctypes.byref(x) for x in v]), "bi_tune_get")
        return {"add_span_digits": v[0].value, "add_tiles_in_flight": v[1].value, "add_nontemporal": v[2].value,
                "mul_waves_per_block": v[3].value, "mul_rows": v[4].value, "csr_lanes": v[5].value,
                "mul_limb_bits": self.tune_get_mul_limb_bits()}

    def tune_get_sub(self, m):
        v = [U32() for _ in range(3)]
        self.check(lib().bi_tune_get_sub(self.h, m, *[ctypes.byref(x) for x in v]), "bi_tune_get_sub")
        return {"sub_span_digits": v[0].value, "sub_tiles_in_flight": v[1].value, "sub_nontemporal": v[2].value}

    def tune_get_csr(self):
        v = [U32() for _ in range(5)]
        self.check(lib().bi_tune_get_csr(self.h, *[ctypes.byref(x) for x in v]), "bi_tune_get_csr")
        return {"csr_add_span_digits": v[0].value, "csr_add_tiles_in_flight": v[1].value,
                "csr_add_nontemporal": v[2].value, "csr_mul_lanes": v[3].value, "csr_mul_rows": v[4].value}

    def tune_set_add(self, m, span, tiles, nt):
        self.check(lib().bi_tune_set_add(self.h, m, span, tiles, nt), "bi_tune_set_add")

    def tune_set_sub(self, m, span, tiles, nt):
        self.check(lib().bi_tune_set_sub(self.h, m, span, tiles, nt), "bi_tune_set_sub")

    def tune_set_csr(self, add_span, add_tiles, add_nt, lanes, rows):
        self.check(lib().bi_tune_set_csr(self.h, add_span, add_tiles, add_nt, lanes, rows), "bi_tune_set_csr")

    def tune_set_mul(self, waves, rows=8):
        self.check(lib().bi_tune_set_mul(self.h, waves, rows), "bi_tune_set_mul")

    def tune_set_mul_limb_bits(self, bits):
        self.check(lib().bi_tune_set_mul_limb_bits(self.h, bits), "bi_tune_set_mul_limb_bits")

    def tune_get_mul_limb_bits(self):
        v = U32()
        self.check(lib().bi_tune_get_mul_limb_bits(self.h, ctypes.byref(v)), "bi_tune_get_mul_limb_bits")
        return v.value

    def tune_save(self, path):
        self.check(lib().bi_tune_save(self.h, path.encode()), "bi_tune_save")

    def tune_load(self, path):
        self.check(lib().bi_tune_load(self.h, path.encode()), "bi_tune_load")

    # ---- host-pointer entry points (drop-in for gpuAdd / gpuMultiply)
    def _host2(self, fn, a, b, dtype=np.uint32):
        a = np.ascontiguousarray(a, dtype)
        b = np.ascontiguousarray(b, dtype)
        n, m = a.shape
        r = np.empty_like(a)
        self.check(getattr(lib(), fn)(self.h, m, n, _ptr(a), _ptr(b), _ptr(r)), fn)
        return r

    def add(self, a, b):
        return self._host2("bi_add_u32", a, b)

    def add_u64(self, a, b):
        return self._host2("bi_add_u64", a, b, np.uint64)

    def mul(self, a, b):
        return self._host2("bi_mul_u32", a, b)

    def mul_u64(self, a, b):
        return self._host2("bi_mul_u64", a, b, np.uint64)

    # ---- fused chains: p*(a+b) and (a*b)^p mod B^m
    def add_chain_dev(self, m, n, p, da, db, dr):
        self.check(lib().bi_add_chain_u32_dev(self.h, m, n, p, da.ptr, db.ptr, dr.ptr), "bi_add_chain_u32_dev")

    def mul_chain_dev(self, m, n, p, da, db, dr):
        self.check(lib().bi_mul_chain_u32_dev(self.h, m, n, p, da.ptr, db.ptr, dr.ptr), "bi_mul_chain_u32_dev")

    def _chain(self, fn, a, b, p):
        a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
        n, m = a.shape
        da, db, dr = self.upload(a), self.upload(b), self.alloc(a.nbytes)
        fn(m, n, p, da, db, dr)
        out = dr.download(np.uint32, n * m).reshape(n, m)
        for x in (da, db, dr):
            self._release(x)
        return out

    def add_chain(self, a, b, p):
        return self._chain(self.add_chain_dev, a, b, p)

    def mul_chain(self, a, b, p):
        return self._chain(self.mul_chain_dev, a, b, p)

    # ---- full products of long operands (algo 0 auto, 1 quadratic, 2 transform)
    def mul_full_dev(self, n, ma, mb, da, db, dr, algo=0):
        self.check(lib().bi_mul_full_u32_dev(self.h, n, ma, mb, da.ptr, db.ptr, dr.ptr, algo), "bi_mul_full_u32_dev")

    def mul_full(self, a, b, algo=0):
        """a: [n][ma], b: [n][mb] uint32 -> [n][ma+mb], the whole product."""
        a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
        n, ma = a.shape
        mb = b.shape[1]
        da, db, dr = self.upload(a), self.upload(b), self.alloc(4 * n * (ma + mb))
        self.mul_full_dev(n, ma, mb, da, db, dr, algo)
        out = dr.download(np.uint32, n * (ma + mb)).reshape(n, ma + mb)
        for x in (da, db, dr):
            self._release(x)
        return out

    # ---- GMP-compatible export / import (order -1, 4- or 8-byte words)
    def export(self, x, wsize=4):
        """x: [n][m] uint32 -> (offs [n+1] uint64 in words, packed words as uint32)."""
        x = np.ascontiguousarray(x, np.uint32)
        n, m = x.shape
        dx, doff = self.upload(x), self.alloc(8 * (n + 1))
        self.check(lib().bi_export_sizes_u32_dev(self.h, m, n, dx.ptr, wsize, doff.ptr), "bi_export_sizes_u32_dev")
        offs = doff.download(np.uint64, n + 1)
        total32 = int(offs[-1]) * (wsize // 4)
        dout = self.alloc(max(4, 4 * total32))
        self.check(lib().bi_export_u32_dev(self.h, m, n, dx.ptr, wsize, doff.ptr, dout.ptr), "bi_export_u32_dev")
        out = dout.download(np.uint32, total32)
        for b in (dx, doff, dout):
            self._release(b)
        return offs, out

    def import_(self, offs, words, m, wsize=4):
        """inverse of export: packed words (uint32 view) + offsets -> [n][m]."""
        offs = np.ascontiguousarray(offs, np.uint64)
        words = np.ascontiguousarray(words, np.uint32)
        n = len(offs) - 1
        din = self.upload(words if words.size else np.zeros(1, np.uint32))
        doff, dx = self.upload(offs), self.alloc(4 * n * m)
        self.check(lib().bi_import_u32_dev(self.h, m, n, din.ptr, wsize, doff.ptr, dx.ptr), "bi_import_u32_dev")
        x = dx.download(np.uint32, n * m).reshape(n, m)
        for b in (din, doff, dx):
            self._release(b)
        return x

    # ---- helpers (host arrays in, host arrays out; device round trip)
    def sub(self, a, b):
        a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
        n, m = a.shape
        da, db, dr, ds = self.upload(a), self.upload(b), self.alloc(a.nbytes), self.alloc(max(n, 4))
        self.check(lib().bi_sub_u32_dev(self.h, m, n, da.ptr, db.ptr, dr.ptr, ds.ptr), "bi_sub_u32_dev")
        out = dr.download(np.uint32, n * m).reshape(n, m), ds.download(np.uint8, n)
        for x in (da, db, dr, ds):
            self._release(x)
        return out

    def cmp(self, a, b):
        a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
        n, m = a.shape
        da, db, do = self.upload(a), self.upload(b), self.alloc(max(n, 4))
        self.check(lib().bi_cmp_u32_dev(self.h, m, n, da.ptr, db.ptr, do.ptr), "bi_cmp_u32_dev")
        out = do.download(np.int8, n)
        for x in (da, db, do):
            self._release(x)
        return out

    def prec(self, a):
        a = np.ascontiguousarray(a, np.uint32)
        n, m = a.shape
        da, do = self.upload(a), self.alloc(4 * n)
        self.check(lib().bi_prec_u32_dev(self.h, m, n, da.ptr, do.ptr), "bi_prec_u32_dev")
        out = do.download(np.uint32, n)
        for x in (da, do):
            self._release(x)
        return out

    def shift(self, s, a, in_place=False):
        a = np.ascontiguousarray(a, np.uint32)
        n, m = a.shape
        da = self.upload(a)
        dr = da if in_place else self.alloc(a.nbytes)
        self.check(lib().bi_shift_u32_dev(self.h, m, n, int(s), da.ptr, dr.ptr), "bi_shift_u32_dev")
        out = dr.download(np.uint32, n * m).reshape(n, m)
        for x in {id(da): da, id(dr): dr}.values():
            self._release(x)
        return out

    def muld(self, a, d):
        a, d = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(d, np.uint32)
        n, m = a.shape
        da, dd, dr = self.upload(a), self.upload(d), self.alloc(a.nbytes)
        self.check(lib().bi_muld_u32_dev(self.h, m, n, da.ptr, dd.ptr, dr.ptr), "bi_muld_u32_dev")
        out = dr.download(np.uint32, n * m).reshape(n, m)
        for x in (da, dd, dr):
            self._release(x)
        return out

    def quod(self, a, d):
        a, d = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(d, np.uint32)
        n, m = a.shape
        da, dd, dr, drem = self.upload(a), self.upload(d), self.alloc(a.nbytes), self.alloc(4 * n)
        self.check(lib().bi_quod_u32_dev(self.h, m, n, da.ptr, dd.ptr, dr.ptr, drem.ptr), "bi_quod_u32_dev")
        out = dr.download(np.uint32, n * m).reshape(n, m), drem.download(np.uint32, n)
        for x in (da, dd, dr, drem):
            self._release(x)
        return out

    # ---- ragged (CSR) batches
    def csr_plan(self, off, mode=0, lane_budget=0):
        off = np.ascontiguousarray(off, np.int64)
        h = ctypes.c_void_p()
        self.check(lib().bi_csr_plan_create(self.h, len(off) - 1, _ptr(off), int(mode), int(lane_budget),
                                            ctypes.byref(h)), "bi_csr_plan_create")
        return CsrPlan(h, len(off) - 1, mode)

    def add_csr_dev(self, n, total, doff, da, db, dr):
        self.check(lib().bi_add_csr_u32_dev(self.h, n, doff.ptr, total, da.ptr, db.ptr, dr.ptr), "bi_add_csr_u32_dev")

    def mul_csr_dev(self, plan, doff, doff_out, da, db, dr):
        self.check(lib().bi_mul_csr_u32_dev(self.h, plan.h, doff.ptr, doff_out.ptr if doff_out else None, da.ptr,
                                            db.ptr, dr.ptr), "bi_mul_csr_u32_dev")

    def addmul_csr_dev(self, plan, doff, doff_out, da, db, dr_add, dr_mul):
        self.check(lib().bi_addmul_csr_u32_dev(self.h, plan.h, doff.ptr, doff_out.ptr if doff_out else None, da.ptr,
                                               db.ptr, dr_add.ptr, dr_mul.ptr), "bi_addmul_csr_u32_dev")

    def add_csr(self, off, a, b):
        off = np.ascontiguousarray(off, np.int64)
        doff, da, db = self.upload(off), self.upload(a), self.upload(b)
        dr = self.alloc(max(4, a.nbytes))
        self.add_csr_dev(len(off) - 1, int(off[-1]), doff, da, db, dr)
        out = dr.download(np.uint32, len(a))
        for x in (doff, da, db, dr):
            self._release(x)
        return out

    def mul_csr(self, off, a, b, mode=0, lane_budget=0, with_add=False):
        off = np.ascontiguousarray(off, np.int64)
        plan = self.csr_plan(off, mode, lane_budget)
        sizes = np.diff(off)
        off_out = np.zeros_like(off)
        np.cumsum(2 * sizes, out=off_out[1:])
        nout = int(off_out[-1]) if mode else len(a)
        doff, doo, da, db = self.upload(off), self.upload(off_out), self.upload(a), self.upload(b)
        dr = self.alloc(max(4, 4 * nout))
        if with_add:
            dra = self.alloc(max(4, a.nbytes))
            self.addmul_csr_dev(plan, doff, doo, da, db, dra, dr)
            radd = dra.download(np.uint32, len(a))
            self._release(dra)
        else:
            self.mul_csr_dev(plan, doff, doo, da, db, dr)
        out = dr.download(np.uint32, nout)
        for x in (doff, doo, da, db, dr):
            self._release(x)
        plan.close()
        if with_add:
            return out, radd
        return (out, off_out) if mode else out

    # ---- division
    def divmod_dev(self, n, mu, mv, du, dv, dq, dr):
        return lib().bi_divmod_u32_dev(self.h, n, mu, mv, du.ptr, dv.ptr, dq.ptr, dr.ptr)

    def divmod(self, u, v, allow_zero=False):
        u, v = np.ascontiguousarray(u, np.uint32), np.ascontiguousarray(v, np.uint32)
        n, mu = u.shape
        mv = v.shape[1]
        du, dv, dq, dr = self.upload(u), self.upload(v), self.alloc(u.nbytes), self.alloc(v.nbytes)
        rc = self.divmod_dev(n, mu, mv, du, dv, dq, dr)
        if not (allow_zero and rc == -5):
            self.check(rc, "bi_divmod_u32_dev")
        q, r = dq.download(np.uint32, n * mu).reshape(n, mu), dr.download(np.uint32, n * mv).reshape(n, mv)
        for x in (du, dv, dq, dr):
            self._release(x)
        return (q, r, rc) if allow_zero else (q, r)

    def tune_set_div_one(self, n_max):
        self.check(lib().bi_tune_set_div_one(self.h, n_max), "bi_tune_set_div_one")

    def tune_get_div_one(self):
        v = U32()
        self.check(lib().bi_tune_get_div_one(self.h, ctypes.byref(v)), "bi_tune_get_div_one")
        return v.value

    def div_info(self):
        P_, st = U32(), U32()
        macs = ctypes.c_double()
        self.check(lib().bi_div_last_info(self.h, ctypes.byref(P_), ctypes.byref(st), ctypes.byref(macs)),
                   "bi_div_last_info")
        vm, tm = ctypes.c_double(), ctypes.c_double()
        self.check(lib().bi_div_last_work(self.h, ctypes.byref(vm), ctypes.byref(tm)), "bi_div_last_work")
        return {"P": P_.value, "newton_steps": st.value, "macs": macs.value, "valu_macs": vm.value,
                "transform_macs": tm.value}

    def _release(self, buf):
        buf.free()
        self._bufs = [b for b in self._bufs if b is not buf]

    # ---- device-pointer entry points
    def add_dev(self, m, n, da, db, dr):
        self.check(lib().bi_add_u32_dev(self.h, m, n, da.ptr, db.ptr, dr.ptr), "bi_add_u32_dev")

    def mul_dev(self, m, n, da, db, dr):
        self.check(lib().bi_mul_u32_dev(self.h, m, n, da.ptr, db.ptr, dr.ptr), "bi_mul_u32_dev")
