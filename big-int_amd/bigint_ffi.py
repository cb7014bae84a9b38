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
    "bi_add_u32": [P, U32, U64, P, P, P],
    "bi_add_u64": [P, U32, U64, P, P, P],
    "bi_add_u32_dev": [P, U32, U64, P, P, P],
    "bi_add_u64_dev": [P, U32, U64, P, P, P],
    "bi_mul_u32": [P, U32, U64, P, P, P],
    "bi_mul_u64": [P, U32, U64, P, P, P],
    "bi_mul_u32_dev": [P, U32, U64, P, P, P],
    "bi_mul_u64_dev": [P, U32, U64, P, P, P],
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

    # ---- device-pointer entry points
    def add_dev(self, m, n, da, db, dr):
        self.check(lib().bi_add_u32_dev(self.h, m, n, da.ptr, db.ptr, dr.ptr), "bi_add_u32_dev")

    def mul_dev(self, m, n, da, db, dr):
        self.check(lib().bi_mul_u32_dev(self.h, m, n, da.ptr, db.ptr, dr.ptr), "bi_mul_u32_dev")
