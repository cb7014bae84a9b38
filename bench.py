#!/usr/bin/env python3
"""Benchmark of the batched big-integer hot path on one MI355X.

Headline (``value``): batched add of 131072 pairs of 512-digit (16384-bit)
integers, 5% of them all-ones + 1 carry chains, in HBM GB/s with
bytes = 3 * n * m * 4 (two operands read, one result written; the
reference's formula, thesis/sections/performance.tex:75-78, cuda/main.cu:201-203).
One step = one batched add over the whole batch, inputs resident in HBM.

Also reported on the same line: the multiply of the same shape with zero top
halves in G u32-MAC/s (m^2/4 digit products per instance) against the
measured v_mad_u64_u32 rate, and the CPU oracle timed on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "big-int_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_I8_PEAK_MACS = 2.5e15   # MI355X dense int8 matrix rate: 5.0 POPS = 2.5e15 multiply-adds/s


def mul_engine(ctx, m):
    """Which multiply kernel the context's setting runs at m digits."""
    bits = ctx.tune_get_mul_limb_bits()
    if bits == 8 and 33 <= m <= 1024:
        return "mfma"
    return "limb27" if bits == 27 and m <= 864 else "valu"


def mfma_count(pa, pb, m):
    """int8 32x32x32 products the matrix-core multiply issues for one
    instance whose operands have pa, pb significant digits (its step and tile
    bounds, byte for byte)."""
    nt = (m + 255) // 256
    nc = nt
    m8 = (m + 7) & ~7
    nta = min(nt, (pa + pb + 2 + 255) // 256) if (pa and pb) else 0
    pb_blocks = min(pb // 8 + 1, m8 // 8)
    qend = min(pb_blocks, 32 * (nta - 1) + 31) + 1 if nta else 0
    steps = 1 + 2 * ((max(qend, 1) - 1 + 1) // 2)
    return steps * nc
METRIC = "add HBM GB/s (% peak); mul Gu32ops/s; ragged-batch vs padded-plain speedup"


def rnd(x, k):
    return None if x is None or x != x else round(x, k)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WARM_MS = 250.0   # GPU time run before each compute-bound measurement
RAGGED_DIGITS = 48286733   # digits per operand of the 2^17-instance ragged batch (seed 31)
# Under rocprofv3 (its ROCPROF_* settings in the environment) the one-launch
# division (k_div_one) is not run: the small-batch timing and the serving
# trace, whose division rounds take it, are skipped.  A kernel trace of the
# whole bench saw its queue abort ("AQL packet is malformed") at the first
# one-launch division after the earlier phases; the same calls pass without
# the profiler (the GPU tests, bench runs), and no cause has been found in
# the kernel or its launch.
PROFILED = any(k.startswith("ROCPROF_") for k in os.environ)


def timed(ctx, fn, steps, warmup, warm_ms=0.0):
    """Average ms of `steps` back-to-back calls after `warmup` calls.  warm_ms:
    keep calling until about that much GPU time has run first -- the clock of a
    compute-bound kernel ramps up over tens of milliseconds of load, and a
    short warm-up would time the ramp, not the sustained rate."""
    for _ in range(warmup):
        fn()
    ctx.sync()
    if warm_ms > 0:
        ctx.timer_start()
        fn()
        one = max(ctx.timer_stop(), 1e-3)
        for _ in range(min(4000, int(warm_ms / one))):
            fn()
        ctx.sync()
    ctx.timer_start()
    for _ in range(steps):
        fn()
    ms = ctx.timer_stop()
    ctx.sync()
    return ms / steps


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def cpu_cores():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def host_info():
    """The GPU box's host CPU: model name and the cores this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    l3 = None
    try:   # last-level cache: sum of the distinct L3 slices
        seen = {}
        base = "/sys/devices/system/cpu"
        for cpu in os.listdir(base):
            p = os.path.join(base, cpu, "cache", "index3")
            if cpu.startswith("cpu") and os.path.isdir(p):
                with open(os.path.join(p, "shared_cpu_list")) as f:
                    key = f.read().strip()
                with open(os.path.join(p, "size")) as f:
                    seen[key] = f.read().strip()
        kib = sum(int(v.rstrip("K")) for v in seen.values() if v.endswith("K"))
        l3 = {"slices": len(seen), "total_mib": round(kib / 1024, 1)} if seen else None
    except (OSError, ValueError):
        pass
    return {"cpu_model": model, "nproc": nproc, "l3": l3}


def cpu_repeat(fn, seconds):
    """Runs fn until `seconds` have passed (at least once); (reps, wall s)."""
    fn()   # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            return reps, el


def cpu_baseline_add(a, b, seconds):
    """The CPU oracle (sequential C restatement, OpenMP over instances) on the
    whole batch (805 MB, larger than the host's last-level cache); GB/s with
    the same byte formula.  Beside it: one core, and GMP (mpz_add per
    instance with import/export, the reference's CPU path, cuda/helpers.h:
    161-187) on all cores over the whole batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = cpu_cores()
    n, m = a.shape
    out = np.empty_like(a)
    reps, el = cpu_repeat(lambda: O.add(a, b, cores, out=out), seconds)
    gbs = 3.0 * a.nbytes * reps / el / 1e9
    # the same restatement on one thread, on the first 4096 pairs
    s1a, s1b = a[:4096], b[:4096]
    o1 = np.empty_like(s1a)
    reps1, el1 = cpu_repeat(lambda: O.add(s1a, s1b, 1, out=o1), 2.0)
    gbs1 = 3.0 * s1a.nbytes * reps1 / el1 / 1e9
    res = {"value": round(gbs, 3), "unit": "GB/s", "cores": cores, "kind": "port",
           "sample": f"{reps} passes of the oracle add over the whole batch ({n} x {m} digits, "
                     f"{3 * a.nbytes / 1e6:.0f} MB a pass), {el:.1f} s wall",
           "one_core": {"value": round(gbs1, 3), "sample": f"{reps1} passes over the first 4096 pairs, "
                                                           f"{el1:.1f} s wall"}}
    if O.gmp_available():
        reps2, el2 = cpu_repeat(lambda: O.gmp_add(a, b, cores, out=out), 3.0)
        res["gmp"] = {"value": round(3.0 * a.nbytes * reps2 / el2 / 1e9, 3), "unit": "GB/s", "cores": cores,
                      "kind": "gmp", "sample": f"{reps2} passes of mpz_import/mpz_add/mpz_export over the whole "
                                               f"batch, {el2:.1f} s wall"}
    return res


def bench_mul(ctx, W, O, args, n, m, da, db, dr, idx):
    a2, b2 = W.plain_pairs(n, m, seed=2025, top_zero=True)
    da.upload(a2)
    db.upload(b2)
    mul_steps = max(5, args.steps // 5)
    mul_ms = timed(ctx, lambda: ctx.mul_dev(m, n, da, db, dr), mul_steps, 2, WARM_MS)
    macs = n * (m * m / 4)
    engine = mul_engine(ctx, m)
    info = {"engine": {"valu": "v_mad_u64_u32 multiply-adds (32-bit digits)",
                       "limb27": "v_mad_u64_u32 multiply-adds (27-bit limbs)",
                       "mfma": "int8 matrix cores"}[engine],
            "roofline": {"bound": "valu", "achieved": rnd(macs / (mul_ms * 1e-3) / 1e12, 4),
                         "peak": rnd(ctx.mac_rate() / 1e12, 4), "unit": "T MAC/s (measured mad+carry rate)",
                         "frac": rnd(macs / (mul_ms * 1e-3) / ctx.mac_rate(), 4)}}
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok = bool((r[idx] == O.mul(a2[idx], b2[idx])).all())
    info["matrix_cores_optin"] = bench_mul_mfma(ctx, O, n, m, a2, b2, da, db, dr, idx)
    cpu = None
    if not args.no_cpu:
        cores = cpu_cores()
        out = np.empty_like(a2)
        reps, el = cpu_repeat(lambda: O.mul(a2, b2, cores, out=out), 3.0)
        s1a, s1b = a2[:1024], b2[:1024]
        reps1, el1 = cpu_repeat(lambda: O.mul(s1a, s1b, 1), 2.0)
        cpu = {"value": round(reps * n * (m * m / 4) / el / 1e9, 3), "unit": "G u32-MAC/s", "cores": cores,
               "kind": "port", "sample": f"{reps} passes of the oracle multiply over the whole batch, "
                                         f"{el:.1f} s wall",
               "one_core": {"value": round(reps1 * 1024 * (m * m / 4) / el1 / 1e9, 3),
                            "sample": f"{reps1} passes over the first 1024 pairs, {el1:.1f} s wall"}}
        if O.gmp_available():
            reps2, el2 = cpu_repeat(lambda: O.gmp_mul(a2, b2, cores, out=out), 3.0)
            cpu["gmp"] = {"value": round(reps2 * n * (m * m / 4) / el2 / 1e9, 3), "unit": "G u32-MAC/s",
                          "cores": cores, "kind": "gmp",
                          "sample": f"{reps2} passes of mpz_import/mpz_mul/mpz_export over the whole batch, "
                                    f"{el2:.1f} s wall (the reference's gmpMultiply, cuda/main.cu:422-442)"}
    return mul_ms, macs / (mul_ms * 1e-3), ok, cpu, info


def bench_mul_mfma(ctx, O, n, m, a2, b2, da, db, dr, idx):
    """The opt-in byte convolution on the int8 matrix cores (limb_bits 8) on
    the same batch: outside the multiply-add design the default and the
    autotuner keep, timed here beside it.  Roofline: the slower of streaming
    12 B per digit through HBM and the pair's (2m)^2 byte products at the
    int8 matrix peak; also the products the kernel issues (tile edges and
    block padding included) against that peak."""
    old = ctx.tune_get_mul_limb_bits()
    ctx.tune_set_mul_limb_bits(8)
    try:
        t = timed(ctx, lambda: ctx.mul_dev(m, n, da, db, dr), 10, 2, WARM_MS)
        r = dr.download(np.uint32, n * m).reshape(n, m)
        ok = bool((r[idx] == O.mul(a2[idx], b2[idx])).all())
    finally:
        ctx.tune_set_mul_limb_bits(old)
    byte_macs = n * (2.0 * m) ** 2
    pa = np.where(a2.any(axis=1), m - np.argmax(a2[:, ::-1] != 0, axis=1), 0)
    pb = np.where(b2.any(axis=1), m - np.argmax(b2[:, ::-1] != 0, axis=1), 0)
    issued = sum(mfma_count(int(x), int(y), m) for x, y in zip(pa[:4096], pb[:4096])) * (n / 4096.0)
    hbm_ms = 12.0 * n * m / (HBM_PEAK_GBS * 1e9) * 1e3
    mfma_ms = byte_macs / MFMA_I8_PEAK_MACS * 1e3
    macs = n * (m * m / 4)
    log(f"mul on the matrix cores (opt-in): {t * 1e3:.1f} us/launch, {macs / (t * 1e-3) / 1e9:.1f} G MAC/s")
    return {"setting": "bi_tune_set_mul_limb_bits(ctx, 8); not the default, not chosen by the tuner",
            "ms": rnd(t, 4), "value": rnd(macs / (t * 1e-3) / 1e9, 2), "unit": "G u32-MAC/s (m^2/4 per pair)",
            "roofline": {"bound": "hbm" if hbm_ms >= mfma_ms else "mfma",
                         "frac": rnd(max(hbm_ms, mfma_ms) / t, 4),
                         "hbm_gbs": rnd(12.0 * n * m / (t * 1e-3) / 1e9, 1),
                         "algorithmic_byte_macs_t": rnd(byte_macs / (t * 1e-3) / 1e12, 2),
                         "peak_mfma_t": MFMA_I8_PEAK_MACS / 1e12, "peak_hbm_gbs": HBM_PEAK_GBS},
            "issued_per_pair": rnd(issued / n, 2),
            "issued_frac_of_mfma_peak": rnd(issued * 32768.0 / (t * 1e-3) / MFMA_I8_PEAK_MACS, 4),
            "parity_sample": ok}


def bench_chain(ctx, O, args, n, m, a, b, da, db, dr, idx):
    """The reference's bench chains, fused: ten additions r = 10(a+b) and six
    products r = (ab)^6 mod B^m (low m digits each, m(m+1)/2 digit products
    per link) on the add batch, operands read once and the result written
    once."""
    da.upload(a)
    db.upload(b)
    t_add = timed(ctx, lambda: ctx.add_chain_dev(m, n, 10, da, db, dr), max(5, args.steps // 5), 2, WARM_MS)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok = bool((r[idx] == O.add_chain(a[idx], b[idx], 10)).all())
    t_mul = timed(ctx, lambda: ctx.mul_chain_dev(m, n, 6, da, db, dr), 3, 1, WARM_MS)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    sub = idx[:128]
    ok &= bool((r[sub] == O.mul_chain(a[sub], b[sub], 6)).all())
    macs = 6.0 * n * m * (m + 1) / 2
    log(f"chains: ten adds {t_add * 1e3:.1f} us ({3.0 * n * m * 4 / (t_add * 1e-3) / 1e9:.0f} GB/s one pass), "
        f"six products {t_mul:.3f} ms ({macs / (t_mul * 1e-3) / 1e9:.0f} G MAC/s)")
    return {"workload": "131072 x 512 digits: r = 10(a+b) as ten chained adds; r = (ab)^6 mod B^512 as six "
                        "chained truncated products, on chip between links",
            "add10_ms": round(t_add, 4), "add10_gbs_one_pass": round(3.0 * n * m * 4 / (t_add * 1e-3) / 1e9, 1),
            "add10_adds_per_s": round(10.0 * n / (t_add * 1e-3), 1),
            "mul6_ms": round(t_mul, 4), "mul6_gmacs": round(macs / (t_mul * 1e-3) / 1e9, 1),
            "parity_sample": ok}


def bench_full(ctx, O, args):
    """Whole products of long operands (bi_mul_full_u32_dev) at the division
    workload's operand sizes: 2048 pairs of 8192 x 4096 digits, the
    number-theoretic transform against the quadratic long product on the same
    pairs."""
    rng = np.random.default_rng(17)
    n, ma, mb = 2048, 8192, 4096
    a = rng.integers(0, 1 << 32, (n, ma), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 1 << 32, (n, mb), dtype=np.uint64).astype(np.uint32)
    da, db, dr = ctx.upload(a), ctx.upload(b), ctx.alloc(4 * n * (ma + mb))
    res, outs = {}, {}
    for algo in (1, 2):
        res[algo] = timed(ctx, lambda: ctx.mul_full_dev(n, ma, mb, da, db, dr, algo), 3, 1, WARM_MS)
        outs[algo] = dr.download(np.uint32, n * (ma + mb)).reshape(n, ma + mb)
    idx = np.array([0, 1, 1023, n - 1])
    ok = bool((outs[2] == outs[1]).all() and (outs[2][idx] == O.mul_full(a[idx], b[idx], threads=4)).all())
    for x in (da, db, dr):
        ctx._release(x)
    work = float(n) * ma * mb
    log(f"full products 8192x4096: transform {res[2]:.3f} ms, quadratic {res[1]:.3f} ms "
        f"({res[1] / res[2]:.2f}x)")
    # crossover: square operands, 1024 pairs (quadratic time / transform time)
    cross = {}
    for m in (1024, 1536, 2048, 3072, 4096):
        x = rng.integers(0, 1 << 32, (1024, m), dtype=np.uint64).astype(np.uint32)
        dx, dy = ctx.upload(x), ctx.alloc(4 * 1024 * 2 * m)
        t = {algo: timed(ctx, lambda: ctx.mul_full_dev(1024, m, m, dx, dx, dy, algo), 3, 1, WARM_MS / 4)
             for algo in (1, 2)}
        cross[str(m)] = round(t[1] / t[2], 3)
        ctx._release(dx)
        ctx._release(dy)
    log("crossover (quadratic / transform time, m x m digits):", cross)
    cpu = None
    if not args.no_cpu and O.gmp_available():
        cores, k = cpu_cores(), 512
        t0 = time.perf_counter()
        gm = O.gmp_mul_full(a[:k], b[:k], cores)
        el = time.perf_counter() - t0
        ok = ok and bool((gm[:4] == outs[2][:4]).all())
        cpu = {"value": round(k / el, 2), "unit": "products/s", "cores": cores, "kind": "gmp",
               "sample": f"{k} products with GMP mpz_mul (the reference's `mult`, prototypes/div.c:184-199), "
                         f"OpenMP over instances, {el:.1f} s wall"}
    return {"workload": "2048 pairs, 8192 x 4096 digits (the division's operand sizes), whole 12288-digit "
                        "products",
            "transform_ms": round(res[2], 4), "quadratic_ms": round(res[1], 4),
            "speedup": round(res[1] / res[2], 3),
            "quadratic_equiv_gmacs": round(work / (res[2] * 1e-3) / 1e9, 1),
            "speedup_square_m": cross,
            "products_per_s": round(n / (res[2] * 1e-3), 1),
            "cpu_gmp": cpu,
            "parity_sample": ok}


def bench_helpers(ctx, F, O, n, m, a, b, da, db, dr, idx):
    """The helpers on the add batch (131072 x 512 digits): |a-b| with sign,
    compare, precision, shift by 7 digits, multiply and divide by one digit,
    GMP-format export and import (4-byte words).  Streaming ops report GB/s
    of algorithmic bytes (operands read once, results written once) against
    HBM; compare and precision stop at the first differing / nonzero digit
    from the top, so they report time only."""
    L = F.lib()
    rng = np.random.default_rng(23)
    d = rng.integers(1, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    da.upload(a)
    db.upload(b)
    dd = ctx.upload(d)
    small = ctx.alloc(8 * (n + 1))
    words = ctx.alloc(n * m * 4)
    digit_bytes = float(n) * m * 4
    out, ok = {}, True

    def run(name, call, nbytes):
        t = timed(ctx, lambda: ctx.check(call(), "bi_" + name), 10, 2)
        out[name] = {"us": round(t * 1e3, 1)}
        if nbytes:
            out[name]["gbs"] = round(nbytes / (t * 1e-3) / 1e9, 1)

    run("sub", lambda: L.bi_sub_u32_dev(ctx.h, m, n, da.ptr, db.ptr, dr.ptr, small.ptr), 3 * digit_bytes + n)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok &= bool((r[idx] == O.sub(a[idx], b[idx])[0]).all())
    run("cmp", lambda: L.bi_cmp_u32_dev(ctx.h, m, n, da.ptr, db.ptr, small.ptr), 0)
    run("prec", lambda: L.bi_prec_u32_dev(ctx.h, m, n, da.ptr, small.ptr), 0)
    run("shift", lambda: L.bi_shift_u32_dev(ctx.h, m, n, 7, da.ptr, dr.ptr), 2 * digit_bytes)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok &= bool((r[idx] == O.shift(7, a[idx])).all())
    run("muld", lambda: L.bi_muld_u32_dev(ctx.h, m, n, da.ptr, dd.ptr, dr.ptr), 2 * digit_bytes + 4 * n)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok &= bool((r[idx] == O.muld(a[idx], d[idx])).all())
    run("quod", lambda: L.bi_quod_u32_dev(ctx.h, m, n, da.ptr, dd.ptr, dr.ptr, None), 2 * digit_bytes + 4 * n)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok &= bool((r[idx] == O.quod(a[idx], d[idx])[0]).all())
    ctx.check(L.bi_export_sizes_u32_dev(ctx.h, m, n, da.ptr, 4, small.ptr), "bi_export_sizes_u32_dev")
    ctx.sync()
    offs = small.download(np.uint64, n + 1)
    run("export", lambda: L.bi_export_u32_dev(ctx.h, m, n, da.ptr, 4, small.ptr, words.ptr),
        digit_bytes + 4.0 * float(offs[-1]))
    run("import", lambda: L.bi_import_u32_dev(ctx.h, m, n, words.ptr, 4, small.ptr, dr.ptr),
        digit_bytes + 4.0 * float(offs[-1]))
    r = dr.download(np.uint32, n * m).reshape(n, m)
    ok &= bool((r == a).all())
    for x in (dd, small, words):
        ctx._release(x)
    log("helpers:", out)
    return {"workload": "131072 x 512 digits (the add batch); one-digit operands uniform nonzero",
            "ops": out, "parity_sample": ok}


def bench_pcie(ctx, W, O, a_big, b_big):
    """Host-pointer add (bi_add_u32: operands in, result out over PCIe,
    pageable host buffers, chunks through pinned staging on two streams) on
    16384 x 512 digits and on the whole 131072 x 512 batch: the rate a caller
    sees with host buffers, 3 n m 4 bytes moved over the link."""
    import time
    out = {"workload": "bi_add_u32, pageable host buffers"}
    ok = True
    for name, (a, b) in (("16384x512", W.plain_pairs(16384, 512, seed=9)), ("131072x512", (a_big, b_big))):
        ctx.add(a, b)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            r = ctx.add(a, b)
        dt = (time.perf_counter() - t0) / reps
        idx = np.arange(0, len(a), max(1, len(a) // 97))
        ok &= bool((r[idx] == O.add(a[idx], b[idx])).all())
        out[name] = {"gbs": round(3.0 * a.nbytes / dt / 1e9, 2), "ms": round(dt * 1e3, 3)}
    out["gbs"] = out["131072x512"]["gbs"]
    out["parity_sample"] = ok
    log("host-pointer add:", out)
    return out


def bench_ragged(ctx, W, O, args):
    """2^17 instances, digits log-uniform in [8, 2048], int64 CSR offsets, 10%
    all-ones/one chains, 5% zero operands.  Ragged add and low-n_i multiply in
    one persistent launch each (and both together), against the plain kernel
    launched once per padded power-of-two size class on the same instances.
    Useful multiply work: sum n_i(n_i+1)/2 digit products (the low n_i digits),
    the same count on both sides."""
    N = 1 << 17
    off, a, b = W.ragged(N, 8, 2048, seed=31)
    sizes = np.diff(off)
    total = int(off[-1])
    useful = float((sizes * (sizes + 1) // 2).sum())
    doff, da, db = ctx.upload(off), ctx.upload(a), ctx.upload(b)
    dr, dr2 = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
    plan = ctx.csr_plan(off, 0)
    steps = max(3, args.steps // 10)
    t_add = timed(ctx, lambda: ctx.add_csr_dev(N, total, doff, da, db, dr), max(20, args.steps), 3, WARM_MS)
    t_mul = timed(ctx, lambda: ctx.mul_csr_dev(plan, doff, None, da, db, dr), steps, 2, WARM_MS)
    t_both = timed(ctx, lambda: ctx.addmul_csr_dev(plan, doff, None, da, db, dr2, dr), steps, 2, WARM_MS)
    rm = dr.download(np.uint32, total)
    ra = dr2.download(np.uint32, total)
    pick = np.random.default_rng(1).choice(N, 256, replace=False)
    ok = True
    for i in pick:
        s, e = off[i], off[i + 1]
        want_a = O.add(a[None, s:e], b[None, s:e])[0]
        want_m = O.mul(a[None, s:e], b[None, s:e])[0]
        ok &= bool((ra[s:e] == want_a).all() and (rm[s:e] == want_m).all())
    pinfo = plan.info()
    plan.close()
    for buf in (doff, da, db, dr, dr2):
        ctx._release(buf)
    # padded plain comparator: one launch per power-of-two class
    cls = np.maximum(8, 2 ** np.ceil(np.log2(sizes)).astype(np.int64))
    t_pad = 0.0
    ok_pad = True
    for c in np.unique(cls):
        idx = np.nonzero(cls == c)[0]
        k, mc = len(idx), int(c)
        pa = np.zeros((k, mc), np.uint32)
        pb = np.zeros((k, mc), np.uint32)
        for j, i in enumerate(idx):
            s, e = off[i], off[i + 1]
            pa[j, : e - s] = a[s:e]
            pb[j, : e - s] = b[s:e]
        qa, qb, qr = ctx.upload(pa), ctx.upload(pb), ctx.alloc(pa.nbytes)
        t_pad += timed(ctx, lambda: ctx.mul_dev(mc, k, qa, qb, qr), steps, 1, WARM_MS)
        r = qr.download(np.uint32, k * mc).reshape(k, mc)
        j = 0
        i = idx[0]
        s, e = off[i], off[i + 1]
        ok_pad &= bool((r[j, : e - s] == rm[s:e]).all())
        for buf in (qa, qb, qr):
            ctx._release(buf)
    rag = useful / (t_mul * 1e-3)
    pad = useful / (t_pad * 1e-3)
    log(f"ragged: add {t_add * 1e3:.0f} us, mul {t_mul:.3f} ms, both {t_both:.3f} ms; padded mul {t_pad:.3f} ms")
    cpu = None
    if not args.no_cpu:
        cores, ns = cpu_cores(), 8192
        so = off[: ns + 1].copy()
        sa, sb = a[: so[-1]].copy(), b[: so[-1]].copy()
        su = float((sizes[:ns] * (sizes[:ns] + 1) // 2).sum())
        reps, el = cpu_repeat(lambda: O.mul_csr(so, sa, sb, 0, cores), 3.0)
        cpu = {"value": round(reps * su / el / 1e9, 3), "unit": "G useful u32-MAC/s", "cores": cores,
               "kind": "port", "sample": f"{reps} passes of the oracle ragged multiply over the first {ns} "
                                         f"instances, {el:.1f} s wall"}
        if O.gmp_available():
            reps2, el2 = cpu_repeat(lambda: O.gmp_csr(1, off, a, b, cores), 3.0)
            cpu["gmp"] = {"value": round(reps2 * useful / el2 / 1e9, 3), "unit": "G useful u32-MAC/s",
                          "cores": cores, "kind": "gmp",
                          "sample": f"{reps2} passes of mpz_mul over all {N} instances (low n_i digits kept), "
                                    f"{el2:.1f} s wall"}
    return {"workload": "2^17 instances, digits log-uniform 8..2048, 10% carry chains, 5% zero operands",
            "total_digits": total, "useful_macs": useful,
            "add_gbs": round(3.0 * total * 4 / (t_add * 1e-3) / 1e9, 2), "add_ms": round(t_add, 4),
            "mul_ms": round(t_mul, 4), "mul_gu32ops": round(rag / 1e9, 2),
            "addmul_one_launch_ms": round(t_both, 4),
            "padded_plain_mul_ms": round(t_pad, 4), "padded_plain_gu32ops": round(pad / 1e9, 2),
            "speedup_vs_padded": round(t_pad / t_mul, 3), "plan": pinfo, "parity_sample": bool(ok and ok_pad),
            "cpu_baseline": cpu}


def bench_div(ctx, W, O, args, mac_peak):
    """16384 divisions of 4096-u64-limb numerators by 2048-u64-limb divisors
    (8192 / 4096 u32 digits, divisor top limb nonzero): one hipGraph per batch."""
    n, mu, mv = 16384, 8192, 4096
    u, v = W.div_pairs(n, mu, mv, seed=41)
    du, dv, dq, dr = ctx.upload(u), ctx.upload(v), ctx.alloc(u.nbytes), ctx.alloc(v.nbytes)
    ctx.check(ctx.divmod_dev(n, mu, mv, du, dv, dq, dr), "divmod")      # builds and captures the graph
    for _ in range(int(WARM_MS / 45) + 1):                               # sustained clock (see timed)
        ctx.check(ctx.divmod_dev(n, mu, mv, du, dv, dq, dr), "divmod")
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.check(ctx.divmod_dev(n, mu, mv, du, dv, dq, dr), "divmod")
    t = (time.perf_counter() - t0) / reps
    info = ctx.div_info()
    q = dq.download(np.uint32, n * mu).reshape(n, mu)
    r = dr.download(np.uint32, n * mv).reshape(n, mv)
    pick = np.random.default_rng(2).choice(n, 16, replace=False)
    gq, gr = (O.Gmp().batch("div", u[pick], v[pick]) if O.Gmp.available() else O.py_div(u[pick], v[pick]))
    ok = bool((q[pick] == gq).all() and (r[pick] == gr).all())
    for buf in (du, dv, dq, dr):
        ctx._release(buf)
    log(f"div: {t * 1e3:.1f} ms per batch, {n / t:.0f} div/s, P={info['P']}, steps={info['newton_steps']}, "
        f"quadratic {info['valu_macs'] / 1e9:.0f} G + transform {info['transform_macs'] / 1e9:.0f} G digit products")
    # roofline: every long product of the pipeline counted as quadratic
    # digit products (the transform products at their quadratic equivalent)
    # against the measured v_mad_u64_u32 + carry pair rate -- the time the
    # schoolbook pipeline would take at the VALU peak, against the batch time
    # (frac > 1 would mean the transforms beat the schoolbook at its peak);
    # beside it the part still done by the quadratic kernels
    vm, tm = info["valu_macs"], info["transform_macs"]
    allm = info["macs"]
    res = {"workload": "16384 x (8192 / 4096 u32 digits) divisions, hipGraph", "ms_per_batch": round(t * 1e3, 2),
           "divisions_per_s": round(n / t, 1), "macs": allm, "valu_macs": vm, "transform_macs": tm,
           "roofline": {"bound": "valu (all products at their quadratic digit-product count)",
                        "achieved": round(allm / t / 1e12, 4), "peak": round(mac_peak / 1e12, 4),
                        "unit": "T digit products/s (quadratic equivalent) per second of batch time",
                        "frac": round(allm / t / mac_peak, 4),
                        "schoolbook_at_peak_ms": round(allm / mac_peak * 1e3, 2),
                        "quadratic_kernels_share_at_peak_ms": round(vm / mac_peak * 1e3, 2)},
           "precision_digits": info["P"], "newton_steps": info["newton_steps"], "parity_sample_gmp": ok}
    if not args.no_cpu:
        cores = cpu_cores()
        k = cores
        t0 = time.perf_counter()
        O.div(u[:k], v[:k], cores)
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(k / el, 3), "unit": "divisions/s", "cores": cores, "kind": "port",
                               "sample": f"{k} divisions (one per core) with the restatement of prototypes/div.c"}
        if O.gmp_available():
            k = 2048
            t0 = time.perf_counter()
            gq, gr = O.gmp_div(u[:k], v[:k], cores)
            el = time.perf_counter() - t0
            res["cpu_gmp"] = {"value": round(k / el, 3), "unit": "divisions/s", "cores": cores, "kind": "gmp",
                              "sample": f"{k} divisions with GMP mpz_tdiv_qr (the check of prototypes/div.c:"
                                        f"470-486), OpenMP over instances, {el:.1f} s wall"}
            res["parity_sample_gmp_all"] = bool((q[:k] == gq).all() and (r[:k] == gr).all())
    if PROFILED:
        res["small_batches"] = {"skipped": "profiled run: the one-launch division is not run under rocprofv3"}
        log("div small: skipped under the profiler")
    else:
        res["small_batches"] = bench_div_small(ctx, W, O)
    return res


def bench_div_small(ctx, W, O):
    """Small division batches (a serving round's shapes): bi_divmod_u32_dev
    per call, host-synchronous, median of 20 calls, on the one-launch path
    (a workgroup per division, up to 256 instances) and forced onto the
    batched pipeline for comparison."""
    out = {"workload": "16 divisions per call, bi_divmod_u32_dev (host-synchronous), median of 20 calls"}
    ok = True
    old = ctx.tune_get_div_one()
    try:
        for mu, mv in ((4096, 2048), (1024, 512), (64, 32)):
            n = 16
            u, v = W.div_pairs(n, mu, mv, seed=mu + 3)
            du, dv, dq, dr = ctx.upload(u), ctx.upload(v), ctx.alloc(u.nbytes), ctx.alloc(v.nbytes)
            row = {}
            for name, lim in (("one_launch_ms", 256), ("batched_pipeline_ms", 0)):
                ctx.tune_set_div_one(lim)
                for _ in range(3):
                    ctx.check(ctx.divmod_dev(n, mu, mv, du, dv, dq, dr), "divmod")
                ts = []
                for _ in range(20):
                    t0 = time.perf_counter()
                    ctx.check(ctx.divmod_dev(n, mu, mv, du, dv, dq, dr), "divmod")
                    ts.append(time.perf_counter() - t0)
                row[name] = round(float(np.median(ts)) * 1e3, 4)
                q = dq.download(np.uint32, n * mu).reshape(n, mu)
                r = dr.download(np.uint32, n * mv).reshape(n, mv)
                gq, gr = O.py_div(u[:4], v[:4])
                ok &= bool((q[:4] == gq).all() and (r[:4] == gr).all())
            for buf in (du, dv, dq, dr):
                ctx._release(buf)
            out[f"{mu}/{mv}"] = row
            log(f"div small: 16 x {mu}/{mv}: one launch {row['one_launch_ms']:.3f} ms, "
                f"batched pipeline {row['batched_pipeline_ms']:.3f} ms per call")
    finally:
        ctx.tune_set_div_one(old)
    out["parity_sample"] = ok
    return out


def bench_sweep(ctx, F, O, args, mac_peak):
    """The reference's size sweep (cuda/main.cu:476-513; published as
    thesis/sections/performance.tex:181-217,307-343): total work fixed at 2^32
    bits per operand batch, n = 2^27 / m instances of m digits, uniform random
    digits.  Add for m = 16..8192 in GB/s (12 B per digit) against HBM; the
    multiply for m = 16..4096 with the top m/2 digits of both operands zero,
    m^2/4 digit products per instance, against the slower of HBM and its
    engine's product rate (multiply-adds: the measured mad+carry rate; matrix
    cores: (2m)^2 byte products per instance at the int8 peak).
    Each add size class is tuned first (its own class; the multiply keeps the
    shape tuned at m = 512).  Also the small plain configuration (1024 x 64
    digits): add and multiply latency in us."""
    import ctypes
    L = F.lib()
    total = 1 << 27

    def rows(buf, m, idx):
        """instances idx of an [n][m] device batch, copied back one by one"""
        out = np.empty((len(idx), m), np.uint32)
        for j, i in enumerate(idx):
            ctx.check(L.bi_copy_d2h(ctx.h, out[j].ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_void_p(buf.ptr.value + 4 * int(i) * m), 4 * m), "bi_copy_d2h")
        return out

    rng = np.random.default_rng(71)
    a = np.frombuffer(rng.bytes(4 * total), np.uint32)
    b = np.frombuffer(rng.bytes(4 * total), np.uint32)
    da, db, dr = ctx.upload(a), ctx.upload(b), ctx.alloc(4 * total)
    za, zb = ctx.alloc(4 * total), ctx.alloc(4 * total)
    adds, muls, ok = [], [], True
    for m in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192):
        n = total // m
        if not args.no_tune:
            ctx.tune(m, n * m * 4, 0)
        t = timed(ctx, lambda: ctx.add_dev(m, n, da, db, dr), 20, 3)
        gbs = 12.0 * total / (t * 1e-3) / 1e9
        pick = np.array([0, 1, n // 2, n - 1])
        ok &= bool((rows(dr, m, pick) == O.add(a.reshape(n, m)[pick], b.reshape(n, m)[pick])).all())
        adds.append({"m": m, "n": n, "us": round(t * 1e3, 2), "gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)})
        if m > 4096:
            continue
        # operands with zero top halves: shift up then down by m/2 digits
        for src, dst in ((da, za), (db, zb)):
            ctx.check(L.bi_shift_u32_dev(ctx.h, m, n, m // 2, src.ptr, dr.ptr), "bi_shift_u32_dev")
            ctx.check(L.bi_shift_u32_dev(ctx.h, m, n, -(m // 2), dr.ptr, dst.ptr), "bi_shift_u32_dev")
        tm = timed(ctx, lambda: ctx.mul_dev(m, n, za, zb, dr), 5, 1, WARM_MS)
        macs = float(n) * m * m / 4
        rate = macs / (tm * 1e-3)
        # roofline of this size: the slower of streaming its 12 B per digit
        # through HBM and the products at the engine's rate (multiply-adds at
        # the measured pair rate; (2m)^2 byte products per pair at the int8
        # matrix peak)
        engine = mul_engine(ctx, m)
        hbm_ms = 12.0 * total / (HBM_PEAK_GBS * 1e9) * 1e3
        comp_ms = (float(n) * (2.0 * m) ** 2 / MFMA_I8_PEAK_MACS if engine == "mfma" else macs / mac_peak) * 1e3
        bound_ms = max(hbm_ms, comp_ms)
        xa, xb = rows(za, m, pick), rows(zb, m, pick)
        ok &= bool((xa[:, m // 2:] == 0).all() and (xa[:, : m // 2] == a.reshape(n, m)[pick, : m // 2]).all()
                   and (rows(dr, m, pick) == O.mul(xa, xb)).all())
        muls.append({"m": m, "n": n, "ms": round(tm, 4), "gmacs": round(rate / 1e9, 1),
                     "frac": round(rate / mac_peak, 4),
                     "engine": engine,
                     "bound": "hbm" if hbm_ms >= comp_ms else ("mfma" if engine == "mfma" else "valu"),
                     "roofline_frac": round(bound_ms / tm, 4)})
    for x in (da, db, dr, za, zb):
        ctx._release(x)
    # the small plain configuration: 1024 pairs of 64 digits
    n1, m1 = 1024, 64
    a1 = a[: n1 * m1].reshape(n1, m1).copy()
    b1 = b[: n1 * m1].reshape(n1, m1).copy()
    c1 = ctx.upload(a1), ctx.upload(b1), ctx.alloc(a1.nbytes)
    t_add = timed(ctx, lambda: ctx.add_dev(m1, n1, *c1), 200, 10)
    ok &= bool((c1[2].download(np.uint32, n1 * m1).reshape(n1, m1) == O.add(a1, b1)).all())
    a1[:, m1 // 2:] = 0
    b1[:, m1 // 2:] = 0
    c1[0].upload(a1)
    c1[1].upload(b1)
    t_mul = timed(ctx, lambda: ctx.mul_dev(m1, n1, *c1), 200, 10)
    ok &= bool((c1[2].download(np.uint32, n1 * m1).reshape(n1, m1) == O.mul(a1, b1)).all())
    for x in c1:
        ctx._release(x)
    log("sweep add GB/s:", {x["m"]: x["gbs"] for x in adds})
    log("sweep mul G MAC/s:", {x["m"]: x["gmacs"] for x in muls})
    log(f"1024 x 64: add {t_add * 1e3:.2f} us, mul {t_mul * 1e3:.2f} us")
    return {"workload": "2^32 bits per operand batch: n = 2^27/m instances of m u32 digits, uniform digits "
                        "(multiply: top m/2 digits zero, m^2/4 digit products each)",
            "add": adds, "mul": muls,
            "small_1024x64": {"add_us": round(t_add * 1e3, 3), "mul_us": round(t_mul * 1e3, 3),
                              "add_gbs": round(12.0 * n1 * m1 / (t_add * 1e-3) / 1e9, 1),
                              "mul_gmacs": round(n1 * m1 * m1 / 4 / (t_mul * 1e-3) / 1e9, 1)},
            "parity_sample": ok}


def bench_serve(ctx, F, W, O, args):
    """Open-loop Poisson stream: 50% add / 40% mul / 10% div, operand digits
    log-uniform in [16, 4096].  Capacity is the best of three bursts (every
    request submitted at once); the measured stream runs at 60% of it.  A
    request's latency runs from its scheduled arrival to its result in host
    memory; throughput and capacity count requests per second up to the last
    completion (arrival + latency), not the trace runner's copy-out after it."""
    rng = np.random.default_rng(51)
    nreq = 4000
    sizes = W.ragged_sizes(nreq, 16, 4096, seed=52)
    ops, As, Bs = [], [], []
    for i in range(nreq):
        x = rng.random()
        op = F.OP_ADD if x < 0.5 else (F.OP_MUL if x < 0.9 else F.OP_DIV)
        mdig = int(sizes[i])
        a = rng.integers(0, 1 << 32, mdig, dtype=np.uint64).astype(np.uint32)
        if op == F.OP_DIV:
            bb = rng.integers(0, 1 << 32, max(1, mdig // 2), dtype=np.uint64).astype(np.uint32)
            bb[-1] |= 1
        else:
            bb = rng.integers(0, 1 << 32, mdig, dtype=np.uint64).astype(np.uint32)
        ops.append(op)
        As.append(a)
        Bs.append(bb)
    srv = F.Server(ctx, arena_bytes=1 << 30, slots=2)   # 2 batch streams + 2 division streams: 4 queues
    try:
        cap = 0.0
        for _ in range(3):
            _, lat0, st0, wall0 = srv.run_trace(ops, As, Bs, np.zeros(nreq))
            cap = max(cap, nreq / float(lat0.max()))
        rate = 0.6 * cap
        arr = np.cumsum(rng.exponential(1.0 / rate, nreq))
        arr -= arr[0]
        outs, lat, st, wall = srv.run_trace(ops, As, Bs, arr)
        span = float((arr + lat).max())
        stats = srv.stats()
    finally:
        srv.close()
    ok = bool((st == 0).all())
    for i in np.random.default_rng(3).choice(nreq, 64, replace=False):
        a, bb = As[i][None], Bs[i][None]
        if ops[i] == F.OP_ADD:
            ok &= bool((outs[i] == O.add(a, bb)[0]).all())
        elif ops[i] == F.OP_MUL:
            ok &= bool((outs[i] == O.mul(a, bb)[0]).all())
        else:
            q, r = O.py_div(a, bb)
            ok &= bool((outs[i][0] == q[0]).all() and (outs[i][1] == r[0]).all())
    log(f"serve: capacity {cap:.0f} req/s, at {rate:.0f} req/s p50 {np.percentile(lat, 50) * 1e3:.2f} ms "
        f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms")
    return {"workload": "open-loop Poisson, 50% add / 40% mul / 10% div, digits log-uniform 16..4096",
            "requests": nreq, "capacity_req_s": round(cap, 1), "rate_req_s": round(rate, 1),
            "throughput_req_s": round(nreq / span, 1), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "batches": stats["batches"],
            "max_batch": stats["max_batch"], "parity_sample": ok}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep default shapes (profiling runs)")
    ap.add_argument("--only", default="", help="comma list of mul,chain,helpers,pcie,ragged,div,full,serve,sweep (for profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every workload here runs on a single MI355X "
              "(instances are independent: N GPUs would each take a contiguous share of the batch).",
              file=sys.stderr)
        sys.exit(2)

    import bigint_ffi as F
    import bigint_workloads as W

    ctx = F.Context(0)
    info = ctx.device_info()
    log("device:", info)

    n, m = 131072, 512
    only = set(x for x in args.only.split(",") if x)
    want = (lambda k: not only or k in only)
    # autotuner: pick the add / |a-b| stream shapes, the multiply workgroup
    # shape and the ragged shapes on this GPU (interleaved median rounds; the
    # default stays unless a candidate is 2% faster)
    if not args.no_tune:
        ctx.tune(m, n * m * 4, 0)
        if want("helpers"):
            ctx.tune(m, n * m * 4, 3)
        if want("mul"):
            ctx.tune(m, 16384 * m * 4, 1)
        if want("ragged"):
            ctx.tune(2048, RAGGED_DIGITS * 4, 2)
    tuned = ctx.tune_get(m)
    tuned.update(ctx.tune_get_sub(m))
    tuned.update(ctx.tune_get_csr())
    log("tuned:", tuned)
    log("generating inputs ...")
    a, b = W.plain_pairs(n, m, seed=2024, chain_frac=0.05)
    da, db, dr = ctx.upload(a), ctx.upload(b), ctx.alloc(a.nbytes)

    add_ms = timed(ctx, lambda: ctx.add_dev(m, n, da, db, dr), args.steps, args.warmup)
    add_bytes = 3.0 * n * m * 4
    add_gbs = add_bytes / (add_ms * 1e-3) / 1e9
    log(f"add: {add_ms * 1e3:.1f} us/launch, {add_gbs:.1f} GB/s")

    # parity spot check of the timed batch (full parity lives in tests/)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    idx = np.random.default_rng(0).choice(n, 512, replace=False)
    add_ok = bool((r[idx] == O.add(a[idx], b[idx])).all())

    # plain copy of the same bytes: the practical HBM ceiling on this box
    cbytes = 2 * n * m * 4
    csrc, cdst = da, dr
    copy_ms = timed(ctx, lambda: ctx.copy_kernel(cdst, csrc, n * m * 4), 20, 3)
    copy_gbs = cbytes / (copy_ms * 1e-3) / 1e9
    log(f"copy: {copy_gbs:.1f} GB/s")

    # multiply, top halves zero
    mul_ms, mul_rate, mul_ok, mul_cpu, mac_peak = float("nan"), float("nan"), None, None, ctx.mac_rate()
    mul_info = {}
    if want("mul"):
        mul_ms, mul_rate, mul_ok, mul_cpu, mul_info = bench_mul(ctx, W, O, args, n, m, da, db, dr, idx)
    log(f"mul: {mul_ms * 1e3:.1f} us/launch, {mul_rate / 1e9:.1f} G MAC/s, mad peak {mac_peak / 1e9:.1f} G/s")
    chain = bench_chain(ctx, O, args, n, m, a, b, da, db, dr, idx) if want("chain") else None
    helpers = bench_helpers(ctx, F, O, n, m, a, b, da, db, dr, idx) if want("helpers") else None
    pcie = bench_pcie(ctx, W, O, a, b) if want("pcie") else None
    for buf in (da, db, dr):
        ctx._release(buf)
    ragged = bench_ragged(ctx, W, O, args) if want("ragged") else None
    div = bench_div(ctx, W, O, args, mac_peak) if want("div") else None
    full = bench_full(ctx, O, args) if want("full") else None
    serve = bench_serve(ctx, F, W, O, args) if want("serve") and not PROFILED else None
    if want("serve") and PROFILED:
        log("serve: skipped under the profiler (its division rounds take the one-launch division)")
    sweep = bench_sweep(ctx, F, O, args, mac_peak) if want("sweep") else None

    traffic = load_traffic().get("add_131072x512", {}).get("traffic_bytes")
    out = {
        "metric": METRIC,
        "value": round(add_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(add_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: seeded uniform u32 digits, 5% of pairs all-ones + 1",
        "config": {"workload": "plain batched add, 131072 pairs x 512 u32 digits (16384-bit), 5% carry chains",
                   "n": n, "m": m, "limb": "u32 (u64 view: m=256, same bytes and kernel)"},
        "roofline": {"bound": "hbm", "achieved": round(add_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(add_gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "copy_kernel_gbs": round(copy_gbs, 2),
        "parity_sample": {"add": add_ok, "mul": mul_ok},
        "mul": {"workload": "plain batched multiply, 131072 pairs x 512 u32 digits, top 256 digits zero",
                "value": rnd(mul_rate / 1e9, 2), "unit": "G u32-MAC/s (m^2/4 = 65536 digit products per pair)",
                "ms": rnd(mul_ms, 4), "engine": mul_info.get("engine"),
                "roofline": mul_info.get("roofline"),
                "matrix_cores_optin": mul_info.get("matrix_cores_optin"),
                "cpu_baseline": mul_cpu},
        "tuned": tuned,
        "ragged": ragged,
        "div": div,
        "full_products": full,
        "serve": serve,
        "sweep": sweep,
        "chain": chain,
        "helpers": helpers,
        "pcie_inclusive_add": pcie,
        "device": info,
        "host": host_info(),
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_add(a, b, args.cpu_seconds)
    ctx.close()
    failed = parity_failures(out)
    if failed:
        # a timed run never reports throughput for wrong results
        out["value"] = None
        out["roofline"]["achieved"] = None
        out["roofline"]["frac"] = None
        out["parity_failed"] = failed
    print(json.dumps(out), flush=True)
    if failed:
        log("PARITY FAILED:", failed)
        sys.exit(1)


def parity_failures(out):
    """Names of every parity spot check in the line that is not true."""
    bad = []

    def walk(x, path):
        if isinstance(x, dict):
            for k, v in x.items():
                if k.startswith("parity_sample"):
                    if isinstance(v, dict):
                        bad.extend(f"{path}{k}.{kk}" for kk, vv in v.items() if vv is False)
                    elif v is False:
                        bad.append(path + k)
                else:
                    walk(v, f"{path}{k}.")

    walk(out, "")
    return bad


if __name__ == "__main__":
    main()
