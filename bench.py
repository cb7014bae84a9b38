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
METRIC = "add HBM GB/s (% peak); mul Gu32ops/s; ragged-batch vs padded-plain speedup"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(ctx, fn, steps, warmup):
    for _ in range(warmup):
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


def cpu_baseline_add(a, b, seconds):
    """The CPU oracle (sequential C restatement, OpenMP over instances) on a
    bounded sample of the same batch; GB/s with the same byte formula."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    ns = min(len(a), 16384)
    sa, sb = np.ascontiguousarray(a[:ns]), np.ascontiguousarray(b[:ns])
    O.add(sa, sb, cores)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        O.add(sa, sb, cores)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    gbs = 3.0 * ns * sa.shape[1] * 4 * reps / el / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"{reps} passes of the oracle add over {ns} x {sa.shape[1]} digits (first {ns} pairs of the "
                      f"batch), {el:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every workload here runs on a single MI355X "
              "(instances are independent; see the sharding notes in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import bigint_ffi as F
    import bigint_workloads as W

    ctx = F.Context(0)
    info = ctx.device_info()
    log("device:", info)

    n, m = 131072, 512
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
    a2, b2 = W.plain_pairs(n, m, seed=2025, top_zero=True)
    da.upload(a2)
    db.upload(b2)
    mul_steps = max(5, args.steps // 5)
    mul_ms = timed(ctx, lambda: ctx.mul_dev(m, n, da, db, dr), mul_steps, 2)
    macs = n * (m * m / 4)
    mul_rate = macs / (mul_ms * 1e-3)
    r = dr.download(np.uint32, n * m).reshape(n, m)
    mul_ok = bool((r[idx] == O.mul(a2[idx], b2[idx])).all())
    mac_peak = ctx.mac_rate()
    log(f"mul: {mul_ms * 1e3:.1f} us/launch, {mul_rate / 1e9:.1f} G MAC/s, mad peak {mac_peak / 1e9:.1f} G/s")

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
                "value": round(mul_rate / 1e9, 2), "unit": "G u32-MAC/s (m^2/4 = 65536 digit products per pair)",
                "ms": round(mul_ms, 4),
                "roofline": {"bound": "valu (v_mad_u64_u32, no MFMA)", "achieved": round(mul_rate / 1e12, 4),
                             "peak": round(mac_peak / 1e12, 4), "unit": "T MAC/s (measured mad+carry rate)",
                             "frac": round(mul_rate / mac_peak, 4)}},
        "device": info,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_add(a, b, args.cpu_seconds)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
