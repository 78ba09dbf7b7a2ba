#!/usr/bin/env python3
"""Benchmark of the gfx950 NTT (one JSON line on stdout).

Headline: butterflies/s of a forward NTT over the 753-bit field MNT6-753 Fr
at n = 2^18 (24 x u32 Montgomery limbs per element, operands resident in
HBM), one transform per step.  The same line carries:
  roofline      the pass kernel against the integer-multiply (v_mad_u64_u32)
                peak, with its measured average launch time;
  cpu_baseline  the oracle's restatement of the reference host FFT
                (test/fft_host.h, log_cpus = 10 as test/main.cpp:71) on this
                host's cores, on a bounded sample;
  batched       256 ragged transforms (75% 753-bit / 25% 32-bit, log2 n in
                [10, 18], forward/inverse 50/50) as fused launches under a
                captured hipGraph, against looping the plain plans;
  serving       Poisson arrivals (70% 753-bit / 30% 32-bit, log2 n in
                [12, 20]) through the continuous-batching server against
                one-at-a-time plain dispatch: p50/p99 latency, throughput;
  fp32          the 32-bit field at n = 2^26 against HBM bandwidth;
  configs       every configured size as a plain transform;
  per_transform median/min/max of individually timed transforms;
  end_to_end    the host-buffer entry point (best_fft) including its copies;
  inner_product the reference's multiexp over field elements at 2^18.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]
"""
import argparse
import ctypes
import json
import os
import random
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X figures: 256 CUs at 2.4 GHz; v_mad_u64_u32 issues at 64 lanes/clk/CU
# (measured 34.7e12/s in a dependent-chain microbenchmark); HBM3E 8 TB/s.
CUS, CLOCK_HZ, MAD_LANES_PER_CLK = 256, 2.4e9, 64
MAD_PEAK = CUS * CLOCK_HZ * MAD_LANES_PER_CLK  # 3.93e13 v_mad_u64_u32 / s
HBM_PEAK = 8.0e12
# one 753-bit butterfly = one Montgomery product of 24-limb operands:
# 2 * 24^2 + 24 = 1176 32x32->64 multiply-adds (add/sub not counted)
MADS_PER_BFLY_753 = 1176
REF_DERIVED_BFLY_S = 1.10e5  # reference README.md:46, 2^18 in 21.42 s, (n/2) log2 n / t
# untimed GPU time before each timed region of plain transforms (engine
# clock ramp under sustained load); --settle-ms overrides it
SETTLE_MS = 100.0


def set_settle_ms(v):
    global SETTLE_MS
    SETTLE_MS = v


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bfly(log_n):
    return (1 << log_n) // 2 * log_n


def inputs(field, log_n, seed):
    import gpusnarks_amd as g
    from gpusnarks_amd import synth
    return synth.uniform_residues(seed, 1 << log_n, g.MODULI[field], g.NLIMBS[field])


def time_plan(field, log_n, steps, warmup, inverse=False, want_output=False, settle_ms=None, info=None):
    """Device-resident transform, timed with HIP events on its own stream:
    `warmup` untimed transforms, then untimed transforms until `settle_ms` of
    GPU time has passed (the engine clock rises over the first ~25 ms of
    sustained load: MNT6-753 Fr 2^18 runs at 0.150 ms per transform after 5
    warm-up transforms and at 0.130 ms after 200), then `steps` timed back to
    back.  With `info`, the same `steps` are also timed right after the
    `warmup` transforms, before the settling, and reported there."""
    import gpusnarks_amd as g
    settle_ms = SETTLE_MS if settle_ms is None else settle_ms
    x = inputs(field, log_n, 0x5EED0000 + log_n)
    plan = g.Plan(field, log_n, inverse=inverse)
    d_in = g.DeviceArray.from_numpy(x)
    d_out = g.DeviceArray(x.nbytes)
    s = g.Stream()
    for _ in range(warmup):
        plan.execute(d_in.ptr, d_out.ptr, s.handle)
    s.synchronize()
    e0, e1 = g.Event(), g.Event()

    def timed(k):
        e0.record(s.handle)
        for _ in range(k):
            plan.execute(d_in.ptr, d_out.ptr, s.handle)
        e1.record(s.handle)
        e1.synchronize()
        return g.Event.elapsed_ms(e0, e1)

    cold = timed(steps) if info is not None else None
    spent, extra = 0.0, 0
    while spent < settle_ms:
        spent += timed(max(steps, 8))
        extra += max(steps, 8)
    ms = timed(steps)
    if info is not None:
        info.update({"warmup_steps": warmup, "settle_steps": extra + steps, "settle_gpu_ms": spent + cold,
                     "ms_per_step_after_warmup_only": cold / steps})
    # spot check against a host round trip of the same input
    out = d_out.download(x.shape)
    ref = g.best_fft(x, None, field, inverse=inverse)
    assert np.array_equal(out, ref), "device-resident result differs from the host path"
    if want_output:
        return ms / steps, plan.num_passes, out
    return ms / steps, plan.num_passes


def step_times(field, log_n, runs=20, warmup=3):
    """Per-transform durations (HIP events around each execute, operands in
    HBM): median and min over `runs` transforms, after `warmup` transforms
    and SETTLE_MS of untimed GPU time."""
    import gpusnarks_amd as g
    x = inputs(field, log_n, 0x5EED0000 + log_n)
    plan = g.Plan(field, log_n)
    d_in = g.DeviceArray.from_numpy(x)
    d_out = g.DeviceArray(x.nbytes)
    s = g.Stream()
    for _ in range(warmup):
        plan.execute(d_in.ptr, d_out.ptr, s.handle)
    e0, e1 = g.Event(), g.Event()
    spent = 0.0
    while spent < SETTLE_MS:
        e0.record(s.handle)
        for _ in range(runs):
            plan.execute(d_in.ptr, d_out.ptr, s.handle)
        e1.record(s.handle)
        e1.synchronize()
        spent += g.Event.elapsed_ms(e0, e1)
    ev = [g.Event() for _ in range(runs + 1)]
    ev[0].record(s.handle)
    for i in range(runs):
        plan.execute(d_in.ptr, d_out.ptr, s.handle)
        ev[i + 1].record(s.handle)
    ev[-1].synchronize()
    t = sorted(g.Event.elapsed_ms(ev[i], ev[i + 1]) for i in range(runs))
    return {"runs": runs, "median_ms": t[runs // 2], "min_ms": t[0], "max_ms": t[-1]}


def end_to_end(field, log_n, runs=5):
    """The host-buffer entry point (gsn_fft_host, the best_fft call): input
    copied to HBM, transformed, copied back; wall clock, median of `runs`."""
    import gpusnarks_amd as g
    x = inputs(field, log_n, 0x5EED0100 + log_n)
    L = g.lib()
    ptr = x.ctypes.data_as(ctypes.c_void_p)
    g.check(L.gsn_fft_host(field, ptr, x.shape[0], None, 0))  # plan, twiddles, staging buffers
    t = []
    for _ in range(runs):
        t0 = time.perf_counter()
        g.check(L.gsn_fft_host(field, ptr, x.shape[0], None, 0))
        t.append((time.perf_counter() - t0) * 1e3)
    t.sort()
    ms = t[runs // 2]
    return {"runs": runs, "median_ms": ms, "bytes_each_way": int(x.nbytes),
            "butterflies_per_s": bfly(log_n) / (ms * 1e-3),
            "note": "pageable host buffer in, transform, host buffer out (PCIe-inclusive); not the headline"}


def bench_inner_product(log_n=18, reps=20):
    """The reference's multiexp over Scalar x Scalar (test/main.cpp:89-132,
    n = 2^18, MNT4-753 Fr) on device buffers: sum_i a_i b_i."""
    import gpusnarks_amd as g
    field = g.FIELD_MNT4753_FR
    a = inputs(field, log_n, 0x5EED0020)
    b = inputs(field, log_n, 0x5EED0021)
    da, db, dout = g.DeviceArray.from_numpy(a), g.DeviceArray.from_numpy(b), g.DeviceArray(96)
    s = g.Stream()
    g.inner_product(field, dout.ptr, da.ptr, db.ptr, a.shape[0], s.handle)
    e0, e1 = g.Event(), g.Event()
    e0.record(s.handle)
    for _ in range(reps):
        g.inner_product(field, dout.ptr, da.ptr, db.ptr, a.shape[0], s.handle)
    e1.record(s.handle)
    e1.synchronize()
    ms = g.Event.elapsed_ms(e0, e1) / reps
    return {"field": "MNT4-753 Fr", "n": a.shape[0], "ms": ms, "products_per_s": a.shape[0] / (ms * 1e-3),
            "note": "general 753-bit Montgomery products (1431 multiply-adds each) plus a two-level sum"}


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


_NATIVE_ORACLE = []


def native_oracle():
    """The oracle compiled for this host's CPU (gcc -O3 -march=native
    -fopenmp) into a temporary directory, as the CPU baseline asks; the
    shipped oracle/liboracle.so (built without -march) if that fails."""
    import subprocess
    import tempfile
    from tests.oracle_lib import ORACLE_SO
    if not _NATIVE_ORACLE:
        path, flags = ORACLE_SO, "oracle/liboracle.so (-O3 -fopenmp)"
        try:
            d = tempfile.mkdtemp(prefix="gsn_oracle_")
            so = os.path.join(d, "liboracle_native.so")
            r = subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-o", so,
                                os.path.join(ROOT, "oracle", "oracle.c")], capture_output=True, timeout=120)
            if r.returncode == 0:
                path, flags = so, "gcc -O3 -march=native -fopenmp (built on this host)"
        except (OSError, subprocess.SubprocessError):
            pass
        _NATIVE_ORACLE.append((path, flags))
    return _NATIVE_ORACLE[0]


def cpu_full_run(log_n, threads, field, gpu_out):
    """The reference harness's host path in full (test/main.cpp:66-84): the
    oracle's _basic_parallel_radix2_FFT_inner with log_cpus = 10 on `threads`
    cores, wall clock, and its output compared element for element with the
    GPU's transform of the same input."""
    from tests.oracle_lib import Oracle
    import gpusnarks_amd as g
    so, build_flags = native_oracle()
    orc = Oracle(so)
    x = inputs(field, log_n, 0x5EED0000 + log_n)
    om = g.root_of_unity(field, log_n)
    t0 = time.perf_counter()
    y = orc.parallel_fft(field, x, om, 10, threads=threads)
    sec = time.perf_counter() - t0
    return {"log2_n": log_n, "seconds": sec, "butterflies_per_s": bfly(log_n) / sec, "cores": threads,
            "bit_exact_vs_gpu": bool(np.array_equal(y, gpu_out)), "build": build_flags,
            "note": "full run, not a sample: phase 1 (1024 partial sums), 1024 sub-transforms, scatter"}


def cpu_baseline(log_n, threads, field=None, serial=True, j_sample=0, gpu_out=None):
    """Oracle (test/fft_host.h restated) on the host: the reference
    harness's parallel radix-2 FFT with log_cpus = 10, bounded sample.  With
    `gpu_out`, the serial transform's output (computed in full) is compared
    element for element with the GPU's (test/main.cpp:80-84)."""
    import ctypes
    from tests.oracle_lib import Oracle
    import gpusnarks_amd as g
    so, build_flags = native_oracle()
    orc = Oracle(so)
    field = g.FIELD_MNT6753_FR if field is None else field
    fname = {g.FIELD_FP32: "Fp32", g.FIELD_MNT4753_FR: "MNT4-753 Fr", g.FIELD_MNT6753_FR: "MNT6-753 Fr"}[field]
    n = 1 << log_n
    x = inputs(field, log_n, 0x5EED0000 + log_n)
    om = g.root_of_unity(field, log_n)
    log_cpus = 10
    rows = n >> log_cpus
    # phase 1 (fft_host.h:76-102): 1024 partial sums, each O(n) products;
    # time a slice of the j range and scale
    # a multiple of the thread count; the headline samples half of phase 1
    j_sample = j_sample or max(threads, 16)
    tmp = np.zeros((j_sample * rows, g.NLIMBS[field]), np.uint32)
    f = orc.field(field)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    t0 = time.perf_counter()
    orc.lib.or_parallel_phase1_sample(f, x.ctypes.data_as(u32p), n, om.ctypes.data_as(u32p), log_cpus, 0,
                                      j_sample, tmp.ctypes.data_as(u32p), threads)
    raw_p1 = time.perf_counter() - t0
    t_p1 = raw_p1 * ((1 << log_cpus) / j_sample)
    # phase 2 (fft_host.h:103-108): 1024 serial FFTs of n/1024, one core each
    w_c = orc.pow(field, om[None], [1 << log_cpus])[0]
    p2 = min(threads, j_sample)  # sub-transforms timed, one after another on one core
    t0 = time.perf_counter()
    for i in range(p2):
        orc.serial_fft(field, tmp[i * rows:(i + 1) * rows], w_c)
    raw_p2 = time.perf_counter() - t0
    t_p2 = raw_p2 / p2 * (1 << log_cpus) / threads
    total = t_p1 + t_p2
    out = {
        "value": bfly(log_n) / total, "unit": "butterflies/s", "cores": threads, "kind": "port",
        "sample": "oracle _basic_parallel_radix2_FFT_inner, %s n=2^%d, log_cpus=10: phase 1 timed "
                  "for j in [0,%d) of 1024 on %d threads and scaled x%.0f; phase 2 timed for %d of 1024 "
                  "size-%d FFTs on one core and scaled to %d threads" % (fname, log_n, j_sample, threads,
                                                                         1024 / j_sample, p2, rows, threads),
        "seconds_extrapolated": total,
        "sample_cpu_seconds": raw_p1 + raw_p2,
    }
    out.update(cpu_info())
    out["build"] = build_flags
    if serial:
        # and the O(n log n) serial transform on one core, in full
        t0 = time.perf_counter()
        y = orc.serial_fft(field, x, om)
        t_serial = time.perf_counter() - t0
        out["serial_radix2_1core_butterflies_per_s"] = bfly(log_n) / t_serial
        out["serial_radix2_1core_seconds"] = t_serial
        if gpu_out is not None:
            out["bit_exact_vs_gpu"] = bool(np.array_equal(y, gpu_out))
    return out


def bench_configs(threads, skip_cpu):
    """Every size the configurations name, plain transforms with operands in
    HBM: butterflies/s, elements/s and the roofline fraction of each."""
    import gpusnarks_amd as g
    rows = []
    for name, field, log_n, inv in [("fp32", g.FIELD_FP32, 16, False), ("fp32", g.FIELD_FP32, 20, False),
                                    ("fp32", g.FIELD_FP32, 24, False), ("fp32", g.FIELD_FP32, 26, False),
                                    ("fp32", g.FIELD_FP32, 28, False), ("mnt4", g.FIELD_MNT4753_FR, 15, False),
                                    ("mnt4", g.FIELD_MNT4753_FR, 15, True), ("mnt6", g.FIELD_MNT6753_FR, 16, False),
                                    ("mnt6", g.FIELD_MNT6753_FR, 16, True), ("mnt6", g.FIELD_MNT6753_FR, 20, False)]:
        steps = 20 if log_n <= 26 else 5
        ms, passes = time_plan(field, log_n, steps, 2, inverse=inv)
        r = {"field": name, "log2_n": log_n, "inverse": inv, "ms": ms, "passes": passes,
             "butterflies_per_s": bfly(log_n) / (ms * 1e-3), "elements_per_s": (1 << log_n) / (ms * 1e-3)}
        if field == g.FIELD_FP32:
            alg = 8 * (1 << log_n) * ((log_n + 14) // 15)
            r["algorithmic_GBps"] = alg / (ms * 1e-3) / 1e9
            r["frac_of_8TBps"] = alg / (ms * 1e-3) / HBM_PEAK
        else:
            r["frac_of_mad_peak"] = bfly(log_n) * MADS_PER_BFLY_753 / (ms * 1e-3) / MAD_PEAK
        rows.append(r)
        log("config %s 2^%d%s: %.4f ms" % (name, log_n, " inverse" if inv else "", ms))
    # MNT4-753 Fr at 2^16, one size past its 2-adicity, over the extended
    # domain H u g*H (two 2^15 transforms and one elementwise kernel)
    for inv in (False, True):
        x = inputs(g.FIELD_MNT4753_FR, 16, 0x5EED0010)
        plan = g.Plan.extended(g.FIELD_MNT4753_FR, 16, inverse=inv)
        d_in, d_out = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
        s = g.Stream()
        for _ in range(2):
            plan.execute(d_in.ptr, d_out.ptr, s.handle)
        e0, e1 = g.Event(), g.Event()
        e0.record(s.handle)
        for _ in range(20):
            plan.execute(d_in.ptr, d_out.ptr, s.handle)
        e1.record(s.handle)
        e1.synchronize()
        ms = g.Event.elapsed_ms(e0, e1) / 20
        rows.append({"field": "mnt4", "log2_n": 16, "inverse": inv, "domain": "extended radix-2 (H u gH)",
                     "ms": ms, "passes": plan.num_passes, "butterflies_per_s": bfly(16) / (ms * 1e-3),
                     "elements_per_s": (1 << 16) / (ms * 1e-3),
                     "frac_of_mad_peak": bfly(16) * MADS_PER_BFLY_753 / (ms * 1e-3) / MAD_PEAK})
        log("config mnt4 2^16 extended%s: %.4f ms" % (" inverse" if inv else "", ms))
    # MNT4-753 Fr past 2^16: n = 525 * 2^15 over the mixed-radix plan (odd
    # DFTs of 525 points in LDS, 525 radix-2 transforms of 2^15, a transpose)
    import math
    n_mix = 525 << 15
    rng = np.random.default_rng(0x5EED0011)
    x = rng.integers(0, 1 << 32, size=(n_mix, 24), dtype=np.uint64).astype(np.uint32)
    x[:, 23] %= np.uint32(g.MODULI[g.FIELD_MNT4753_FR] >> (32 * 23))
    d = g.DeviceArray.from_numpy(x)
    for inv in (False, True):
        plan = g.Plan.mixed(g.FIELD_MNT4753_FR, n_mix, inverse=inv)
        s = g.Stream()
        plan.execute(d.ptr, d.ptr, s.handle)
        e0, e1 = g.Event(), g.Event()
        e0.record(s.handle)
        for _ in range(5):
            plan.execute(d.ptr, d.ptr, s.handle)
        e1.record(s.handle)
        e1.synchronize()
        ms = g.Event.elapsed_ms(e0, e1) / 5
        rows.append({"field": "mnt4", "n": n_mix, "log2_n": math.log2(n_mix), "inverse": inv,
                     "domain": "mixed radix 525 * 2^15", "ms": ms, "passes": plan.num_passes,
                     "elements_per_s": n_mix / (ms * 1e-3),
                     "radix2_equivalent_butterflies_per_s": n_mix / 2 * math.log2(n_mix) / (ms * 1e-3)})
        log("config mnt4 525*2^15 mixed%s: %.3f ms" % (" inverse" if inv else "", ms))
        plan.close()
    d.free()
    out = {"plain": rows}
    if not skip_cpu:
        out["cpu_fp32"] = [cpu_baseline(l, threads, g.FIELD_FP32, serial=(l == 20)) for l in (16, 20, 22)]
        out["cpu_mnt6_n16"] = cpu_baseline(16, threads, g.FIELD_MNT6753_FR, serial=False)
    return out


def bench_batched(reps=10, warmup=3):
    import gpusnarks_amd as g
    from gpusnarks_amd import synth
    items = synth.ragged_workload()
    bufs = []
    for k, (field, log_n, inv) in enumerate(items):
        x = inputs(field, log_n, synth.RAGGED_INPUT_SEED + k)
        bufs.append((g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes), x))
    total_bfly = sum(bfly(l) for _, l, _ in items)
    s = g.Stream()
    # plain: one plan execution per transform, looped on one stream
    plans = {}
    for field, log_n, inv in items:
        if (field, log_n, inv) not in plans:
            plans[(field, log_n, inv)] = g.Plan(field, log_n, inverse=inv)

    def run_plain():
        for (field, log_n, inv), (di, do, _) in zip(items, bufs):
            plans[(field, log_n, inv)].execute(di.ptr, do.ptr, s.handle)

    batch = g.Batch([(f, l, inv, di.ptr, do.ptr) for (f, l, inv), (di, do, _) in zip(items, bufs)], use_graph=True)
    run_batch = lambda: batch.launch(s.handle)

    def timed(fn):
        e0, e1 = g.Event(), g.Event()
        e0.record(s.handle)
        for _ in range(reps):
            fn()
        e1.record(s.handle)
        e1.synchronize()
        return g.Event.elapsed_ms(e0, e1) / reps

    # both forms warmed up (the engine clock settles over the first
    # milliseconds of load), then timed alternately, twice each
    for _ in range(warmup):
        run_plain()
        run_batch()
    s.synchronize()
    ms_plain = timed(run_plain)
    outs_plain = [do.download(x.shape) for _, do, x in bufs]
    ms_batch = timed(run_batch)
    ms_plain = (ms_plain + timed(run_plain)) / 2
    ms_batch = (ms_batch + timed(run_batch)) / 2
    same = all(np.array_equal(do.download(x.shape), ref) for (_, do, x), ref in zip(bufs, outs_plain))
    return {
        "transforms": len(items), "mix": "192 x MNT6-753 Fr + 64 x Fp32, log2 n uniform in [10,18], fwd/inv 50/50",
        "butterflies": total_bfly, "launches_batched": batch.num_launches,
        "launches_plain": sum(plans[k].num_passes for k in items),
        "ms_batched": ms_batch, "ms_plain_loop": ms_plain,
        "butterflies_per_s_batched": total_bfly / (ms_batch * 1e-3),
        "butterflies_per_s_plain_loop": total_bfly / (ms_plain * 1e-3),
        "speedup": ms_plain / ms_batch, "bit_exact_vs_plain": bool(same),
    }


def bench_batched_same(log_n=18, count=16, reps=10, warmup=3):
    """`count` MNT6-753 Fr transforms of one size as one fused schedule
    (hipGraph), against looping the plain plan over the same operands."""
    import gpusnarks_amd as g
    field = g.FIELD_MNT6753_FR
    bufs = []
    for k in range(count):
        x = inputs(field, log_n, 0xB18 + k)
        bufs.append((g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes), x))
    plan = g.Plan(field, log_n)
    s = g.Stream()
    batch = g.Batch([(field, log_n, False, di.ptr, do.ptr) for di, do, _ in bufs], use_graph=True)

    def timed(fn):
        e0, e1 = g.Event(), g.Event()
        e0.record(s.handle)
        for _ in range(reps):
            fn()
        e1.record(s.handle)
        e1.synchronize()
        return g.Event.elapsed_ms(e0, e1) / reps

    def run_plain():
        for di, do, _ in bufs:
            plan.execute(di.ptr, do.ptr, s.handle)

    run_batch = lambda: batch.launch(s.handle)
    for _ in range(warmup):
        run_plain()
        run_batch()
    s.synchronize()
    ms_plain = timed(run_plain)
    ref = [do.download(x.shape) for _, do, x in bufs]
    ms_batch = timed(run_batch)
    ms_plain = (ms_plain + timed(run_plain)) / 2
    ms_batch = (ms_batch + timed(run_batch)) / 2
    same = all(np.array_equal(do.download(x.shape), r) for (_, do, x), r in zip(bufs, ref))
    total = count * bfly(log_n)
    return {"transforms": count, "log2_n": log_n, "field": "MNT6-753 Fr",
            "ms_batched": ms_batch, "ms_plain_loop": ms_plain,
            "butterflies_per_s_batched": total / (ms_batch * 1e-3),
            "butterflies_per_s_plain_loop": total / (ms_plain * 1e-3),
            "roofline_frac_batched": total * MADS_PER_BFLY_753 / (ms_batch * 1e-3) / MAD_PEAK,
            "launches_batched": batch.num_launches, "bit_exact_vs_plain": bool(same)}


def bench_serving(requests=1000, rates=(0.3, 0.7, 1.1), seed=0x5EED0004):
    import gpusnarks_amd as g
    from gpusnarks_amd import synth
    reqs = synth.serving_requests(seed, requests)
    rng = random.Random(seed + 1)  # arrival times
    # one input per (field, size) class, one output buffer per request (the
    # transforms run out of place, so requests never share a buffer they write)
    ins = {}
    for k, (field, log_n, inv) in enumerate(reqs):
        if (field, log_n) not in ins:
            ins[(field, log_n)] = g.DeviceArray.from_numpy(inputs(field, log_n, 0x5E4E + k))
    bufs = [(ins[(field, log_n)], g.DeviceArray(ins[(field, log_n)].nbytes)) for field, log_n, _ in reqs]
    plans = {key: g.Plan(*key[:2], inverse=key[2]) for key in set(reqs)}
    s = g.Stream()
    # plain capacity: back-to-back one-at-a-time service time
    t0 = time.perf_counter()
    for key, (d, o) in zip(reqs, bufs):
        plans[key].execute(d.ptr, o.ptr, s.handle)
        s.synchronize()
    busy = time.perf_counter() - t0
    capacity = requests / busy  # requests/s
    total_bfly = sum(bfly(l) for _, l, _ in reqs)
    out = {"requests_per_rate": requests, "mix": "70% MNT6-753 Fr / 30% Fp32, log2 n uniform in [12,20], fwd/inv",
           "plain_capacity_req_per_s": capacity, "rates": []}
    srv = g.Server(num_streams=4, max_batch=64)
    # warm-up: one request per (field, size, direction) builds the plans'
    # twiddle tables before anything is timed
    seen = {}
    for key, (d, o) in zip(reqs, bufs):
        if key not in seen:
            seen[key] = srv.submit_device(key[0], key[1], d.ptr, o.ptr, inverse=key[2])
    for t in seen.values():
        srv.wait(t)
    for frac in rates:
        lam = frac * capacity
        arrivals = np.cumsum([rng.expovariate(lam) for _ in range(requests)])
        # one-at-a-time plain dispatch (FIFO, each request runs alone)
        lat_plain = []
        t_start = time.perf_counter()
        for a, key, (d, o) in zip(arrivals, reqs, bufs):
            now = time.perf_counter() - t_start
            if now < a:
                time.sleep(a - now)
            plans[key].execute(d.ptr, o.ptr, s.handle)
            s.synchronize()
            lat_plain.append(time.perf_counter() - t_start - a)
        span_plain = time.perf_counter() - t_start
        # served: continuous batching
        tickets = []
        t_start = time.perf_counter()
        for a, (field, log_n, inv), (d, o) in zip(arrivals, reqs, bufs):
            now = time.perf_counter() - t_start
            if now < a:
                time.sleep(a - now)
            tickets.append(srv.submit_device(field, log_n, d.ptr, o.ptr, inverse=inv))
        for t in tickets:
            srv.wait(t)
        span_srv = time.perf_counter() - t_start
        lat_srv = []
        for t in tickets:
            ts, tb, td = srv.times(t)
            lat_srv.append(td - ts)
        p = lambda v, q: float(np.percentile(np.array(v) * 1e3, q))
        out["rates"].append({
            "offered_fraction_of_plain_capacity": frac, "arrival_rate_req_per_s": lam,
            "served": {"p50_ms": p(lat_srv, 50), "p99_ms": p(lat_srv, 99),
                       "throughput_req_per_s": requests / span_srv,
                       "butterflies_per_s": total_bfly / span_srv},
            "plain_one_at_a_time": {"p50_ms": p(lat_plain, 50), "p99_ms": p(lat_plain, 99),
                                    "throughput_req_per_s": requests / span_plain,
                                    "butterflies_per_s": total_bfly / span_plain},
        })
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log-n", type=int, default=18)
    ap.add_argument("--field", choices=["mnt6", "mnt4", "fp32"], default="mnt6",
                    help="field of the headline transform (profiling runs)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run the u32 copy kernel over this many MiB (counter calibration)")
    ap.add_argument("--settle-ms", type=float, default=SETTLE_MS,
                    help="untimed GPU time after the warm-up transforms, before each timed region")
    ap.add_argument("--section", choices=["batched", "batched_same", "serving"], default=None,
                    help="only run one workload of the line and print its JSON (profiling runs)")
    args = ap.parse_args()
    set_settle_ms(args.settle_ms)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import gpusnarks_amd as g
    if args.copy_calibration_mib:
        nbytes = args.copy_calibration_mib << 20
        src, dst = g.DeviceArray(nbytes), g.DeviceArray(nbytes)
        for _ in range(max(1, args.steps)):
            g.check(g.lib().gsn_debug_copy_u32(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes // 4, None))
        g.synchronize()
        log("copied %d MiB x %d" % (args.copy_calibration_mib, args.steps))
        return
    if args.section:
        fn = {"batched": bench_batched, "batched_same": bench_batched_same, "serving": bench_serving}[args.section]
        print(json.dumps({args.section: fn()}))
        return
    field = {"mnt6": g.FIELD_MNT6753_FR, "mnt4": g.FIELD_MNT4753_FR, "fp32": g.FIELD_FP32}[args.field]
    log_n = args.log_n
    log("headline: field %s n=2^%d, %d steps after %d warmup" % (args.field, log_n, args.steps, args.warmup))
    if field != g.FIELD_MNT6753_FR:
        ms, passes = time_plan(field, log_n, args.steps, args.warmup)
        print(json.dumps({"field": args.field, "log2_n": log_n, "ms": ms, "passes": passes,
                          "butterflies_per_s": bfly(log_n) / (ms * 1e-3)}))
        return
    settle = {}
    ms, passes, gpu_out = time_plan(field, log_n, args.steps, args.warmup, want_output=True, info=settle)
    value = bfly(log_n) / (ms * 1e-3)
    ms_launch = ms / passes
    mads_per_launch = bfly(log_n) / passes * MADS_PER_BFLY_753
    achieved = mads_per_launch / (ms_launch * 1e-3)
    roofline = {
        "bound": "valu-int-mul", "unit": "Tmad/s",
        "achieved": achieved / 1e12, "peak": MAD_PEAK / 1e12, "frac": achieved / MAD_PEAK,
        "traffic": None,
        "kernel": "ntt753_kernel (one Stockham pass, radix 2^%d)" % (log_n // passes),
        "launches_per_step": passes, "avg_launch_ms": ms_launch,
        "work_per_launch": "%d butterflies x 1176 v_mad_u64_u32" % (bfly(log_n) // passes),
        "peak_note": "256 CU x 64 lanes/clk x 2.4 GHz (peak engine clock); clock_measured_ghz is the engine "
                     "clock while the 753-bit product runs on every CU (gsn_debug_clock_ghz)",
    }
    ghz = ctypes.c_double(0.0)
    g.check(g.lib().gsn_debug_clock_ghz(ctypes.byref(ghz)))
    roofline["clock_measured_ghz"] = ghz.value
    roofline["peak_at_measured_clock"] = MAD_PEAK / 1e12 * ghz.value * 1e9 / CLOCK_HZ
    roofline["frac_at_measured_clock"] = achieved / (MAD_PEAK * ghz.value * 1e9 / CLOCK_HZ)
    # the denominator's instruction rate, measured in this run: independent
    # v_mad_u64_u32 chains on every CU (gsn_debug_mad_rate)
    rate, rate_ghz = ctypes.c_double(0.0), ctypes.c_double(0.0)
    g.check(g.lib().gsn_debug_mad_rate(ctypes.byref(rate), ctypes.byref(rate_ghz)))
    roofline["peak_probe"] = {
        "tmad_per_s": rate.value / 1e12, "clock_ghz": rate_ghz.value,
        "lanes_per_clk_per_cu": rate.value / (CUS * rate_ghz.value * 1e9),
        "frac_of_probe": achieved / rate.value,
        "kernel": "mad_rate_kernel: 8 independent v_mad_u64_u32 chains per thread, 8 waves per SIMD"}
    line_timing = step_times(field, log_n)
    traffic_file = os.path.join(ROOT, "profiles", "ntt753_n18_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tr = json.load(fh)
        roofline["traffic"] = tr.get("hbm_bytes_per_launch")
        roofline["traffic_source"] = "%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this kernel, committed " \
                                     "summary; not measured in this run)" % tr.get("source")
    line = {
        "metric": "NTT butterflies/s, 753-bit n=2^18, batched vs plain; serving p99 latency",
        "value": value, "unit": "butterflies/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": value / REF_DERIVED_BFLY_S, "dtype": "u32",
        "data": "synthetic (SplitMix64 residues, seeded)",
        "config": {"workload": "forward NTT, MNT6-753 Fr (753-bit, 24 x u32 Montgomery limbs), n=2^%d, "
                               "operands resident in HBM, one transform per step" % log_n,
                   "field": "MNT6-753 Fr", "log2_n": log_n, "elements_per_s": (1 << log_n) / (ms * 1e-3)},
        "roofline": roofline,
        "per_transform": line_timing,
        "clock_settle": dict(settle, note="after the `warmup` transforms, untimed transforms run until %g ms of "
                             "GPU time has passed; ms_per_step_after_warmup_only times the same steps before "
                             "that" % SETTLE_MS),
    }
    if not args.skip_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        log("cpu baseline on %d threads" % threads)
        line["cpu_baseline"] = cpu_baseline(log_n, threads, j_sample=512, gpu_out=gpu_out)
        line["bit_exact_vs_cpu"] = line["cpu_baseline"].get("bit_exact_vs_gpu")
        # the same harness path run in full at 2^16, beside its sampled estimate
        _, _, out16 = time_plan(field, 16, 3, 1, want_output=True, settle_ms=0.0)
        full16 = cpu_full_run(16, threads, field, out16)
        est16 = cpu_baseline(16, threads, field, serial=False)
        full16["sampled_estimate_seconds"] = est16["seconds_extrapolated"]
        line["cpu_baseline"]["full_run_n16"] = full16
    if not args.skip_extras:
        log("fp32 n=2^26")
        ms32, p32 = time_plan(g.FIELD_FP32, 26, 20, 3)
        bytes_alg = 8 * (1 << 26) * ((26 + 14) // 15)  # 8 n per pass of a 2^15 on-chip tile
        line["fp32_n26"] = {
            "ms": ms32, "butterflies_per_s": bfly(26) / (ms32 * 1e-3), "passes": p32,
            "algorithmic_hbm_bytes": bytes_alg, "achieved_GBps": bytes_alg / (ms32 * 1e-3) / 1e9,
            "frac_of_8TBps": bytes_alg / (ms32 * 1e-3) / HBM_PEAK,
            "actual_pass_bytes_GBps": 8 * (1 << 26) * p32 / (ms32 * 1e-3) / 1e9,
            "per_transform": step_times(g.FIELD_FP32, 26),
        }
        valu_file = os.path.join(ROOT, "profiles", "ntt_fp32_n26_valu.json")
        if os.path.exists(valu_file):
            # the time the passes would take with the VALU issuing every cycle
            # (their counted instructions at 2 or 4 cycles each, committed
            # counter summary): the ceiling this butterfly leaves for HBM use
            with open(valu_file) as fh:
                vb = json.load(fh)["valu_bound_ms"]
            line["fp32_n26"]["valu_ceiling"] = {
                "valu_bound_ms": vb, "valu_ceiling_frac": vb / ms32,
                "max_frac_of_8TBps_at_valu_bound": bytes_alg / (vb * 1e-3) / HBM_PEAK,
                "source": "profiles/ntt_fp32_n26_counters.txt (SQ_INSTS_VALU per pass x the kernels' cycles per "
                          "instruction; not measured in this run)"}
        log("inner product")
        line["inner_product"] = bench_inner_product()
        log("end to end (host buffers)")
        line["end_to_end"] = {"mnt6_n18": end_to_end(g.FIELD_MNT6753_FR, log_n),
                              "fp32_n26": end_to_end(g.FIELD_FP32, 26)}
        log("batched ragged workload")
        line["batched"] = bench_batched()
        log("batched 16 x 2^%d" % log_n)
        line["batched_same_size"] = bench_batched_same(log_n)
        line["configs"] = bench_configs(min(16, len(os.sched_getaffinity(0))), args.skip_cpu)
        log("serving")
        line["serving"] = bench_serving()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
