"""Autotuner for pass schedules.

For each (field, log2 n) it enumerates pass schedules -- how log2 n splits
into passes (the radix of each pass), how many columns a tile holds (tile
depth; for the 753-bit field also the workgroup size: 64, 128 or 256
threads) and the register radix -- times every candidate on the GPU with
HIP events, checks that its output is bit-identical to the default
schedule's, and writes the fastest per size to tuning.txt beside the
library, which the library reads at start-up.

    python3 -m gpusnarks_amd.autotune [--fields mnt6,fp32] [--out PATH]
"""
import argparse
import itertools
import os
import sys
import time

import numpy as np

import gpusnarks_amd as g
from gpusnarks_amd import synth

FIELD_NAMES = {"fp32": g.FIELD_FP32, "mnt4": g.FIELD_MNT4753_FR, "mnt6": g.FIELD_MNT6753_FR}
DEFAULT_SIZES = {"fp32": range(10, 29), "mnt4": range(10, 16), "mnt6": range(10, 21)}

# compiled 32-bit tile shapes (gpusnarks_amd/csrc/kernels_fp32.hip)
F32_SINGLE = {r: (r, 0, min(4, r)) for r in range(1, 13)}
F32_MULTI = {r: (r, 4, 4) for r in range(5, 11)}
F32_WIDE = {r: (r, 5, 4) for r in range(5, 10)}
F32_WIDE32 = {r: (r, 5, 5) for r in (9, 10)}  # 32 elements per thread


def compositions(total, parts, lo, hi):
    """Ordered splits of `total` into `parts` integers in [lo, hi]."""
    if parts == 1:
        if lo <= total <= hi:
            yield (total,)
        return
    for first in range(lo, min(hi, total) + 1):
        for rest in compositions(total - first, parts - 1, lo, hi):
            yield (first,) + rest


def candidates(field, log_n):
    out = []
    if field == g.FIELD_FP32:
        if log_n <= 12:
            out.append([F32_SINGLE[log_n]])
        for P in (2, 3, 4):
            for comp in compositions(log_n, P, 5, 10):
                # each pass on its own: 16- or 32-column tiles
                shapes = [[s for s in (F32_MULTI.get(r), F32_WIDE.get(r), F32_WIDE32.get(r))
                           if s and log_n - r >= s[1]] for r in comp]
                for sched in itertools.product(*shapes):
                    out.append(list(sched))
        return out
    for P in (1, 2, 3, 4):
        for comp in compositions(log_n, P, 3 if P > 1 else 1, 9):
            for tile in (7, 8, 9, 10):
                if max(comp) > tile:
                    continue
                sched = [(r, min(tile - r, log_n - r), 1) for r in comp]
                if sched not in out:
                    out.append(sched)
    return out


# GPU time of untimed transforms before each candidate is timed: the engine
# clock rises over the first ~25 ms of sustained load and falls again while
# the host compares the previous candidate's output
SETTLE_MS = 50.0


def time_schedule(field, log_n, sched, x_dev, y_dev, stream, reps, settle_ms=SETTLE_MS):
    """Mean time of one transform under `sched`: untimed transforms until
    `settle_ms` of GPU time has passed, then at least `reps` timed back to
    back (more for short transforms, so the timed region lasts ~10 ms)."""
    g.tuning_set(field, log_n, sched)
    plan = g.Plan(field, log_n)
    e0, e1 = g.Event(), g.Event()

    def timed(k):
        e0.record(stream.handle)
        for _ in range(k):
            plan.execute(x_dev.ptr, y_dev.ptr, stream.handle)
        e1.record(stream.handle)
        e1.synchronize()
        return g.Event.elapsed_ms(e0, e1)

    per = timed(1)
    spent = per
    while spent < settle_ms:
        k = max(1, min(1000, int(5.0 / max(per, 1e-3))))
        spent += timed(k)
    k = max(reps, min(2000, int(10.0 / max(per, 1e-3))))
    ms = timed(k) / k
    plan.close()
    return ms


def tune(fields, sizes=None, budget_s=60.0, log=print):
    results = {}
    stream = g.Stream()
    for name in fields:
        field = FIELD_NAMES[name]
        for log_n in (sizes or DEFAULT_SIZES[name]):
            n = 1 << log_n
            x = synth.uniform_residues(0x7E57 + log_n, n, g.MODULI[field], g.NLIMBS[field])
            x_dev = g.DeviceArray.from_numpy(x)
            y_dev = g.DeviceArray(x.nbytes)
            g.tuning_clear()
            ref = g.best_fft(x, None, field)
            default = g.tuning_get(field, log_n)
            cands = candidates(field, log_n)
            if default not in cands:
                cands.insert(0, default)
            reps = max(3, min(50, int(2e7 // (n * (24 if field else 1)))))
            best = None
            t_start = time.time()
            for sched in cands:
                if time.time() - t_start > budget_s:
                    break
                try:
                    ms = time_schedule(field, log_n, sched, x_dev, y_dev, stream, reps)
                except g.GsnError:
                    continue
                got = y_dev.download(x.shape)
                if not np.array_equal(got, ref):
                    raise AssertionError("schedule %s for %s 2^%d changed the result" % (sched, name, log_n))
                if best is None or ms < best[0]:
                    best = (ms, sched)
            g.tuning_clear()
            ms_default = time_schedule(field, log_n, default, x_dev, y_dev, stream, reps)
            g.tuning_clear()
            results[(field, log_n)] = (best[1], best[0], ms_default)
            log("%s 2^%d: best %s %.4f ms (default %s %.4f ms, %d candidates)"
                % (name, log_n, best[1], best[0], default, ms_default, len(cands)))
    return results


def batch_lines(path):
    """The 'batch ...' lines of an existing table (schedules of transforms
    fused into batches, which this tuner, timing transforms alone, keeps)."""
    if not os.path.exists(path):
        return []
    with open(path) as fh:
        return [l.rstrip("\n") for l in fh if l.split("#", 1)[0].split()[:1] == ["batch"]]


def kept_lines(path, results):
    """The lines of an existing table for sizes this run did not tune, by
    (field, log2 n)."""
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as fh:
        for l in fh:
            tok = l.split("#", 1)[0].split()
            if len(tok) >= 3 and tok[0] != "batch" and (int(tok[0]), int(tok[1])) not in results:
                out[(int(tok[0]), int(tok[1]))] = l.rstrip("\n")
    return out


def write_table(results, path, keep_from=None):
    """Tuned sizes from `results`, the other sizes and the batch lines as
    the existing table (`keep_from`, else `path`) has them."""
    lines = ["# field log_n  log_r,log_w,log_e per pass  (written by gpusnarks_amd.autotune)",
             "# 'batch field log_n ...': the schedule of transforms fused into batches, where it differs"]
    rows = kept_lines(keep_from or path, results)
    for (field, log_n), (sched, ms, ms_default) in results.items():
        rows[(field, log_n)] = ("%d %d %s  # %.4f ms (default %.4f ms)"
                                % (field, log_n, " ".join("%d,%d,%d" % s for s in sched), ms, ms_default))
    lines += [rows[k] for k in sorted(rows)]
    lines += batch_lines(keep_from or path)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fields", default="mnt6,mnt4,fp32")
    ap.add_argument("--sizes", default="", help="comma-separated log2 n (default: per field)")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.txt"))
    ap.add_argument("--budget-s", type=float, default=60.0, help="time budget per size")
    args = ap.parse_args()
    sizes = [int(s) for s in args.sizes.split(",") if s] or None
    res = tune(args.fields.split(","), sizes, args.budget_s, log=lambda *a: print(*a, flush=True))
    write_table(res, args.out, keep_from=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.txt"))
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    sys.exit(main())
