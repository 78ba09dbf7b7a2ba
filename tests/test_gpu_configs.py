"""GPU parity at the configured sizes and workloads, against the CPU oracle.

* the headline transform (MNT6-753 Fr, n = 2^18), forward and inverse,
  bit-exact against the oracle's full serial transform (test/fft_host.h:18-54)
  -- the reference harness's device-versus-host comparison (test/main.cpp:80-84);
* the sizes too large for the O(n log n) oracle in seconds (MNT6-753 Fr 2^22,
  the 32-bit field at 2^24, 2^26, 2^28): X_k at k = 0, 1, n/2, n-1 and 64
  seeded k from the DFT's definition, O(n) each -- which a wrong root or a
  permuted output fails even though it inverts itself;
* the batched ragged workload exactly as benchmarked (256 transforms, 192 x
  MNT6-753 Fr + 64 x 32-bit, log2 n in [10, 18], forward/inverse) under a
  captured hipGraph, every output against the oracle, on three replays;
* the serving mix (70% MNT6-753 Fr / 30% 32-bit, log2 n in [12, 20],
  forward/inverse, Poisson arrivals) through the continuous-batching server,
  every output against the oracle (spot values above 2^18);
* every schedule in the autotuner's table, against the oracle.
"""
import os
import random
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gpusnarks_amd as g
from gpusnarks_amd import synth
from tests.oracle_lib import Oracle, oracle_transform, primitive_root, spot_indices, spot_values

pytestmark = pytest.mark.gpu
NAMES = {g.FIELD_FP32: "fp32", g.FIELD_MNT4753_FR: "mnt4", g.FIELD_MNT6753_FR: "mnt6"}
WORKERS = min(16, len(os.sched_getaffinity(0)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def inputs(field, log_n, seed):
    return synth.uniform_residues(seed, 1 << log_n, g.MODULI[field], g.NLIMBS[field])


def test_headline_size_matches_oracle(orc):
    """MNT6-753 Fr n = 2^18: the GPU transform equals the oracle's serial
    radix-2 transform element for element, forward and inverse (ctypes
    releases the GIL, so the two oracle runs overlap)."""
    field, log_n = g.FIELD_MNT6753_FR, 18
    x = inputs(field, log_n, 0x5EED0000 + log_n)  # the benchmark's headline input
    om = primitive_root(field, log_n)
    say("headline 2^18: oracle forward + inverse on two threads")
    with ThreadPoolExecutor(2) as ex:
        f_fwd = ex.submit(oracle_transform, orc, field, x, False, om)
        f_inv = ex.submit(oracle_transform, orc, field, x, True, om)
        got_fwd = g.best_fft(x, om, field)
        got_inv = g.best_ifft(x, om, field)
        assert np.array_equal(got_fwd, f_fwd.result())
        assert np.array_equal(got_inv, f_inv.result())


@pytest.mark.parametrize("field,log_n", [(g.FIELD_MNT6753_FR, 22), (g.FIELD_FP32, 24), (g.FIELD_FP32, 26),
                                         (g.FIELD_FP32, 28)])
def test_full_size_spot_values(orc, field, log_n):
    """Outputs at 68 indices against the DFT's definition forward, and at 20
    of them inverse, transforming device buffers."""
    n = 1 << log_n
    if field == g.FIELD_FP32:
        x = hashed_fp32(log_n, 0x5B07 + log_n)
    else:
        x = inputs(field, log_n, 0x5B07 + log_n)
    om = primitive_root(field, log_n)
    ks = spot_indices(n, 64, log_n)
    d_in, d_out = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
    for inverse, kk in ((False, ks), (True, ks[:20])):
        say("spot values", NAMES[field], log_n, "inverse" if inverse else "forward", len(kk))
        plan = g.Plan(field, log_n, om, inverse=inverse)
        with ThreadPoolExecutor(1) as ex:
            fut = ex.submit(spot_values, orc, field, x, om, kk, inverse, WORKERS)
            plan.execute(d_in.ptr, d_out.ptr)
            g.synchronize()
            out = d_out.download(x.shape)
            assert np.array_equal(out[kk.astype(np.int64)], fut.result())
        plan.close()
        del out


def _oracle_all(orc, items, xs):
    """Oracle transforms of every item, WORKERS at a time, largest first."""
    order = sorted(range(len(items)), key=lambda k: -items[k][1])
    res = [None] * len(items)
    with ThreadPoolExecutor(WORKERS) as ex:
        futs = {k: ex.submit(oracle_transform, orc, items[k][0], xs[k], items[k][2],
                             primitive_root(items[k][0], items[k][1])) for k in order}
        for k, f in futs.items():
            res[k] = f.result()
    return res


def test_ragged_batch_workload_matches_oracle(orc):
    """The benchmark's 256-transform ragged workload as one fused schedule
    under a captured hipGraph: every output equals the oracle's transform,
    on each of three replays with the outputs clobbered in between."""
    items = synth.ragged_workload()
    assert len(items) == 256
    assert sum(1 for f, _, _ in items if f == g.FIELD_MNT6753_FR) == 192
    xs = [inputs(f, l, synth.RAGGED_INPUT_SEED + k) for k, (f, l, _) in enumerate(items)]
    say("ragged 256: oracle on %d threads" % WORKERS)
    with ThreadPoolExecutor(1) as bg:
        fut = bg.submit(_oracle_all, orc, items, xs)
        d_in = [g.DeviceArray.from_numpy(x) for x in xs]
        d_out = [g.DeviceArray(x.nbytes) for x in xs]
        batch = g.Batch([(f, l, inv, di.ptr, do.ptr) for (f, l, inv), di, do in zip(items, d_in, d_out)],
                        use_graph=True)
        say("ragged 256: %d launches" % batch.num_launches)
        stream = g.Stream()
        first = None
        for rep in range(3):
            for x, do in zip(xs, d_out):
                do.upload(np.zeros_like(x))
            batch.launch(stream.handle)
            stream.synchronize()
            outs = [do.download(x.shape) for x, do in zip(xs, d_out)]
            if first is None:
                first = outs
            else:  # every replay rewrites every output identically
                for k, (a, b) in enumerate(zip(outs, first)):
                    assert np.array_equal(a, b), (rep, k)
        t0 = time.time()
        want = fut.result()
    say("ragged 256: oracle done %.1f s after the GPU" % (time.time() - t0))
    for k, ((f, l, inv), got) in enumerate(zip(items, first)):
        assert np.array_equal(got, want[k]), (k, NAMES[f], l, inv)
    batch.close()


def test_serving_mix_matches_oracle(orc):
    """48 requests of the serving mix (log2 n in [12, 20]) arriving as a
    Poisson stream, host and device operands alternating: each result equals
    the oracle's (the full transform up to 2^18, spot values at 36 indices
    above), i.e. the same alone and in the mix."""
    reqs = synth.serving_requests(0x5EED0044, 48)
    assert {l for _, l, _ in reqs} >= {12, 20}
    xs = [inputs(f, l, 0x5E4E00 + k) for k, (f, l, _) in enumerate(reqs)]
    small = [k for k, (_, l, _) in enumerate(reqs) if l <= 18]
    large = [k for k, (_, l, _) in enumerate(reqs) if l > 18]
    say("serving 48: %d full oracle transforms, %d spot-checked" % (len(small), len(large)))

    def checker():
        res = dict(zip(small, _oracle_all(orc, [reqs[k] for k in small], [xs[k] for k in small])))
        for k in large:
            f, l, inv = reqs[k]
            ks = spot_indices(1 << l, 32, k)
            res[k] = (ks, spot_values(orc, f, xs[k], primitive_root(f, l), ks, inv, WORKERS))
        return res

    with ThreadPoolExecutor(1) as bg:
        fut = bg.submit(checker)
        srv = g.Server(num_streams=4, max_batch=16)
        rng = random.Random(5)
        tickets, dev = [], {}
        for k, ((f, l, inv), x) in enumerate(zip(reqs, xs)):
            if k % 2:
                d = g.DeviceArray.from_numpy(x)
                o = g.DeviceArray(x.nbytes)
                dev[k] = (d, o)
                tickets.append((srv.submit_device(f, l, d.ptr, o.ptr, inverse=inv), None))
            else:
                tickets.append(srv.submit(x, f, inverse=inv))
            time.sleep(rng.expovariate(2000.0))
        got = []
        for k, (t, out) in enumerate(tickets):
            srv.wait(t)
            got.append(out if out is not None else dev[k][1].download(xs[k].shape))
        srv.close()
        want = fut.result()
    for k in small:
        assert np.array_equal(got[k], want[k]), (k, reqs[k])
    for k in large:
        ks, v = want[k]
        assert np.array_equal(got[k][ks.astype(np.int64)], v), (k, reqs[k])


def _tuning_entries():
    path = os.path.join(os.path.dirname(g.library_path), "tuning.txt")
    out = []
    with open(path) as fh:
        for line in fh:
            line = line.split("#", 1)[0].split()
            # 'batch' lines (schedules of fused transforms) run in the ragged
            # batch test above
            if len(line) >= 3 and line[0] != "batch":
                out.append((int(line[0]), int(line[1])))
    return out


def hashed_fp32(log_n, salt):
    """i * 2654435761 + salt mod p: 32-bit inputs of any size, cheaply."""
    n = 1 << log_n
    return ((np.arange(n, dtype=np.uint64) * 2654435761 + salt) % g.MODULI[g.FIELD_FP32]).astype(
        np.uint32).reshape(-1, 1)


def test_tuned_schedules_match_oracle(orc):
    """Every (field, size) the autotuner's table schedules: the full oracle
    transform up to 2^16, spot values at 36 indices above."""
    entries = _tuning_entries()
    assert len(entries) >= 30
    say("tuned schedules: %d entries" % len(entries))
    small = [(f, l) for f, l in entries if l <= 16]
    large = [(f, l) for f, l in entries if l > 16]

    def x_of(f, l):
        return hashed_fp32(l, 0x70E0 + l) if (f == g.FIELD_FP32 and l > 20) else inputs(f, l, 0x70E0 + l)

    xs_small = [x_of(f, l) for f, l in small]
    with ThreadPoolExecutor(WORKERS) as ex:
        futs = [ex.submit(orc.serial_fft, f, x, primitive_root(f, l)) for (f, l), x in zip(small, xs_small)]
        got = [g.best_fft(x, None, f) for (f, _), x in zip(small, xs_small)]
        for (f, l), out, fu in zip(small, got, futs):
            assert np.array_equal(out, fu.result()), (NAMES[f], l)
    for f, l in large:
        x = x_of(f, l)
        ks = spot_indices(1 << l, 32, l)
        with ThreadPoolExecutor(1) as ex:
            fu = ex.submit(spot_values, orc, f, x, primitive_root(f, l), ks, False, WORKERS)
            out = g.best_fft(x, None, f)
            assert np.array_equal(out[ks.astype(np.int64)], fu.result()), (NAMES[f], l)
        del out, x
