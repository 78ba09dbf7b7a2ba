"""GPU: batched ragged transforms (fused launches, captured hipGraph) and
the continuous-batching server -- every result bit-identical to the same
transform run alone."""
import random
import sys
import time

import numpy as np
import pytest

import gpusnarks_amd as g
from gpusnarks_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ragged_items(seed, count, sizes753=(1, 13), sizes32=(1, 16), mnt4=True):
    rng = random.Random(seed)
    out = []
    for i in range(count):
        r = rng.random()
        if r < 0.5:
            field, lo, hi = g.FIELD_MNT6753_FR, *sizes753
        elif r < 0.65 and mnt4:
            field, lo, hi = g.FIELD_MNT4753_FR, sizes753[0], min(sizes753[1], 15)
        else:
            field, lo, hi = g.FIELD_FP32, *sizes32
        log_n = rng.randint(lo, hi)
        inverse = rng.random() < 0.5
        x = synth.uniform_residues(1000 * seed + i, 1 << log_n, g.MODULI[field], g.NLIMBS[field])
        out.append((field, log_n, inverse, x))
    return out


def solo(field, x, inverse):
    return g.best_fft(x, None, field, inverse=inverse)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_solo(use_graph):
    items = ragged_items(11 + use_graph, 40)
    want = [solo(f, x, inv) for f, _, inv, x in items]
    d_in = [g.DeviceArray.from_numpy(x) for _, _, _, x in items]
    d_out = [g.DeviceArray(x.nbytes) for _, _, _, x in items]
    batch = g.Batch([(f, l, inv, di.ptr, do.ptr) for (f, l, inv, _), di, do in zip(items, d_in, d_out)],
                    use_graph=use_graph)
    say("batch launches", batch.num_launches, "for", len(items), "transforms")
    assert batch.num_launches < len(items)
    stream = g.Stream()
    for rep in range(3):
        # clobber outputs between replays, so each replay must rewrite them
        for (_, _, _, x), do in zip(items, d_out):
            do.upload(np.zeros_like(x))
        batch.launch(stream.handle)
        stream.synchronize()
        for k, ((f, l, inv, x), do) in enumerate(zip(items, d_out)):
            got = do.download(x.shape)
            assert np.array_equal(got, want[k]), (rep, k, f, l, inv)
    batch.close()


def test_batch_in_place_and_large_mix():
    # the ragged mix of the batched workload at small scale: 753-bit 2^10..2^16, Fp32 2^10..2^18
    items = ragged_items(5, 24, sizes753=(10, 16), sizes32=(10, 18), mnt4=False)
    want = [solo(f, x, inv) for f, _, inv, x in items]
    d = [g.DeviceArray.from_numpy(x) for _, _, _, x in items]
    batch = g.Batch([(f, l, inv, di.ptr, di.ptr) for (f, l, inv, _), di in zip(items, d)], use_graph=True)
    stream = g.Stream()
    batch.launch(stream.handle)
    stream.synchronize()
    for k, ((f, l, inv, x), di) in enumerate(zip(items, d)):
        assert np.array_equal(di.download(x.shape), want[k]), (k, f, l, inv)


def test_server_mix_matches_solo():
    items = ragged_items(21, 60, sizes753=(1, 12), sizes32=(1, 16))
    want = [solo(f, x, inv) for f, _, inv, x in items]
    srv = g.Server(hbm_budget_bytes=1 << 30, num_streams=4, max_batch=16)
    tickets = []
    rng = random.Random(3)
    for f, l, inv, x in items:
        tickets.append(srv.submit(x, f, inverse=inv))
        if rng.random() < 0.3:
            time.sleep(0.0005)  # arrivals spread while earlier waves run
    for k, (t, out) in enumerate(tickets):
        srv.wait(t)
        assert np.array_equal(out, want[k]), k
        ts, tb, td = srv.times(t)
        assert ts <= tb <= td
    srv.close()


def test_server_device_operands():
    items = ragged_items(31, 20, sizes753=(8, 14), sizes32=(8, 18))
    want = [solo(f, x, inv) for f, _, inv, x in items]
    d = [g.DeviceArray.from_numpy(x) for _, _, _, x in items]
    srv = g.Server(num_streams=3)
    tickets = [srv.submit_device(f, l, di.ptr, di.ptr, inverse=inv) for (f, l, inv, _), di in zip(items, d)]
    for k, t in enumerate(tickets):
        srv.wait(t)
        f, l, inv, x = items[k]
        assert np.array_equal(d[k].download(x.shape), want[k]), k
    srv.close()


def test_server_admission_budget():
    # 4 MiB budget: a 2^16 753-bit request (6 MiB operand) is refused outright;
    # 2^12 requests fit one or a few at a time and all complete
    srv = g.Server(hbm_budget_bytes=4 << 20, num_streams=2)
    big = np.zeros((1 << 16, 24), np.uint32)
    with pytest.raises(g.GsnError) as e:
        srv.submit(big, g.FIELD_MNT6753_FR)
    assert e.value.status == 8
    items = ragged_items(41, 12, sizes753=(12, 12), sizes32=(12, 14), mnt4=False)
    want = [solo(f, x, inv) for f, _, inv, x in items]
    tickets = [srv.submit(x, f, inverse=inv) for f, _, inv, x in items]
    for k, (t, out) in enumerate(tickets):
        srv.wait(t)
        assert np.array_equal(out, want[k])
    srv.close()
