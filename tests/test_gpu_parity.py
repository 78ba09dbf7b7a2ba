"""GPU parity: the HIP transform against the golden DFT vectors and the CPU
oracle (bit-exact), and size-independent properties at full sizes."""
import os
import sys

import numpy as np
import pytest

import gpusnarks_amd as g
from gpusnarks_amd import synth
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = {g.FIELD_FP32: "fp32", g.FIELD_MNT4753_FR: "mnt4753_fr", g.FIELD_MNT6753_FR: "mnt6753_fr"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def inputs(field, log_n, seed):
    return synth.uniform_residues(seed, 1 << log_n, g.MODULI[field], g.NLIMBS[field])


@pytest.mark.parametrize("field", [g.FIELD_FP32, g.FIELD_MNT4753_FR, g.FIELD_MNT6753_FR])
def test_golden_dft(field):
    gv = np.load(os.path.join(GOLDEN, "ntt_%s.npz" % NAMES[field]))
    sizes = sorted(int(k[1:].split("_")[0]) for k in gv.files if k.endswith("_input"))
    for n in sizes:
        say("golden", NAMES[field], n)
        x, om = gv["n%d_input" % n], gv["n%d_omega" % n]
        assert np.array_equal(g.best_fft(x, om, field), gv["n%d_forward" % n]), n
        assert np.array_equal(g.best_ifft(x, om, field), gv["n%d_inverse" % n]), n


@pytest.mark.parametrize("field,log_n", [
    (g.FIELD_FP32, 4), (g.FIELD_FP32, 11), (g.FIELD_FP32, 12), (g.FIELD_FP32, 13), (g.FIELD_FP32, 16),
    (g.FIELD_FP32, 17), (g.FIELD_FP32, 20), (g.FIELD_FP32, 21),
    (g.FIELD_MNT6753_FR, 1), (g.FIELD_MNT6753_FR, 5), (g.FIELD_MNT6753_FR, 9), (g.FIELD_MNT6753_FR, 10),
    (g.FIELD_MNT6753_FR, 13), (g.FIELD_MNT6753_FR, 16),
    (g.FIELD_MNT4753_FR, 7), (g.FIELD_MNT4753_FR, 12), (g.FIELD_MNT4753_FR, 15),
])
def test_matches_oracle(orc, field, log_n):
    say("oracle parity", NAMES[field], log_n)
    x = inputs(field, log_n, 0x5EED0001 + log_n)
    om = g.root_of_unity(field, log_n)
    want = orc.serial_fft(field, x, om)
    got = g.best_fft(x, om, field)
    assert np.array_equal(got, want)
    back = g.best_ifft(got, om, field)
    assert np.array_equal(back, x)


@pytest.mark.parametrize("field,log_n", [(g.FIELD_MNT6753_FR, 18), (g.FIELD_MNT6753_FR, 22), (g.FIELD_FP32, 24),
                                         (g.FIELD_FP32, 26)])
def test_roundtrip_and_linearity_full_size(field, log_n):
    say("full-size round trip", NAMES[field], log_n)
    p = g.MODULI[field]
    x = inputs(field, log_n, 0xA11CE + log_n)
    X = g.best_fft(x, None, field)
    assert np.array_equal(g.best_ifft(X, None, field), x)
    # determinism: a second run gives the same bits
    assert np.array_equal(g.best_fft(x, None, field), X)
    # outputs are canonical
    if field == g.FIELD_FP32:
        assert (X[:, 0] < p).all()
        # linearity on a few lanes: F(x + y) == F(x) + F(y)
        y = inputs(field, log_n, 0xB0B + log_n)
        s = ((x.astype(np.uint64) + y) % p).astype(np.uint32)
        lhs = g.best_fft(s, None, field)
        rhs = ((X.astype(np.uint64) + g.best_fft(y, None, field)) % p).astype(np.uint32)
        assert np.array_equal(lhs, rhs)
    else:
        top = X[:, -1].astype(np.uint64) << 32 | X[:, -2]
        ptop = p >> (32 * 22)
        assert (top <= ptop).all()
        # delta at 0 -> all ones (the Montgomery one everywhere)
        one = g.root_of_unity(field, 0)
        d = np.zeros_like(x)
        d[0] = one
        assert (g.best_fft(d, None, field) == one).all()


def test_fp32_largest_configured_size():
    """n = 2^28 (the largest 32-bit configuration, 1 GiB) on device buffers:
    forward then inverse returns the input, and X_0 = sum of the inputs."""
    field, log_n = g.FIELD_FP32, 28
    p = g.MODULI[field]
    say("fp32 2^28 round trip")
    # a multiplicative hash of the index, reduced mod p (uniform enough, and
    # cheap at 2^28)
    x = ((np.arange(1 << log_n, dtype=np.uint64) * 2654435761 + 0x2828) % p).astype(np.uint32).reshape(-1, 1)
    d = g.DeviceArray.from_numpy(x)
    e = g.DeviceArray(x.nbytes)
    fwd, inv = g.Plan(field, log_n), g.Plan(field, log_n, inverse=True)
    fwd.execute(d.ptr, e.ptr)
    g.synchronize()
    X0 = e.download((1, 1))[0, 0]
    # the inputs are Montgomery forms; the transform is linear, so X_0 is
    # their sum mod p in the same representation
    assert int(X0) == int(x[:, 0].astype(np.uint64).sum() % p)
    inv.execute(e.ptr, e.ptr)
    g.synchronize()
    assert np.array_equal(e.download(x.shape), x)


@pytest.mark.parametrize("field,log_n", [(g.FIELD_MNT6753_FR, 17), (g.FIELD_FP32, 23), (g.FIELD_FP32, 21)])
def test_host_entry_staging_matches_device_path(field, log_n):
    """The host entry point moves buffers above one 8 MiB staging chunk
    through pinned double buffering (12 MiB: a ragged last chunk; 32 MiB: four
    chunks; 8 MiB: the plain copy): its result equals the plan on device
    buffers, and the caller's other memory is untouched."""
    say("host staging", NAMES[field], log_n)
    x = inputs(field, log_n, 0x57A6 + log_n)
    plan = g.Plan(field, log_n)
    d_in, d_out = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
    plan.execute(d_in.ptr, d_out.ptr, 0)
    g.synchronize()
    want = d_out.download(x.shape)
    L = g.NLIMBS[field]
    guard = np.full((x.shape[0] + 64, L), 0xDEADBEEF, np.uint32)
    guard[32:-32] = x
    view = guard[32:-32]
    assert view.flags["C_CONTIGUOUS"]
    import ctypes
    g.check(g.lib().gsn_fft_host(field, view.ctypes.data_as(ctypes.c_void_p), x.shape[0], None, 0))
    assert np.array_equal(view, want)
    assert (guard[:32] == 0xDEADBEEF).all() and (guard[-32:] == 0xDEADBEEF).all()


def test_plan_api_device_buffers(orc):
    field, log_n = g.FIELD_MNT6753_FR, 12
    x = inputs(field, log_n, 77)
    plan = g.Plan(field, log_n)
    assert plan.num_passes == 2
    d_in = g.DeviceArray.from_numpy(x)
    d_out = g.DeviceArray(x.nbytes)
    stream = g.Stream()
    plan.execute(d_in.ptr, d_out.ptr, stream.handle)
    stream.synchronize()
    want = orc.serial_fft(field, x, g.root_of_unity(field, log_n))
    assert np.array_equal(d_out.download(x.shape), want)
    # in place
    plan.execute(d_in.ptr, d_in.ptr, stream.handle)
    stream.synchronize()
    assert np.array_equal(d_in.download(x.shape), want)
    plan.close()


def test_errors_on_gpu():
    with pytest.raises(g.GsnError) as e:
        g.best_fft(np.zeros((1 << 16, 24), np.uint32), None, g.FIELD_MNT4753_FR)
    assert e.value.status == 3  # 2^16 exceeds MNT4-753 Fr's 2-adicity 15
    bad = g.root_of_unity(g.FIELD_FP32, 9)  # order 2^9, not 2^10
    with pytest.raises(g.GsnError) as e:
        g.best_fft(np.zeros(1 << 10, np.uint32), bad, g.FIELD_FP32)
    assert e.value.status == 4
    g.inject_faults(1)
    with pytest.raises(g.GsnError) as e:
        g.best_fft(np.zeros(16, np.uint32), None, g.FIELD_FP32)
    assert e.value.status == 10
    # and the next call works again
    g.best_fft(np.zeros(16, np.uint32), None, g.FIELD_FP32)


def test_cpp_drop_in_gpu(tmp_path):
    """The reference's call, best_fft<FieldT>(vector, omega), from a plain
    C++ program linked to the library: each field against an O(n^2) DFT
    computed with the host field type, and best_ifft back to the input."""
    import subprocess
    from tests.test_library import build_dropin
    exe = build_dropin(tmp_path)
    say("c++ drop-in (subprocess)", exe)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gpu: 0 failure(s)" in r.stdout


SCHEDULES = [
    (g.FIELD_MNT6753_FR, 12, [[(6, 1, 1), (6, 1, 1)], [(7, 1, 1), (5, 2, 1)], [(8, 1, 1), (4, 4, 1)],
                              [(9, 1, 1), (3, 7, 1)], [(10, 0, 1), (2, 8, 1)], [(4, 3, 1)] * 3, [(2, 5, 1)] * 6]),
    (g.FIELD_MNT4753_FR, 13, [[(9, 0, 1), (4, 5, 1)], [(5, 5, 1), (8, 2, 1)]]),
    (g.FIELD_FP32, 16, [[(8, 4, 4), (8, 4, 4)], [(8, 5, 4), (8, 5, 4)], [(6, 5, 4), (5, 5, 4), (5, 5, 4)],
                        [(9, 4, 4), (7, 4, 4)], [(10, 4, 4), (6, 4, 4)], [(10, 5, 5), (6, 5, 4)],
                        [(6, 5, 4), (10, 5, 5)], [(9, 5, 5), (7, 5, 4)], [(7, 5, 4), (9, 5, 5)]]),
]


@pytest.mark.parametrize("field,log_n,schedules", SCHEDULES)
def test_every_tile_shape_matches_oracle(orc, field, log_n, schedules):
    """Explicit pass schedules covering every 753-bit workgroup size (64..512
    threads), wide and one-column tiles, and the 16- and 32-column 32-bit
    shapes (16 or 32 elements per thread, first or twiddled pass): each
    bit-exact against the oracle, forward and inverse."""
    x = inputs(field, log_n, 0x5C4ED + log_n)
    om = g.root_of_unity(field, log_n)
    want = orc.serial_fft(field, x, om)
    before = g.tuning_get(field, log_n)
    try:
        for sched in schedules:
            say("schedule", NAMES[field], log_n, sched)
            g.tuning_set(field, log_n, sched)
            got = g.best_fft(x, om, field)
            assert np.array_equal(got, want), sched
            assert np.array_equal(g.best_ifft(got, om, field), x), sched
    finally:
        g.tuning_set(field, log_n, before)


@pytest.mark.parametrize("log_n,schedule", [(22, [(9, 4, 4), (8, 4, 4), (5, 4, 4)])])
def test_fp32_direct_pass_table_matches_oracle(orc, log_n, schedule):
    """32-bit passes past 2^17 columns read the pass-twiddle base w^(jj k)
    from a direct [jj][k] table: a last pass at Ns = 2^17, bit-exact against
    the oracle, forward and inverse."""
    field = g.FIELD_FP32
    say("direct pass table", log_n, schedule)
    x = inputs(field, log_n, 0xD1EC7 + log_n)
    om = g.root_of_unity(field, log_n)
    want = orc.serial_fft(field, x, om)
    before = g.tuning_get(field, log_n)
    try:
        g.tuning_set(field, log_n, schedule)
        got = g.best_fft(x, om, field)
        assert np.array_equal(got, want)
        assert np.array_equal(g.best_ifft(got, om, field), x)
    finally:
        g.tuning_set(field, log_n, before)


def test_host_entry_from_several_threads(orc):
    """The best_fft entry point called from 6 host threads at once (ctypes
    releases the GIL): two threads share each (field, size) plan, and one
    of them retunes that size halfway, which drops the cached plan while the
    other may still be using it.  Every output equals the oracle's."""
    from concurrent.futures import ThreadPoolExecutor
    cases = [(g.FIELD_MNT6753_FR, 12), (g.FIELD_FP32, 16), (g.FIELD_MNT4753_FR, 11)]
    jobs = []
    for c, (field, log_n) in enumerate(cases):
        for t in range(2):
            x = inputs(field, log_n, 0x7A4E + 16 * c + t)
            om = g.root_of_unity(field, log_n)
            jobs.append((field, log_n, x, om, orc.serial_fft(field, x, om)))
    befores = {(f, l): g.tuning_get(f, l) for f, l in cases}

    def work(k):
        field, log_n, x, om, want = jobs[k]
        for rep in range(4):
            if k % 2 == 1 and rep == 2:
                g.tuning_set(field, log_n, befores[(field, log_n)])  # same schedule, fresh plans
            got = g.best_fft(x, om, field)
            if not np.array_equal(got, want):
                return (k, rep)
        return None

    say("host entry: %d threads" % len(jobs))
    try:
        with ThreadPoolExecutor(len(jobs)) as ex:
            bad = [r for r in ex.map(work, range(len(jobs))) if r is not None]
    finally:
        for (f, l), sched in befores.items():
            g.tuning_set(f, l, sched)
    assert not bad, bad
