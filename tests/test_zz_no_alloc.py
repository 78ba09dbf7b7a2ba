"""The allocation-free plan entry point (gsn_plan_execute_noalloc) against
gsn_plan_execute, both called through the C ABI: bit-identical outputs on two
calls in a row, the second with a smaller input (the shared per-device
workspace keeps the first call's contents), forward and inverse; sizes
beyond the workspace limit and plans of other kinds are refused with an
error before anything is launched."""
import ctypes
import sys

import numpy as np
import pytest

import gpusnarks_amd as g
from gpusnarks_amd import synth

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    lib = g.lib()
    lib.gsn_plan_execute_noalloc.argtypes = [vp, vp, vp, vp]
    lib.gsn_plan_workspace_bytes.restype = ctypes.c_size_t
    lib.gsn_plan_workspace_bytes.argtypes = [vp]
    g.check(lib.gsn_device_init(0))
    return lib


LIMIT = 64 << 20  # GSN_NOALLOC_WORKSPACE_BYTES


def run_both(L, field, log_n, inverse, seed):
    x = synth.uniform_residues(seed, 1 << log_n, g.MODULI[field], g.NLIMBS[field])
    plan = g.Plan(field, log_n, inverse=inverse)
    assert 0 < L.gsn_plan_workspace_bytes(plan.handle) <= LIMIT
    d_in = g.DeviceArray.from_numpy(x)
    d_ref, d_out = g.DeviceArray(x.nbytes), g.DeviceArray(x.nbytes)
    s = g.Stream()
    g.check(L.gsn_plan_execute(plan.handle, vp(d_in.ptr), vp(d_ref.ptr), vp(s.handle)))
    g.check(L.gsn_plan_execute_noalloc(plan.handle, vp(d_in.ptr), vp(d_out.ptr), vp(s.handle)))
    s.synchronize()
    ref, got = d_ref.download(x.shape), d_out.download(x.shape)
    plan.close()
    return x, ref, got


@pytest.mark.parametrize("field,big,small", [(g.FIELD_MNT6753_FR, 18, 13), (g.FIELD_FP32, 23, 17),
                                             (g.FIELD_MNT4753_FR, 15, 11)])
@pytest.mark.parametrize("inverse", [False, True])
def test_noalloc_matches_plan_execute(L, field, big, small, inverse):
    for log_n in (big, small):  # the second call reuses the workspace with a smaller input
        say("noalloc field", field, "log_n", log_n, "inverse", inverse)
        x, ref, got = run_both(L, field, log_n, inverse, 0x0A110C + log_n + (100 if inverse else 0))
        assert np.array_equal(got, ref), (field, log_n, inverse)
        assert not np.array_equal(got, x)


def test_noalloc_in_place_and_single_pass(L):
    """d_out == d_in, and a one-pass size (no workspace at all)."""
    for field, log_n in ((g.FIELD_MNT6753_FR, 16), (g.FIELD_FP32, 10)):
        say("noalloc in place field", field, "log_n", log_n)
        x = synth.uniform_residues(0x1A2B + log_n, 1 << log_n, g.MODULI[field], g.NLIMBS[field])
        plan = g.Plan(field, log_n)
        d = g.DeviceArray.from_numpy(x)
        g.check(L.gsn_plan_execute_noalloc(plan.handle, vp(d.ptr), vp(d.ptr), None))
        g.synchronize()
        assert np.array_equal(d.download(x.shape), g.best_fft(x, None, field))
        plan.close()


def test_noalloc_refuses_what_does_not_fit(L):
    d = g.DeviceArray(96)
    big = g.Plan(g.FIELD_FP32, 24)  # 3 passes: two 64 MiB workspaces
    assert L.gsn_plan_workspace_bytes(big.handle) > LIMIT
    st = L.gsn_plan_execute_noalloc(big.handle, vp(d.ptr), vp(d.ptr), None)
    assert st == 1, st  # GSN_ERR_INVALID_ARGUMENT, nothing launched
    assert b"workspace" in g.lib().gsn_last_error()
    big.close()
    ext = g.Plan.extended(g.FIELD_MNT4753_FR, 10)
    assert L.gsn_plan_execute_noalloc(ext.handle, vp(d.ptr), vp(d.ptr), None) == 1
    ext.close()
    assert L.gsn_plan_execute_noalloc(None, vp(d.ptr), vp(d.ptr), None) == 1
