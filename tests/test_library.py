"""CPU-side checks of the C-ABI library: it loads, exports every symbol
include/gpusnarks.h declares, and its host-only logic behaves (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import gpusnarks_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpusnarks.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = g.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_cpp_dropin_symbols_exported():
    # best_fft / best_ifft explicit instantiations (mangled C++ names)
    out = os.popen("nm -D --defined-only %s" % g.library_path).read()
    for name in ("best_fftIN6fields10PrimeField", "best_ifftIN6fields10PrimeField"):
        assert name in out


def test_host_only_queries():
    assert g.lib().gsn_element_bytes(0) == 4 and g.lib().gsn_element_bytes(1) == 96
    assert g.lib().gsn_element_bytes(7) == 0
    assert g.two_adicity(0) == 30 and g.two_adicity(1) == 15 and g.two_adicity(2) == 30
    assert g.lib().gsn_strerror(3).decode().startswith("size exceeds")


@pytest.mark.parametrize("field,name", [(0, "fp32"), (1, "mnt4753_fr"), (2, "mnt6753_fr")])
def test_root_of_unity_matches_golden(field, name):
    gv = np.load(os.path.join(ROOT, "tests", "golden", "ntt_%s.npz" % name))
    for n in (2, 8, 64, 256):
        assert np.array_equal(g.root_of_unity(field, n.bit_length() - 1), gv["n%d_omega" % n])


def test_argument_errors_before_device():
    with pytest.raises(g.GsnError) as e:
        g.best_fft(np.zeros(3, np.uint32), None, g.FIELD_FP32)
    assert e.value.status == 2
    with pytest.raises(g.GsnError) as e:
        g.root_of_unity(g.FIELD_MNT4753_FR, 16)
    assert e.value.status == 3
    with pytest.raises(ValueError):
        g.best_fft(np.zeros((4, 3), np.uint32), None, g.FIELD_MNT6753_FR)


def build_dropin(tmp_path):
    """Compiles tests/cpp/dropin_main.cpp as a caller of the reference's
    entry point would be: g++, include/, libgpusnarks.so, no HIP headers."""
    exe = str(tmp_path / "dropin_main")
    libdir = os.path.dirname(g.library_path)
    cmd = ["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"), "-o", exe,
           "-L" + libdir, "-l:" + os.path.basename(g.library_path), "-Wl,-rpath," + libdir]
    import subprocess
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_drop_in(tmp_path):
    """best_fft<fields::Scalar> etc. compile and link from plain C++; host
    identities hold and bad arguments throw std::runtime_error."""
    import subprocess
    exe = build_dropin(tmp_path)
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host: 0 failure(s)" in r.stdout
