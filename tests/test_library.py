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
    # best_fft / best_ifft / multiexp explicit instantiations (mangled C++ names)
    out = os.popen("nm -D --defined-only %s" % g.library_path).read()
    for name in ("best_fftIN6fields10PrimeField", "best_ifftIN6fields10PrimeField",
                 "multiexpIN6fields10PrimeField"):
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


def _header_array(struct, name):
    text = open(os.path.join(ROOT, "include", "gsn_field_params.h")).read()
    body = text[text.index("struct %s {" % struct):]
    body = body[:body.index("\n};")]
    m = re.search(r"\b%s\[\d+\] = \{(.*?)\}" % name, body, re.S)
    return [int(v.strip().rstrip("u"), 16) for v in m.group(1).split(",")]


@pytest.mark.parametrize("struct", ["Mnt4753FrParams", "Mnt6753FrParams"])
def test_quotient_estimation_constants(struct):
    """The constants of the device's 753-bit product (field_core.h): 3p,
    C783 = 2^783 mod p (limbs and 29-bit digits), its wq = floor(C783 2^783 /
    p), and p^-1 mod 2^783; checked with Python integers from the modulus."""
    limbs = lambda v, bits: sum(x << (bits * i) for i, x in enumerate(v))
    p = limbs(_header_array(struct, "MOD"), 32)
    assert limbs(_header_array(struct, "MOD29"), 29) == p
    assert limbs(_header_array(struct, "THREE_P"), 32) == 3 * p
    c = limbs(_header_array(struct, "C783"), 32)
    assert c == (1 << 783) % p and limbs(_header_array(struct, "C783_29"), 29) == c
    assert limbs(_header_array(struct, "C783_WQ"), 29) == (c << 783) // p
    pinv = limbs(_header_array(struct, "PINV783"), 29)
    assert (pinv * p) % (1 << 783) == 1
    # the exact-division identity the table builder uses for every twiddle
    for w in (1, 2, p - 1, 12345678901234567890 % p, (p * 7) // 11):
        rem = (w << 783) % p
        assert ((((1 << 783) - rem) * pinv) % (1 << 783)) == (w << 783) // p
    # the value bound the kernels rely on: 91 p < 2^761 (30 stages of +3p)
    assert 91 * p < (1 << 761) and 3 * p < (1 << 755)


def test_host_entry_rejects_non_canonical_input():
    """gsn_fft_host refuses elements at or above p before any device work."""
    for field in (g.FIELD_FP32, g.FIELD_MNT6753_FR):
        p = g.MODULI[field]
        L = g.NLIMBS[field]
        x = np.zeros((8, L), np.uint32)
        x[3] = [(p >> (32 * k)) & 0xFFFFFFFF for k in range(L)]  # exactly p
        with pytest.raises(g.GsnError) as e:
            g.best_fft(x, None, field)
        assert e.value.status == 1 and "not below p" in g.lib().gsn_last_error().decode()


def test_missing_library_fails_loudly(monkeypatch):
    """No CPU fallback: without libgpusnarks.so the package refuses to run."""
    monkeypatch.setattr(g, "library_path", "/nonexistent/libgpusnarks.so")
    monkeypatch.setattr(g, "_lib", None)
    with pytest.raises(ImportError):
        g.lib()


def test_transform_without_device_is_an_error():
    """On a host with no GPU the transform returns GSN_ERR_NO_DEVICE rather
    than computing anything on the CPU."""
    if g.lib().gsn_device_init(0) == 0:
        pytest.skip("a GPU is present")
    x = np.zeros((16, 1), np.uint32)
    with pytest.raises(g.GsnError) as e:
        g.best_fft(x, None, g.FIELD_FP32)
    assert e.value.status == 5  # GSN_ERR_NO_DEVICE


def test_tuning_table_alone_and_batch_lines(tmp_path):
    """tuning.txt holds the schedule of a transform that runs alone and,
    on 'batch' lines, the one of transforms fused into batches: loading a
    table with both keeps the first for plain plans; a misplaced 'batch'
    word is an error."""
    shipped = os.path.join(os.path.dirname(g.library_path), "tuning.txt")
    try:
        g.tuning_clear()
        g.tuning_load(shipped)
        assert g.tuning_get(g.FIELD_MNT6753_FR, 18) == [(9, 1, 1), (9, 1, 1)]
        t = tmp_path / "t.txt"
        t.write_text("2 14 7,0,1 7,0,1\nbatch 2 14 6,1,1 8,0,1  # fused\n")
        g.tuning_load(str(t))
        assert g.tuning_get(g.FIELD_MNT6753_FR, 14) == [(7, 0, 1), (7, 0, 1)]
        bad = tmp_path / "bad.txt"
        bad.write_text("2 14 batch 7,0,1 7,0,1\n")
        with pytest.raises(g.GsnError):
            g.tuning_load(str(bad))
    finally:
        g.tuning_clear()
        g.tuning_load(shipped)


def test_autotuner_candidates_and_table_merge(tmp_path):
    """The autotuner's 32-bit candidates mix 16- and 32-column passes and
    all pass the library's schedule check; writing a table for a few sizes
    keeps every other size and the batch lines of the existing table."""
    from gpusnarks_amd import autotune as at
    try:
        for log_n in (18, 26, 28):
            cands = at.candidates(g.FIELD_FP32, log_n)
            widths = {tuple(s[1] for s in c) for c in cands}
            assert any(len(set(w)) > 1 for w in widths), log_n  # mixed tile widths
            for c in cands:
                g.tuning_set(g.FIELD_FP32, log_n, c)  # raises on a shape the library refuses
        old = tmp_path / "old.txt"
        old.write_text("# header\n0 20 6,4,4 6,4,4 8,4,4  # kept\n0 26 8,5,4 9,5,4 9,5,4  # replaced\n"
                       "batch 2 18 9,0,1 9,0,1  # fused\n")
        new = tmp_path / "new.txt"
        at.write_table({(g.FIELD_FP32, 26): ([(9, 5, 4), (8, 5, 4), (9, 5, 4)], 0.4, 0.5)}, str(new),
                       keep_from=str(old))
        rows = [l for l in new.read_text().splitlines() if not l.startswith("#")]
        assert rows[0] == "0 20 6,4,4 6,4,4 8,4,4  # kept"
        assert rows[1].startswith("0 26 9,5,4 8,5,4 9,5,4  #")
        assert rows[2] == "batch 2 18 9,0,1 9,0,1  # fused"
        g.tuning_clear()
        g.tuning_load(str(new))
        assert g.tuning_get(g.FIELD_FP32, 26) == [(9, 5, 4), (8, 5, 4), (9, 5, 4)]
    finally:
        g.tuning_clear()
        g.tuning_load(os.path.join(os.path.dirname(g.library_path), "tuning.txt"))
