"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible slip fails it:
  * golden vectors from tests/golden/gen_golden.py (Python big ints, the DFT
    definition evaluated in O(n^2)) -- pins the field ops and the transform;
  * the known-answer values of the reference's own operator test
    (cuda/device_field_operator_test.cpp:88-100,222-320);
  * the reference's test/fft_host.h, compiled unmodified over the oracle's
    arithmetic (oracle/_ref), against the oracle run with the reference's
    literal butterfly -- pins the loop structure (bit reversal, stage order,
    running twiddle, C-way split, scatter) of the restatement;
  * the reference's Scalar add/subtract (cuda/device_field_operators.h:190-204);
  * closed forms: delta, constants, linearity, convolution theorem, inverse.
"""
import os

import numpy as np
import pytest

from gpusnarks_amd import synth
from tests.oracle_lib import FIELD_IDS, NLIMBS, Oracle, RefHarness, REF_SO

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def golden(kind, name):
    return np.load(os.path.join(GOLDEN, "%s_%s.npz" % (kind, name)))


def ints(arr):
    return [synth.to_int(r) for r in np.asarray(arr).reshape(arr.shape[0], -1)]


def limbs(vals, L):
    return np.stack([synth.from_int(v, L) for v in vals])


FIELDS = list(FIELD_IDS.items())


@pytest.mark.parametrize("name,fid", FIELDS)
def test_constants_match_golden(orc, name, fid):
    g = golden("field", name)
    p = orc.params(fid)
    assert np.array_equal(p["mod"], g["modulus"])
    assert np.array_equal(p["one"], g["one"])
    assert np.array_equal(p["r2"], g["r2"])
    assert p["mprime"] == int(g["mprime"][0])
    assert p["two_adicity"] == {0: 30, 1: 15, 2: 30}[fid]


def test_mprime_closed_form(orc):
    # m' * p == -1 mod 2^32; the reference's 0xFFFFFFFF (device_field_operators.h:47) is not it
    for fid in (0, 1, 2):
        p = orc.params(fid)
        assert (int(p["mprime"]) * int(p["mod"][0])) % 2**32 == 2**32 - 1
    assert orc.params(1)["mprime"] == 0xE45E7FFF


@pytest.mark.parametrize("name,fid", FIELDS)
def test_field_ops_match_golden(orc, name, fid):
    g = golden("field", name)
    a, b = g["a"], g["b"]
    assert np.array_equal(orc.add(fid, a, b), g["add"])
    assert np.array_equal(orc.sub(fid, a, b), g["sub"])
    assert np.array_equal(orc.mul(fid, a, b), g["mont_mul"])
    assert np.array_equal(orc.pow(fid, a, g["exp"]), g["mont_pow"])


@pytest.mark.parametrize("fid", [1, 2])
def test_reference_known_answers(orc, fid):
    """KATs of cuda/device_field_operator_test.cpp, in the Montgomery domain."""
    L = NLIMBS[fid]

    def m(v):  # to Montgomery form
        return orc.unary("or_to_mont", fid, synth.from_int(v, L))

    def val(x):  # out of Montgomery form, as an int
        return synth.to_int(orc.unary("or_from_mont", fid, x)[0])

    one = orc.params(fid)["one"][None, :]
    # testAdd :222-231
    assert val(orc.add(fid, m(1234), m(1234))) == 2468
    # test_subtract :233-244
    assert val(orc.sub(fid, m(1234), m(1234))) == 0
    assert np.array_equal(orc.sub(fid, m(1235), m(1234)), one)
    # testMultiply :246-257
    f1 = orc.mul(fid, m(1234), m(1234))
    assert val(f1) == 1522756
    assert val(orc.mul(fid, f1, m(1234))) == 1879080904
    # testPow :278-297 (2^35 lands in limb 1, not limb 22: limbs are little-endian)
    assert np.array_equal(orc.pow(fid, m(2), [0]), one)
    assert val(orc.pow(fid, m(2), [2])) == 4
    assert val(orc.pow(fid, m(4), [10])) == 1048576
    assert val(orc.pow(fid, m(2), [20])) == 1048576
    r35 = orc.unary("or_from_mont", fid, orc.pow(fid, m(2), [35]))[0]
    assert r35[1] == 8 and r35[0] == 0 and not r35[2:].any()
    # testConstructor :301-320: Scalar(1) == one()
    assert np.array_equal(m(1), one)
    # calculateModPrime :88-100: R * R^-1 == 1
    assert val(one) == 1


def test_bitreverse(orc):
    for l in range(0, 13):
        seen = set()
        for k in range(1 << l):
            r = orc.lib.or_bitreverse(k, l)
            assert r == int(format(k, "0%db" % l)[::-1] or "0", 2) if l else r == 0
            assert orc.lib.or_bitreverse(r, l) == k
            seen.add(r)
        assert len(seen) == 1 << l


@pytest.mark.parametrize("name,fid", FIELDS)
def test_serial_fft_matches_dft_golden(orc, name, fid):
    g = golden("ntt", name)
    sizes = sorted(int(k[1:].split("_")[0]) for k in g.files if k.endswith("_input"))
    assert sizes
    for n in sizes:
        x, om = g["n%d_input" % n], g["n%d_omega" % n]
        assert np.array_equal(orc.serial_fft(fid, x, om), g["n%d_forward" % n]), n
        assert np.array_equal(orc.inverse_fft(fid, x, om), g["n%d_inverse" % n]), n


@pytest.mark.parametrize("name,fid", FIELDS)
def test_parallel_equals_serial(orc, name, fid):
    g = golden("ntt", name)
    for n in (8, 64, 256):
        if "n%d_input" % n not in g.files:
            continue
        x, om = g["n%d_input" % n], g["n%d_omega" % n]
        ref = g["n%d_forward" % n]
        for log_cpus in (0, 1, 2, 3, 10):
            out = orc.parallel_fft(fid, x, om, log_cpus)
            assert np.array_equal(out, ref), (n, log_cpus)


def test_parallel_log_cpus_10_full_split(orc):
    # n >= 2^10 so the harness setting log_cpus = 10 (test/main.cpp:71) takes the split path
    fid = 0
    g = golden("ntt", "fp32")
    x, om = g["n1024_input"], g["n1024_omega"]
    assert np.array_equal(orc.parallel_fft(fid, x, om, 10), g["n1024_forward"])
    assert np.array_equal(orc.parallel_fft(fid, x, om, 10, threads=1), g["n1024_forward"])


ref_missing = not os.path.exists(REF_SO)


@pytest.mark.skipif(ref_missing, reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("name,fid", FIELDS)
def test_literal_butterfly_matches_reference_source(orc, name, fid):
    """The oracle with the reference's butterfly equals test/fft_host.h itself."""
    ref = RefHarness()
    g = golden("ntt", name)
    for n in (2, 8, 64, 256):
        if "n%d_input" % n not in g.files:
            continue
        x, om = g["n%d_input" % n], g["n%d_omega" % n]
        assert np.array_equal(orc.serial_fft(fid, x, om, literal=True), ref.serial_fft(fid, x, om)), n
        for log_cpus in (1, 3):
            assert np.array_equal(orc.parallel_fft(fid, x, om, log_cpus, literal=True),
                                  ref.parallel_fft(fid, x, om, log_cpus)), (n, log_cpus)
    # and the corrected butterfly is a different (and the only DFT-correct) transform
    x, om = g["n8_input"], g["n8_omega"]
    assert not np.array_equal(orc.serial_fft(fid, x, om, literal=True), g["n8_forward"])


@pytest.mark.skipif(ref_missing, reason="oracle/_ref not built (reference tree absent)")
def test_reference_scalar_add_sub(orc):
    """Reference Scalar add/sub (device_field_operators.h:190-204) == oracle, MNT4-753 Fr."""
    ref = RefHarness()
    fid = 1
    assert np.array_equal(ref.modulus(), orc.params(fid)["mod"])
    p = synth.to_int(orc.params(fid)["mod"])
    a = synth.uniform_residues(11, 300, p, 24)
    b = synth.uniform_residues(12, 300, p, 24)
    assert np.array_equal(ref.scalar_op("add", a, b), orc.add(fid, a, b))
    assert np.array_equal(ref.scalar_op("sub", a, b), orc.sub(fid, a, b))


@pytest.mark.parametrize("fid,n", [(0, 64), (0, 256), (2, 16), (1, 16)])
def test_closed_forms(orc, fid, n):
    L = NLIMBS[fid]
    p = synth.to_int(orc.params(fid)["mod"])
    R = 1 << (32 * L)
    qnr = {0: 5, 1: 13, 2: 11}[fid]
    w = pow(qnr, (p - 1) // n, p)
    om = synth.from_int(w * R % p, L)
    one = orc.params(fid)["one"]
    # omega^n == 1 and omega^(n/2) == -1
    assert np.array_equal(orc.pow(fid, om[None], [n])[0], one)
    assert synth.to_int(orc.pow(fid, om[None], [n // 2])[0]) == p - R % p
    # delta -> all equal (the representation of 1 maps to itself everywhere)
    delta = np.zeros((n, L), np.uint32)
    delta[0] = one
    assert (orc.serial_fft(fid, delta, om) == one).all()
    # constant c -> (n*c, 0, ..., 0)
    c = synth.uniform_residues(5, 1, p, L)[0]
    const = np.tile(c, (n, 1))
    out = ints(orc.serial_fft(fid, const, om))
    assert out[0] == n * synth.to_int(c) % p and all(v == 0 for v in out[1:])
    # linearity
    x = synth.uniform_residues(6, n, p, L)
    y = synth.uniform_residues(7, n, p, L)
    lhs = orc.serial_fft(fid, orc.add(fid, x, y), om)
    rhs = orc.add(fid, orc.serial_fft(fid, x, om), orc.serial_fft(fid, y, om))
    assert np.array_equal(lhs, rhs)
    # convolution theorem: NTT(x (*) y) == NTT(x) . NTT(y) (Montgomery product)
    xi, yi = ints(x), ints(y)
    rinv = pow(R, -1, p)
    conv = [sum(xi[j] * yi[(k - j) % n] for j in range(n)) * rinv % p for k in range(n)]
    lhs = orc.serial_fft(fid, limbs(conv, L), om)
    rhs = orc.mul(fid, orc.serial_fft(fid, x, om), orc.serial_fft(fid, y, om))
    assert np.array_equal(lhs, rhs)
    # inverse(forward(x)) == x
    assert np.array_equal(orc.inverse_fft(fid, orc.serial_fft(fid, x, om), om), x)


def _eval_poly(coeffs, x, p):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % p
    return acc


@pytest.mark.parametrize("fid,m", [(0, 8), (0, 32), (1, 16), (2, 16)])
def test_extended_domain_matches_evaluation(orc, fid, m):
    """The extended radix-2 domain H u g*H against direct evaluation of the
    polynomial at the 2m points (w^k, then g w^k, g the field's QNR), and the
    inverse against the coefficients."""
    L = NLIMBS[fid]
    p = synth.to_int(orc.params(fid)["mod"])
    R = 1 << (32 * L)
    qnr = {0: 5, 1: 13, 2: 11}[fid]
    w = pow(qnr, (p - 1) // m, p)
    mont = lambda v: synth.from_int(v * R % p, L)
    val = lambda a: synth.to_int(a) * pow(R, -1, p) % p
    c = synth.uniform_residues(0xE1 + m, 2 * m, p, L)
    cv = [val(r) for r in c]
    out = orc.extended_fft(fid, c, mont(w), mont(qnr))
    points = [pow(w, k, p) for k in range(m)] + [qnr * pow(w, k, p) % p for k in range(m)]
    assert [val(r) for r in out] == [_eval_poly(cv, x, p) for x in points]
    assert np.array_equal(orc.extended_fft(fid, out, mont(w), mont(qnr), inverse=True), c)


def test_extended_domain_mnt4_full_size(orc):
    """MNT4-753 Fr (2-adicity 15) on 2^16 points: spot evaluations of the
    2^16-coefficient polynomial, and the round trip."""
    fid, m = 1, 1 << 15
    L = NLIMBS[fid]
    p = synth.to_int(orc.params(fid)["mod"])
    R = 1 << (32 * L)
    w = pow(13, (p - 1) // m, p)
    assert pow(w, m // 2, p) == p - 1
    mont = lambda v: synth.from_int(v * R % p, L)
    rinv = pow(R, -1, p)
    c = synth.uniform_residues(0xE4, 2 * m, p, L)
    out = orc.extended_fft(fid, c, mont(w), mont(13))
    cv = [synth.to_int(r) * rinv % p for r in c]
    for k in (0, 1, 12345, m - 1):
        assert synth.to_int(out[k]) * rinv % p == _eval_poly(cv, pow(w, k, p), p)
        assert synth.to_int(out[m + k]) * rinv % p == _eval_poly(cv, 13 * pow(w, k, p) % p, p)
    assert np.array_equal(orc.extended_fft(fid, out, mont(w), mont(13), inverse=True), c)


# ---------------------------------------------------------------- G1 curves
from tests import g1_affine as aff  # noqa: E402

# (name, base-field id, coeff a, scalar-field id): MNT4-753 G1 lies over
# MNT6-753 Fr (= MNT4-753 Fq) with a = 2, MNT6-753 G1 over MNT4-753 Fr with
# a = 11 (the reference's coeff_a, cuda/device_field.h:300)
G1_CURVES = [("mnt4753_g1", 2, 2, 1), ("mnt6753_g1", 1, 11, 2)]


def _g1_setup(orc, fid, a, seed):
    q = synth.to_int(orc.params(fid)["mod"])
    R = 1 << 768
    b, P0 = aff.curve_with_point(q, a, seed)
    a_m = synth.from_int(a * R % q, 24)
    return q, R, b, P0, a_m


@pytest.mark.parametrize("name,fid,a,sfid", G1_CURVES)
def test_g1_group_law_matches_affine(orc, name, fid, a, sfid):
    """The oracle's projective add / dbl / [k]P / MSM against textbook affine
    arithmetic, including P + P, P + (-P), the identity, and results
    re-checked on the curve."""
    q, R, b, P0, a_m = _g1_setup(orc, fid, a, 0x61 + fid)
    Ps = [aff.mul(k, P0, a, q) for k in (1, 2, 3, 12345, 2 ** 200 + 7)]
    proj = [aff.to_projective(P, q, R) for P in Ps]
    # scale one representative: (X, Y, Z) and (cX, cY, cZ) are the same point
    c = 0xC0FFEE
    proj[3] = orc.mul(fid, proj[3], np.tile(synth.from_int(c * R % q, 24), (3, 1)))
    for i in range(len(Ps)):
        for j in range(len(Ps)):
            got = aff.to_affine(orc.g1_add(fid, a_m, proj[i], proj[j]), q, R)
            want = aff.add(Ps[i], Ps[j], a, q)
            assert got == want and aff.on_curve(got, a, b, q), (i, j)
        assert aff.to_affine(orc.g1_dbl(fid, a_m, proj[i]), q, R) == aff.add(Ps[i], Ps[i], a, q)
    neg = proj[2].copy()
    neg[1] = orc.sub(fid, np.zeros((1, 24), np.uint32), neg[1:2])[0]
    assert aff.to_affine(orc.g1_add(fid, a_m, proj[2], neg), q, R) is aff.INF
    zero = np.zeros((3, 24), np.uint32)
    assert aff.to_affine(orc.g1_add(fid, a_m, zero, proj[1]), q, R) == Ps[1]
    assert aff.to_affine(orc.g1_add(fid, a_m, zero, zero), q, R) is aff.INF
    r = synth.to_int(orc.params(sfid)["mod"])
    for k in (0, 1, 2, 3, 5, r - 1, 0x1234567890ABCDEF1234567890ABCDEF % r):
        got = aff.to_affine(orc.g1_mul(fid, a_m, proj[0], synth.from_int(k, 24)), q, R)
        assert got == aff.mul(k, Ps[0], a, q), k


@pytest.mark.parametrize("name,fid,a,sfid", G1_CURVES)
def test_g1_msm_matches_affine(orc, name, fid, a, sfid):
    q, R, b, P0, a_m = _g1_setup(orc, fid, a, 0x71 + fid)
    r = synth.to_int(orc.params(sfid)["mod"])
    import random
    rng = random.Random(5)
    ks = [rng.randrange(r) for _ in range(5)] + [0]
    Ps = [aff.mul(rng.randrange(1, 2 ** 64), P0, a, q) for _ in range(6)]
    want = aff.INF
    for k, P in zip(ks, Ps):
        want = aff.add(want, aff.mul(k, P, a, q), a, q)
    pts = np.stack([aff.to_projective(P, q, R) for P in Ps])
    got = orc.g1_msm(fid, a_m, pts, np.stack([synth.from_int(k, 24) for k in ks]))
    assert aff.to_affine(got, q, R) == want


# ---------------------------------------------------------------- mixed radix
def _odd_root(fid):
    from tests.oracle_lib import odd_root
    return odd_root(fid)


@pytest.mark.parametrize("fid", [0, 1, 2])
def test_odd_roots_of_unity(fid):
    """ROOT_ODD has order exactly ODD_MAX (the 3/5/7-smooth part of p - 1)."""
    orc = Oracle()
    p = synth.to_int(orc.params(fid)["mod"])
    M, w = _odd_root(fid)
    assert (p - 1) % M == 0 and pow(w, M, p) == 1
    for l in (3, 5, 7):
        if M % l == 0:
            assert pow(w, M // l, p) != 1
    assert M == {0: 3, 1: 3 * 25 * 7, 2: 9 * 625 * 49}[fid]


@pytest.mark.parametrize("fid,n", [(0, 3), (0, 12), (1, 15), (1, 21), (2, 9), (2, 25)])
def test_dft_direct_matches_definition(orc, fid, n):
    L = NLIMBS[fid]
    p = synth.to_int(orc.params(fid)["mod"])
    R = 1 << (32 * L)
    M, wo = _odd_root(fid)
    m = n
    while m % 2 == 0:
        m //= 2
    n2 = n // m
    w2 = pow({0: 5, 1: 13, 2: 11}[fid], (p - 1) // n2, p)
    w = w2 * pow(wo, M // m, p) % p
    x = synth.uniform_residues(0xD0 + n, n, p, L)
    xv = [synth.to_int(r) * pow(R, -1, p) % p for r in x]
    out = orc.dft_direct(fid, x, synth.from_int(w * R % p, L))
    want = [sum(xv[j] * pow(w, j * k, p) for j in range(n)) % p for k in range(n)]
    assert [synth.to_int(r) * pow(R, -1, p) % p for r in out] == want


@pytest.mark.parametrize("fid,n1,n2", [(0, 3, 16), (1, 15, 8), (1, 105, 4), (1, 525, 2), (2, 9, 8), (2, 245, 4),
                                       (1, 7, 1)])
def test_mixed_fft_matches_direct_dft(orc, fid, n1, n2):
    """The four-step restatement equals the O(n^2) definition, and its
    inverse undoes it."""
    L = NLIMBS[fid]
    p = synth.to_int(orc.params(fid)["mod"])
    R = 1 << (32 * L)
    M, wo = _odd_root(fid)
    w1 = pow(wo, M // n1, p)
    w2 = pow({0: 5, 1: 13, 2: 11}[fid], (p - 1) // n2, p)
    mont = lambda v: synth.from_int(v * R % p, L)
    n = n1 * n2
    x = synth.uniform_residues(0xE0 + n, n, p, L)
    got = orc.mixed_fft(fid, x, n1, mont(w2), mont(w1))
    assert np.array_equal(got, orc.dft_direct(fid, x, mont(w1 * w2 % p)))
    assert np.array_equal(orc.mixed_fft(fid, got, n1, mont(w2), mont(w1), inverse=True), x)


@pytest.mark.parametrize("fid,log_n", [(0, 1), (0, 6), (0, 9), (1, 5), (2, 7)])
@pytest.mark.parametrize("threads", [1, 3])
def test_dft_points_matches_dft_golden_and_direct(orc, fid, log_n, threads):
    """or_dft_points (X_k at chosen k, O(n) each, Horner over chunks of j)
    equals the O(n^2) definition at every k; one and three threads cut j
    into 4 and 12 chunks (one element per chunk at n = 2)."""
    from tests.oracle_lib import primitive_root
    n = 1 << log_n
    import gpusnarks_amd as g
    x = synth.uniform_residues(0xD0D0 + log_n, n, g.MODULI[fid], NLIMBS[fid])
    om = primitive_root(fid, log_n)
    want = orc.dft_direct(fid, x, om)
    got = orc.dft_points(fid, x, om, np.arange(n), threads)
    assert np.array_equal(got, want)
    # indices past n wrap (omega^n = 1)
    assert np.array_equal(orc.dft_points(fid, x, om, [n + 3 % n], threads)[0], want[3 % n])


@pytest.mark.parametrize("name,fid", FIELDS)
def test_dft_points_against_golden(orc, name, fid):
    gv = golden("ntt", name)
    n = max(int(k[1:].split("_")[0]) for k in gv.files if k.endswith("_input"))
    x, om, X = gv["n%d_input" % n], gv["n%d_omega" % n], gv["n%d_forward" % n]
    ks = np.array([0, 1, n // 2, n - 1, 5 % n], np.uint64)
    assert np.array_equal(orc.dft_points(fid, x, om, ks), X[ks.astype(np.int64)])


def test_spot_check_catches_what_a_round_trip_cannot(orc):
    """Two wrong transforms that invert themselves exactly -- a transform
    over w^3 instead of w, and the right outputs in a permuted order -- pass
    inverse(forward(x)) == x but fail the spot evaluation; the true one
    passes it (the mistakes are planted in the oracle's output, on the CPU)."""
    from tests.oracle_lib import primitive_root, spot_indices, spot_values
    import gpusnarks_amd as g
    fid, log_n = 0, 12
    n = 1 << log_n
    x = synth.uniform_residues(0x5B07, n, g.MODULI[fid], 1)
    om = primitive_root(fid, log_n)
    ks = spot_indices(n, 16, 1)
    want = spot_values(orc, fid, x, om, ks)
    good = orc.serial_fft(fid, x, om)
    assert np.array_equal(good[ks.astype(np.int64)], want)
    om3 = orc.pow(fid, om[None], [3])[0]          # also a primitive 2^12-th root
    wrong_root = orc.serial_fft(fid, x, om3)
    assert np.array_equal(orc.inverse_fft(fid, wrong_root.copy(), om3), x)
    assert not np.array_equal(wrong_root[ks.astype(np.int64)], want)
    perm = np.roll(good, 1, axis=0)                 # a permutation its own inverse would undo
    assert not np.array_equal(perm[ks.astype(np.int64)], want)
    # inverse spot values: n^-1 sum x_j w^-jk equals the inverse transform
    inv = orc.inverse_fft(fid, x.copy(), om)
    assert np.array_equal(inv[ks.astype(np.int64)], spot_values(orc, fid, x, om, ks, inverse=True))


def test_primitive_root_helper_matches_library_table():
    """The checker's roots (Python integers from the smallest non-residue)
    have exact order 2^k and equal the library's root table."""
    import gpusnarks_amd as g
    from tests.oracle_lib import primitive_root
    for fid, k in ((0, 26), (2, 22), (1, 15)):
        p, L = g.MODULI[fid], NLIMBS[fid]
        w = synth.to_int(primitive_root(fid, k)) * pow(1 << (32 * L), -1, p) % p
        assert pow(w, 1 << k, p) == 1 and pow(w, 1 << (k - 1), p) == p - 1
        assert np.array_equal(primitive_root(fid, k), g.root_of_unity(fid, k))
