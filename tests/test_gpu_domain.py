"""GPU parity of the domain helpers (pointwise product, scaling, coset shift,
Montgomery conversions) and of the vector field operators (add, sub, pow)
against the CPU oracle's field operations, bit-exact, and the coset
transform's round trip and operator identities at full size."""
import sys

import numpy as np
import pytest

import gpusnarks_amd as g
from gpusnarks_amd import synth
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu
FIELDS = [g.FIELD_FP32, g.FIELD_MNT4753_FR, g.FIELD_MNT6753_FR]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def residues(field, n, seed):
    x = synth.uniform_residues(seed, n, g.MODULI[field], g.NLIMBS[field])
    # edge values: 0, 1, p - 1, R mod p (Montgomery one)
    p = g.MODULI[field]
    L = g.NLIMBS[field]
    edge = [0, 1, p - 1, (1 << (32 * L)) % p]
    for i, v in enumerate(edge):
        x[i] = [(v >> (32 * k)) & 0xFFFFFFFF for k in range(L)]
    return x


@pytest.mark.parametrize("field", FIELDS)
def test_pointwise_and_scale(orc, field):
    n = 1000  # any n: elementwise
    a, b = residues(field, n, 0xA1 + field), residues(field, n, 0xB2 + field)
    say("pointwise/scale field", field)
    da, db, dout = g.DeviceArray.from_numpy(a), g.DeviceArray.from_numpy(b), g.DeviceArray(a.nbytes)
    g.pointwise_mul(field, dout.ptr, da.ptr, db.ptr, n)
    g.synchronize()
    assert np.array_equal(dout.download(a.shape), orc.mul(field, a, b))
    s = b[7]
    g.scale(field, dout.ptr, da.ptr, n, s)
    g.synchronize()
    assert np.array_equal(dout.download(a.shape), orc.mul(field, a, np.repeat(s[None], n, 0)))
    # in place
    g.pointwise_mul(field, da.ptr, da.ptr, db.ptr, n)
    g.synchronize()
    assert np.array_equal(da.download(a.shape), orc.mul(field, a, b))


@pytest.mark.parametrize("field", FIELDS)
def test_vector_operators(orc, field):
    """gsn_vec_add / gsn_vec_sub / gsn_pointwise_mul / gsn_vec_pow against the
    oracle's operators, element by element, edge values included."""
    n = 1000
    a, b = residues(field, n, 0x1A1 + field), residues(field, n, 0x2B2 + field)
    b[4:8] = a[0:4]  # equal operands: a - a = 0, a + a
    say("vector operators field", field)
    da, db, dout = g.DeviceArray.from_numpy(a), g.DeviceArray.from_numpy(b), g.DeviceArray(a.nbytes)
    g.vec_add(field, dout.ptr, da.ptr, db.ptr, n)
    g.synchronize()
    assert np.array_equal(dout.download(a.shape), orc.add(field, a, b))
    g.vec_sub(field, dout.ptr, da.ptr, db.ptr, n)
    g.synchronize()
    assert np.array_equal(dout.download(a.shape), orc.sub(field, a, b))
    g.vec_sub(field, dout.ptr, db.ptr, da.ptr, n)
    g.synchronize()
    assert np.array_equal(dout.download(a.shape), orc.sub(field, b, a))
    g.pointwise_mul(field, dout.ptr, da.ptr, da.ptr, n)  # squares
    g.synchronize()
    assert np.array_equal(dout.download(a.shape), orc.mul(field, a, a))
    for e in [0, 1, 2, 3, 0xFFFFFFFF, 0x8000000000000001, 0xFFFFFFFFFFFFFFFF, 0x5EED5EED1234]:
        g.vec_pow(field, dout.ptr, da.ptr, n, e)
        g.synchronize()
        assert np.array_equal(dout.download(a.shape), orc.pow(field, a, [e] * n)), hex(e)


def test_vector_identities_full_size():
    """At 2^20 elements, without the oracle: (a + b) - b = a and, for the
    32-bit field, Fermat's x^(p-1) = 1 for x != 0 (0^(p-1) = 0)."""
    field = g.FIELD_FP32
    n = 1 << 20
    p = g.MODULI[field]
    a = synth.uniform_residues(0x3C3, n, p, 1)
    b = synth.uniform_residues(0x4D4, n, p, 1)
    say("vector identities n", n)
    da, db, ds = g.DeviceArray.from_numpy(a), g.DeviceArray.from_numpy(b), g.DeviceArray(a.nbytes)
    g.vec_add(field, ds.ptr, da.ptr, db.ptr, n)
    g.vec_sub(field, ds.ptr, ds.ptr, db.ptr, n)
    g.vec_pow(field, db.ptr, da.ptr, n, p - 1)
    g.synchronize()
    assert np.array_equal(ds.download(a.shape), a)
    one = int(orc_params_one(field))
    expect = np.where(a[:, 0] == 0, 0, one).astype(np.uint32)
    assert np.array_equal(db.download(a.shape)[:, 0], expect)


def orc_params_one(field):
    return (1 << (32 * g.NLIMBS[field])) % g.MODULI[field]


@pytest.mark.parametrize("field", FIELDS)
def test_montgomery_conversions(orc, field):
    n = 777
    x = residues(field, n, 0xC3 + field)  # read as plain residues
    say("montgomery conversions field", field)
    prm = orc.params(field)
    dx, dy = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
    g.to_montgomery(field, dy.ptr, dx.ptr, n)
    g.synchronize()
    mont = dy.download(x.shape)
    # x R = mont(x, R^2) with the oracle's Montgomery product
    assert np.array_equal(mont, orc.mul(field, x, np.repeat(prm["r2"][None], n, 0)))
    g.from_montgomery(field, dx.ptr, dy.ptr, n)
    g.synchronize()
    assert np.array_equal(dx.download(x.shape), x)


@pytest.mark.parametrize("field,log_n", [(g.FIELD_FP32, 10), (g.FIELD_MNT4753_FR, 9), (g.FIELD_MNT6753_FR, 9)])
def test_coset_shift_matches_powers(orc, field, log_n):
    n = 1 << log_n
    x = residues(field, n, 0xD4 + field)
    gen = residues(field, 8, 0xE5 + field)[5]
    say("coset shift field", field, "n", n)
    powers = orc.pow(field, np.repeat(gen[None], n, 0), list(range(n)))
    dx, dy = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
    g.coset_shift(field, dy.ptr, dx.ptr, n, gen)
    g.synchronize()
    shifted = dy.download(x.shape)
    assert np.array_equal(shifted, orc.mul(field, x, powers))
    g.coset_shift(field, dy.ptr, dy.ptr, n, gen, inverse=True)
    g.synchronize()
    assert np.array_equal(dy.download(x.shape), x)


@pytest.mark.parametrize("field,log_n", [(g.FIELD_FP32, 20), (g.FIELD_MNT6753_FR, 16)])
def test_coset_transform_round_trip(field, log_n):
    """Evaluate on the coset g*H (shift, forward NTT), interpolate back
    (inverse NTT, unshift): the identity, at full size."""
    n = 1 << log_n
    x = synth.uniform_residues(0xF6 + log_n, n, g.MODULI[field], g.NLIMBS[field])
    gen = synth.uniform_residues(0x77, 1, g.MODULI[field], g.NLIMBS[field])[0]  # a random nonzero shift
    say("coset round trip field", field, "n", n)
    d = g.DeviceArray.from_numpy(x)
    fwd, inv = g.Plan(field, log_n), g.Plan(field, log_n, inverse=True)
    s = g.Stream()
    g.coset_shift(field, d.ptr, d.ptr, n, gen, stream=s.handle)
    fwd.execute(d.ptr, d.ptr, s.handle)
    s.synchronize()
    evals = d.download(x.shape)
    assert not np.array_equal(evals, x)
    inv.execute(d.ptr, d.ptr, s.handle)
    g.coset_shift(field, d.ptr, d.ptr, n, gen, inverse=True, stream=s.handle)
    s.synchronize()
    assert np.array_equal(d.download(x.shape), x)


def test_domain_argument_errors():
    d = g.DeviceArray(4096)
    with pytest.raises(g.GsnError):
        g.coset_shift(g.FIELD_FP32, d.ptr, d.ptr, 1000, np.array([3], np.uint32))  # n not a power of two
    with pytest.raises(g.GsnError):
        g.coset_shift(g.FIELD_FP32, d.ptr, d.ptr, 1024, np.array([0], np.uint32))  # g = 0
    with pytest.raises(g.GsnError):
        g.pointwise_mul(7, d.ptr, d.ptr, d.ptr, 16)
    with pytest.raises(g.GsnError):
        g.scale(g.FIELD_FP32, d.ptr, d.ptr, 0, np.array([1], np.uint32))
    with pytest.raises(g.GsnError):
        g.vec_add(g.FIELD_FP32, d.ptr, d.ptr, None, 16)
    with pytest.raises(g.GsnError):
        g.vec_pow(5, d.ptr, d.ptr, 16, 3)


def mont(field, v):
    L = g.NLIMBS[field]
    return synth.from_int(v * (1 << (32 * L)) % g.MODULI[field], L)


QNR = {g.FIELD_FP32: 5, g.FIELD_MNT4753_FR: 13, g.FIELD_MNT6753_FR: 11}


@pytest.mark.parametrize("field,log_n,custom_shift", [
    (g.FIELD_FP32, 10, False), (g.FIELD_MNT4753_FR, 9, False), (g.FIELD_MNT6753_FR, 9, True),
    (g.FIELD_FP32, 21, True), (g.FIELD_MNT4753_FR, 16, False)])
def test_extended_domain_matches_oracle(orc, field, log_n, custom_shift):
    """Plan.extended (H u g*H) against the oracle's extended-domain FFT,
    forward and inverse, bit-exact; MNT4-753 Fr at 2^16 is one size past its
    2-adicity."""
    n = 1 << log_n
    x = synth.uniform_residues(0xEE0 + log_n, n, g.MODULI[field], g.NLIMBS[field])
    shift = residues(field, 8, 0x5F + log_n)[6] if custom_shift else mont(field, QNR[field])
    say("extended domain field", field, "log_n", log_n, "custom shift", custom_shift)
    fwd = g.Plan.extended(field, log_n, shift if custom_shift else None)
    inv = g.Plan.extended(field, log_n, shift if custom_shift else None, inverse=True)
    d, e = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
    fwd.execute(d.ptr, e.ptr)
    g.synchronize()
    got = e.download(x.shape)
    omega_half = g.root_of_unity(field, log_n - 1)
    if n <= (1 << 16):
        want = orc.extended_fft(field, x, omega_half, shift)
        assert np.array_equal(got, want)
    inv.execute(e.ptr, e.ptr)  # in place
    g.synchronize()
    assert np.array_equal(e.download(x.shape), x)


def test_extended_domain_errors():
    one = mont(g.FIELD_MNT6753_FR, 1)
    with pytest.raises(g.GsnError):
        g.Plan.extended(g.FIELD_MNT4753_FR, 17)  # 2-adicity 15: at most 2^16 points
    with pytest.raises(g.GsnError):
        g.Plan.extended(g.FIELD_MNT6753_FR, 8, one)  # shift in H: g*H = H
    with pytest.raises(g.GsnError):
        g.Plan.extended(g.FIELD_FP32, 0)


@pytest.mark.parametrize("field,n", [(g.FIELD_FP32, 3001), (g.FIELD_MNT4753_FR, 1), (g.FIELD_MNT4753_FR, 2049),
                                     (g.FIELD_MNT6753_FR, 700)])
def test_inner_product_matches_oracle(orc, field, n):
    """gsn_inner_product (the reference's multiexp over Scalar x Scalar)
    against the oracle's products folded with its additions."""
    a, b = residues(field, max(n, 4), 0x1D0 + n)[:n], residues(field, max(n, 4), 0x2E0 + n)[:n]
    say("inner product field", field, "n", n)
    da, db, dout = g.DeviceArray.from_numpy(a), g.DeviceArray.from_numpy(b), g.DeviceArray(a[:1].nbytes)
    g.inner_product(field, dout.ptr, da.ptr, db.ptr, n)
    g.synchronize()
    prods = orc.mul(field, a, b)
    acc = prods[:1].copy()
    for i in range(1, n):
        acc = orc.add(field, acc, prods[i:i + 1])
    assert np.array_equal(dout.download(acc.shape), acc)


def test_inner_product_reference_harness_shape():
    """The reference's test_multiexp (test/main.cpp:89-132): 2^18 copies of
    Scalar(1234) against 2^18 copies of Scalar(1234); the sum is
    2^18 * 1234^2, in Montgomery form."""
    field, n = g.FIELD_MNT4753_FR, 1 << 18
    p, L = g.MODULI[field], g.NLIMBS[field]
    v = np.tile(mont(field, 1234), (n, 1))
    say("inner product 2^18 x Scalar(1234)")
    da, dout = g.DeviceArray.from_numpy(v), g.DeviceArray(96)
    g.inner_product(field, dout.ptr, da.ptr, da.ptr, n)
    g.synchronize()
    assert np.array_equal(dout.download((1, L))[0], mont(field, n * 1234 * 1234 % p))


def cheap_residues(field, n, seed):
    """Uniform-looking canonical elements without rejection sampling: random
    limbs below p's top limb (for the large sizes)."""
    rng = np.random.default_rng(seed)
    L = g.NLIMBS[field]
    x = rng.integers(0, 1 << 32, size=(n, L), dtype=np.uint64).astype(np.uint32)
    top = g.MODULI[field] >> (32 * (L - 1))
    x[:, L - 1] = (x[:, L - 1].astype(np.uint64) % top).astype(np.uint32)
    return x


def mixed_roots(field, n1, a):
    from tests.oracle_lib import odd_root
    M, w = odd_root(field)
    p = g.MODULI[field]
    return g.root_of_unity(field, a), mont(field, pow(w, M // n1, p))


@pytest.mark.parametrize("field,n1,a", [(g.FIELD_FP32, 3, 10), (g.FIELD_MNT4753_FR, 105, 6),
                                        (g.FIELD_MNT4753_FR, 525, 2), (g.FIELD_MNT4753_FR, 15, 9),
                                        (g.FIELD_MNT4753_FR, 7, 0), (g.FIELD_MNT6753_FR, 245, 3),
                                        (g.FIELD_MNT6753_FR, 9, 5), (g.FIELD_MNT6753_FR, 735, 2),
                                        (g.FIELD_MNT4753_FR, 75, 6)])
def test_mixed_radix_matches_oracle(orc, field, n1, a):
    """Plan.mixed (n = n1 2^a over w = w_{2^a} ROOT_ODD^(ODD_MAX/n1)) against
    the oracle's four-step restatement (itself pinned to the O(n^2) DFT),
    forward and inverse, bit-exact."""
    n = n1 << a
    x = residues(field, n, 0x3A0 + n)
    w2, w1 = mixed_roots(field, n1, a)
    say("mixed radix field", field, "n1", n1, "a", a)
    want = orc.mixed_fft(field, x, n1, w2, w1)
    fwd, inv = g.Plan.mixed(field, n), g.Plan.mixed(field, n, inverse=True)
    d, e = g.DeviceArray.from_numpy(x), g.DeviceArray(x.nbytes)
    fwd.execute(d.ptr, e.ptr)
    g.synchronize()
    assert np.array_equal(e.download(x.shape), want)
    inv.execute(e.ptr, e.ptr)  # in place
    g.synchronize()
    assert np.array_equal(e.download(x.shape), x)


def test_mixed_radix_mnt4_largest():
    """MNT4-753 Fr at n = 2^15 * 525 (17 million points, past its 2-adicity
    by a factor 1050): the inverse undoes the forward, and X_0 = sum x."""
    field, n = g.FIELD_MNT4753_FR, 525 << 15
    p = g.MODULI[field]
    x = cheap_residues(field, n, 0x525)
    say("mixed radix MNT4 n", n)
    fwd, inv = g.Plan.mixed(field, n), g.Plan.mixed(field, n, inverse=True)
    d = g.DeviceArray.from_numpy(x)
    fwd.execute(d.ptr, d.ptr)
    g.synchronize()
    X0 = sum(int(v) << (32 * i) for i, v in enumerate(d.download((1, 24))[0]))
    s = 0
    for i in range(24):  # sum of the representations mod p, limb column by column
        s += int(x[:, i].astype(np.uint64).sum()) << (32 * i)
    assert X0 == s % p
    inv.execute(d.ptr, d.ptr)
    g.synchronize()
    assert np.array_equal(d.download(x.shape), x)


def test_mixed_radix_errors():
    with pytest.raises(g.GsnError):
        g.Plan.mixed(g.FIELD_MNT4753_FR, 1 << 10)  # no odd factor
    with pytest.raises(g.GsnError):
        g.Plan.mixed(g.FIELD_MNT4753_FR, 11 * 4)  # 11 is not a supported factor
    with pytest.raises(g.GsnError):
        g.Plan.mixed(g.FIELD_MNT4753_FR, 1575 * 2)  # 3^2 does not divide p - 1 of MNT4-753 Fr
    with pytest.raises(g.GsnError):
        g.Plan.mixed(g.FIELD_MNT4753_FR, 3 << 16)  # 2^16 exceeds the 2-adicity
