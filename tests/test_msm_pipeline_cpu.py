"""The G1 bucket MSM's pipeline (gpusnarks_amd/csrc/gsn_msm.hip) replayed on
the CPU, stage by stage with the same index maps, at the shape of the GPU
run that did not finish (MNT4-753 G1, 13 bases, automatic window c = 4, 189
windows): every loop's trip count is bounded and small, every index stays in
its buffer, and the result equals the oracle's MSM (test/multiexp.h:3-14
with a correct scalar multiplication) and textbook affine arithmetic.  Point
operations are the oracle's (or_g1_add / or_g1_dbl, pinned against affine
arithmetic in tests/test_oracle.py)."""
import random

import numpy as np
import pytest

import gpusnarks_amd as g
from gpusnarks_amd import synth
from tests import g1_affine as aff
from tests.oracle_lib import Oracle

R = 1 << 768
SCALAR_BITS = 753
BUCKET_SLICES = 32  # K in gsn_msm.hip
ZERO = np.zeros((3, 24), np.uint32)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def default_window(n):
    """gsn_msm.hip default_window: max(4, min(16, floor(log2 n) - 3))."""
    lg = n.bit_length() - 1
    return max(4, min(16, lg - 3))


def replay_msm(orc, bf, a_m, pts, scalars, c, counts):
    """Each stage of MsmBasesT::build and ::run, in order, asserting the
    bounds the kernels rely on; counts[] collects loop trip counts."""
    n = pts.shape[0]
    windows = (SCALAR_BITS + c - 1) // c
    total = windows * n
    buckets, H = 1 << c, 1 << (c - 1)
    assert total <= 0xFFFFFFFF
    # msm_table_kernel: table[w n + i] = [2^(c w)] P_i
    table = np.zeros((total, 3, 24), np.uint32)
    for i in range(n):
        q = pts[i].copy()
        for w in range(windows):
            table[w * n + i] = q
            if w + 1 < windows:
                for _ in range(c):
                    q = orc.g1_dbl(bf, a_m, q)
                    counts["dbl_table"] += 1
    # msm_digits_kernel: key = digit w of the plain scalar, val = w n + i
    keys = np.zeros(total, np.uint32)
    vals = np.zeros(total, np.uint32)
    mask = (1 << c) - 1
    for i in range(n):
        s = scalars[i]
        for w in range(windows):
            bit = w * c
            keys[w * n + i] = (s >> bit) & mask if bit < 768 else 0
            vals[w * n + i] = w * n + i
    # hipcub radix sort of (key, val) on bits [0, c): stable
    order = np.argsort(keys, kind="stable")
    keys2, vals2 = keys[order], vals[order]
    assert np.all(np.diff(keys2.astype(np.int64)) >= 0)
    # msm_bounds_kernel: start[d] = first position with key >= d, d <= buckets
    start = np.searchsorted(keys2, np.arange(buckets + 1), side="left")
    assert start[0] == 0 and start[buckets] == total and np.all(np.diff(start) >= 0)
    # msm_bucket_kernel: part[d K + k] = sum of table[vals2[j]], j = start[d] + k, + K, ...
    part = np.zeros((buckets * BUCKET_SLICES, 3, 24), np.uint32)
    for t in range(buckets * BUCKET_SLICES):
        d, k = divmod(t, BUCKET_SLICES)
        acc = ZERO.copy()
        if d != 0:
            trips = 0
            for j in range(start[d] + k, start[d + 1], BUCKET_SLICES):
                assert 0 <= vals2[j] < total
                acc = orc.g1_add(bf, a_m, acc, table[vals2[j]])
                trips += 1
            counts["bucket_adds"] += trips
            counts["max_slice_trips"] = max(counts["max_slice_trips"], trips)
        part[t] = acc

    def tree(a, groups, stride):
        half = stride // 2
        while half >= 1:  # the host loop of launches: half = stride/2 .. 1
            for t in range(groups * half):
                base = (t // half) * stride + t % half
                assert base + half < a.shape[0]
                a[base] = orc.g1_add(bf, a_m, a[base], a[base + half])
                counts["tree_adds"] += 1
            half //= 2
            counts["tree_launches"] += 1

    tree(part, buckets, BUCKET_SLICES)
    # msm_gather_bits_kernel: T[b H + e] = B[d], d = e with a 1 inserted at bit b
    T = np.zeros((c * H, 3, 24), np.uint32)
    for t in range(c * H):
        b, e = divmod(t, H)
        d = ((e >> b) << (b + 1)) | (1 << b) | (e & ((1 << b) - 1))
        assert (d >> b) & 1 and d < buckets
        T[t] = part[d * BUCKET_SLICES]
    tree(T, c, H)
    # msm_final_kernel: Horner over the bits, one thread
    r = ZERO.copy()
    for b in range(c - 1, -1, -1):
        r = orc.g1_dbl(bf, a_m, r)
        r = orc.g1_add(bf, a_m, r, T[b * H])
        counts["final_ops"] += 2
    return r


@pytest.mark.parametrize("window", [0, 2, 5])
def test_msm_pipeline_replay_matches_oracle(orc, window):
    """MNT4-753 G1 over MNT6-753 Fr (a = 2), the failing test's points and
    scalars (random points, the identity, a repeated base, P and -P, the
    scalars 0, 1 and r - 1)."""
    curve = g.CURVE_MNT4753_G1
    bf, sf = g.CURVE_FIELDS[curve]
    q, r = g.MODULI[bf], g.MODULI[sf]
    a = 2
    b, P0 = aff.curve_with_point(q, a, 0x91 + curve)
    a_m = synth.from_int(a * R % q, 24)
    rng = random.Random(17 + window)
    aff_pts = [aff.mul(rng.randrange(1, 2 ** 40), P0, a, q) for _ in range(9)]
    aff_pts += [aff.INF, aff_pts[0], aff_pts[0], (aff_pts[1][0], (-aff_pts[1][1]) % q)]
    ks = [rng.randrange(r) for _ in range(len(aff_pts) - 3)] + [0, 1, r - 1]
    rng.shuffle(ks)
    pts = np.stack([aff.to_projective(P, q, R) for P in aff_pts])
    n = len(ks)
    c = window or default_window(n)
    assert (window, c) != (0, 0) and (window != 0 or c == 4)
    counts = dict(dbl_table=0, bucket_adds=0, max_slice_trips=0, tree_adds=0, tree_launches=0, final_ops=0)
    got = replay_msm(orc, bf, a_m, pts, ks, c, counts)
    want = aff.INF
    for k, P in zip(ks, aff_pts):
        want = aff.add(want, aff.mul(k, P, a, q), a, q)
    assert aff.to_affine(got, q, R) == want
    orc_out = orc.g1_msm(bf, a_m, pts, np.stack([synth.from_int(k, 24) for k in ks]))
    assert aff.to_affine(orc_out, q, R) == want
    windows = (SCALAR_BITS + c - 1) // c
    # every loop is short: the table's doubling chain is (windows - 1) c per
    # base, each bucket slice adds at most ceil(n windows / K) points
    assert counts["dbl_table"] == n * (windows - 1) * c
    assert counts["max_slice_trips"] <= -(-n * windows // BUCKET_SLICES)
    assert counts["bucket_adds"] <= n * windows
    assert counts["final_ops"] == 2 * c
    if window == 0:
        # the shape of the GPU run that did not finish
        assert (n, c, windows) == (13, 4, 189)
        assert counts["dbl_table"] == 9776 and counts["tree_launches"] == 5 + 3
