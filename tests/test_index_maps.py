"""The kernels' index maps and twiddle schedules, restated in Python over a
small prime field and checked against the O(n^2) DFT (CPU only).

* 753-bit tiles (gpusnarks_amd/csrc/ntt_kernels.h, tile753_body): in-place
  radix-2 decimation in time inside an R x W tile (bit-reversed rows in,
  natural rows out) with the pass twiddle folded into the steps -- step s
  multiplies x1 by w_{Ns 2^(s+1)}^(k + Ns kk), kk = jj mod 2^s -- the
  exchange between steps (threads t and t ^ 2^tau trade through LDS slots,
  within one wavefront when tau < 6), and the per-pass step-table layout of
  build_steps753_kernel.
* 32-bit tiles (pass32_body): radix-E register steps with a first step of
  radix E0 and pass twiddles formed as A B^g C^e.
"""
import random

import pytest

P = 3 * 2**30 + 1
G = 5


def dft(a, w):
    n = len(a)
    return [sum(a[j] * pow(w, j * k, P) for j in range(n)) % P for k in range(n)]


def step_table(log_n, log_ns, log_r, w):
    """build_steps753_kernel: record i of the pass's table."""
    ns, nrec = 1 << log_ns, ((1 << log_r) - 1) << log_ns
    tab = []
    for i in range(nrec):
        s = ((i >> log_ns) + 1).bit_length() - 1
        jx = i - (((1 << s) - 1) << log_ns)
        k, kk = jx >> s, jx & ((1 << s) - 1)
        e = (k + (kk << log_ns)) << (log_n - log_ns - s - 1)
        assert e < (1 << log_n) // 2
        tab.append(pow(w, e, P))
    return tab


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def folded_radix2(a, w, passes, log_w):
    n = len(a)
    log_n = n.bit_length() - 1
    cur, log_ns = list(a), 0
    for log_r in passes:
        R, Ns, nR = 1 << log_r, 1 << log_ns, n >> log_r
        lw = min(log_w, log_n - log_r)
        W = 1 << lw
        active = R * W // 2  # threads t = jj W + c
        tab = step_table(log_n, log_ns, log_r, w)
        out = [None] * n
        for blk in range(nR // W):
            j0 = blk * W
            x = []
            for t in range(active):
                c, jr = t & (W - 1), brev(t >> lw, log_r - 1)
                x.append([cur[j0 + c + jr * nR], cur[j0 + c + (jr + R // 2) * nR]])
            for s in range(log_r):
                if s:
                    # output e of thread t -> LDS element owner, or active +
                    # owner when bit tau of t is set
                    tau = lw + s - 1
                    m = 1 << tau
                    where = lambda o, slot: active + o if slot else o
                    lds = [None] * (2 * active)
                    for t in range(active):
                        # the partner is in the same 64-lane wavefront unless tau >= 6
                        assert ((t ^ m) >> 6 == t >> 6) == (tau < 6)
                        for e, owner in enumerate((t & ~m, t | m)):
                            i = where(owner, t & m)
                            # an exchange without a barrier touches only its
                            # wavefront's slots: o and active + o for o of the
                            # same wavefront (64 threads)
                            if tau < 6:
                                assert (i if i < active else i - active) >> 6 == t >> 6
                            assert lds[i] is None
                            lds[i] = x[t][e]
                    x = [[lds[where(t, 0)], lds[where(t, 1)]] for t in range(active)]
                for t in range(active):
                    c, jj = t & (W - 1), t >> lw
                    k = (j0 + c) & (Ns - 1)
                    kk = jj & ((1 << s) - 1)
                    if s == 0:
                        tw = tab[k] if log_ns else 1          # record k of step 0
                    else:
                        tw = tab[(((1 << s) - 1) << log_ns) + (k << s) + kk]
                    x0, x1 = x[t][0], x[t][1] * tw % P
                    x[t] = [(x0 + x1) % P, (x0 - x1) % P]
            # thread t holds output rows jj and jj + R/2 of column c
            for t in range(active):
                c, jj = t & (W - 1), t >> lw
                j = j0 + c
                k = j & (Ns - 1)
                for e in range(2):
                    out[(j // Ns) * Ns * R + k + (jj + e * (R // 2)) * Ns] = x[t][e]
        cur, log_ns = out, log_ns + log_r
    return cur


def radix_e(a, w, passes, log_w, log_e):
    n = len(a)
    cur, log_ns = list(a), 0
    for log_r in passes:
        R, Ns, nR = 1 << log_r, 1 << log_ns, n >> log_r
        W = min(1 << log_w, nR)
        E = 1 << min(log_e, log_r)
        log_e0 = (log_r % log_e) or log_e if log_r >= log_e else log_r
        E0, G0 = 1 << log_e0, E >> log_e0
        wp = pow(w, n // (Ns * R), P)                       # w_{Ns R}
        out = [None] * n
        for blk in range(nR // W):
            j0 = blk * W
            tile = [[cur[j0 + c + r * nR] for c in range(W)] for r in range(R)]
            # step 0 with pass twiddles A B^g C^e, then radix-E0 DFTs
            new = [[None] * W for _ in range(R)]
            for c in range(W):
                k = (j0 + c) & (Ns - 1)
                for jj in range(R // E):
                    A, B, C = pow(wp, jj * k, P), pow(wp, (R // E) * k, P), pow(wp, (R // E0) * k, P)
                    for g in range(G0):
                        tg = A * pow(B, g, P) % P
                        col = jj + g * (R // E)
                        xs = [tile[col + e * (R // E0)][c] * tg * pow(C, e, P) % P for e in range(E0)]
                        X = dft(xs, pow(w, n // E0, P))
                        for e in range(E0):
                            new[col * E0 + e][c] = X[e]
            tile, ns2 = new, E0
            while ns2 < R:                                  # radix-E steps, Stockham in the tile
                new = [[None] * W for _ in range(R)]
                for c in range(W):
                    for jj in range(R // E):
                        kk = jj & (ns2 - 1)
                        xs = [tile[jj + e * (R // E)][c] * pow(w, e * kk * (n // (ns2 * E)), P) % P
                              for e in range(E)]
                        X = dft(xs, pow(w, n // E, P))
                        for e in range(E):
                            new[(jj // ns2) * ns2 * E + kk + e * ns2][c] = X[e]
                tile, ns2 = new, ns2 * E
            for r in range(R):
                for c in range(W):
                    j = j0 + c
                    k = j & (Ns - 1)
                    out[(j // Ns) * Ns * R + k + r * Ns] = tile[r][c]
        cur, log_ns = out, log_ns + log_r
    return cur


@pytest.mark.parametrize("log_n,passes,log_w", [(4, [4], 0), (6, [3, 3], 1), (7, [3, 4], 2), (8, [4, 4], 0),
                                                (8, [2, 3, 3], 1), (9, [5, 4], 3), (9, [3, 3, 3], 0),
                                                (1, [1], 0), (10, [9, 1], 0), (10, [6, 4], 2)])
def test_folded_radix2_schedule(log_n, passes, log_w):
    rng = random.Random(log_n * 100 + len(passes))
    n = 1 << log_n
    a = [rng.randrange(P) for _ in range(n)]
    w = pow(G, (P - 1) // n, P)
    assert folded_radix2(a, w, passes, log_w) == dft(a, w)


@pytest.mark.parametrize("log_n,passes,log_w,log_e", [(6, [6], 0, 2), (7, [3, 4], 1, 2), (8, [5, 3], 2, 2),
                                                      (9, [4, 5], 1, 2), (8, [2, 3, 3], 1, 2)])
def test_radix_e_schedule_with_product_twiddles(log_n, passes, log_w, log_e):
    rng = random.Random(log_n * 7 + log_e)
    n = 1 << log_n
    a = [rng.randrange(P) for _ in range(n)]
    w = pow(G, (P - 1) // n, P)
    assert radix_e(a, w, passes, log_w, log_e) == dft(a, w)
