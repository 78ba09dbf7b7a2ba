"""The kernels' index maps and twiddle schedules, restated in Python over a
small prime field and checked against the O(n^2) DFT (CPU only).

* 753-bit tiles (gpusnarks_amd/csrc/ntt_kernels.h, tile753_body): radix-2
  Stockham steps inside an R x W tile with the pass twiddle folded into the
  steps -- step s multiplies x1 by w_{Ns 2^(s+1)}^(k + Ns kk), kk = jj mod 2^s
  -- and the per-pass step-table layout of build_steps753_kernel.
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


def folded_radix2(a, w, passes, log_w):
    n = len(a)
    log_n = n.bit_length() - 1
    cur, log_ns = list(a), 0
    for log_r in passes:
        R, Ns, nR = 1 << log_r, 1 << log_ns, n >> log_r
        W = min(1 << log_w, nR)
        tab = step_table(log_n, log_ns, log_r, w)
        out = [None] * n
        for blk in range(nR // W):
            j0 = blk * W
            tile = [[cur[j0 + c + r * nR] for c in range(W)] for r in range(R)]
            for s in range(log_r):
                ns2 = 1 << s
                new = [[None] * W for _ in range(R)]
                for c in range(W):
                    k = (j0 + c) & (Ns - 1)
                    for jj in range(R // 2):
                        kk = jj & (ns2 - 1)
                        if s == 0:
                            t = tab[k] if log_ns else 1          # record k of step 0
                        else:
                            t = tab[(((1 << s) - 1) << log_ns) + (k << s) + kk]
                        x0, x1 = tile[jj][c], tile[jj + R // 2][c] * t % P
                        row = (jj // ns2) * ns2 * 2 + kk
                        new[row][c] = (x0 + x1) % P
                        new[row + ns2][c] = (x0 - x1) % P
                tile = new
            for r in range(R):
                for c in range(W):
                    j = j0 + c
                    k = j & (Ns - 1)
                    out[(j // Ns) * Ns * R + k + r * Ns] = tile[r][c]
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
                                                (8, [2, 3, 3], 1), (9, [5, 4], 3), (9, [3, 3, 3], 0)])
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
