#!/usr/bin/env python3
"""Writes tests/golden/*.npz: field-operation and NTT golden vectors.

Pure-Python big-integer arithmetic; nothing from the reference tree is
imported or run.  The NTT vectors are O(n^2) evaluations of the DFT
definition X_k = sum_j a_j * omega^(j*k) mod p, i.e. the transform that
best_fft (cuda/fft_kernel.h:24-25) is meant to compute.  Inputs are the
seeded residues of gpusnarks_amd.synth, interpreted as Montgomery-form
representations; because the DFT is linear, the expected outputs are the
DFT of those representations (again Montgomery form).

    python3 tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from gpusnarks_amd import synth  # noqa: E402

FIELDS = {
    # id: (name, p, nlimbs, qnr) -- moduli from cuda/device_field.h:443-446
    0: ("fp32", 3 * 2**30 + 1, 1, 5),
    1: ("mnt4753_fr", int(
        "41898490967918953402344214791240637128170709919953949071783502921025352812571106773058893763790338921418070971888253786114353726529584385201591605722013126468931404347949840543007986327743462853720628051692141265303114721689601"), 24, 13),
    2: ("mnt6753_fr", int(
        "41898490967918953402344214791240637128170709919953949071783502921025352812571106773058893763790338921418070971888458477323173057491593855069696241854796396165721416325350064441470418137846398469611935719059908164220784476160001"), 24, 11),
}

NTT_SIZES = {0: [1, 2, 4, 8, 64, 256, 1024], 1: [2, 8, 64, 256], 2: [2, 8, 64, 256]}


def arr(vals, nlimbs):
    return np.stack([synth.from_int(v, nlimbs) for v in vals]) if vals else np.zeros((0, nlimbs), np.uint32)


def dft(vals, omega, p):
    n = len(vals)
    out = []
    for k in range(n):
        wk = pow(omega, k, p)
        acc, w = 0, 1
        for j in range(n):
            acc += vals[j] * w
            w = w * wk % p
        out.append(acc % p)
    return out


def field_vectors(fid):
    name, p, L, g = FIELDS[fid]
    R = 1 << (32 * L)
    rinv = pow(R, -1, p)
    rnd = [synth.to_int(r) for r in synth.uniform_residues(0xF1E1D000 + fid, 200, p, L)]
    edge = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, R % p, (R * R) % p]
    a = edge + rnd[:100]
    b = list(reversed(edge)) + rnd[100:200]
    exps = [0, 1, 2, 3, 31, 32, 65535, 2**20 + 7] + [int(x) % (2**32) for x in synth.splitmix64_stream(77 + fid, 100)]
    d = {
        "modulus": synth.from_int(p, L),
        "one": synth.from_int(R % p, L),
        "r2": synth.from_int(R * R % p, L),
        "mprime": np.array([(-pow(p, -1, 1 << 32)) % (1 << 32)], np.uint32),
        "a": arr(a, L), "b": arr(b, L),
        "add": arr([(x + y) % p for x, y in zip(a, b)], L),
        "sub": arr([(x - y) % p for x, y in zip(a, b)], L),
        # Montgomery product a*b*R^-1 mod p
        "mont_mul": arr([x * y * rinv % p for x, y in zip(a, b)], L),
        "exp": np.array(exps, dtype=np.uint64),
        # a is a Montgomery representation: pow in the domain = (a/R)^e * R
        "mont_pow": arr([pow(x * rinv % p, e, p) * R % p for x, e in zip(a, exps)], L),
    }
    return d


def ntt_vectors(fid):
    name, p, L, g = FIELDS[fid]
    R = 1 << (32 * L)
    d = {}
    for n in NTT_SIZES[fid]:
        omega = pow(g, (p - 1) // n, p) if n > 1 else 1
        if n > 1:
            assert pow(omega, n // 2, p) == p - 1
        x = [synth.to_int(r) for r in synth.uniform_residues(0x5EED0000 + 97 * fid + n, n, p, L)]
        fwd = dft(x, omega, p)
        oinv = pow(omega, -1, p)
        ninv = pow(n, -1, p)
        inv = [v * ninv % p for v in dft(x, oinv, p)]
        d["n%d_omega" % n] = synth.from_int(omega * R % p, L)   # Montgomery form
        d["n%d_input" % n] = arr(x, L)
        d["n%d_forward" % n] = arr(fwd, L)
        d["n%d_inverse" % n] = arr(inv, L)
    return d


def main():
    for fid, (name, p, L, g) in FIELDS.items():
        np.savez_compressed(os.path.join(HERE, "field_%s.npz" % name), **field_vectors(fid))
        np.savez_compressed(os.path.join(HERE, "ntt_%s.npz" % name), **ntt_vectors(fid))
        print("wrote", name)


if __name__ == "__main__":
    main()
