#!/usr/bin/env python3
"""Generate ``include/gsn_field_params.h``: the constants of the three NTT fields.

Pure-Python big-integer arithmetic only.  Run from anywhere:

    python3 gpusnarks_amd/csrc/gen_field_params.py

Fields (moduli as printed in the reference):
  * MNT4-753 Fr  -- cuda/device_field.h:443-444 (decimal), limbs :62-65.
                    2-adicity 15, so radix-2 NTTs only up to n = 2^15.
  * MNT6-753 Fr  -- cuda/device_field.h:445-446 (decimal).  2-adicity 30.
  * Fp32         -- p = 3*2^30 + 1, the build's own 32-bit NTT prime
                    (the reference's 32-bit type, fields/dummy_field.h:24-62,
                    is wrap-around uint32 arithmetic, not a prime field).

Per 753-bit field:
  * R_API = 2^768 is the Montgomery radix of the 24 x u32 element
    representation that the public API and the CPU oracle use
    (cuda/device_field_operators.h:151 multiplies 24 limbs, so its radix is
    2^768).
  * The device multiplies data by constant twiddles with a quotient
    estimated from a precomputed wq = floor(w 2^783 / p) (field_core.h);
    C783 = 2^783 mod p and PINV783 = p^-1 mod 2^783 let the device compute wq
    for every entry of a twiddle table (wq = (2^783 - (w 2^783 mod p)) p^-1
    mod 2^783, an exact division).
"""
import os
import sys

MNT4_FR = int(
    "41898490967918953402344214791240637128170709919953949071783502921025352812571106773058893763790338921418070971888253786114353726529584385201591605722013126468931404347949840543007986327743462853720628051692141265303114721689601")
MNT6_FR = int(
    "41898490967918953402344214791240637128170709919953949071783502921025352812571106773058893763790338921418070971888458477323173057491593855069696241854796396165721416325350064441470418137846398469611935719059908164220784476160001")
FP32 = 3 * 2**30 + 1

# little-endian limbs of MNT4-753 Fr exactly as stored in cuda/device_field.h:62-65
REF_MNT4_LIMBS = [610172929, 1586521054, 752685471, 3818738770, 2596546032, 1669861489,
                  1987204260, 1750781161, 3411246648, 3087994277, 4061660573, 2971133814,
                  2707093405, 2580620505, 3902860685, 134068517, 1821890675, 1589111033,
                  1536143341, 3086587728, 4007841197, 270700578, 764593169, 115910]


def is_probable_prime(n, rounds=40):
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for q in small:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = 0x9E3779B97F4A7C15
    for i in range(rounds):
        x = (x * 6364136223846793005 + 1442695040888963407) % 2**64
        a = 2 + x % (n - 3)
        y = pow(a, d, n)
        if y in (1, n - 1):
            continue
        for _ in range(s - 1):
            y = y * y % n
            if y == n - 1:
                break
        else:
            return False
    return True


def two_adicity(p):
    s, q = 0, p - 1
    while q % 2 == 0:
        q //= 2
        s += 1
    return s


def smallest_qnr(p):
    g = 2
    while pow(g, (p - 1) // 2, p) != p - 1:
        g += 1
    return g


def odd_smooth_root(p):
    """(M, root): M = the largest divisor of p - 1 made of the primes 3, 5, 7,
    and a primitive M-th root of unity (g^((p-1)/M) for the smallest g that
    is no l-th power for any such prime l dividing M)."""
    q, m, primes = p - 1, 1, []
    for l in (3, 5, 7):
        while q % l == 0:
            q //= l
            m *= l
            if l not in primes:
                primes.append(l)
    g = 2
    while any(pow(g, (p - 1) // l, p) == 1 for l in primes):
        g += 1
    root = pow(g, (p - 1) // m, p)
    assert pow(root, m, p) == 1 and all(pow(root, m // l, p) != 1 for l in primes)
    return m, root


def limbs(x, bits, count):
    mask = (1 << bits) - 1
    out = []
    for _ in range(count):
        out.append(x & mask)
        x >>= bits
    assert x == 0, "value does not fit"
    return out


def carr(vals):
    return "{" + ", ".join("0x%08xu" % v for v in vals) + "}"


def field753(name, p):
    assert p.bit_length() == 753 and is_probable_prime(p)
    s = two_adicity(p)
    g = smallest_qnr(p)
    r_api = 1 << 768
    root = pow(g, (p - 1) >> s, p)           # primitive 2^s-th root of unity
    assert pow(root, 1 << (s - 1), p) == p - 1
    d = dict(
        name=name, p=p, s=s, g=g,
        mod32=limbs(p, 32, 24), mod29=limbs(p, 29, 26),
        mp32=(-pow(p, -1, 1 << 32)) % (1 << 32),
        one_api=limbs(r_api % p, 32, 24),                # R_API mod p = Montgomery one
        r2_api=limbs(r_api * r_api % p, 32, 24),         # R_API^2 mod p (to_monty factor)
        threep32=limbs(3 * p, 32, 24),
        c783=(1 << 783) % p,
        pinv783=pow(p, -1, 1 << 783),
        root_api=limbs(root * r_api % p, 32, 24),        # primitive 2^s-th root, API Montgomery form
        root=root,
    )
    d["odd_max"], odd_root = odd_smooth_root(p)
    d["root_odd_api"] = limbs(odd_root * r_api % p, 32, 24)
    return d


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    out_path = os.path.join(here, "..", "..", "include", "gsn_field_params.h")

    assert sum(v << (32 * i) for i, v in enumerate(REF_MNT4_LIMBS)) == MNT4_FR
    f4 = field753("Mnt4753Fr", MNT4_FR)
    f6 = field753("Mnt6753Fr", MNT6_FR)
    assert f4["s"] == 15 and f6["s"] == 30
    assert is_probable_prime(FP32) and two_adicity(FP32) == 30
    g32 = smallest_qnr(FP32)
    mp32 = (-pow(FP32, -1, 1 << 32)) % (1 << 32)
    root32 = pow(g32, (FP32 - 1) >> 30, FP32)

    L = []
    L.append("// Generated by gen_field_params.py -- do not edit by hand.")
    L.append("// Field constants for the NTT fields.  All multi-limb values are little-endian.")
    L.append("#pragma once")
    L.append("#include <cstdint>")
    L.append("")
    L.append("namespace gsn {")
    L.append("")
    L.append("// p = 3*2^30 + 1; Montgomery radix 2^32.")
    L.append("struct Fp32Params {")
    L.append("  static constexpr uint32_t P = 0x%08xu;" % FP32)
    L.append("  static constexpr uint32_t MPRIME = 0x%08xu;        // -p^-1 mod 2^32" % mp32)
    L.append("  static constexpr uint32_t PINV = 0x%08xu;          // p^-1 mod 2^32" % pow(FP32, -1, 1 << 32))
    L.append("  static constexpr uint32_t ONE = 0x%08xu;           // 2^32 mod p" % ((1 << 32) % FP32))
    L.append("  static constexpr uint32_t R2 = 0x%08xu;            // 2^64 mod p" % ((1 << 64) % FP32))
    L.append("  static constexpr uint32_t TWO_ADICITY = 30;")
    L.append("  static constexpr uint32_t GENERATOR = %du;" % g32)
    L.append("  static constexpr uint32_t ROOT_MONT = 0x%08xu;     // primitive 2^30-th root, Montgomery form"
             % (root32 * (1 << 32) % FP32))
    odd32, oroot32 = odd_smooth_root(FP32)
    L.append("  static constexpr uint32_t ODD_MAX = %du;  // 3/5/7-smooth part of p - 1" % odd32)
    L.append("  static constexpr uint32_t ROOT_ODD = 0x%08xu;   // primitive ODD_MAX-th root, Montgomery form"
             % (oroot32 * (1 << 32) % FP32))
    L.append("};")
    L.append("")
    for f in (f4, f6):
        L.append("// %s: p = %d" % (f["name"], f["p"]))
        L.append("struct %sParams {" % f["name"])
        L.append("  static constexpr uint32_t MOD[24] = %s;" % carr(f["mod32"]))
        L.append("  static constexpr uint32_t MOD29[26] = %s;" % carr(f["mod29"]))
        L.append("  static constexpr uint32_t THREE_P[24] = %s;" % carr(f["threep32"]))
        L.append("  static constexpr uint32_t MPRIME = 0x%08xu;   // -p^-1 mod 2^32" % f["mp32"])
        L.append("  static constexpr uint32_t ONE[24] = %s;  // 2^768 mod p" % carr(f["one_api"]))
        L.append("  static constexpr uint32_t R2[24] = %s;  // 2^1536 mod p" % carr(f["r2_api"]))
        L.append("  static constexpr uint32_t C783[24] = %s;  // 2^783 mod p" % carr(limbs(f["c783"], 32, 24)))
        L.append("  static constexpr uint32_t C783_29[26] = %s;  // the same, 29-bit digits" % carr(limbs(f["c783"], 29, 26)))
        L.append("  static constexpr uint32_t C783_WQ[27] = %s;  // floor(C783 2^783 / p), 29-bit digits"
                 % carr(limbs((f["c783"] << 783) // f["p"], 29, 27)))
        L.append("  static constexpr uint32_t PINV783[27] = %s;  // p^-1 mod 2^783, 29-bit digits"
                 % carr(limbs(f["pinv783"], 29, 27)))
        L.append("  static constexpr uint32_t ROOT_MONT[24] = %s;  // primitive 2^%d-th root, Montgomery form"
                 % (carr(f["root_api"]), f["s"]))
        L.append("  static constexpr uint32_t ODD_MAX = %du;  // 3/5/7-smooth part of p - 1" % f["odd_max"])
        L.append("  static constexpr uint32_t ROOT_ODD[24] = %s;  // primitive ODD_MAX-th root, Montgomery form"
                 % carr(f["root_odd_api"]))
        L.append("  static constexpr uint32_t TWO_ADICITY = %du;" % f["s"])
        L.append("  static constexpr uint32_t QNR = %du;" % f["g"])
        L.append("};")
        L.append("")
    L.append("}  // namespace gsn")
    L.append("")
    with open(out_path, "w") as fh:
        fh.write("\n".join(L))
    print("wrote", out_path)


if __name__ == "__main__":
    sys.exit(main())
