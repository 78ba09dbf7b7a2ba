"""Textbook affine arithmetic on y^2 = x^3 + a x + b over F_q with Python
integers (chord and tangent with field inversions): the independent check
of the projective G1 formulas of the oracle and of the device.  Also the
synthetic curves the tests use: b is fixed by a random point, since the
formulas never read b."""
import random

INF = None


def add(P, Q, a, q):
    if P is INF:
        return Q
    if Q is INF:
        return P
    (x1, y1), (x2, y2) = P, Q
    if x1 == x2:
        if (y1 + y2) % q == 0:
            return INF
        lam = (3 * x1 * x1 + a) * pow(2 * y1, -1, q) % q
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, q) % q
    x3 = (lam * lam - x1 - x2) % q
    return (x3, (lam * (x1 - x3) - y1) % q)


def mul(k, P, a, q):
    R = INF
    for bit in bin(k)[2:] if k else "":
        R = add(R, R, a, q)
        if bit == "1":
            R = add(R, P, a, q)
    return R


def curve_with_point(q, a, seed):
    """(b, P0): a random point P0 = (x0, y0) and the b that puts it on the
    curve."""
    rng = random.Random(seed)
    x0, y0 = rng.randrange(1, q), rng.randrange(1, q)
    b = (y0 * y0 - x0 ** 3 - a * x0) % q
    return b, (x0, y0)


def on_curve(P, a, b, q):
    if P is INF:
        return True
    x, y = P
    return (y * y - x ** 3 - a * x - b) % q == 0


def to_projective(P, q, R, L=24):
    """(3, L) limbs of (x R, y R, R) mod q (Z = 1), or the identity (0, 0, 0)."""
    import numpy as np
    from gpusnarks_amd import synth
    if P is INF:
        return np.zeros((3, L), np.uint32)
    return np.stack([synth.from_int(v * R % q, L) for v in (P[0], P[1], 1)])


def to_affine(arr, q, R):
    """Projective limbs (Montgomery form) back to an affine point."""
    from gpusnarks_amd import synth
    rinv = pow(R, -1, q)
    X, Y, Z = (synth.to_int(arr[i]) * rinv % q for i in range(3))
    if Z == 0:
        return INF
    zi = pow(Z, -1, q)
    return (X * zi % q, Y * zi % q)
