"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the
reference-source harness (oracle/_ref/libref_fft_host.so).  Test-side only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_fft_host.so")

FIELD_IDS = {"fp32": 0, "mnt4753_fr": 1, "mnt6753_fr": 2}
NLIMBS = {0: 1, 1: 24, 2: 24}

_u32p = ctypes.POINTER(ctypes.c_uint32)


class OrField(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int), ("nlimbs", ctypes.c_int),
                ("mod", ctypes.c_uint32 * 24), ("mprime", ctypes.c_uint32),
                ("one", ctypes.c_uint32 * 24), ("r2", ctypes.c_uint32 * 24),
                ("two_adicity", ctypes.c_uint)]


def _p(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError("%s is not built (run: make -C oracle)" % path)
        L = ctypes.CDLL(path)
        L.or_field_get.restype = ctypes.POINTER(OrField)
        L.or_field_get.argtypes = [ctypes.c_int]
        for nm in ("or_add", "or_sub", "or_mul"):
            getattr(L, nm).argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, _u32p]
        L.or_pow.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, ctypes.c_uint64]
        L.or_inverse.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p]
        L.or_to_mont.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p]
        L.or_from_mont.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p]
        L.or_bitreverse.restype = ctypes.c_size_t
        L.or_bitreverse.argtypes = [ctypes.c_size_t, ctypes.c_uint]
        L.or_serial_radix2_fft.argtypes = [ctypes.POINTER(OrField), _u32p, ctypes.c_size_t, _u32p, _u32p, ctypes.c_int]
        L.or_parallel_radix2_fft_inner.argtypes = [ctypes.POINTER(OrField), _u32p, ctypes.c_size_t, _u32p,
                                                   ctypes.c_uint, _u32p, ctypes.c_int, ctypes.c_int]
        L.or_parallel_phase1_sample.argtypes = [ctypes.POINTER(OrField), _u32p, ctypes.c_size_t, _u32p,
                                                ctypes.c_uint, ctypes.c_size_t, ctypes.c_size_t, _u32p, ctypes.c_int]
        L.or_inverse_fft.argtypes = [ctypes.POINTER(OrField), _u32p, ctypes.c_size_t, _u32p]
        L.or_dft_direct.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, ctypes.c_size_t, _u32p]
        L.or_dft_points.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, ctypes.c_size_t, _u32p,
                                    ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t, ctypes.c_int]
        L.or_mixed_fft.argtypes = [ctypes.POINTER(OrField), _u32p, ctypes.c_size_t, ctypes.c_size_t, _u32p, _u32p,
                                   ctypes.c_int]
        L.or_g1_add.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, _u32p, _u32p]
        L.or_g1_dbl.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, _u32p]
        L.or_g1_mul.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, _u32p, _u32p, ctypes.c_int]
        L.or_g1_msm.argtypes = [ctypes.POINTER(OrField), _u32p, _u32p, _u32p, _u32p, ctypes.c_int, ctypes.c_size_t]
        L.or_extended_radix2_fft.argtypes = [ctypes.POINTER(OrField), _u32p, ctypes.c_size_t, _u32p, _u32p,
                                             ctypes.c_int]
        self.lib = L

    def field(self, fid):
        return self.lib.or_field_get(fid)

    def params(self, fid):
        f = self.field(fid).contents
        L = f.nlimbs
        return dict(nlimbs=L, mod=np.array(f.mod[:L], np.uint32), mprime=f.mprime,
                    one=np.array(f.one[:L], np.uint32), r2=np.array(f.r2[:L], np.uint32),
                    two_adicity=f.two_adicity)

    def _binop(self, name, fid, a, b):
        a = np.ascontiguousarray(a, np.uint32)
        b = np.ascontiguousarray(b, np.uint32)
        out = np.zeros_like(a)
        f = self.field(fid)
        fn = getattr(self.lib, name)
        for i in range(a.shape[0]):
            fn(f, _p(out[i]), _p(a[i]), _p(b[i]))
        return out

    def add(self, fid, a, b): return self._binop("or_add", fid, a, b)
    def sub(self, fid, a, b): return self._binop("or_sub", fid, a, b)
    def mul(self, fid, a, b): return self._binop("or_mul", fid, a, b)

    def pow(self, fid, a, exps):
        a = np.ascontiguousarray(a, np.uint32)
        out = np.zeros_like(a)
        f = self.field(fid)
        for i in range(a.shape[0]):
            self.lib.or_pow(f, _p(out[i]), _p(a[i]), int(exps[i]))
        return out

    def unary(self, name, fid, a):
        a = np.ascontiguousarray(a, np.uint32).reshape(-1, NLIMBS[fid])
        out = np.zeros_like(a)
        f = self.field(fid)
        for i in range(a.shape[0]):
            getattr(self.lib, name)(f, _p(out[i]), _p(a[i]))
        return out

    def serial_fft(self, fid, x, omega, literal=False):
        a = np.ascontiguousarray(x, np.uint32).copy()
        one = self.params(fid)["one"]
        self.lib.or_serial_radix2_fft(self.field(fid), _p(a), a.shape[0],
                                      _p(np.ascontiguousarray(omega, np.uint32)), _p(one), int(literal))
        return a

    def parallel_fft(self, fid, x, omega, log_cpus, literal=False, threads=0):
        a = np.ascontiguousarray(x, np.uint32).copy()
        one = self.params(fid)["one"]
        self.lib.or_parallel_radix2_fft_inner(self.field(fid), _p(a), a.shape[0],
                                              _p(np.ascontiguousarray(omega, np.uint32)), log_cpus,
                                              _p(one), int(literal), threads)
        return a

    def dft_direct(self, fid, x, omega):
        a = np.ascontiguousarray(x, np.uint32)
        out = np.zeros_like(a)
        self.lib.or_dft_direct(self.field(fid), _p(out), _p(a), a.shape[0], _p(np.ascontiguousarray(omega, np.uint32)))
        return out

    def dft_points(self, fid, x, omega, ks, threads=0):
        """sum_j x[j] omega^(j k) for each k in ks (O(n) per index)."""
        a = np.ascontiguousarray(x, np.uint32)
        k = np.ascontiguousarray(np.asarray(ks, dtype=np.uint64).reshape(-1))
        out = np.zeros((k.shape[0], NLIMBS[fid]), np.uint32)
        self.lib.or_dft_points(self.field(fid), _p(out), _p(a), a.shape[0], _p(np.ascontiguousarray(omega, np.uint32)),
                               k.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), k.shape[0], threads)
        return out

    def mixed_fft(self, fid, x, n1, w2, w1, inverse=False):
        a = np.ascontiguousarray(x, np.uint32).copy()
        self.lib.or_mixed_fft(self.field(fid), _p(a), n1, a.shape[0] // n1, _p(np.ascontiguousarray(w2, np.uint32)),
                              _p(np.ascontiguousarray(w1, np.uint32)), int(inverse))
        return a

    def extended_fft(self, fid, x, omega_half, shift, inverse=False):
        a = np.ascontiguousarray(x, np.uint32).copy()
        self.lib.or_extended_radix2_fft(self.field(fid), _p(a), a.shape[0],
                                        _p(np.ascontiguousarray(omega_half, np.uint32)),
                                        _p(np.ascontiguousarray(shift, np.uint32)), int(inverse))
        return a

    # G1 points: (3, 24) uint32 arrays (X, Y, Z), Montgomery form over the
    # base field `fid`; coeff_a one Montgomery-form element
    def g1_add(self, fid, coeff_a, p, q):
        r = np.zeros((3, 24), np.uint32)
        self.lib.or_g1_add(self.field(fid), _p(np.ascontiguousarray(coeff_a, np.uint32)), _p(r),
                           _p(np.ascontiguousarray(p, np.uint32)), _p(np.ascontiguousarray(q, np.uint32)))
        return r

    def g1_dbl(self, fid, coeff_a, p):
        r = np.zeros((3, 24), np.uint32)
        self.lib.or_g1_dbl(self.field(fid), _p(np.ascontiguousarray(coeff_a, np.uint32)), _p(r),
                           _p(np.ascontiguousarray(p, np.uint32)))
        return r

    def g1_mul(self, fid, coeff_a, p, k_limbs):
        r = np.zeros((3, 24), np.uint32)
        k = np.ascontiguousarray(k_limbs, np.uint32)
        self.lib.or_g1_mul(self.field(fid), _p(np.ascontiguousarray(coeff_a, np.uint32)), _p(r),
                           _p(np.ascontiguousarray(p, np.uint32)), _p(k), k.shape[-1])
        return r

    def g1_msm(self, fid, coeff_a, points, k_limbs):
        r = np.zeros((3, 24), np.uint32)
        pts = np.ascontiguousarray(points, np.uint32)
        k = np.ascontiguousarray(k_limbs, np.uint32)
        self.lib.or_g1_msm(self.field(fid), _p(np.ascontiguousarray(coeff_a, np.uint32)), _p(r), _p(pts), _p(k),
                           k.shape[-1], pts.shape[0])
        return r

    def inverse_fft(self, fid, x, omega):
        a = np.ascontiguousarray(x, np.uint32).copy()
        self.lib.or_inverse_fft(self.field(fid), _p(a), a.shape[0], _p(np.ascontiguousarray(omega, np.uint32)))
        return a


class RefHarness:
    """The reference's test/fft_host.h compiled over the oracle's arithmetic."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.ref_serial_radix2_fft.argtypes = [ctypes.c_int, _u32p, ctypes.c_size_t, _u32p]
        L.ref_parallel_radix2_fft_inner.argtypes = [ctypes.c_int, _u32p, ctypes.c_size_t, _u32p, ctypes.c_uint]
        L.ref_scalar_add.argtypes = [_u32p, _u32p, _u32p]
        L.ref_scalar_sub.argtypes = [_u32p, _u32p, _u32p]
        L.ref_scalar_modulus.argtypes = [_u32p]
        self.lib = L

    def serial_fft(self, fid, x, omega):
        a = np.ascontiguousarray(x, np.uint32).copy()
        assert self.lib.ref_serial_radix2_fft(fid, _p(a), a.shape[0], _p(np.ascontiguousarray(omega, np.uint32))) == 0
        return a

    def parallel_fft(self, fid, x, omega, log_cpus):
        a = np.ascontiguousarray(x, np.uint32).copy()
        assert self.lib.ref_parallel_radix2_fft_inner(fid, _p(a), a.shape[0],
                                                      _p(np.ascontiguousarray(omega, np.uint32)), log_cpus) == 0
        return a

    def scalar_op(self, name, a, b):
        a = np.ascontiguousarray(a, np.uint32)
        b = np.ascontiguousarray(b, np.uint32)
        out = np.zeros_like(a)
        fn = getattr(self.lib, "ref_scalar_" + name)
        for i in range(a.shape[0]):
            fn(_p(out[i]), _p(a[i]), _p(b[i]))
        return out

    def modulus(self):
        out = np.zeros(24, np.uint32)
        self.lib.ref_scalar_modulus(_p(out))
        return out


def odd_root(fid):
    """(M, w): the 3/5/7-smooth part M of p - 1 and the primitive M-th root of
    unity w (a plain integer) that include/gsn_field_params.h holds as
    ODD_MAX / ROOT_ODD."""
    import re
    text = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include",
                             "gsn_field_params.h")).read()
    struct = {0: "Fp32Params", 1: "Mnt4753FrParams", 2: "Mnt6753FrParams"}[fid]
    body = text[text.index("struct %s {" % struct):]
    body = body[:body.index("\n};")]
    M = int(re.search(r"ODD_MAX = (\d+)u", body).group(1))
    m = re.search(r"ROOT_ODD(?:\[24\])? = \{?([^;}]*)\}?;", body).group(1)
    vals = [int(v.strip().rstrip("u"), 16) for v in m.split(",")]
    L = NLIMBS[fid]
    mod = re.search(r"MOD\[24\] = \{([^}]*)\}", body) if L > 1 else None
    p = (sum(int(v.strip().rstrip("u"), 16) << (32 * i) for i, v in enumerate(mod.group(1).split(",")))
         if mod else int(re.search(r"P = (0x[0-9a-f]+)u", body).group(1), 16))
    rinv = pow(1 << (32 * L), -1, p)
    return M, sum(v << (32 * i) for i, v in enumerate(vals)) * rinv % p


# ---------------------------------------------------------------- checkers

def mont(fid, v):
    """The Montgomery form (limbs) of the plain integer v."""
    from gpusnarks_amd import synth
    import gpusnarks_amd as g
    p, L = g.MODULI[fid], NLIMBS[fid]
    return synth.from_int(v * (1 << (32 * L)) % p, L)


def primitive_root(fid, log_n):
    """A primitive 2^log_n-th root of unity in Montgomery form, from the
    field's smallest quadratic non-residue (5, 13, 11) with Python integers:
    independent of the library's own root table."""
    import gpusnarks_amd as g
    p = g.MODULI[fid]
    qnr = {0: 5, 1: 13, 2: 11}[fid]
    return mont(fid, pow(qnr, (p - 1) >> log_n, p))


def oracle_transform(orc, fid, x, inverse, omega):
    """The oracle's forward (serial radix-2, test/fft_host.h:18-54) or
    inverse (omega^-1 then n^-1) transform of x."""
    if inverse:
        return orc.inverse_fft(fid, x, omega)
    return orc.serial_fft(fid, x, omega)


def spot_indices(n, count, seed):
    """k = 0, 1, n/2, n-1 and `count` further indices drawn from a seed."""
    rng = np.random.default_rng(seed)
    ks = [0, 1 % n, n // 2, n - 1] + [int(v) for v in rng.integers(0, n, size=count)]
    return np.array(ks, dtype=np.uint64)


def spot_values(orc, fid, x, omega, ks, inverse=False, threads=0):
    """X_k for k in ks from the definition: sum_j x_j omega^(jk) forward,
    n^-1 sum_j x_j omega^(-jk) inverse (O(n) per index)."""
    n = x.shape[0]
    w = orc.unary("or_inverse", fid, omega)[0] if inverse else np.ascontiguousarray(omega, np.uint32)
    v = orc.dft_points(fid, x, w, ks, threads)
    if inverse:
        nm = mont(fid, n)
        ninv = orc.unary("or_inverse", fid, nm)[0]
        v = orc.mul(fid, v, np.broadcast_to(ninv, v.shape).copy())
    return v
