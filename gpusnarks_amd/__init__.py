"""gpusnarks_amd: MI355X (gfx950) finite-field NTT.

Python mirror of the C ABI in include/gpusnarks.h, which sits beneath the
reference's entry point ``best_fft(std::vector<FieldT>&, const FieldT&)``
(cuda/fft_kernel.h:24-25).  Elements are numpy uint32 arrays of shape
(n, limbs): 1 limb for FIELD_FP32, 24 little-endian limbs in Montgomery form
for the 753-bit fields (the layout of fields::Scalar, cuda/device_field.h:75).

All compute runs in the HIP library libgpusnarks.so; importing this module
without the built library raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "FIELD_FP32", "FIELD_MNT4753_FR", "FIELD_MNT6753_FR", "NLIMBS", "ELEMENT_BYTES", "GsnError",
    "lib", "library_path", "best_fft", "best_ifft", "root_of_unity", "two_adicity", "Plan",
    "inject_faults", "MODULI", "DeviceArray", "Stream", "Event", "Batch", "Server",
    "tuning_set", "tuning_get", "tuning_clear", "tuning_load", "device_memory", "synchronize",
]

FIELD_FP32 = 0
FIELD_MNT4753_FR = 1
FIELD_MNT6753_FR = 2
NLIMBS = {FIELD_FP32: 1, FIELD_MNT4753_FR: 24, FIELD_MNT6753_FR: 24}
ELEMENT_BYTES = {f: 4 * l for f, l in NLIMBS.items()}
MODULI = {
    FIELD_FP32: 3 * 2**30 + 1,
    # cuda/device_field.h:443-446
    FIELD_MNT4753_FR: int("41898490967918953402344214791240637128170709919953949071783502921025352812571106773"
                          "05889376379033892141807097188825378611435372652958438520159160572201312646893140434"
                          "7949840543007986327743462853720628051692141265303114721689601"),
    FIELD_MNT6753_FR: int("41898490967918953402344214791240637128170709919953949071783502921025352812571106773"
                          "05889376379033892141807097188845847732317305749159385506969624185479639616572141632"
                          "5350064441470418137846398469611935719059908164220784476160001"),
}

_PKG = os.path.dirname(os.path.abspath(__file__))
# $GSN_LIBRARY selects another build of the library (same-box comparisons)
library_path = os.environ.get("GSN_LIBRARY") or os.path.join(_PKG, "libgpusnarks.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


class BatchItem(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("log_n", ctypes.c_uint), ("inverse", ctypes.c_int),
                ("omega", ctypes.c_void_p), ("d_in", ctypes.c_void_p), ("d_out", ctypes.c_void_p)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_inflight", ctypes.c_int),
                ("num_streams", ctypes.c_int), ("max_batch", ctypes.c_int)]


class GsnError(RuntimeError):
    def __init__(self, status, name, detail):
        super().__init__("%s (status %d): %s" % (name, status, detail))
        self.status = status


def lib():
    """The loaded libgpusnarks.so (ctypes.CDLL)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise ImportError("libgpusnarks.so is not built; run `python3 -m gpusnarks_amd.build`")
    L = ctypes.CDLL(library_path)
    L.gsn_strerror.restype = ctypes.c_char_p
    L.gsn_strerror.argtypes = [ctypes.c_int]
    L.gsn_last_error.restype = ctypes.c_char_p
    L.gsn_element_bytes.restype = ctypes.c_size_t
    L.gsn_element_bytes.argtypes = [ctypes.c_int]
    L.gsn_two_adicity.argtypes = [ctypes.c_int]
    L.gsn_root_of_unity.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_void_p]
    L.gsn_fft_host.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int]
    L.gsn_plan_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_uint,
                                  ctypes.c_void_p, ctypes.c_int]
    L.gsn_plan_create_ext.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_uint,
                                      ctypes.c_void_p, ctypes.c_int]
    L.gsn_plan_create_mixed.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_size_t,
                                        ctypes.c_int]
    L.gsn_plan_destroy.argtypes = [ctypes.c_void_p]
    L.gsn_msm_bases_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_uint, ctypes.c_void_p]
    L.gsn_msm_bases_destroy.argtypes = [ctypes.c_void_p]
    L.gsn_msm_bases_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint),
                                     ctypes.POINTER(ctypes.c_size_t)]
    L.gsn_msm.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.gsn_plan_execute.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.gsn_plan_execute_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.gsn_plan_workspace_bytes.restype = ctypes.c_size_t
    L.gsn_plan_workspace_bytes.argtypes = [ctypes.c_void_p]
    L.gsn_plan_device_bytes.restype = ctypes.c_size_t
    L.gsn_plan_device_bytes.argtypes = [ctypes.c_void_p]
    L.gsn_plan_num_passes.argtypes = [ctypes.c_void_p]
    L.gsn_device_init.argtypes = [ctypes.c_int]
    L.gsn_debug_inject_faults.argtypes = [ctypes.c_int]
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.gsn_device_malloc.argtypes = [ctypes.POINTER(vp), sz]
    L.gsn_device_free.argtypes = [vp]
    L.gsn_device_memory.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
    for nm in ("gsn_memcpy_htod", "gsn_memcpy_dtoh", "gsn_memcpy_dtod"):
        getattr(L, nm).argtypes = [vp, vp, sz, vp]
    L.gsn_stream_create.argtypes = [ctypes.POINTER(vp)]
    L.gsn_stream_destroy.argtypes = [vp]
    L.gsn_stream_synchronize.argtypes = [vp]
    L.gsn_event_create.argtypes = [ctypes.POINTER(vp)]
    L.gsn_event_destroy.argtypes = [vp]
    L.gsn_event_record.argtypes = [vp, vp]
    L.gsn_event_synchronize.argtypes = [vp]
    L.gsn_event_elapsed_ms.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    L.gsn_debug_copy_u32.argtypes = [vp, vp, sz, vp]
    L.gsn_debug_clock_ghz.argtypes = [ctypes.POINTER(ctypes.c_double)]
    L.gsn_debug_mad_rate.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.gsn_pointwise_mul.argtypes = [ctypes.c_int, vp, vp, vp, sz, vp]
    L.gsn_inner_product.argtypes = [ctypes.c_int, vp, vp, vp, sz, vp]
    L.gsn_vec_add.argtypes = [ctypes.c_int, vp, vp, vp, sz, vp]
    L.gsn_vec_sub.argtypes = [ctypes.c_int, vp, vp, vp, sz, vp]
    L.gsn_vec_pow.argtypes = [ctypes.c_int, vp, vp, sz, ctypes.c_uint64, vp]
    L.gsn_scale.argtypes = [ctypes.c_int, vp, vp, sz, vp, vp]
    L.gsn_coset_shift.argtypes = [ctypes.c_int, vp, vp, sz, vp, ctypes.c_int, vp]
    L.gsn_to_montgomery.argtypes = [ctypes.c_int, vp, vp, sz, vp]
    L.gsn_from_montgomery.argtypes = [ctypes.c_int, vp, vp, sz, vp]
    L.gsn_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BatchItem), sz, ctypes.c_int]
    L.gsn_batch_launch.argtypes = [vp, vp]
    L.gsn_batch_destroy.argtypes = [vp]
    L.gsn_batch_execute.argtypes = [ctypes.POINTER(BatchItem), sz, vp]
    L.gsn_batch_num_launches.argtypes = [vp]
    L.gsn_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ServerConfig)]
    L.gsn_server_submit.argtypes = [vp, ctypes.c_int, ctypes.c_uint, ctypes.c_int, vp, vp, vp,
                                    ctypes.POINTER(ctypes.c_uint64)]
    L.gsn_server_submit_device.argtypes = L.gsn_server_submit.argtypes
    L.gsn_server_poll.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    L.gsn_server_wait.argtypes = [vp, ctypes.c_uint64]
    L.gsn_server_ticket_times.argtypes = [vp, ctypes.c_uint64] + [ctypes.POINTER(ctypes.c_double)] * 3
    L.gsn_server_destroy.argtypes = [vp]
    L.gsn_tuning_set.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.gsn_tuning_get.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)]
    L.gsn_tuning_load.argtypes = [ctypes.c_char_p]
    _lib = L
    return L


def check(status):
    if status != 0:
        L = lib()
        raise GsnError(status, L.gsn_strerror(status).decode(), L.gsn_last_error().decode())


def _as_elems(a, field):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    L = NLIMBS[field]
    if a.ndim == 1 and L == 1:
        a = a.reshape(-1, 1)
    if a.ndim != 2 or a.shape[1] != L:
        raise ValueError("expected an (n, %d) uint32 array" % L)
    return a


def _omega_ptr(omega, field):
    if omega is None:
        return None, None
    w = np.ascontiguousarray(np.asarray(omega, dtype=np.uint32).reshape(-1))
    if w.shape[0] != NLIMBS[field]:
        raise ValueError("omega must have %d limbs" % NLIMBS[field])
    return w, w.ctypes.data_as(ctypes.c_void_p)


def best_fft(a, omega=None, field=FIELD_MNT6753_FR, inverse=False):
    """Transform of the rows of `a` (returns a new array); host buffers,
    synchronous -- the semantics of best_fft (cuda/fft_kernel.cu:117-147)."""
    x = _as_elems(a, field).copy()
    w, wp = _omega_ptr(omega, field)
    check(lib().gsn_fft_host(field, x.ctypes.data_as(ctypes.c_void_p), x.shape[0], wp, int(bool(inverse))))
    return x


def best_ifft(a, omega=None, field=FIELD_MNT6753_FR):
    return best_fft(a, omega, field, inverse=True)


# ---- domain helpers on device buffers (pointers), asynchronous on `stream`

def _elem_ptr(v, field):
    v = np.ascontiguousarray(np.asarray(v, dtype=np.uint32).reshape(-1))
    if v.shape[0] != NLIMBS[field]:
        raise ValueError("expected one element of %d limbs" % NLIMBS[field])
    return v, v.ctypes.data_as(ctypes.c_void_p)


def pointwise_mul(field, d_out, d_a, d_b, n, stream=None):
    """d_out[i] = d_a[i] * d_b[i] (Montgomery forms)."""
    check(lib().gsn_pointwise_mul(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_a), ctypes.c_void_p(d_b), n,
                                  stream))


def inner_product(field, d_out, d_a, d_b, n, stream=None):
    """*d_out = sum_i d_a[i] * d_b[i] on device (the reference's multiexp over
    Scalar x Scalar, cuda/multi_exp.cu:104-135)."""
    check(lib().gsn_inner_product(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_a), ctypes.c_void_p(d_b), n,
                                  stream))


def vec_add(field, d_out, d_a, d_b, n, stream=None):
    """d_out[i] = d_a[i] + d_b[i] (Scalar operator+, cuda/device_field.h:124-131)."""
    check(lib().gsn_vec_add(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_a), ctypes.c_void_p(d_b), n,
                            stream))


def vec_sub(field, d_out, d_a, d_b, n, stream=None):
    """d_out[i] = d_a[i] - d_b[i] (Scalar operator-, cuda/device_field.h:133-140)."""
    check(lib().gsn_vec_sub(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_a), ctypes.c_void_p(d_b), n,
                            stream))


def vec_pow(field, d_out, d_in, n, exponent, stream=None):
    """d_out[i] = d_in[i] ** exponent (Scalar operator^, cuda/device_field.h:151-158)."""
    check(lib().gsn_vec_pow(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_in), n, int(exponent), stream))


def scale(field, d_out, d_in, n, s, stream=None):
    """d_out[i] = d_in[i] * s, s one element (Montgomery form)."""
    v, vp_ = _elem_ptr(s, field)
    check(lib().gsn_scale(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_in), n, vp_, stream))


def coset_shift(field, d_out, d_in, n, g, inverse=False, stream=None):
    """d_out[i] = d_in[i] * g^i (g^-i with inverse=True)."""
    v, vp_ = _elem_ptr(g, field)
    check(lib().gsn_coset_shift(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_in), n, vp_, int(bool(inverse)),
                                stream))


def to_montgomery(field, d_out, d_in, n, stream=None):
    check(lib().gsn_to_montgomery(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_in), n, stream))


def from_montgomery(field, d_out, d_in, n, stream=None):
    check(lib().gsn_from_montgomery(field, ctypes.c_void_p(d_out), ctypes.c_void_p(d_in), n, stream))


def root_of_unity(field, log_n):
    out = np.zeros(NLIMBS[field], dtype=np.uint32)
    check(lib().gsn_root_of_unity(field, log_n, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def two_adicity(field):
    return lib().gsn_two_adicity(field)


def inject_faults(count):
    """The next `count` transform launches fail with GSN_ERR_INJECTED_FAULT."""
    check(lib().gsn_debug_inject_faults(count))


class Plan:
    """Device-resident twiddles + workspace for one (field, size, direction)."""

    def __init__(self, field, log_n, omega=None, inverse=False):
        self.field, self.log_n, self.inverse = field, log_n, bool(inverse)
        self._omega, wp = _omega_ptr(omega, field)
        h = ctypes.c_void_p()
        check(lib().gsn_plan_create(ctypes.byref(h), field, log_n, wp, int(self.inverse)))
        self.handle = h

    @classmethod
    def extended(cls, field, log_n, shift=None, inverse=False):
        """Plan over the extended domain H u shift*H (|H| = 2^(log_n - 1)),
        which reaches log_n = 2-adicity + 1 (MNT4-753 Fr at 2^16)."""
        self = cls.__new__(cls)
        self.field, self.log_n, self.inverse = field, log_n, bool(inverse)
        self._omega, sp = _omega_ptr(shift, field)
        h = ctypes.c_void_p()
        check(lib().gsn_plan_create_ext(ctypes.byref(h), field, log_n, sp, int(self.inverse)))
        self.handle = h
        return self

    @classmethod
    def mixed(cls, field, n, inverse=False):
        """Plan for n = n1 * 2^a points (n1 > 1 odd, from the 3/5/7-smooth part
        of p - 1) over w = w_{2^a} * ROOT_ODD^(ODD_MAX / n1)."""
        self = cls.__new__(cls)
        self.field, self.inverse, self._omega = field, bool(inverse), None
        self.log_n, self._n = None, n
        h = ctypes.c_void_p()
        check(lib().gsn_plan_create_mixed(ctypes.byref(h), field, n, int(self.inverse)))
        self.handle = h
        return self

    @property
    def n(self):
        return getattr(self, "_n", None) or 1 << self.log_n

    @property
    def device_bytes(self):
        return lib().gsn_plan_device_bytes(self.handle)

    @property
    def num_passes(self):
        return lib().gsn_plan_num_passes(self.handle)

    def execute(self, d_in, d_out, stream=0):
        """Asynchronous transform of device buffer d_in into d_out (ints)."""
        check(lib().gsn_plan_execute(self.handle, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                                     ctypes.c_void_p(stream)))

    def execute_tensor(self, t_in, t_out=None):
        """Transform torch device tensors (int32 storage of the limbs) on
        torch's current stream."""
        import torch
        t_out = t_in if t_out is None else t_out
        stream = torch.cuda.current_stream().cuda_stream
        self.execute(t_in.data_ptr(), t_out.data_ptr(), stream)
        return t_out

    def close(self):
        if getattr(self, "handle", None):
            lib().gsn_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ G1 MSM

CURVE_MNT4753_G1, CURVE_MNT6753_G1 = 0, 1
# (base field, scalar field) of each curve
CURVE_FIELDS = {CURVE_MNT4753_G1: (FIELD_MNT6753_FR, FIELD_MNT4753_FR),
                CURVE_MNT6753_G1: (FIELD_MNT4753_FR, FIELD_MNT6753_FR)}


class MsmBases:
    """Fixed bases prepared for repeated multi-scalar multiplications (the
    curve form of the reference's multiexp, cuda/multi_exp.cu:140):
    d_points holds n projective points (X, Y, Z), 3 x 24 u32 each."""

    def __init__(self, curve, d_points, n, window_bits=0, stream=None):
        h = ctypes.c_void_p()
        check(lib().gsn_msm_bases_create(ctypes.byref(h), curve, ctypes.c_void_p(d_points), n, window_bits,
                                         stream))
        self.handle, self.curve, self.n = h, curve, n

    def info(self):
        c, b = ctypes.c_uint(), ctypes.c_size_t()
        check(lib().gsn_msm_bases_info(self.handle, ctypes.byref(c), ctypes.byref(b)))
        return c.value, b.value

    def msm(self, d_out, d_scalars, stream=None):
        """*d_out = sum_i [s_i] P_i (one projective point, device memory)."""
        check(lib().gsn_msm(self.handle, ctypes.c_void_p(d_out), ctypes.c_void_p(d_scalars), stream))

    def close(self):
        if getattr(self, "handle", None):
            lib().gsn_msm_bases_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ plumbing

def device_memory():
    """(free, total) device bytes."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().gsn_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def synchronize():
    check(lib().gsn_device_synchronize())


class DeviceArray:
    """A device allocation of `nbytes`, with numpy upload/download."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().gsn_device_malloc(ctypes.byref(p), nbytes))
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a, stream=0):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().gsn_memcpy_htod(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes,
                                    ctypes.c_void_p(stream)))

    def download(self, shape, dtype=np.uint32, stream=0):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().gsn_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), out.nbytes,
                                    ctypes.c_void_p(stream)))
        return out

    def copy_from(self, other, nbytes=None, stream=0):
        nbytes = other.nbytes if nbytes is None else nbytes
        check(lib().gsn_memcpy_dtod(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), nbytes,
                                    ctypes.c_void_p(stream)))

    def free(self):
        if self.ptr:
            lib().gsn_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().gsn_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self):
        check(lib().gsn_stream_synchronize(ctypes.c_void_p(self.handle)))

    def close(self):
        if self.handle:
            lib().gsn_stream_destroy(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """HIP event; elapsed_ms(start, stop) times work between two records."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().gsn_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream=0):
        check(lib().gsn_event_record(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream)))

    def synchronize(self):
        check(lib().gsn_event_synchronize(ctypes.c_void_p(self.handle)))

    @staticmethod
    def elapsed_ms(start, stop):
        ms = ctypes.c_float()
        check(lib().gsn_event_elapsed_ms(ctypes.byref(ms), ctypes.c_void_p(start.handle),
                                         ctypes.c_void_p(stop.handle)))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().gsn_event_destroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass


# ------------------------------------------------------------------ batches

class Batch:
    """Many transforms on device buffers; items are (field, log_n, inverse,
    d_in, d_out[, omega]) with d_in/d_out device pointers (ints)."""

    def __init__(self, items, use_graph=True):
        arr = (BatchItem * len(items))()
        self._omegas = []
        for i, it in enumerate(items):
            field, log_n, inverse, d_in, d_out = it[:5]
            omega = it[5] if len(it) > 5 else None
            w, wp = _omega_ptr(omega, field)
            self._omegas.append(w)
            arr[i] = BatchItem(field, log_n, int(bool(inverse)), wp, d_in, d_out)
        self._items = arr
        h = ctypes.c_void_p()
        check(lib().gsn_batch_create(ctypes.byref(h), arr, len(items), int(bool(use_graph))))
        self.handle = h

    @property
    def num_launches(self):
        return lib().gsn_batch_num_launches(self.handle)

    def launch(self, stream=0):
        check(lib().gsn_batch_launch(self.handle, ctypes.c_void_p(stream)))

    def close(self):
        if getattr(self, "handle", None):
            lib().gsn_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ serving

class Server:
    """Continuous-batching NTT server (host or device operands)."""

    def __init__(self, hbm_budget_bytes=0, max_inflight=0, num_streams=0, max_batch=0):
        cfg = ServerConfig(hbm_budget_bytes, max_inflight, num_streams, max_batch)
        h = ctypes.c_void_p()
        check(lib().gsn_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        self.handle = h
        self._keep = {}

    def submit(self, a, field, inverse=False, omega=None, out=None):
        """Queues the transform of host array `a`; returns (ticket, out)."""
        x = _as_elems(a, field)
        out = x.copy() if out is None else out
        w, wp = _omega_ptr(omega, field)
        t = ctypes.c_uint64()
        check(lib().gsn_server_submit(self.handle, field, x.shape[0].bit_length() - 1, int(bool(inverse)), wp,
                                      x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.byref(t)))
        self._keep[t.value] = (x, out, w)
        return t.value, out

    def submit_device(self, field, log_n, d_in, d_out, inverse=False, omega=None):
        w, wp = _omega_ptr(omega, field)
        t = ctypes.c_uint64()
        check(lib().gsn_server_submit_device(self.handle, field, log_n, int(bool(inverse)), wp,
                                             ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), ctypes.byref(t)))
        self._keep[t.value] = (w,)
        return t.value

    def poll(self, ticket):
        done = ctypes.c_int()
        check(lib().gsn_server_poll(self.handle, ticket, ctypes.byref(done)))
        return bool(done.value)

    def wait(self, ticket):
        st = lib().gsn_server_wait(self.handle, ticket)
        self._keep.pop(ticket, None)
        check(st)

    def times(self, ticket):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().gsn_server_ticket_times(self.handle, ticket, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "handle", None):
            lib().gsn_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ tuning

def tuning_set(field, log_n, shapes):
    """shapes: list of (log_r, log_w, log_e) per pass."""
    flat = (ctypes.c_int * (3 * len(shapes)))(*[v for s in shapes for v in s])
    check(lib().gsn_tuning_set(field, log_n, flat, len(shapes)))


def tuning_get(field, log_n):
    buf = (ctypes.c_int * 48)()
    n = ctypes.c_int()
    check(lib().gsn_tuning_get(field, log_n, buf, 16, ctypes.byref(n)))
    return [tuple(buf[3 * i:3 * i + 3]) for i in range(n.value)]


def tuning_clear():
    check(lib().gsn_tuning_clear())


def tuning_load(path):
    check(lib().gsn_tuning_load(path.encode()))
