"""ctypes binding of liblifted_hip.so (the C ABI declared in include/lifted_hip.h).

The library is the product: there is no CPU fallback. Importing fails loudly
when the library has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblifted_hip.so")

LFX_F32, LFX_F64, LFX_BF16, LFX_F16 = 0, 1, 2, 3
LFX_FLAG_PLAIN = 1
LFX_FLAG_LDS_ONLY = 2


class LfxError(RuntimeError):
    """Error reported by the HIP library (carries the LFX_E* code)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class LfxDesc(ctypes.Structure):
    _fields_ = [
        ("wavelet", ctypes.c_uint32),
        ("transform", ctypes.c_uint32),
        ("bc", ctypes.c_uint32),
        ("ndim", ctypes.c_int32),
        ("shape", ctypes.POINTER(ctypes.c_int64)),
        ("stride_in", ctypes.POINTER(ctypes.c_int64)),
        ("stride_out", ctypes.POINTER(ctypes.c_int64)),
        ("naxes", ctypes.c_int32),
        ("axes", ctypes.POINTER(ctypes.c_int64)),
        ("separable", ctypes.c_int32),
        ("levels", ctypes.POINTER(ctypes.c_int64)),
        ("storage", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
    ]


class LfxStep(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("target", ctypes.c_int32),
        ("offset", ctypes.c_int32),
        ("ntaps", ctypes.c_int32),
        ("sign", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("taps", ctypes.c_double * 16),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m lifted_amd.build` "
                "(lifted_amd has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(LfxDesc)
        L.lfx_lwt.argtypes = [dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.lfx_lwt.restype = ctypes.c_int
        L.lfx_lwt_host.argtypes = [dp, ctypes.c_void_p, ctypes.c_void_p]
        L.lfx_lwt_host.restype = ctypes.c_int
        L.lfx_last_error.restype = ctypes.c_char_p
        L.lfx_max_level.argtypes = [ctypes.c_uint32, ctypes.c_int64]
        L.lfx_max_level.restype = ctypes.c_int64
        L.lfx_wavelet_width.argtypes = [ctypes.c_uint32]
        L.lfx_wavelet_width.restype = ctypes.c_int64
        L.lfx_program_create.argtypes = [ctypes.POINTER(LfxStep), ctypes.c_int32, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint32)]
        L.lfx_program_create.restype = ctypes.c_int
        L.lfx_program_dump.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_int32]
        L.lfx_program_dump.restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise LfxError(rc, lib().lfx_last_error().decode())
    return rc


def i64(vals):
    a = (ctypes.c_int64 * max(len(vals), 1))()
    for i, v in enumerate(vals):
        a[i] = int(v)
    return a


def make_desc(wavelet, transform, bc, shape, stride_in, stride_out, axes, levels, separable, storage, flags=0):
    keep = [i64(shape), i64(stride_in), i64(stride_out), i64(axes), i64(levels)]
    d = LfxDesc()
    d.wavelet = int(wavelet)
    d.transform = int(transform)
    d.bc = int(bc)
    d.ndim = len(shape)
    d.shape, d.stride_in, d.stride_out, d.axes, d.levels = keep
    d.naxes = len(axes)
    d.separable = 1 if separable else 0
    d.storage = int(storage)
    d.flags = int(flags)
    d._keep = keep  # keep the arrays alive with the struct
    return d


CPU_LIB_PATH = os.path.join(HERE, "liblifted_cpu.so")
_cpu = None


def cpu_lib():
    """liblifted_cpu.so (include/lifted_cpu.h): the CPU product path."""
    global _cpu
    if _cpu is None:
        if not os.path.exists(CPU_LIB_PATH):
            raise ImportError(f"{CPU_LIB_PATH} is missing: build it with `python -m lifted_amd.build`")
        L = ctypes.CDLL(CPU_LIB_PATH)
        L.lfx_cpu_lwt.argtypes = [ctypes.POINTER(LfxDesc), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
        L.lfx_cpu_lwt.restype = ctypes.c_int
        L.lfx_cpu_last_error.restype = ctypes.c_char_p
        L.lfx_cpu_max_level.argtypes = [ctypes.c_uint32, ctypes.c_int64]
        L.lfx_cpu_max_level.restype = ctypes.c_int64
        _cpu = L
    return _cpu


def cpu_check(rc):
    if rc < 0:
        raise LfxError(rc, cpu_lib().lfx_cpu_last_error().decode())
    return rc
