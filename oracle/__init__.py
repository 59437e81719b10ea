"""ctypes binding of the CPU restatement (``oracle/lifted_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker / CPU baseline.
The product package ``lifted_amd`` never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblifted_oracle.so")
_lib = None

FORWARD, INVERSE, FORWARD_ADJOINT, INVERSE_ADJOINT = 0, 1, 2, 3
ZERO, CONSTANT, PERIODIC, SYMMETRIC, REFLECT = 0, 1, 2, 3, 4


class OracleStep(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("target", ctypes.c_int32),
        ("offset", ctypes.c_int32),
        ("ntaps", ctypes.c_int32),
        ("sign", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("taps", ctypes.c_double * 16),
    ]


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.oracle_lwt.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                   i64p, i64p, i64p, ctypes.c_int32, i64p, ctypes.c_int32, i64p,
                                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
        lib.oracle_lwt.restype = ctypes.c_int
        lib.oracle_lwt_custom.argtypes = [ctypes.POINTER(OracleStep), ctypes.c_int32, ctypes.c_uint32] + \
            lib.oracle_lwt.argtypes[1:]
        lib.oracle_lwt_custom.restype = ctypes.c_int
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_max_level.argtypes = [ctypes.c_uint32, ctypes.c_int64]
        lib.oracle_max_level.restype = ctypes.c_long
        lib.oracle_wavelet_dump.argtypes = [ctypes.c_uint32, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_int32]
        lib.oracle_wavelet_dump.restype = ctypes.c_int
        lib.oracle_batch.argtypes = [i64p, i64p, ctypes.POINTER(ctypes.c_uint32), i64p, ctypes.c_int64,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int32]
        lib.oracle_batch.restype = ctypes.c_int
        _lib = lib
    return _lib


def _arr(vals):
    a = (ctypes.c_int64 * max(len(vals), 1))()
    for i, v in enumerate(vals):
        a[i] = int(v)
    return a


def make_steps(steps):
    """steps: list of dicts {kind, target, offset, ntaps, sign, taps} -> ctypes array."""
    arr = (OracleStep * max(len(steps), 1))()
    kinds = {"generic": 0, "unit": 1, "repeat": 2, "scale": 3}
    for i, s in enumerate(steps):
        arr[i].kind = kinds[s["kind"]] if isinstance(s["kind"], str) else int(s["kind"])
        arr[i].target = {"s": 0, "d": 1}.get(s.get("target", "s"), s.get("target", 0))
        arr[i].offset = int(s.get("offset", 0))
        arr[i].ntaps = int(s.get("ntaps", len(s.get("taps", [])) or 1))
        arr[i].sign = int(s.get("sign", 1))
        for k, t in enumerate(s.get("taps", [])):
            arr[i].taps[k] = float(t)
    return arr, len(steps)


def lwt(x, wavelet, transform, bc=ZERO, axes=None, level=1, out=None, n_threads=1, steps=None, width=2):
    """Run the restatement on a numpy array (float32 / float64, any strides).

    ``level`` int -> level-by-level (non-separable) path, sequence -> separable.
    ``steps`` (optional) -> a caller-defined step program instead of ``wavelet``.
    """
    lib = _load()
    if x.dtype not in (np.float32, np.float64):
        raise TypeError("oracle supports float32/float64 only")
    ndim = x.ndim
    if axes is None:
        axes = list(range(ndim))
    elif np.isscalar(axes):
        axes = [int(axes)]
    axes = [int(a) for a in axes]
    if out is None:
        out = np.empty_like(x)
    sep = not np.isscalar(level)
    levels = list(level) if sep else [int(level)]
    es = x.itemsize
    args = [int(transform), int(bc), ndim, _arr(x.shape), _arr([s // es for s in x.strides]),
            _arr([s // es for s in out.strides]), len(axes), _arr(axes), int(sep), _arr(levels),
            0 if x.dtype == np.float32 else 1, ctypes.c_void_p(x.ctypes.data),
            ctypes.c_void_p(out.ctypes.data), int(n_threads)]
    if steps is None:
        rc = lib.oracle_lwt(int(wavelet), *args)
    else:
        arr, n = make_steps(steps)
        rc = lib.oracle_lwt_custom(arr, n, int(width), *args)
    if rc != 0:
        raise ValueError(lib.oracle_last_error().decode())
    return out


def max_level(wavelet, n):
    return _load().oracle_max_level(int(wavelet), int(n))


def wavelet_dump(wavelet, dtype):
    """Resolved per-type coefficients: list of rows [kind,target,offset,ntaps,sign,taps...]."""
    buf = (ctypes.c_double * (21 * 32))()
    n = _load().oracle_wavelet_dump(int(wavelet), 0 if dtype == np.float32 else 1, buf, 32)
    if n < 0:
        raise ValueError("unknown wavelet")
    return np.array(buf[: 21 * n]).reshape(n, 21)


def batch(x, signals, transform, bc=ZERO, out=None, n_threads=1):
    """Packed 1-D signals [(offset, length, wavelet, levels)] of a flat array."""
    lib = _load()
    if out is None:
        out = np.empty_like(x)
    n = len(signals)
    offs = _arr([s[0] for s in signals])
    lens = _arr([s[1] for s in signals])
    wv = (ctypes.c_uint32 * max(n, 1))(*[int(s[2]) for s in signals])
    lv = _arr([s[3] for s in signals])
    rc = lib.oracle_batch(offs, lens, wv, lv, n, int(transform), int(bc), 0 if x.dtype == np.float32 else 1,
                          ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(out.ctypes.data), int(n_threads))
    if rc != 0:
        raise ValueError(lib.oracle_last_error().decode())
    return out
