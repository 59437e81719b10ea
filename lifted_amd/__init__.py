"""lifted_amd: the lifting-scheme DWT of jcapriot/lifted on AMD Instinct MI355X.

Mirrors the pylifted API (src/pylifted/_wavelets.pyx:93-268 of the reference):

    lwt_forward(x, wavelet, mode=BoundaryCondition.Zero, axes=None, level=1,
                n_threads=1, out=None)
    lwt_inverse(...), lwt_forward_adjoint(...), lwt_inverse_adjoint(...)

``x`` may be a torch tensor on a ROCm device (transformed in place on the GPU,
stream-ordered on torch's current stream; float32, float64, bfloat16 and
float16 -- the 16-bit types are computed in float32) or a numpy array
(float32 / float64; copied to the GPU and back, the reference's calling
convention). ``level`` int -> level-by-level transform, sequence -> separable
with one level count per axis.

``device='cpu'`` runs host arrays (numpy, or torch CPU tensors) through the
CPU product path (liblifted_cpu.so: the same line bodies compiled for the
host, lines split over ``n_threads`` threads, 0 = all cores); results equal
the GPU's bit for bit. It is chosen explicitly: the GPU path never falls back
to it, and a missing HIP library fails loudly.
"""
import enum
import numbers

from . import _lib
from ._lib import LfxError, LFX_FLAG_PLAIN


class Wavelet(enum.IntEnum):
    """Values are the reference's ABI (src/lifted-common.hpp:33-69)."""
    Lazy = 0
    Daubechies1 = 1
    Daubechies2 = 2
    Daubechies3 = 3
    Daubechies4 = 4
    Daubechies5 = 5
    Daubechies6 = 6
    Daubechies7 = 7
    Daubechies8 = 8
    Daubechies9 = 9
    Daubechies10 = 10
    Symlet4 = 104
    Symlet5 = 105
    Symlet6 = 106
    Coiflet2 = 202
    Coiflet3 = 203
    Bior1_3 = 10103
    Bior1_5 = 10105
    Bior2_2 = 10202
    Bior2_4 = 10204
    Bior2_6 = 10206
    Bior2_8 = 10208
    Bior3_1 = 10301
    Bior3_3 = 10303
    Bior3_5 = 10305
    Bior3_7 = 10307
    Bior3_9 = 10309
    Bior4_2 = 10402
    Bior4_4 = 10404
    Bior4_6 = 10406
    Bior5_5 = 10505
    Bior6_8 = 10608
    CDF5_3 = 100503
    CDF9_7 = 100907


class BoundaryCondition(enum.IntEnum):
    """src/lifted-common.hpp:74-81"""
    Zero = 0
    Constant = 1
    Periodic = 2
    Symmetric = 3
    Reflect = 4


class Transform(enum.IntEnum):
    """src/lifted-common.hpp:89-95"""
    Forward = 0
    Inverse = 1
    ForwardAdjoint = 2
    InverseAdjoint = 3


# pylifted's string names (src/pylifted/_wavelets.pyx:93-132)
_MAP_STR_TO_WVLT = {
    "lazy": Wavelet.Lazy, "haar": Wavelet.Daubechies1,
    "db1": Wavelet.Daubechies1, "db2": Wavelet.Daubechies2, "db3": Wavelet.Daubechies3,
    "db4": Wavelet.Daubechies4, "db5": Wavelet.Daubechies5, "db6": Wavelet.Daubechies6,
    "db7": Wavelet.Daubechies7, "db8": Wavelet.Daubechies8, "db9": Wavelet.Daubechies9,
    "db10": Wavelet.Daubechies10,
    "sym1": Wavelet.Daubechies1, "sym2": Wavelet.Daubechies2, "sym3": Wavelet.Daubechies3,
    "sym4": Wavelet.Symlet4, "sym5": Wavelet.Symlet5, "sym6": Wavelet.Symlet6,
    "coif2": Wavelet.Coiflet2, "coif3": Wavelet.Coiflet3,
    "bior1.3": Wavelet.Bior1_3, "bior1.5": Wavelet.Bior1_5,
    "bior2.2": Wavelet.Bior2_2, "bior2.4": Wavelet.Bior2_4, "bior2.6": Wavelet.Bior2_6,
    "bior2.8": Wavelet.Bior2_8,
    "bior3.1": Wavelet.Bior3_1, "bior3.3": Wavelet.Bior3_3, "bior3.5": Wavelet.Bior3_5,
    "bior3.7": Wavelet.Bior3_7, "bior3.9": Wavelet.Bior3_9,
    "bior4.2": Wavelet.Bior4_2, "bior4.4": Wavelet.Bior4_4, "bior4.6": Wavelet.Bior4_6,
    "bior5.5": Wavelet.Bior5_5, "bior6.8": Wavelet.Bior6_8,
    "cdf5.3": Wavelet.CDF5_3, "cdf9.7": Wavelet.CDF9_7,
}

_custom_ids = set()


def _wavelet_id(wavelet):
    if isinstance(wavelet, str):
        try:
            return int(_MAP_STR_TO_WVLT[wavelet])
        except KeyError:
            raise KeyError(wavelet + " is not a valid Wavelet.") from None
    if isinstance(wavelet, CustomWavelet):
        return wavelet.id
    if isinstance(wavelet, numbers.Integral):
        w = int(wavelet)
        if w in Wavelet._value2member_map_ or w in _custom_ids:
            return w
        raise KeyError(f"{wavelet} is not a valid Wavelet.")
    raise TypeError("wavelet must be a Wavelet, an int or a name string")


def _mode_id(mode):
    if isinstance(mode, str):
        return int(BoundaryCondition[mode])
    if isinstance(mode, numbers.Integral) and int(mode) in BoundaryCondition._value2member_map_:
        return int(mode)
    raise ValueError(f"{mode!r} is not a valid BoundaryCondition")


def get_axes(axes, ndim):
    """pylifted get_axes (_wavelets.pyx:134-158): None -> all, int or iterable,
    negatives allowed, >= ndim rejected."""
    if axes is None:
        axes = tuple(range(ndim))
    else:
        try:
            axes = tuple(axes)
        except TypeError:
            axes = (axes,)
    out = []
    for a in axes:
        if not isinstance(a, numbers.Integral):
            raise TypeError("Axes must be integers")
        if a >= ndim:
            raise ValueError(f"Axis {a} is beyond the array's dimensions, {ndim}.")
        if a < -ndim:
            raise ValueError(f"Axis {a} is beyond the array's dimensions, {ndim}.")
        out.append(int(a) % ndim)
    return out


def get_level(level, axes):
    """pylifted get_level (_wavelets.pyx:160-175): int -> (level, None);
    sequence -> (None, levels) with len(levels) == len(axes)."""
    if isinstance(level, numbers.Integral):
        if level < 0:
            raise ValueError("level must be non-negative")
        return int(level), None
    try:
        levels = [int(v) for v in level]
    except TypeError:
        raise TypeError("Level must be a single number, or a list of numbers") from None
    if len(levels) == 0:
        raise TypeError("levels must be a non-empty list")
    if len(levels) != len(axes):
        raise TypeError("levels must be the same length as axes")
    if any(v < 0 for v in levels):
        raise ValueError("levels must be non-negative")
    return None, levels


def _storage_torch(dtype):
    import torch
    m = {torch.float32: _lib.LFX_F32, torch.float64: _lib.LFX_F64,
         torch.bfloat16: _lib.LFX_BF16, torch.float16: _lib.LFX_F16}
    if dtype not in m:
        raise TypeError(f"unsupported dtype {dtype}: float32, float64, bfloat16, float16")
    return m[dtype]


def _is_torch(x):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor)


def _lwt_cpu(x, wid, bc, op, axes_, level_, levels_, n_threads, out):
    """device='cpu': numpy arrays (or torch CPU tensors) through
    liblifted_cpu.so, the lines of each axis split over n_threads threads
    (0: all cores), as pylifted's n_threads does (_wavelets.pyx:177-268)."""
    import numpy as np
    sep = levels_ is not None
    if _is_torch(x):
        if x.is_cuda:
            raise ValueError("device='cpu' takes host arrays (numpy, or torch CPU tensors)")
        import torch
        if x.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"unsupported dtype {x.dtype}: float32 or float64")
        if out is None:
            out = torch.empty_like(x)
        if out.dtype != x.dtype or out.is_cuda or tuple(out.shape) != tuple(x.shape):
            raise ValueError("`out` must be a CPU tensor of the input's dtype and shape")
        storage = _lib.LFX_F32 if x.dtype == torch.float32 else _lib.LFX_F64
        d = _lib.make_desc(wid, op, bc, list(x.shape), list(x.stride()), list(out.stride()), axes_,
                           levels_ if sep else [level_], sep, storage)
        _lib.cpu_check(_lib.cpu_lib().lfx_cpu_lwt(d, x.data_ptr(), out.data_ptr(), int(n_threads)))
        return out
    if not isinstance(x, np.ndarray):
        raise TypeError("x must be a numpy array or a torch tensor")
    if x.dtype not in (np.float32, np.float64):
        raise TypeError(f"unsupported dtype {x.dtype}: float32 or float64")
    if out is None:
        out = np.empty(x.shape, dtype=x.dtype, order="F" if (x.flags.f_contiguous and not
                                                              x.flags.c_contiguous) else "C")
    if out.dtype != x.dtype:
        raise ValueError(f"`out.dtype`:{out.dtype} does not match `input.dtype`:{x.dtype}.")
    for i in range(x.ndim):
        if x.shape[i] != out.shape[i]:
            raise ValueError(f"`out.shape` does not match input along axis {i}.")
    es = x.itemsize
    storage = _lib.LFX_F32 if x.dtype == np.float32 else _lib.LFX_F64
    d = _lib.make_desc(wid, op, bc, list(x.shape), [s // es for s in x.strides], [s // es for s in out.strides],
                       axes_, levels_ if sep else [level_], sep, storage)
    _lib.cpu_check(_lib.cpu_lib().lfx_cpu_lwt(d, x.ctypes.data, out.ctypes.data, int(n_threads)))
    return out


def _lwt(x, wavelet, mode, op, axes, level, n_threads=1, out=None, plain=False, device=None):
    wid = _wavelet_id(wavelet)
    bc = _mode_id(mode)
    axes_ = get_axes(axes, x.ndim)
    level_, levels_ = get_level(level, axes_)
    if device == "cpu":
        if plain:
            raise ValueError("plain applies to the GPU kernels only")
        return _lwt_cpu(x, wid, bc, op, axes_, level_, levels_, n_threads, out)
    if device not in (None, "gpu", "cuda"):
        raise ValueError(f"device must be None / 'gpu' (the HIP kernels) or 'cpu', not {device!r}")
    sep = levels_ is not None
    flags = LFX_FLAG_PLAIN if plain else 0
    L = _lib.lib()
    if _is_torch(x):
        import torch
        storage = _storage_torch(x.dtype)
        if not x.is_cuda:
            raise ValueError("torch tensors must live on a ROCm device (use numpy for host arrays)")
        if out is None:
            out = torch.empty_like(x)
        if out.dtype != x.dtype:
            raise ValueError(f"`out.dtype`:{out.dtype} does not match `input.dtype`:{x.dtype}.")
        if out.device != x.device:
            raise ValueError("`out` must live on the input's device")
        for i in range(x.ndim):
            if x.shape[i] != out.shape[i]:
                raise ValueError(f"`out.shape` does not match input along axis {i}.")
        d = _lib.make_desc(wid, op, bc, list(x.shape), list(x.stride()), list(out.stride()), axes_,
                           levels_ if sep else [level_], sep, storage, flags)
        with torch.cuda.device(x.device):
            stream = torch.cuda.current_stream(x.device).cuda_stream
            _lib.check(L.lfx_lwt(d, x.data_ptr(), out.data_ptr(), stream))
        return out
    import numpy as np
    if not isinstance(x, np.ndarray):
        raise TypeError("x must be a numpy array or a torch tensor")
    if x.dtype not in (np.float32, np.float64):
        raise TypeError(f"unsupported dtype {x.dtype}: float32 or float64")
    if out is None:
        out = np.empty(x.shape, dtype=x.dtype, order="F" if (x.flags.f_contiguous and not
                                                              x.flags.c_contiguous) else "C")
    if out.dtype != x.dtype:
        raise ValueError(f"`out.dtype`:{out.dtype} does not match `input.dtype`:{x.dtype}.")
    for i in range(x.ndim):
        if x.shape[i] != out.shape[i]:
            raise ValueError(f"`out.shape` does not match input along axis {i}.")
    es = x.itemsize
    storage = _lib.LFX_F32 if x.dtype == np.float32 else _lib.LFX_F64
    d = _lib.make_desc(wid, op, bc, list(x.shape), [s // es for s in x.strides], [s // es for s in out.strides],
                       axes_, levels_ if sep else [level_], sep, storage, flags)
    _lib.check(L.lfx_lwt_host(d, x.ctypes.data, out.ctypes.data))
    return out


def lwt_forward(x, wavelet, mode=BoundaryCondition.Zero, axes=None, level=1, n_threads=1, out=None,
        plain=False, device=None):
    return _lwt(x, wavelet, mode, Transform.Forward, axes, level, n_threads, out=out, plain=plain,
                device=device)


def lwt_inverse(x, wavelet, mode=BoundaryCondition.Zero, axes=None, level=1, n_threads=1, out=None,
        plain=False, device=None):
    return _lwt(x, wavelet, mode, Transform.Inverse, axes, level, n_threads, out=out, plain=plain,
                device=device)


def lwt_forward_adjoint(x, wavelet, mode=BoundaryCondition.Zero, axes=None, level=1, n_threads=1, out=None,
        plain=False, device=None):
    return _lwt(x, wavelet, mode, Transform.ForwardAdjoint, axes, level, n_threads, out=out, plain=plain,
                device=device)


def lwt_inverse_adjoint(x, wavelet, mode=BoundaryCondition.Zero, axes=None, level=1, n_threads=1, out=None,
        plain=False, device=None):
    return _lwt(x, wavelet, mode, Transform.InverseAdjoint, axes, level, n_threads, out=out, plain=plain,
                device=device)


def lwt(x, wavelet, transform, mode=BoundaryCondition.Zero, axes=None, level=1, out=None, plain=False, device=None,
        n_threads=1):
    """Any of the four transforms by Transform value. device='cpu' runs the
    CPU product path on host arrays with n_threads threads (0: all cores)."""
    return _lwt(x, wavelet, mode, int(transform), axes, level, n_threads, out=out, plain=plain, device=device)


def max_level(wavelet, n):
    """max_level(width, n) of src/lifted-common.hpp:1099-1108."""
    return _lib.check(_lib.lib().lfx_max_level(_wavelet_id(wavelet), int(n)))


def wavelet_width(wavelet):
    return _lib.check(_lib.lib().lfx_wavelet_width(_wavelet_id(wavelet)))


class CustomWavelet:
    """A caller-defined lifting scheme (runtime step program).

    steps: list of dicts, e.g.
        {"kind": "generic", "target": "d", "offset": 0, "taps": [-0.5, -0.5]}
        {"kind": "unit", "target": "s", "offset": -1, "sign": +1}
        {"kind": "repeat", "target": "s", "offset": -1, "n": 2, "value": 0.25}
        {"kind": "scale", "values": [k]}  or  {"kind": "scale", "values": [s_mul, d_div]}
    Coefficients are doubles cast to the transform type.
    """

    def __init__(self, steps, width=2):
        import ctypes
        arr = (_lib.LfxStep * max(len(steps), 1))()
        kinds = {"generic": 0, "unit": 1, "repeat": 2, "scale": 3}
        for i, s in enumerate(steps):
            k = kinds[s["kind"]]
            arr[i].kind = k
            arr[i].target = 1 if s.get("target", "s") in ("d", 1) else 0
            arr[i].offset = int(s.get("offset", 0))
            arr[i].sign = int(s.get("sign", 1))
            if k == 0:
                taps = list(s["taps"])
                if not 1 <= len(taps) <= 16:
                    raise LfxError(-1, "generic steps take 1..16 taps")
                arr[i].ntaps = len(taps)
                for j, t in enumerate(taps):
                    arr[i].taps[j] = float(t)
            elif k == 1:
                arr[i].ntaps = 1
            elif k == 2:
                arr[i].ntaps = int(s["n"])
                arr[i].taps[0] = float(s["value"])
            else:
                vals = list(s["values"])
                if len(vals) == 1:
                    vals = vals * 2
                arr[i].taps[0], arr[i].taps[1] = float(vals[0]), float(vals[1])
        wid = ctypes.c_uint32()
        _lib.check(_lib.lib().lfx_program_create(arr, len(steps), int(width), ctypes.byref(wid)))
        self.id = int(wid.value)
        self.steps = steps
        self.width = width
        _custom_ids.add(self.id)


def program_dump(wavelet, transform, storage=_lib.LFX_F64):
    """Resolved per-step data the GPU uses (for tests)."""
    import ctypes
    import numpy as np
    buf = (ctypes.c_double * (25 * 16))()
    n = _lib.check(_lib.lib().lfx_program_dump(_wavelet_id(wavelet), int(transform), int(storage), buf, 16))
    return np.array(buf[: 25 * n]).reshape(n, 25)


__all__ = [
    "Wavelet", "BoundaryCondition", "Transform", "CustomWavelet", "LfxError",
    "lwt_forward", "lwt_inverse", "lwt_forward_adjoint", "lwt_inverse_adjoint", "lwt",
    "max_level", "wavelet_width", "program_dump",
]
