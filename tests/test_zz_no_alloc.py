"""lfx_lwt_noalloc (the entry point that allocates nothing: its scratch and
step programs live in fixed device arrays) against lfx_lwt, through the C
ABI, on the same inputs: two calls in a row per case, the second with a
smaller input, each bit for bit equal to lfx_lwt's result and to the oracle;
and a clear error when a call would need more scratch than the fixed array."""
import ctypes

import numpy as np
import pytest

from helpers import assert_parity, rng

pytestmark = pytest.mark.gpu


def _bind():
    from lifted_amd import _lib
    L = _lib.lib()
    dp = ctypes.POINTER(_lib.LfxDesc)
    L.lfx_lwt_noalloc.argtypes = [dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.lfx_lwt_noalloc.restype = ctypes.c_int
    return L, _lib


def _storage(t):
    import torch
    return {torch.float32: 0, torch.float64: 1, torch.bfloat16: 2, torch.float16: 3}[t.dtype]


def _run(entry, x, out, w, tf, bc, axes, level):
    """One call of lfx_lwt or lfx_lwt_noalloc on device tensors."""
    import torch
    L, _lib = _bind()
    sep = not isinstance(level, int)
    d = _lib.make_desc(w, tf, bc, list(x.shape), list(x.stride()), list(out.stride()), list(axes),
                       list(level) if sep else [level], sep, _storage(x))
    fn = L.lfx_lwt_noalloc if entry == "noalloc" else L.lfx_lwt
    _lib.check(fn(d, x.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream))
    return out


CASES = [
    # (name, (shape, shape of the second, smaller call), dtype, wavelet, bc, axes, level)
    ("fp64 long lines, tiled passes", ((3, 70001), (2, 30011)), "float64", 100907, 3, [1], [7]),
    ("fp32 long lines, tiled passes", ((2, 200003), (3, 40000)), "float32", 2, 0, [1], [9]),
    ("fp32 separable whole lines", ((37, 300), (20, 129)), "float32", 10204, 1, [1, 0], [3, 4]),
    ("fp64 level by level", ((40, 52), (33, 20)), "float64", 3, 4, [0, 1], 2),
    ("bf16 level by level, one axis", ((64, 96), (40, 48)), "bfloat16", 100907, 0, [0], 3),
    ("fp16 long line", ((1, 60001), (1, 30001)), "float16", 10, 3, [1], [6]),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_lwt_two_calls(gpu, oracle, case):
    import torch
    name, shapes, dt, w, bc, axes, level = case
    tdt = getattr(torch, dt)
    for k, shape in enumerate(shapes):
        xh = rng(500 + 7 * k + len(name)).standard_normal(shape)
        x = torch.from_numpy(xh).to(tdt).to(gpu)
        for tf in (0, 1, 2, 3):
            a = _run("lwt", x, torch.empty_like(x), w, tf, bc, axes, level)
            b = _run("noalloc", x, torch.full_like(x, float("nan")), w, tf, bc, axes, level)
            torch.cuda.synchronize()
            assert torch.equal(a, b), f"{name} call {k} tf={tf}: lfx_lwt_noalloc != lfx_lwt"
            if tdt in (torch.float32, torch.float64):
                xn = x.cpu().numpy()
                ref = oracle.lwt(xn, w, tf, bc, axes=axes, level=level)
                assert_parity(b.cpu().numpy(), ref, f"{name} call {k} tf={tf}")
        # in place: the input is copied into the fixed scratch first
        xi = x.clone()
        ref = _run("lwt", x, torch.empty_like(x), w, 0, bc, axes, level)
        _run("noalloc", xi, xi, w, 0, bc, axes, level)
        torch.cuda.synchronize()
        assert torch.equal(xi, ref), f"{name} call {k}: in place"


def test_noalloc_rejects_oversized_scratch(gpu):
    """An in-place call of long lines needs a copy of the whole input in the
    scratch: 80 MB of float64 exceeds the fixed 64 MiB and is refused with a
    clear message, leaving the array untouched. Caller-defined programs, and
    16-bit storage over several axes (rounded once per level through float32
    workspaces the fixed memory does not hold), are refused."""
    import torch
    import lifted_amd as la
    from helpers import MOCK_WAVELETS
    L, _lib = _bind()
    x = torch.randn(1, 10_000_000, dtype=torch.float64, device=gpu)
    x0 = x.clone()
    with pytest.raises(_lib.LfxError) as e:
        _run("noalloc", x, x, 2, 0, 0, [1], [5])
    assert "scratch" in str(e.value) and "at most" in str(e.value)
    torch.cuda.synchronize()
    assert torch.equal(x, x0)
    cw = la.CustomWavelet(MOCK_WAVELETS["MultiStep"])
    y = torch.randn(16, 64, device=gpu)
    with pytest.raises(_lib.LfxError):
        _run("noalloc", y, torch.empty_like(y), cw.id, 0, 0, [1], [1])
    h = y.to(torch.bfloat16)
    with pytest.raises(_lib.LfxError):
        _run("noalloc", h, torch.empty_like(h), 100907, 0, 0, [1, 0], 2)
