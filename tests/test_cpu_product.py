"""The CPU product path (device='cpu', liblifted_cpu.so: the GPU kernels'
line bodies compiled for the host, lines split over a thread pool) against
the oracle, bit for bit: every table wavelet x boundary x transform on odd
and even lengths, the reference's ND shapes and axes (separable and level by
level), layouts (Fortran order, strided, negative strides, in place), one and
several threads, long lines (no LDS limit on the CPU), and argument errors.
It never loads the oracle: the test compares the two."""
import itertools

import numpy as np
import pytest

from helpers import BCS, WAVELETS, assert_parity, band_slices, fill_sin, rng

F, I, FA, IA = 0, 1, 2, 3


def _ok(bc, n, L, w):
    if bc == 0 or w == 0:
        return True
    m = n
    for _ in range(L):
        nd, ns = m // 2, m - m // 2
        if nd == 0 or (bc == 4 and (nd < 2 or ns < 2)):
            return False
        m = ns
    return True


@pytest.mark.parametrize("w", WAVELETS)
def test_cpu_path_1d_all_wavelets(oracle, w):
    import lifted_amd as la
    g = rng(1000 + w)
    for n, L, bc, tf, dt in itertools.product((37, 64), (1, 3), BCS, (F, I, FA, IA), (np.float32, np.float64)):
        if not _ok(bc, n, L, w):
            continue
        x = g.standard_normal(n).astype(dt)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        got = la.lwt(x, w, tf, mode=bc, level=[L], device="cpu")
        assert_parity(got, ref, f"w={w} n={n} L={L} bc={bc} tf={tf} {dt.__name__}", bands=band_slices((n,), [0], [L]))


SHAPES = [((34, 31), (1, 0)), ((23, 41), (0, 1)), ((25, 24, 31), (2, 1)), ((24, 23, 41), (0, 2)),
          ((23, 24, 31), (2, 1, 0))]


@pytest.mark.parametrize("shape,axes", SHAPES)
def test_cpu_path_nd_and_threads(oracle, shape, axes):
    import lifted_amd as la
    x = fill_sin(int(np.prod(shape)), -1000.0, 1001.0).reshape(shape)
    for (w, bc), level, tf, sep, th in itertools.product(((100907, 3), (3, 2), (10204, 1), (2, 0)), (1, 3),
                                                         (F, I, FA, IA), (True, False), (1, 4)):
        lvl = [level] * len(axes) if sep else level
        if any(not _ok(bc, shape[a], level, w) for a in axes):
            continue
        ref = oracle.lwt(x, w, tf, bc, axes=list(axes), level=lvl)
        got = la.lwt(x, w, tf, mode=bc, axes=list(axes), level=lvl, device="cpu", n_threads=th)
        assert_parity(got, ref, f"shape={shape} axes={axes} w={w} bc={bc} L={level} tf={tf} sep={sep} th={th}",
                      bands=band_slices(shape, list(axes), lvl))


def test_cpu_path_layouts_in_place_and_long_lines(oracle):
    import lifted_amd as la
    x = rng(7).standard_normal((40, 52)).astype(np.float32)
    for tf in (F, I, FA, IA):
        ref = oracle.lwt(x, 100907, tf, 3, axes=[0, 1], level=[2, 3])
        xf = np.asfortranarray(x)
        got = la.lwt(xf, 100907, tf, mode=3, axes=[0, 1], level=[2, 3], device="cpu")
        assert got.flags.f_contiguous
        assert_parity(got, ref, f"F-order tf={tf}")
        neg = np.ascontiguousarray(x[::-1])[::-1]
        assert_parity(la.lwt(neg, 100907, tf, mode=3, axes=[0, 1], level=[2, 3], device="cpu"), ref,
                      f"negative stride tf={tf}")
        big = np.zeros((80, 104), np.float32)
        big[::2, ::2] = x
        out = np.zeros((40, 156), np.float32)
        got = la.lwt(big[::2, ::2], 100907, tf, mode=3, axes=[0, 1], level=[2, 3], out=out[:, ::3], device="cpu")
        assert_parity(got, ref, f"strided tf={tf}")
        xi = x.copy()
        la.lwt(xi, 100907, tf, mode=3, axes=[0, 1], level=[2, 3], out=xi, device="cpu")
        assert_parity(xi, ref, f"in place tf={tf}")
        ref2 = oracle.lwt(x, 100907, tf, 3, axes=[1, 0], level=2)
        xi = x.copy()
        la.lwt(xi, 100907, tf, mode=3, axes=[1, 0], level=2, out=xi, device="cpu")
        assert_parity(xi, ref2, f"in place level-by-level tf={tf}")
    # lines far longer than a GPU workgroup's LDS, every boundary, 16 threads over 3 lines
    y = rng(8).standard_normal((3, 70001))
    for bc, tf in itertools.product(BCS, (F, FA)):
        ref = oracle.lwt(y, 10, tf, bc, axes=[1], level=[9])
        got = la.lwt(y, 10, tf, mode=bc, axes=[1], level=[9], device="cpu", n_threads=16)
        assert_parity(got, ref, f"long lines bc={bc} tf={tf}")


def test_cpu_path_torch_cpu_tensor_and_errors(oracle):
    import torch
    import lifted_amd as la
    x = rng(9).standard_normal((16, 300))
    t = torch.from_numpy(x)
    got = la.lwt_forward(t, "cdf9.7", mode="Symmetric", axes=1, level=[4], device="cpu", n_threads=2)
    assert isinstance(got, torch.Tensor) and not got.is_cuda
    assert_parity(got.numpy(), oracle.lwt(x, 100907, F, 3, axes=[1], level=[4]), "torch CPU tensor")
    with pytest.raises(la.LfxError):  # Reflect on a band of length 1
        la.lwt_forward(np.zeros(3), 2, mode="Reflect", level=[2], device="cpu")
    with pytest.raises(TypeError):
        la.lwt_forward(np.zeros(8, np.int32), 2, device="cpu")
    with pytest.raises(ValueError):
        la.lwt_forward(np.zeros(8), 2, device="tpu")
