"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs.

Both paths perform the same operations in the same order with the same fused
multiply-adds, so every comparison here is bit for bit (the contract's bar,
4 ulp in float32 and 1e-12 relative in float64, is looser). A failure reports
the worst error in units of the last place of each coefficient band's own
max |ref| (bands from lifted_amd.layout), and the session summary prints the
measured worst case of every test file.
"""
import itertools
import os

import numpy as np
import pytest

from helpers import (BCS, MOCK_WAVELETS, WAVELETS, WIDTHS, assert_parity, band_slices, fill_sin,
                     oracle_level_by_level_rounded, oracle_steps, rng)

pytestmark = pytest.mark.gpu
F, I, FA, IA = 0, 1, 2, 3


def _check(got, ref, dtype, what, axes=None, level=None):
    """Bit for bit; on failure the message gives the worst per-band error
    (bands of the transform given by axes / level, else the whole array)."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.dtype == ref.dtype == dtype, f"{what}: dtype {got.dtype} / {ref.dtype}"
    bands = band_slices(ref.shape, axes, level) if level is not None else None
    assert_parity(got, ref, what, bands=bands)


def _gpu(x_np, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x_np)).to(dev)


def _ok_bands(bc, n, L, steps=True):
    if bc == 0 or not steps:
        return True
    m = n
    for _ in range(L):
        nd = m // 2
        ns = m - nd
        if nd == 0 or (bc == 4 and (nd < 2 or ns < 2)):
            return False
        m = ns
    return True


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("w", WAVELETS)
def test_1d_all_wavelets_bcs_transforms(gpu, oracle, w, dtype):
    import lifted_amd as la
    g = rng(w)
    for n, L, bc, tf in itertools.product((37, 64), (1, 3), BCS, (F, I, FA, IA)):
        if not _ok_bands(bc, n, L, w != 0):
            continue
        x = g.standard_normal(n).astype(dtype)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        got = la.lwt(_gpu(x, gpu), w, tf, mode=bc, level=[L]).cpu().numpy()
        _check(got, ref, dtype, f"w={w} n={n} L={L} bc={bc} tf={tf}", axes=[0], level=[L])


@pytest.mark.parametrize("name", list(MOCK_WAVELETS))
def test_mock_wavelets(gpu, oracle, name):
    """The reference's mock step programs (tests/test_helpers.h:102-360), all BCs,
    lengths 153 / 510 (tests/test_single_levels.cpp:395-405)."""
    import lifted_amd as la
    cw = la.CustomWavelet(MOCK_WAVELETS[name])
    st = oracle_steps(MOCK_WAVELETS[name])
    for n, bc, tf, dtype in itertools.product((153, 510), BCS, (F, I, FA, IA), (np.float64, np.float32)):
        x = fill_sin(n, dtype=dtype)
        ref = oracle.lwt(x, 0, tf, bc, level=1, steps=st)
        got = la.lwt(_gpu(x, gpu), cw, tf, mode=bc, level=1).cpu().numpy()
        _check(got, ref, dtype, f"{name} n={n} bc={bc} tf={tf}")


REF_SHAPES = [((31,), (0,)), ((32,), (0,)), ((32, 1), (0,)), ((1, 32), (1,)), ((34, 31), (1,)),
              ((23, 41), (0,)), ((34, 31), (1, 0)), ((23, 41), (0, 1)), ((25, 24, 31), (2,)),
              ((24, 23, 41), (1,)), ((24, 23, 41), (0,)), ((25, 24, 31), (2, 1)), ((24, 23, 41), (1, 0)),
              ((24, 23, 41), (0, 2)), ((23, 24, 31), (2, 1, 0))]


@pytest.mark.parametrize("shape,axes", REF_SHAPES)
def test_nd_shapes_of_reference_tests(gpu, oracle, shape, axes):
    """tests/test_dwt.cpp shapes/axes, separable and level-by-level, levels 1 and 3."""
    import lifted_amd as la
    n = int(np.prod(shape))
    for (w, bc), level, tf, sep in itertools.product(((2, 0), (100907, 3), (3, 2), (10204, 1)), (1, 3),
                                                      (F, I, FA, IA), (True, False)):
        if any(not _ok_bands(bc, shape[a], level) for a in axes):
            continue
        x = fill_sin(n, -1000.0, 1001.0).reshape(shape)
        lvl = [level] * len(axes) if sep else level
        ref = oracle.lwt(x, w, tf, bc, axes=list(axes), level=lvl)
        got = la.lwt(_gpu(x, gpu), w, tf, mode=bc, axes=list(axes), level=lvl).cpu().numpy()
        _check(got, ref, np.float64, f"shape={shape} axes={axes} w={w} L={level} tf={tf} sep={sep}", axes=list(axes),
               level=lvl)


def test_layouts_fortran_strided_inplace(gpu, oracle):
    import torch
    import lifted_amd as la
    x = rng(7).standard_normal((40, 52))
    for tf in (F, I, FA, IA):
        ref = oracle.lwt(x, 100907, tf, 3, axes=[0, 1], level=[2, 3])
        xt = torch.from_numpy(x).to(gpu)
        # Fortran-ordered input and output
        xf = xt.t().contiguous().t()
        got = la.lwt(xf, 100907, tf, mode=3, axes=[0, 1], level=[2, 3])
        assert got.stride() == xf.stride()
        _check(got.cpu().numpy(), ref, np.float64, f"F-order tf={tf}")
        # strided view in, contiguous out
        big = torch.zeros((80, 104), dtype=torch.float64, device=gpu)
        big[::2, ::2] = xt
        got = la.lwt(big[::2, ::2], 100907, tf, mode=3, axes=[0, 1], level=[2, 3],
                     out=torch.empty((40, 52), dtype=torch.float64, device=gpu))
        _check(got.cpu().numpy(), ref, np.float64, f"strided tf={tf}")
        # in place
        xi = xt.clone()
        la.lwt(xi, 100907, tf, mode=3, axes=[0, 1], level=[2, 3], out=xi)
        _check(xi.cpu().numpy(), ref, np.float64, f"in-place tf={tf}")
        # level by level in place
        ref2 = oracle.lwt(x, 100907, tf, 3, axes=[1, 0], level=2)
        xi = xt.clone()
        la.lwt(xi, 100907, tf, mode=3, axes=[1, 0], level=2, out=xi)
        _check(xi.cpu().numpy(), ref2, np.float64, f"in-place non-sep tf={tf}")


def test_numpy_host_path(gpu, oracle):
    """numpy in/out goes through lfx_lwt_host (the reference's calling convention)."""
    import lifted_amd as la
    x = rng(3).standard_normal((33, 20)).astype(np.float32)
    for tf in (F, I, FA, IA):
        ref = oracle.lwt(x, 3, tf, 4, level=1)
        got = la.lwt(x, 3, tf, mode=4, level=1)
        assert isinstance(got, np.ndarray)
        _check(got, ref, np.float32, f"host tf={tf}")
    xf = np.asfortranarray(x)
    got = la.lwt_forward(xf, "db3", mode="Reflect", level=1)
    assert got.flags.f_contiguous
    _check(got, oracle.lwt(x, 3, F, 4, level=1), np.float32, "host F-order")


def test_kat_mixed_bc_2d_dot_on_gpu(gpu):
    """<fwd_sym(y2), inv_adj_zero(y2)> of scratch/test_python_package.ipynb (1974.8783750525286)."""
    import torch
    import lifted_amd as la
    x = np.linspace(-10, 10, 1024)
    y2 = 1 / ((x - 2)[:, None] ** 2 + (x + 3) ** 2 + 4)
    t = torch.from_numpy(y2).to(gpu)
    a = la.lwt_forward(t, "db3", mode=la.BoundaryCondition.Symmetric, level=1)
    b = la.lwt_inverse_adjoint(t, "db3", mode=la.BoundaryCondition.Zero, level=1)
    v = float((a * b).sum())
    assert abs(v - 1974.8783750525286) <= 1e-12 * 1974.9


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_half_storage_single_axis_matches_fp32_rounded(gpu, oracle, storage):
    """16-bit storage, fp32 arithmetic: one axis pass equals the fp32 oracle on
    the 16-bit input, rounded once to the storage type."""
    import torch
    import lifted_amd as la
    tdt = getattr(torch, storage)
    x = torch.from_numpy(rng(11).standard_normal((16, 300)).astype(np.float32)).to(tdt)
    x32 = x.float().cpu().numpy()
    for tf in (F, I, FA, IA):
        ref = torch.from_numpy(oracle.lwt(x32, 100907, tf, 0, axes=[1], level=[4])).to(tdt).float().numpy()
        got = la.lwt(x.to(gpu), 100907, tf, axes=[1], level=[4]).float().cpu().numpy()
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_half_storage_separable_3d_matches_fp32_rounded(gpu, oracle, storage):
    """16-bit storage over every axis of a 3-D box (the strided axes of at
    most 512 samples read in place by the across-line kernel into float32
    workspaces, the contiguous one through a row plan): equals the fp32 oracle
    on the 16-bit input, rounded once to the storage type at the end, with the
    across-line kernel on and off."""
    import torch
    import lifted_amd as la
    from lifted_amd import tune
    tdt = getattr(torch, storage)
    x = torch.from_numpy(rng(12).standard_normal((40, 48, 64)).astype(np.float32)).to(tdt)
    x32 = x.float().cpu().numpy()
    for tf, across in itertools.product((F, I, FA, IA), (1, 0)):
        ref = torch.from_numpy(oracle.lwt(x32, 100907, tf, 0, axes=[0, 1, 2], level=[2, 3, 2])).to(tdt).float().numpy()
        with tune.tuning(across=across):
            got = la.lwt(x.to(gpu), 100907, tf, axes=[0, 1, 2], level=[2, 3, 2]).float().cpu().numpy()
        np.testing.assert_array_equal(got, ref, err_msg=f"{storage} tf={tf} across={across}")


def test_config1_db2_65536_fp64_8_levels(gpu, oracle):
    """1-D float64 signal of 65536 N(0,1) samples, Daubechies2, 8 levels,
    forward then inverse (the reference's CPU-runnable workload): coefficients
    equal to the oracle and perfect reconstruction, for all five boundaries
    (Periodic runs the whole line with its bands in global scratch), separable
    and level-by-level."""
    import lifted_amd as la
    x = np.random.default_rng(20261201).standard_normal(65536)
    t = _gpu(x, gpu)
    for bc, sep in itertools.product((0, 1, 2, 3, 4), (True, False)):
        lvl = [8] if sep else 8
        ref = oracle.lwt(x, 2, F, bc, level=lvl)
        w = la.lwt(t, 2, F, mode=bc, level=lvl)
        _check(w.cpu().numpy(), ref, np.float64, f"fwd bc={bc} sep={sep}")
        r = la.lwt(w, 2, I, mode=bc, level=lvl).cpu().numpy()
        _check(r, oracle.lwt(ref, 2, I, bc, level=lvl), np.float64, f"inv bc={bc} sep={sep}")
        assert np.abs(r - x).max() <= 1e-13 * np.abs(x).max()


@pytest.mark.parametrize("strip", [0, 1024])
@pytest.mark.parametrize("tf", [F, I, FA, IA])
def test_long_lines_tiled(gpu, oracle, tf, strip):
    """Lines longer than one workgroup's LDS take the tiled multi-level passes
    (float32: register windows of 2048 samples, or 1024 with the strip knob)."""
    import lifted_amd as la
    from lifted_amd import tune
    g = rng(100 + tf)
    for (w, bc, n, L, dtype) in [(100907, 0, 200003, 9, np.float32), (100907, 3, 65536 * 3, 12, np.float32),
                                 (3, 4, 70001, 5, np.float64), (10208, 1, 40000, 7, np.float32),
                                 (10, 3, 1 << 18, 13, np.float32), (100503, 0, (1 << 20) + 7, 17, np.float32)]:
        x = g.standard_normal(n).astype(dtype)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        with tune.tuning(strip=strip):
            got = la.lwt(_gpu(x, gpu), w, tf, mode=bc, level=[L]).cpu().numpy()
        _check(got, ref, dtype, f"w={w} bc={bc} n={n} L={L} tf={tf} strip={strip}", axes=[0], level=[L])


@pytest.mark.parametrize("tf", [F, I, FA, IA])
def test_long_lines_periodic(gpu, oracle, tf):
    """Periodic lines longer than one workgroup's LDS: one workgroup per line
    with its bands in global scratch, so the wrap can read the far end of
    every band (with the reference's unsigned modulo for negative indices at
    lengths that are not powers of two). Power-of-two and odd lengths, up to
    2^20 + 7 samples, float32 and float64, several lines along a strided axis,
    level-by-level, and in place; all against the oracle."""
    import torch
    import lifted_amd as la
    g = rng(200 + tf)
    for (w, n, L, dtype) in [(100907, (1 << 20) + 7, 6, np.float32), (2, 65536, 14, np.float64),
                             (10608, 40001, 5, np.float32), (3, 1 << 17, 10, np.float32),
                             (100503, 30011, 7, np.float64)]:
        x = g.standard_normal(n).astype(dtype)
        ref = oracle.lwt(x, w, tf, 2, level=[L])
        got = la.lwt(_gpu(x, gpu), w, tf, mode=2, level=[L]).cpu().numpy()
        _check(got, ref, dtype, f"w={w} n={n} L={L} tf={tf}", axes=[0], level=[L])
    x = g.standard_normal((30000, 5)).astype(np.float32)
    for lvl in ([4], 3):
        ref = oracle.lwt(x, 10204, tf, 2, axes=[0], level=lvl)
        t = torch.from_numpy(x).to(gpu)
        _check(la.lwt(t, 10204, tf, mode=2, axes=[0], level=lvl).cpu().numpy(), ref, np.float32,
               f"strided tf={tf} level={lvl}", axes=[0], level=lvl)
        la.lwt(t, 10204, tf, mode=2, axes=[0], level=lvl, out=t)
        _check(t.cpu().numpy(), ref, np.float32, f"in place tf={tf} level={lvl}", axes=[0], level=lvl)


def test_long_lines_2d_strided_axis(gpu, oracle):
    """Tiled passes along a strided axis, and in place (input copied first)."""
    import torch
    import lifted_amd as la
    x = rng(9).standard_normal((40000, 3)).astype(np.float32)
    for tf in (F, I, FA, IA):
        ref = oracle.lwt(x, 100907, tf, 3, axes=[0], level=[6])
        t = torch.from_numpy(x).to(gpu)
        _check(la.lwt(t, 100907, tf, mode=3, axes=[0], level=[6]).cpu().numpy(), ref, np.float32, f"tf={tf}")
        la.lwt(t, 100907, tf, mode=3, axes=[0], level=[6], out=t)
        _check(t.cpu().numpy(), ref, np.float32, f"in-place tf={tf}")


@pytest.mark.gpu
def test_cpp_dropin_header_round_trips(gpu):
    """The four lifted::lwt overloads of include/lifted.hpp on host arrays:
    Inverse(Forward(x)) == x for both paths and types, and the dot-product
    identity of Forward / ForwardAdjoint."""
    import subprocess
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp_shim")
    subprocess.run(["make", "-s", "-C", d], check=True)
    r = subprocess.run([os.path.join(d, "build", "shim_main"), "all"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_dropin_overloads_match_oracle(gpu, oracle, tmp_path):
    """One call of each of the four lifted::lwt overloads of include/lifted.hpp
    (lfx_lwt_f32 / _f64 / _sep_f32 / _sep_f64 underneath) on seeded host arrays,
    written out by the shim and compared here with the oracle bit for bit:
    level-by-level float CDF 9/7 Symmetric Forward, level-by-level double
    Daubechies4 Zero Inverse, separable float Bior4_4 Constant ForwardAdjoint,
    separable double Symlet5 Reflect InverseAdjoint from a Fortran-ordered
    input into a C-ordered output."""
    import subprocess
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp_shim")
    subprocess.run(["make", "-s", "-C", d], check=True)
    r = subprocess.run([os.path.join(d, "build", "shim_main"), "dump", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape, axes = (33, 40, 27), [0, 2, 1]
    cases = [("lbl_f32", np.float32, 100907, F, 3, 2, "C"), ("lbl_f64", np.float64, 4, I, 0, 2, "C"),
             ("sep_f32", np.float32, 10404, FA, 1, [2, 3, 1], "C"), ("sep_f64", np.float64, 105, IA, 4, [2, 3, 1], "F")]
    for name, dt, w, tf, bc, lvl, order in cases:
        x = np.fromfile(tmp_path / f"x_{name}.bin", dtype=dt).reshape(shape, order=order)
        got = np.fromfile(tmp_path / f"y_{name}.bin", dtype=dt).reshape(shape)
        ref = oracle.lwt(x, w, tf, bc, axes=axes, level=lvl, out=np.empty(shape, dt))
        _check(got, ref, dt, f"lifted::lwt {name}", axes=axes, level=lvl)


@pytest.mark.parametrize("shape,axes,levels", [((37, 300), (1, 0), (3, 4)), ((64, 48, 20), (0, 2, 1), (4, 3, 2)),
                                               ((64, 64, 64), (0, 1, 2), (4, 4, 4)), ((9, 512), (1,), (4,))])
def test_nd_float32_short_lines(gpu, oracle, shape, axes, levels):
    """float32 ND arrays whose lines run as batched plans of equal lines; with
    the Zero boundary, bands of at most 512 samples run four lines per
    wavefront (row items). Separable and level-by-level, all transforms,
    against the oracle (4 ulp of max |coefficient|)."""
    import lifted_amd as la
    g = rng(sum(shape))
    x = g.standard_normal(shape).astype(np.float32)
    for (w, bc), tf, sep in itertools.product(((100907, 0), (10, 0), (10204, 0), (2, 3)), (F, I, FA, IA),
                                              (True, False)):
        lvl = list(levels) if sep else min(levels)
        if any(not _ok_bands(bc, shape[a], max(levels)) for a in axes):
            continue
        ref = oracle.lwt(x, w, tf, bc, axes=list(axes), level=lvl)
        got = la.lwt(_gpu(x, gpu), w, tf, mode=bc, axes=list(axes), level=lvl).cpu().numpy()
        _check(got, ref, np.float32, f"shape={shape} axes={axes} w={w} bc={bc} tf={tf} sep={sep}", axes=list(axes),
               level=lvl)


def test_nd_float32_whole_tile_transposes(gpu, oracle):
    """Boxes whose transposed dims are multiples of 64 with aligned rows take
    the vector transposes (k_permute_v4); C- and F-ordered inputs, separable
    and level-by-level (whose coarser sub-boxes fall back to the scalar
    transpose), all transforms, against the oracle; and 16-bit storage."""
    import torch
    import lifted_amd as la
    x = rng(21).standard_normal((128, 64, 192)).astype(np.float32)
    xt = torch.from_numpy(x).to(gpu)
    xf = xt.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    for tf, (axes, lvl) in itertools.product((F, I, FA, IA), (((0, 2, 1), [3, 2, 4]), ((2, 0), 2))):
        ref = oracle.lwt(x, 100907, tf, 0, axes=list(axes), level=lvl)
        for src, name in ((xt, "C"), (xf, "F")):
            got = la.lwt(src, 100907, tf, axes=list(axes), level=lvl)
            _check(got.cpu().numpy(), ref, np.float32, f"{name}-order tf={tf} axes={axes} level={lvl}")
    # 16-bit storage, separable over strided first and last axes: 16-bit ->
    # float32 transposes in, float32 -> 16-bit out; the float32 oracle on the
    # 16-bit input, rounded once to the storage type, bit for bit
    for storage, tf in itertools.product(("bfloat16", "float16"), (F, FA)):
        tdt = getattr(torch, storage)
        h = xt.to(tdt)
        ref = torch.from_numpy(oracle.lwt(h.float().cpu().numpy(), 100907, tf, 0, axes=[0, 2, 1],
                                          level=[3, 2, 4])).to(tdt).float().numpy()
        got = la.lwt(h, 100907, tf, axes=[0, 2, 1], level=[3, 2, 4]).float().cpu().numpy()
        np.testing.assert_array_equal(got, ref, err_msg=f"{storage} tf={tf}")


def test_level_by_level_inverse_copies_in_to_out(gpu, oracle):
    """Level-by-level Inverse / ForwardAdjoint with a separate output of the
    input's strides first copy the input to the output. Gap-free boxes copy as
    one flat range: C and F order, a size-1 dim, byte counts that are not a
    multiple of 16 (the tail is copied separately), and a base that is not
    16-byte aligned (no streaming part); all against the oracle."""
    import torch
    import lifted_amd as la
    cases = [((3, 37, 5), torch.float32), ((40, 1, 52), torch.float32), ((3, 37, 5), torch.bfloat16),
             ((64, 48), torch.float16), ((33, 20), torch.float64)]
    for shape, dt in cases:
        x = torch.from_numpy(rng(len(shape) + shape[0]).standard_normal(shape).astype(np.float32)).to(dt)
        xs = x.float().numpy() if dt != torch.float64 else x.numpy()
        axes = [0, len(shape) - 1]
        for tf, order in itertools.product((I, FA), ("C", "F", "offset")):
            if order == "C":
                src = x.to(gpu)
                out = torch.empty_like(src)
            elif order == "F":
                src = x.to(gpu).permute(*reversed(range(len(shape)))).contiguous().permute(
                    *reversed(range(len(shape))))
                out = torch.empty_like(src)
                assert out.stride() == src.stride()
            else:
                # a view one element into a larger buffer: base off the 16-byte grid
                buf = torch.empty(x.numel() + 1, dtype=dt, device=gpu)
                src = buf[1:].view(shape)
                src.copy_(x.to(gpu))
                obuf = torch.empty(x.numel() + 1, dtype=dt, device=gpu)
                out = obuf[1:].view(shape)
            ref = oracle.lwt(xs.astype(np.float64) if dt == torch.float64 else xs, 100907, tf, 0, axes=axes, level=2)
            got = la.lwt(src, 100907, tf, axes=axes, level=2, out=out)
            what = f"shape={shape} dtype={dt} tf={tf} order={order}"
            if dt in (torch.float32, torch.float64):
                _check(got.cpu().numpy(), ref, np.float64 if dt == torch.float64 else np.float32, what)
            else:
                # 16-bit storage is rounded once per level: the oracle one level
                # at a time, rounded to the storage type after each level
                ref16 = oracle_level_by_level_rounded(oracle, xs, 100907, tf, 0, axes, 2, str(dt).split(".")[1])
                _check(got.float().cpu().numpy(), ref16, np.float32, what, axes=axes, level=2)


def test_layouts_float32_batched_lines(gpu, oracle):
    """float32 layouts through the batched-lines path: F-order, strided views
    (in and out), a negative stride, in place, separable and level-by-level,
    a size-1 dim, and 16-bit storage in F-order; all against the oracle."""
    import torch
    import lifted_amd as la
    x = rng(12).standard_normal((40, 52, 3)).astype(np.float32)
    xt = torch.from_numpy(x).to(gpu)
    for (w, bc), tf in itertools.product(((100907, 0), (3, 3)), (F, I, FA, IA)):
        for lvl, axes in (([2, 3], [0, 1]), (2, [1, 0]), ([1], [2])):
            ref = oracle.lwt(x, w, tf, bc, axes=axes, level=lvl)
            what = f"w={w} bc={bc} tf={tf} axes={axes} level={lvl}"
            # Fortran-ordered in and out
            xf = xt.permute(2, 1, 0).contiguous().permute(2, 1, 0)
            got = la.lwt(xf, w, tf, mode=bc, axes=axes, level=lvl)
            assert got.stride() == xf.stride()
            _check(got.cpu().numpy(), ref, np.float32, "F-order " + what)
            # strided view in, strided view out
            big = torch.zeros((80, 104, 3), dtype=torch.float32, device=gpu)
            big[::2, ::2] = xt
            outb = torch.zeros((40, 156, 3), dtype=torch.float32, device=gpu)
            got = la.lwt(big[::2, ::2], w, tf, mode=bc, axes=axes, level=lvl, out=outb[:, ::3])
            _check(got.cpu().numpy(), ref, np.float32, "strided " + what)
            # negative stride along the first axis (a numpy view: host path)
            neg = np.ascontiguousarray(x[::-1])[::-1]
            assert neg.strides[0] < 0
            got = la.lwt(neg, w, tf, mode=bc, axes=axes, level=lvl)
            _check(got, ref, np.float32, "negative stride " + what)
            # in place
            xi = xt.clone()
            la.lwt(xi, w, tf, mode=bc, axes=axes, level=lvl, out=xi)
            _check(xi.cpu().numpy(), ref, np.float32, "in place " + what)
    # size-1 dimension and 16-bit storage in F-order
    y = rng(13).standard_normal((1, 300, 5)).astype(np.float32)
    for tf in (F, FA):
        ref = oracle.lwt(y, 100907, tf, 0, axes=[1, 0], level=[3, 0])
        got = la.lwt(torch.from_numpy(y).to(gpu), 100907, tf, axes=[1, 0], level=[3, 0]).cpu().numpy()
        _check(got, ref, np.float32, f"size-1 dim tf={tf}")
        h = torch.from_numpy(y).to(torch.bfloat16)
        y16 = h.float().numpy()
        ref16 = torch.from_numpy(oracle.lwt(y16, 100907, tf, 0, axes=[1], level=[3])).to(torch.bfloat16).float()
        hf = h.to(gpu).permute(2, 1, 0).contiguous().permute(2, 1, 0)
        got16 = la.lwt(hf, 100907, tf, axes=[1], level=[3]).float().cpu()
        assert torch.equal(got16, ref16), f"bf16 F-order tf={tf}"


def test_two_threads_two_streams_share_the_workspace(gpu, oracle):
    """Two host threads transform float64 arrays of different sizes on two
    streams at once, repeatedly. Long lines take the tiled passes, whose band
    scratch lives in the library's one workspace per device (the second size
    needs a larger one, so it is grown while the other thread may be using
    it). Every result equals the oracle bit for bit."""
    import concurrent.futures
    import torch
    import lifted_amd as la
    cases = [((3, 70001), 100907, 3, [6]), ((2, 150001), 3, 1, [9])]
    data = []
    for k, (shape, w, bc, lvl) in enumerate(cases):
        x = rng(300 + k).standard_normal(shape)
        data.append((x, w, bc, lvl, oracle.lwt(x, w, F, bc, axes=[1], level=lvl),
                     oracle.lwt(x, w, FA, bc, axes=[1], level=lvl)))

    def worker(k):
        x, w, bc, lvl, ref_f, ref_a = data[k]
        s = torch.cuda.Stream()
        outs = []
        with torch.cuda.stream(s):
            xd = torch.from_numpy(x).to(gpu, non_blocking=False)
            for _ in range(6):
                yf = la.lwt(xd, w, F, mode=bc, axes=[1], level=lvl)
                ya = la.lwt(xd, w, FA, mode=bc, axes=[1], level=lvl)
                outs.append((yf, ya))
        s.synchronize()
        return [(yf.cpu().numpy(), ya.cpu().numpy()) for yf, ya in outs]

    with concurrent.futures.ThreadPoolExecutor(2) as ex:
        results = list(ex.map(worker, [0, 1]))
    for k, res in enumerate(results):
        _, _, _, lvl, ref_f, ref_a = data[k]
        for i, (gf, ga) in enumerate(res):
            _check(gf, ref_f, np.float64, f"thread {k} call {i} forward", axes=[1], level=lvl)
            _check(ga, ref_a, np.float64, f"thread {k} call {i} adjoint", axes=[1], level=lvl)


@pytest.mark.parametrize("storage", ["float32", "bfloat16", "float16"])
def test_two_axis_level_by_level(gpu, oracle, storage):
    """Level-by-level transforms over two axes whose second axis is contiguous
    run one two-axis register launch per level (Zero boundary, table
    wavelets): batched and plain 2-D arrays, odd and even sizes, a strided
    register axis, all transforms, against the oracle bit for bit (16-bit
    storage: the oracle one level at a time, rounded to the storage type after
    each level). In-place calls, the other axis order and non-Zero boundaries
    take the chained path and must agree too."""
    import torch
    import lifted_amd as la
    tdt = getattr(torch, storage)
    g = rng(77)
    cases = [((300, 200), [0, 1], 100907, 3), ((3, 129, 257), [1, 2], 2, 4), ((2, 500, 260), [1, 2], 10608, 2),
             ((97, 131), [0, 1], 3, 3), ((4, 64, 4096), [1, 2], 100907, 5), ((250, 333), [0, 1], 105, 3)]
    for shape, axes, w, L in cases:
        x = torch.from_numpy(g.standard_normal(shape).astype(np.float32)).to(tdt)
        x32 = x.float().numpy()
        xd = x.to(gpu)
        for tf in (F, I, FA, IA):
            got = la.lwt(xd, w, tf, axes=axes, level=L).float().cpu().numpy()
            if storage == "float32":
                ref = oracle.lwt(x32, w, tf, 0, axes=axes, level=L)
            else:
                ref = oracle_level_by_level_rounded(oracle, x32, w, tf, 0, axes, L, storage)
            _check(got, ref, np.float32, f"{storage} shape={shape} w={w} L={L} tf={tf}", axes=axes, level=L)
    # a strided register axis (every other row of a larger array)
    big = torch.from_numpy(g.standard_normal((400, 200)).astype(np.float32)).to(tdt).to(gpu)
    view = big[::2]
    xs = view.float().cpu().numpy()
    for tf in (F, I):
        got = la.lwt(view, 100907, tf, axes=[0, 1], level=3).float().cpu().numpy()
        ref = (oracle.lwt(xs, 100907, tf, 0, axes=[0, 1], level=3) if storage == "float32" else
               oracle_level_by_level_rounded(oracle, xs, 100907, tf, 0, [0, 1], 3, storage))
        _check(got, ref, np.float32, f"{storage} strided rows tf={tf}", axes=[0, 1], level=3)
    # in place (chained path) agrees with the out-of-place two-axis path
    x = torch.from_numpy(g.standard_normal((256, 192)).astype(np.float32)).to(tdt).to(gpu)
    for tf in (F, I):
        a = la.lwt(x, 100907, tf, axes=[0, 1], level=3)
        xi = x.clone()
        la.lwt(xi, 100907, tf, axes=[0, 1], level=3, out=xi)
        assert torch.equal(a, xi), f"{storage} in place tf={tf}"


@pytest.mark.parametrize("across", [1, 0])
def test_across_strided_axes(gpu, oracle, across):
    """Strided axes of at most 512 samples (Zero boundary) run through the
    across-line kernel (the default, knob across=1), which reads 32
    neighbouring lines per workgroup where they lie, or through a transpose
    and a row plan (across=0): single strided axes (partial blocks of lines,
    several batch dims, row and column ranges of a larger array, in place,
    more than four levels in two passes), separable and level-by-level
    transforms whose strided axes chain through it, all transforms, bit for
    bit against the oracle; 16-bit storage rounded once (single axis) against
    the float32 oracle on the 16-bit input."""
    from lifted_amd import tune
    with tune.tuning(across=across):
        _across_cases(gpu, oracle)


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_across_and_two_axis_every_wavelet(gpu, oracle, w):
    """Every table wavelet through the two kernels that float32 ND calls take
    by default under the Zero boundary: the across-line kernel (a strided axis
    of 130 samples read where it lies, 48 lines along the contiguous dim) and
    the two-axis level kernel (level by level over the last two axes, the
    second contiguous), all four transforms, bit for bit against the oracle."""
    import lifted_amd as la
    g = rng(1000 + w)
    x = g.standard_normal((130, 3, 48)).astype(np.float32)
    y = g.standard_normal((2, 96, 160)).astype(np.float32)
    for tf in (F, I, FA, IA):
        ref = oracle.lwt(x, w, tf, 0, axes=[0], level=[3])
        got = la.lwt(_gpu(x, gpu), w, tf, axes=[0], level=[3]).cpu().numpy()
        _check(got, ref, np.float32, f"across w={w} tf={tf}", axes=[0], level=[3])
        ref = oracle.lwt(y, w, tf, 0, axes=[1, 2], level=2)
        got = la.lwt(_gpu(y, gpu), w, tf, axes=[1, 2], level=2).cpu().numpy()
        _check(got, ref, np.float32, f"two-axis w={w} tf={tf}", axes=[1, 2], level=2)


def _across_cases(gpu, oracle):
    import torch
    import lifted_amd as la
    g = rng(512)
    cases = [((3, 300, 45), [1], [6], 100907), ((512, 40), [0], [8], 2), ((5, 2, 130, 7), [2], [3], 10),
             ((64, 96, 80), [0, 1, 2], [3, 3, 3], 100907), ((40, 50, 36), [1, 0], [2, 3], 10204),
             ((96, 70, 33), [0, 2, 1], 3, 100907)]
    for (shape, axes, lvl, w), tf in itertools.product(cases, (F, I, FA, IA)):
        x = g.standard_normal(shape).astype(np.float32)
        level = lvl if isinstance(lvl, int) else list(lvl)
        ref = oracle.lwt(x, w, tf, 0, axes=axes, level=level)
        got = la.lwt(_gpu(x, gpu), w, tf, axes=axes, level=level).cpu().numpy()
        _check(got, ref, np.float32, f"shape={shape} axes={axes} level={level} w={w} tf={tf}", axes=axes,
               level=level)
    # views: a row range, a column range (row pitch above the line count), and in place
    x = g.standard_normal((256, 70)).astype(np.float32)
    xt = _gpu(x, gpu)
    for tf in (F, FA):
        for sl in (np.s_[10:200, :], np.s_[:, 3:50]):
            v = x[sl]
            ref = oracle.lwt(np.ascontiguousarray(v), 100907, tf, 0, axes=[0], level=[4])
            got = la.lwt(xt[sl], 100907, tf, axes=[0], level=[4])
            _check(got.cpu().numpy(), ref, np.float32, f"view {sl} tf={tf}")
        y = xt.clone()
        la.lwt(y, 100907, tf, axes=[0], level=[4], out=y)
        _check(y.cpu().numpy(), oracle.lwt(x, 100907, tf, 0, axes=[0], level=[4]), np.float32, f"in place tf={tf}")
    # 16-bit storage over one strided axis: read and written in 16 bits by the kernel
    for storage, tf in itertools.product(("bfloat16", "float16"), (F, I, FA)):
        tdt = getattr(torch, storage)
        h = _gpu(g.standard_normal((4, 200, 90)).astype(np.float32), gpu).to(tdt)
        ref = torch.from_numpy(oracle.lwt(h.float().cpu().numpy(), 100907, tf, 0, axes=[1],
                                          level=[5])).to(tdt).float().numpy()
        got = la.lwt(h, 100907, tf, axes=[1], level=[5]).float().cpu().numpy()
        np.testing.assert_array_equal(got, ref, err_msg=f"{storage} tf={tf}")
