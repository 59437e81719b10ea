"""Full-size workloads on the GPU: parity against the oracle at the sizes the
benchmark uses, and the size-independent properties (round trips)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_headline_workload_parity(gpu, oracle):
    """The benchmark's batched workload (4096 signals, ~6.3e8 samples, random
    wavelets and levels): GPU forward and forward-adjoint equal the oracle per
    signal within 4 ulp of the signal's max |coefficient|."""
    import torch
    import bench
    from lifted_amd.batch import BatchPlan
    sigs, total = bench.make_workload()
    g = torch.Generator(device=gpu)
    g.manual_seed(bench.SEED)
    x = torch.randn(total, generator=g, device=gpu, dtype=torch.float32)
    y = BatchPlan(sigs, 0).execute(x)
    z = BatchPlan(sigs, 2).execute(y)
    xh, yh, zh = x.cpu().numpy(), y.cpu().numpy(), z.cpu().numpy()
    yr = oracle.batch(xh, sigs, oracle.FORWARD, oracle.ZERO, n_threads=16)
    zr = oracle.batch(yh, sigs, oracle.FORWARD_ADJOINT, oracle.ZERO, n_threads=16)
    from helpers import assert_parity, band_ulp_error
    from lifted_amd.layout import line_bands
    same = [bool(np.array_equal(yh, yr)), bool(np.array_equal(zh, zr))]
    if not all(same):  # name the first signal that differs, with its per-band error
        for off, n, w, L in sigs:
            for got, ref, name in ((yh, yr, "forward"), (zh, zr, "adjoint")):
                a, b = got[off:off + n], ref[off:off + n]
                if not np.array_equal(a, b):
                    e = band_ulp_error(a, b, list(line_bands(n, L).values()))
                    pytest.fail(f"{name} signal at {off} (n={n}, wavelet {w}, {L} levels): {e:.3g} ulp")
    assert_parity(yh, yr, f"forward, {len(sigs)} signals")
    assert_parity(zh, zr, f"forward-adjoint, {len(sigs)} signals")
    print(f"{len(sigs)} signals, {total} samples: forward and forward-adjoint bit-exact")


def _c2_image():
    """4096^2 float32: 32 Gaussians (centres U[0,4096)^2, sigma log-U[64,1024],
    amplitudes N(0,1)) plus 0.01 N(0,1), seed 20261201 + 1."""
    rng = np.random.default_rng(20261201 + 1)
    n = 4096
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float32)
    img = np.zeros((n, n), np.float32)
    for _ in range(32):
        cy, cx = rng.uniform(0, n, 2)
        sig = np.exp(rng.uniform(np.log(64), np.log(1024)))
        amp = rng.standard_normal()
        img += (amp * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig * sig))).astype(np.float32)
    img += 0.01 * rng.standard_normal((n, n)).astype(np.float32)
    return img


@pytest.mark.parametrize("level", [[5, 5], 5], ids=["separable", "level-by-level"])
def test_c2_image_parity(gpu, oracle, level):
    """4096^2 float32, CDF 9/7, 5 levels, F / I / FA, fused and plain kernels,
    against the oracle bit for bit (a failure reports the worst error per
    coefficient band), and I(F(x)) ~ x."""
    import torch
    import lifted_amd as la
    from helpers import assert_parity, band_slices
    img = _c2_image()
    x = torch.tensor(img, device=gpu)
    bands = band_slices(img.shape, [0, 1], level)
    for tf in (0, 1, 2):
        ref = oracle.lwt(img, 100907, tf, oracle.ZERO, level=level, n_threads=16)
        for plain in (False, True):
            got = la._lwt(x, 100907, 0, tf, (0, 1), level, plain=plain).cpu().numpy()
            assert_parity(got, ref, f"tf={tf} plain={plain}", bands=bands)
    y = la.lwt_forward(x, 100907, 0, axes=(0, 1), level=level)
    z = la.lwt_inverse(y, 100907, 0, axes=(0, 1), level=level)
    assert float((z - x).abs().max()) <= 1e-4 * float(x.abs().max())


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_c5_reduced_16bit_level_by_level_bit_exact(gpu, oracle, storage):
    """The 16-bit workload's path at a reduced size: 2 x 1024^2 1/f^2 fields
    stored in 16 bits, CDF 9/7, 6 levels level by level over the last two
    axes, Forward then Inverse of its output, float32 arithmetic. Every level
    rounds its sub-box once to the storage type, so the oracle is applied one
    level at a time with the same rounding after each level; both transforms
    must match it bit for bit. Also float32 storage against the plain oracle."""
    import torch
    import lifted_amd as la
    from helpers import assert_parity, band_slices, field_1f2, oracle_level_by_level_rounded, round_to
    tdt = getattr(torch, storage)
    x = round_to(field_1f2((2, 1024, 1024), 20261201 + 4), storage)
    axes, L = [1, 2], 6
    xd = torch.from_numpy(x).to(tdt).to(gpu)
    y = la.lwt_forward(xd, 100907, 0, axes=axes, level=L)
    yh = y.float().cpu().numpy()
    ref_y = oracle_level_by_level_rounded(oracle, x, 100907, 0, 0, axes, L, storage)
    assert_parity(yh, ref_y, f"{storage} forward", bands=band_slices(x.shape, axes, L))
    z = la.lwt_inverse(y, 100907, 0, axes=axes, level=L)
    ref_z = oracle_level_by_level_rounded(oracle, yh, 100907, 1, 0, axes, L, storage)
    zh = z.float().cpu().numpy()
    assert_parity(zh, ref_z, f"{storage} inverse", bands=band_slices(x.shape, axes, L))
    u = float(torch.finfo(tdt).eps) / 2
    rt = float(np.linalg.norm(zh - x) / np.linalg.norm(x))
    print(f"{storage}: round trip relative L2 error {rt:.3g} (unit roundoff {u:.3g})")
    assert rt <= 2 * u
    # float32 storage of the same fields: the oracle level by level, unrounded
    x32 = torch.from_numpy(x).to(gpu)
    for tf in (0, 1):
        got = la.lwt(x32, 100907, tf, axes=axes, level=L).cpu().numpy()
        ref = oracle.lwt(x, 100907, tf, 0, axes=axes, level=L)
        assert_parity(got, ref, f"float32 tf={tf}", bands=band_slices(x.shape, axes, L))


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_c5_full_size_round_trip(gpu, storage):
    """The full 16-bit workload: 16 x 8192^2 1/f^2 fields stored in 16 bits,
    CDF 9/7, 6 levels level by level, Forward then Inverse. Every level
    rounds its sub-box once in each direction (12 roundings of relative size
    at most u, the unit roundoff, on a near-orthogonal transform), so the
    reconstruction's relative L2 error against the 16-bit input is bounded by
    2 u; the largest single error by 16 u of max |x|. The forward also stays
    within 8 u (relative L2) of the same transform of the same input in
    float64."""
    import torch
    import lifted_amd as la
    tdt = getattr(torch, storage)
    B, n, L, axes = 16, 8192, 6, (1, 2)
    g = torch.Generator(device=gpu).manual_seed(20261201 + 4)
    ky = torch.fft.fftfreq(n, device=gpu).view(n, 1)
    kx = torch.fft.rfftfreq(n, device=gpu).view(1, n // 2 + 1)
    k2 = ky * ky + kx * kx
    filt = torch.where(k2 > 0, 1.0 / torch.where(k2 > 0, k2, torch.ones_like(k2)), torch.zeros_like(k2))
    x = torch.empty(B, n, n, dtype=tdt, device=gpu)
    for b in range(B):
        f = torch.fft.irfft2(torch.fft.rfft2(torch.randn(n, n, device=gpu, generator=g)) * filt, s=(n, n))
        x[b] = ((f - f.mean()) / f.std()).to(tdt)
        del f
    u = float(torch.finfo(tdt).eps) / 2
    y = la.lwt_forward(x, 100907, 0, axes=axes, level=L)
    z = la.lwt_inverse(y, 100907, 0, axes=axes, level=L)
    xf = x.float()
    rt = float((z.float() - xf).norm() / xf.norm())
    mx = float((z.float() - xf).abs().max() / xf.abs().max())
    del z
    ref = la.lwt_forward(x.double(), 100907, 0, axes=axes, level=L)
    cf = float((y.double() - ref).norm() / ref.norm())
    print(f"{storage}: round trip rel L2 {rt:.3g}, max {mx:.3g} of max|x|; forward vs float64 rel L2 {cf:.3g} "
          f"(u = {u:.3g})")
    assert rt <= 2 * u and mx <= 16 * u and cf <= 8 * u
