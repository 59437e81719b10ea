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
    eps = np.finfo(np.float32).eps
    worst = 0.0
    for off, n, w, L in sigs:
        for got, ref in ((yh, yr), (zh, zr)):
            a, b = got[off:off + n], ref[off:off + n]
            e = float(np.abs(a - b).max() / (eps * max(np.abs(b).max(), 1e-30)))
            worst = max(worst, e)
            assert e <= 4.0, (off, n, w, L, e)
    print(f"max error {worst:.3f} ulp of max|coef| over {len(sigs)} signals")


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
    against the oracle (4 ulp of max |coefficient|), and I(F(x)) ~ x."""
    import torch
    import lifted_amd as la
    from helpers import max_ulp_diff
    img = _c2_image()
    x = torch.tensor(img, device=gpu)
    for tf in (0, 1, 2):
        ref = oracle.lwt(img, 100907, tf, oracle.ZERO, level=level, n_threads=16)
        for plain in (False, True):
            got = la._lwt(x, 100907, 0, tf, (0, 1), level, plain=plain).cpu().numpy()
            e = max_ulp_diff(got, ref)
            assert e <= 4.0, (tf, plain, e)
    y = la.lwt_forward(x, 100907, 0, axes=(0, 1), level=level)
    z = la.lwt_inverse(y, 100907, 0, axes=(0, 1), level=level)
    assert float((z - x).abs().max()) <= 1e-4 * float(x.abs().max())
